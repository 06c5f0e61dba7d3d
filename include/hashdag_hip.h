/* hashdag_hip.h - C ABI of the MI355X HashDAG hot path (tracing, edits, hash table, GC).
 *
 * Plain C types only: device buffers are passed as raw device pointers, streams as the
 * opaque hipStream_t handle cast to void* (NULL = the HIP default stream). Every entry
 * point returns an hd_status (0 = OK, < 0 = error); hd_last_error() returns the message.
 * Contexts are independent; calls on one context are stream-ordered and must not race.
 *
 * Each declaration names the reference interface it replaces (paths relative to the
 * reference's src/ directory).
 */
#ifndef HASHDAG_HIP_H
#define HASHDAG_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t hd_status;
#define HD_OK 0
#define HD_ERR_INVALID (-1)          /* bad argument or state */
#define HD_ERR_HIP (-2)              /* HIP runtime error */
#define HD_ERR_BUCKET_FULL (-3)      /* a hash bucket overflowed (the reference does not check: hash_table.h:393,461) */
#define HD_ERR_POOL_FULL (-4)        /* the physical page budget is exhausted (hash_table.h:794-809) */
#define HD_ERR_OOM (-5)              /* device allocation failed */
#define HD_ERR_COLOR_FULL (-6)       /* colour arena / colour tree capacity exhausted */
#define HD_ERR_HEADER_OVERFLOW (-7)  /* voxel count >= 2^24 in a node header (hash_dag_edits.h:605) */

typedef struct hd_ctx hd_ctx;

/* Hash table geometry and capacities. Replaces the compile-time knobs of typedefs.h:202-238
 * and hash_dag_globals.h:7-40 (SCENE_DEPTH, PAGE_SIZE, BUCKETS_*, C_colorTreeLevels) and the
 * pool size given to HashTable::create (hash_table.cpp:5, main.cpp:46). */
typedef struct hd_config {
    uint32_t levels;              /* scene depth D: extent 2^D, leaf level D-2 */
    uint32_t page_size;           /* words per page (512) */
    uint32_t num_top_levels;      /* levels using the top-level bucket geometry (9) */
    uint32_t bucket_bits_top;     /* log2 buckets per top level (10) */
    uint32_t bucket_bits_low;     /* log2 buckets per low level (16) */
    uint32_t bucket_size_top;     /* words per top-level bucket (1024) */
    uint32_t bucket_size_low;     /* words per low-level bucket (4096) */
    uint32_t color_tree_levels;   /* levels covered by the colour tree (10) */
    uint32_t pool_pages;          /* physical page budget */
    uint32_t device;              /* HIP device ordinal */
    uint64_t color_arena_words;   /* 64-bit words reserved for colour leaves */
    uint32_t color_nodes_cap;     /* colour tree entries capacity */
    uint32_t color_leaves_cap;    /* unique colour leaf slots capacity */
} hd_config;

/* Fills the reference defaults for depth `levels`. */
void hd_default_config(uint32_t levels, hd_config* out);

/* HashTable::create (hash_table.cpp:5-71) + DAGTracer setup (dag_tracer.cu:9-45). */
hd_status hd_create(const hd_config* cfg, hd_ctx** out);
/* HashTable::destroy (hash_table.cpp:73-109), HashDAGColors::free (hash_dag_colors.h:122-134). */
hd_status hd_destroy(hd_ctx* ctx);
/* Last error message of ctx (or of the calling thread when ctx is NULL). */
const char* hd_last_error(hd_ctx* ctx);

/* HashDAGFactory::load_from_DAG (hash_dag_factory.cpp:152-184): BasicDAG words (root at 0,
 * basic_dag.h:11-45) -> full nodes first, then every node appended, uploaded to the device. */
hd_status hd_load_basic_dag(hd_ctx* ctx, const uint32_t* data, uint64_t n_words);

/* HashDAGFactory::load_colors_from_DAG (hash_dag_factory.cpp:186-198) after
 * HashColorsBuilder::build (hash_dag_colors.h:215-256): the colour tree `nodes` (8 entries
 * per node; level color_tree_levels-1 entries are leaf handles), the shared-leaf `offsets`
 * and the main compressed leaf (blocks / byte-swapped weights / macro pairs). */
hd_status hd_load_colors(hd_ctx* ctx, const uint32_t* nodes, uint64_t n_nodes, const uint64_t* offsets,
                         uint64_t n_offsets, const uint64_t* blocks, uint64_t n_blocks, const uint32_t* weights,
                         uint64_t n_weights, const uint64_t* macro, uint64_t n_macro);

/* Sizes of the reference-layout image (HashTable members, hash_table.h:818-861). */
typedef struct hd_image_info {
    uint32_t root;           /* HashDAG::firstNodeIndex (hash_dag.h:221) */
    uint32_t pool_top;       /* HashTable::poolTop */
    uint32_t pool_pages;     /* CpuData::poolMaxSize */
    uint32_t levels;
    uint64_t total_pages;    /* C_totalPages (page table length) */
    uint64_t total_buckets;  /* C_totalNumberOfBuckets */
    uint64_t total_virtual;  /* C_totalVirtualAddresses */
} hd_image_info;
hd_status hd_image_info_get(hd_ctx* ctx, hd_image_info* out);

/* Device -> host copy in the reference's CPU layout: pool of physical pages (pool_pages *
 * page_size words), page table, bucket sizes, full node indices (hash_table.h:829-843). */
hd_status hd_download_image(hd_ctx* ctx, uint32_t* pool, uint32_t* page_table, uint32_t* bucket_sizes,
                            uint32_t* full_nodes);
/* HashTable::full_upload_to_gpu (hash_table.cpp:186-206) of a reference-layout image. */
hd_status hd_upload_image(hd_ctx* ctx, const uint32_t* pool, const uint32_t* page_table, const uint32_t* bucket_sizes,
                          const uint32_t* full_nodes, uint32_t pool_top, uint32_t root);

/* Checks a reference-layout image the way hd_upload_image does before any copy (it rejects
 * a failing image with HD_ERR_INVALID): every page-table entry maps a distinct physical page
 * below pool_top, every bucket size fits its bucket with its used pages mapped, and every node
 * reachable from the root or from the full nodes of levels 1 .. leaf level lies in its level's
 * used bucket range with child pointers inside the table. `pool` holds pages 0 .. pool_top-1.
 * Needs no device. why (may be NULL): the reason, NUL-terminated within why_len bytes.
 * No counterpart in the reference, whose loader reads a dump as is (hash_dag_factory.cpp:225-246). */
hd_status hd_validate_image(const hd_config* cfg, const uint32_t* pool, const uint32_t* page_table,
                            const uint32_t* bucket_sizes, const uint32_t* full_nodes, uint32_t pool_top, uint32_t root,
                            char* why, uint32_t why_len);

/* HashDAG::firstNodeIndex get/set (undo/redo = root history, hash_dag.h:68-120). */
hd_status hd_get_root(hd_ctx* ctx, uint32_t* root);
hd_status hd_set_root(hd_ctx* ctx, uint32_t root);

/* Colour state download (HashDAGColors members, hash_dag_colors.h:65-73). */
hd_status hd_colors_info(hd_ctx* ctx, uint64_t* n_nodes, uint64_t* n_leaves, uint64_t* n_offsets);
hd_status hd_download_color_nodes(hd_ctx* ctx, uint32_t* nodes);
/* sizes[0..3] = n_blocks, n_weights, n_macro, offset (UINT64_MAX = unique) of the leaf a
 * handle resolves to (HashDAGColors::get_leaf, hash_dag_colors.h:37-50) */
hd_status hd_color_leaf_sizes(hd_ctx* ctx, uint32_t handle, uint64_t sizes[4]);
hd_status hd_download_color_leaf(hd_ctx* ctx, uint32_t handle, uint64_t* blocks, uint32_t* weights, uint64_t* macro);

/* Camera ray parameters: get_trace_params (dag_tracer.cu:71-113), FP64 on the host.
 * rotation = CameraView::rotation rows (camera_view.h:8-27); 60 degree vertical fov. */
typedef struct hd_ray_params {
    double camera_position[3];
    double ray_min[3];
    double ray_ddx[3];
    double ray_ddy[3];
} hd_ray_params;
hd_status hd_ray_params_from_camera(const double position[3], const double rotation[9], const double bounds_min[3],
                                    const double bounds_max[3], uint32_t levels, uint32_t width, uint32_t height,
                                    hd_ray_params* out);

/* DAGTracer::resolve_paths -> Tracer::trace_paths (dag_tracer.cu:116-143, tracer.cu:145-252).
 * d_paths: width*height uint4 (x, y, z, 0); row r holds the ray of pixel y = height-1-r. */
hd_status hd_trace_paths(hd_ctx* ctx, const hd_ray_params* params, uint32_t width, uint32_t height, uint32_t* d_paths,
                         void* stream);
/* DAGTracer::resolve_colors -> Tracer::trace_colors (dag_tracer.cu:145-179, tracer.cu:254-451),
 * debug colours off. d_colors: width*height RGBA8 words. */
hd_status hd_trace_colors(hd_ctx* ctx, uint32_t width, uint32_t height, const uint32_t* d_paths, uint32_t* d_colors,
                          void* stream);
/* DAGTracer::resolve_shadows -> Tracer::trace_shadows (dag_tracer.cu:181-219, tracer.cu:589-697),
 * EXACT_SHADOWS; read-modify-write of d_colors. */
hd_status hd_trace_shadows(hd_ctx* ctx, const hd_ray_params* params, uint32_t width, uint32_t height,
                           const uint32_t* d_paths, uint32_t* d_colors, float shadow_bias, float fog_density,
                           void* stream);

/* Tracer tuning (tile shape); 0 = default. Not in the reference (its launch shape is fixed
 * at (4,64) blocks, dag_tracer.cu:121-122). */
hd_status hd_set_trace_tile(hd_ctx* ctx, uint32_t tile_w, uint32_t tile_h);
/* Tracer tuning, not in the reference: node fetch strategy (0 = child pointer then header,
 * 1 = whole node). Both give the same images. */
hd_status hd_set_trace_fetch(hd_ctx* ctx, uint32_t fetch_mode);
/* out: tile_w, tile_h, fetch mode */
hd_status hd_get_trace_launch(hd_ctx* ctx, uint32_t out[3]);
/* Tracer schedule, not in the reference (one thread per pixel, dag_tracer.cu:121-122), for
 * paths and shadows; every schedule gives the same images:
 *   0  one wavefront per pixel tile, each lane walking its ray to the end;
 *   1  persistent wavefronts taking tiles from per-region queues, handing new pixels to their
 *      idle lanes once `param` (1..64) of them are idle;
 *   2  one wavefront per pixel tile, walks stopped after `param` steps; those rays are then
 *      searched a wavefront per ray (64 lanes expanding the ray's next nodes in DFS order);
 *   3  as 2 with a resume step (node fetch 0 only; with fetch 1 it runs as 2): walks stopped
 *      after param & 0xFFFF steps save their state, wavefronts of 64 such walks continue them
 *      up to param >> 16 steps in all, and the rest are searched a wavefront per ray. */
hd_status hd_set_trace_schedule(hd_ctx* ctx, uint32_t schedule, uint32_t param);
/* out: schedule, param */
hd_status hd_get_trace_schedule(hd_ctx* ctx, uint32_t out[2]);

/* Autotuner for the tracer launch shape: times full frames (paths + colours + shadows) over
 * n_params cameras for each wavefront tile shape (8x8, 16x4, 4x16, 32x2, 2x32, 64x1), node
 * fetch strategy and schedule (per tile; persistent refilled at 8 or 24 idle lanes; per tile
 * with walks beyond 96, 128 or 192 steps searched a wavefront per ray), keeps
 * the fastest in ctx and returns it with its ms per frame. No counterpart in the reference
 * (fixed (4,64) blocks, dag_tracer.cu:121-122). */
hd_status hd_autotune_trace(hd_ctx* ctx, const hd_ray_params* params, uint32_t n_params, uint32_t width,
                            uint32_t height, uint32_t* d_paths, uint32_t* d_colors, uint32_t reps, uint32_t* best_w,
                            uint32_t* best_h, float* best_ms);

/* Algorithmic HBM bytes of the last traced frame per kernel (paths, colors, shadows), filled
 * by an instrumented launch: hd_trace_count_bytes runs the three kernels once with counters.
 * No counterpart in the reference (a measurement hook for the roofline of the three tracer
 * kernels, tracer.cu:165-423). */
hd_status hd_trace_count_bytes(hd_ctx* ctx, const hd_ray_params* params, uint32_t width, uint32_t height,
                               uint32_t* d_paths, uint32_t* d_colors, float shadow_bias, float fog_density,
                               uint64_t bytes[3]);

/* Editors (hash_dag_editors.h): SphereEditor<true/false> (272-313), BoxEditor<true/false>
 * (223-270), SpherePaintEditor (315-330); centre and radius in voxels. */
#define HD_EDIT_SPHERE_ADD 0u
#define HD_EDIT_SPHERE_CARVE 1u
#define HD_EDIT_BOX_ADD 2u
#define HD_EDIT_BOX_CARVE 3u
#define HD_EDIT_SPHERE_PAINT 4u
#define HD_EDIT_COPY 5u  /* CopyEditorWithoutDecompression with its transform and swirl (hash_dag_editors.h:332-435) */
#define HD_EDIT_FILL 6u  /* FillEditorColors: flood fill of the empty voxels around center (hash_dag_editors.h:624-768) */
/* CopyEditorWithDecompression (hash_dag_editors.h:437-616, the copy the checked-in
 * script_definitions.h:2 selects): the box of radius around center - source + dest reads the
 * voxels of the box of radius around center through a buffer of 2 ceil(radius) voxels per axis
 * from truncate(center - radius); nodes above levels - 4 whose source region is empty are left
 * unedited (COPY_EMPTY_CHECKS). No transform or swirl; radius below 65536. */
#define HD_EDIT_COPY_DEC 7u
/* SphereNoiseEditor(isAdding = true / false) (hash_dag_editors.h:762-809, engine.cpp:688): in
 * the box of half side 1.5 radius, voxels within radius * (1 + FastNoise Perlin fractal) of the
 * centre are set (adding); the carving form leaves values unchanged, as the reference's
 * BufferEditor::get_new_value does. A voxel that becomes set takes the editor's rgb565 min/max
 * colour with a 3-bit weight. Each starts its own pass; 1.5 radius at most 512. */
#define HD_EDIT_NOISE_ADD 8u
#define HD_EDIT_NOISE_CARVE 9u
typedef struct hd_edit {
    uint32_t kind;
    float center[3];
    float radius;
    float source[3];  /* copy: the box around center is copied from source to dest */
    float dest[3];
    uint32_t swirl;       /* copy: 1 = swirl about y with period swirl_period (hash_dag_editors.h:366-372) */
    float swirl_period;
    double transform[9];  /* copy: Matrix3x3 transform, row-major; all zeros = identity */
} hd_edit;

typedef struct hd_edit_stats {
    uint64_t num_voxels;     /* "num voxels" (hash_dag.h:374-378): voxels the editors touched */
    uint64_t new_words;      /* words appended to the hash table */
    uint64_t leaves_rebuilt; /* "num color leaves rebuilt" (hash_dag.h:379) */
    uint32_t root;           /* new firstNodeIndex */
    uint32_t passes;         /* device passes run */
    uint64_t colors_encoded; /* colours written into the rebuilt colour leaves */
} hd_edit_stats;

/* Engine::edit<T> -> HashDAG::edit_threads<T> (engine.h:77-103, hash_dag.h:255-405) for each
 * edit in order. fused = 0: one device pass per edit (the plain form); fused = 1: the whole
 * batch in one pass, canonically identical to applying the edits one by one. On error the
 * previous root is kept. One undo frame per pass. */
hd_status hd_edit_batch(hd_ctx* ctx, const hd_edit* edits, uint32_t n, uint32_t fused, hd_edit_stats* stats,
                        void* stream);
/* No counterpart in the reference: hd_edit_batch(fused = 1) for callers that cannot allocate
 * device memory (code under graph capture, serving loops, harnesses that forbid hipMalloc).
 * One fused pass of at most HD_STATIC_MAX_EDITS sphere / box / paint editors, run in a 32 MiB
 * file-scope device array of the library (one per GPU, held until the process ends) instead
 * of the context's growable scratch; nothing is allocated. Limits per pass: 4096 work items
 * per level, 32768 edit-list words per level, 2^20 staged colours, 16384 rebuilt colour
 * leaves; a pass that needs more, or more colour arena / colour tree room than the context
 * holds, returns HD_ERR_OOM / HD_ERR_COLOR_FULL with the DAG as it was. Needs a colour tree
 * (hd_load_colors, or any earlier hd_edit_batch). The pass is not recorded in the undo
 * history, which it clears. Calls that overlap on one GPU share its array and race. */
#define HD_STATIC_MAX_EDITS 256u
hd_status hd_edit_batch_static(hd_ctx* ctx, const hd_edit* edits, uint32_t n, hd_edit_stats* stats, void* stream);
/* Launch form of edit passes, no counterpart in the reference (its edit_threads recurses on 8
 * host threads, hash_dag.h:256-405). A pass is a fixed sequence of steps (expand per level, the
 * leaves, build / dedup / find / add per level, the colour steps, the commit), run as:
 *   - one cooperative launch of blocks_per_cu (1..8, capped at what fits) resident blocks per
 *     CU, a grid-wide barrier between steps, when the pass holds at most persist_max_edits
 *     editors (not inside a frame graph);
 *   - else, when the pass has at most workgroup_max_items work items on every level (bounded on
 *     the host by the nodes of each level the editors' boxes overlap; inside a frame graph by
 *     the scratch capacity), as one launch of a single workgroup with a block barrier between
 *     steps;
 *   - else as a launch per step, over a grid sized by that bound for the step's level.
 * workgroup_max_items = 0 and persist_max_edits = 0 (the defaults after hd_create, unless
 * hd_autotune_edit_launch picked others): a launch per step. Results are the same in every
 * form. */
hd_status hd_set_edit_launch(hd_ctx* ctx, uint32_t workgroup_max_items, uint32_t persist_max_edits,
                             uint32_t blocks_per_cu);
hd_status hd_get_edit_launch(hd_ctx* ctx, uint32_t* workgroup_max_items, uint32_t* persist_max_edits,
                             uint32_t* blocks_per_cu);
/* Edit launch autotuner, no counterpart in the reference: times one pass per sample editor
 * (sphere / box / paint; each applied, timed, undone; reps rounds after one warm-up round) with
 * workgroup_max_items = 0, 32, 128, 512 and 2048, keeps the fastest in the context (out:
 * the pick and the mean milliseconds per edit of each candidate). The trial passes leave their
 * nodes for the next GC and clear the redo history. */
hd_status hd_autotune_edit_launch(hd_ctx* ctx, const hd_edit* sample, uint32_t n, uint32_t reps,
                                  uint32_t* workgroup_max_items, float ms_per_edit[5]);
/* Edit pass measurement, no counterpart in the reference (its "edits" stat is a time,
 * hash_dag.h:262,368): on = 1 zeroes the counters and counts the algorithmic bytes of every
 * following pass (the counting kernels are slower; time passes without it); on = 0 stops and
 * returns out[0] node words read by the descent (header and child pointers of every node the
 * passes visit, 8 B per leaf), out[1] bucket bytes the finds read (tag bytes and compared
 * words), out[2] node words written, out[3] staged colours: every 64-colour chunk's marker and
 * every colour of an explicit chunk (8 B each), written once and read twice by the encoder,
 * out[4] colour leaf words written. Old colours decoded are not counted. */
hd_status hd_edit_count_bytes(hd_ctx* ctx, uint32_t on, uint64_t out[5]);
/* HashDAGUndoRedo::undo / redo (hash_dag.h:68-120) and free (133-145). */
hd_status hd_undo(hd_ctx* ctx);
hd_status hd_redo(hd_ctx* ctx);
hd_status hd_undo_clear(hd_ctx* ctx);
hd_status hd_undo_depth(hd_ctx* ctx, uint32_t* undo, uint32_t* redo);
/* Pre-sizes the edit work-item scratch (items per level, edit-list words per level,
 * colours per pass); grown automatically when a pass overflows. Not in the reference. */
hd_status hd_reserve_edit_scratch(hd_ctx* ctx, uint32_t items_per_level, uint32_t edit_list_words, uint64_t colors);

/* HashDAG::remove_stale_nodes(maxLevel) (hash_dag.h:407-547), called with levels - 2 by the
 * engine (engine.cpp:178-185): keeps only nodes reachable from the root at levels 0..max_level,
 * re-appended level by level bottom-up with their children remapped, full nodes first.
 * Clears the undo history. stats (may be NULL): live nodes, words before, words after, 0.
 * Device scratch: 4 B + 1 bit per virtual word of the table, plus about 64 B per node of the
 * largest level; allocated by the first call, kept in the context for later calls (grown when
 * one needs more) and freed by hd_destroy. */
hd_status hd_gc(hd_ctx* ctx, uint32_t max_level, uint64_t stats[4], void* stream);
/* GC that also keeps alive the nodes reachable from extra level-0 roots (rewritten in place to
 * their addresses after GC) and, with keep_undo = 1, from the roots of every undo / redo
 * frame, which are rewritten so the history survives the GC. No counterpart in the reference:
 * its GC marks from firstNodeIndex only and the engine frees the undo history with it
 * (hash_dag.h:426, engine.cpp:178-185). hd_gc is this call with no extra roots. */
hd_status hd_gc_roots(hd_ctx* ctx, uint32_t max_level, uint32_t keep_undo, uint32_t* roots, uint32_t n_roots,
                      uint64_t stats[4], void* stream);

/* Frame graph: one fused edit pass (up to max_edits editors) followed by the trace of the
 * resulting DAG (paths, colours, shadows into d_paths / d_colors), captured once as a HIP graph
 * and replayed per call with new edits and camera. The replay is committed like
 * hd_edit_batch(fused = 1) + hd_trace_*; a replay that runs out of a capacity is rolled back
 * and run eagerly instead (*replayed = 0), and the graph is re-captured when a buffer it uses
 * has been reallocated. No counterpart in the reference (one edit per UI event, then a frame:
 * engine.cpp:575-718). width = height = 0 with null d_paths / d_colors builds an edit-only
 * graph (the pass and its readback; params may then be null). */
typedef struct hd_frame_graph hd_frame_graph;
hd_status hd_frame_graph_create(hd_ctx* ctx, uint32_t max_edits, uint32_t width, uint32_t height, uint32_t* d_paths,
                                uint32_t* d_colors, float shadow_bias, float fog_density, hd_frame_graph** out);
hd_status hd_frame_graph_launch(hd_frame_graph* graph, const hd_edit* edits, uint32_t n, const hd_ray_params* params,
                                hd_edit_stats* stats, uint32_t* replayed);
hd_status hd_frame_graph_destroy(hd_frame_graph* graph);

/* Request serving: trace and edit requests executed in arrival order with continuous
 * batching on one device-resident DAG (runs of edit requests fuse into one pass; runs of trace
 * requests run concurrently on trace slots; admission by free page budget, GC when short).
 * No counterpart in the reference, whose engine serves one frame loop from the UI thread
 * (engine.cpp:575-718); the work per request is hd_trace_* (DAGTracer::resolve_*) or
 * hd_edit_batch (HashDAG::edit_threads). */
#define HD_REQ_TRACE 0u
#define HD_REQ_EDIT 1u
#define HD_REQ_UNDO 2u  /* HashDAGUndoRedo::undo after every earlier request (replay.cpp:7-12) */
#define HD_REQ_REDO 3u
typedef struct hd_request {
    double arrival_s;        /* seconds after the start of the run, non-decreasing */
    uint32_t kind;           /* HD_REQ_TRACE / HD_REQ_EDIT / HD_REQ_UNDO / HD_REQ_REDO */
    uint32_t width, height;  /* trace: frame size */
    uint32_t first_edit;     /* edit: range in the edits array */
    uint32_t n_edits;
    uint32_t pad;
    hd_ray_params params;    /* trace: camera rays */
} hd_request;
typedef struct hd_request_result {
    double start_s, done_s;  /* when the request's batch started / the request completed */
    uint64_t checksum;       /* trace: frame checksum (hd_frame_checksum); edit: root after its batch */
    uint64_t voxels;         /* trace: pixels; edit: voxels edited by its whole batch */
    uint32_t batch;          /* batch the request ran in */
    int32_t status;
    uint32_t gc;             /* edit: 1 if stale nodes were collected before its batch */
    uint32_t pad;
} hd_request_result;
typedef struct hd_serve_config {
    uint32_t trace_slots;      /* concurrent traces (streams with frame buffers), 1..16 */
    uint32_t max_batch_edits;  /* cap on edits fused into one pass */
    uint32_t realtime;         /* 1: requests arrive at arrival_s; 0: all at once */
    uint32_t keep_undo;        /* 0: clear the undo history after every edit batch; 1: keep it, also across GC */
    uint64_t reserve_pages;    /* GC watermark: collect when free pages < estimate + reserve
                                  (a batch is rejected only if free pages < estimate after GC) */
    float shadow_bias, fog_density;
} hd_serve_config;
hd_status hd_serve_run(hd_ctx* ctx, const hd_request* reqs, uint32_t n, const hd_edit* edits, uint32_t n_edits,
                       const hd_serve_config* cfg, hd_request_result* out);
/* Memory state for admission control (the proposed hd_mem_info of a serving layer; the
 * reference has none: HashTable::allocate_page aborts when the pool is exhausted,
 * hash_table.h:794-809). free_pages = pages never handed out + mapped pages beyond the used
 * part of their bucket (reusable after GC). */
typedef struct hd_mem_info_t {
    uint64_t pool_pages, pool_top;
    uint64_t reusable_pages;  /* mapped pages past their bucket's used words */
    uint64_t free_pages;      /* pool_pages - pool_top + reusable_pages */
    uint64_t used_words;      /* sum of bucket sizes */
    uint64_t arena_words, arena_top;  /* colour arena capacity / used */
} hd_mem_info_t;
hd_status hd_mem_info(hd_ctx* ctx, hd_mem_info_t* out);
/* Order-independent 64-bit checksum of a width*height RGBA8 frame (sum over pixels of a
 * 64-bit mix of pixel and index). No counterpart in the reference (a test/serving aid). */
hd_status hd_frame_checksum(hd_ctx* ctx, const uint32_t* d_colors, uint32_t width, uint32_t height, uint64_t* out,
                            void* stream);

/* Streaming device copy of `bytes` (multiple of 16) from d_src to d_dst: measures the
 * achievable HBM bandwidth next to the tracer (not in the reference). */
hd_status hd_stream_copy(const void* d_src, void* d_dst, uint64_t bytes, void* stream);

/* Synthetic terrain scene (not in the reference, whose scenes are offline .bin files,
 * basic_dag.cpp:55-117): generates a BasicDAG plus colour tree and main leaf on the host. */
typedef struct hd_terrain_params {
    uint32_t levels;
    uint32_t footprint;        /* 0 = 2^levels */
    uint64_t seed;
    uint32_t octaves;
    double amplitude;
    double frequency;
    uint32_t color_tree_levels;
    uint32_t colors;           /* 0/1 */
    uint32_t color_seed;
    uint32_t threads;          /* 0 = all cores */
} hd_terrain_params;
typedef struct hd_scene hd_scene;
hd_status hd_scene_generate(const hd_terrain_params* p, hd_scene** out);
void hd_scene_free(hd_scene* s);
/* sizes[0..5] = dag words, color nodes, color offsets, blocks, weights, macro; sizes[6] = voxels */
hd_status hd_scene_sizes(const hd_scene* s, uint64_t sizes[7]);
hd_status hd_scene_copy(const hd_scene* s, uint32_t* dag, uint32_t* color_nodes, uint64_t* color_offsets,
                        uint64_t* blocks, uint32_t* weights, uint64_t* macro);
/* unique node counts per level (levels-1 entries), then BasicDAG words per level (levels-1) */
hd_status hd_scene_level_nodes(const hd_scene* s, uint64_t* out);
int32_t hd_terrain_height(const hd_terrain_params* p, uint32_t x, uint32_t z);

/* ---- files ---- (errors without a context: hd_last_error(NULL))
 * BasicDAGFactory::save_dag_to_file / load_dag_from_file (basic_dag.cpp:89-117, serializer.h:11-122):
 * DAGInfo (6 doubles), levels (uint32), uint64 word count, the BasicDAG words. Reading with
 * words = NULL returns the sizes only. */
hd_status hd_write_basic_dag_file(const char* path, const double bounds_min[3], const double bounds_max[3],
                                  uint32_t levels, const uint32_t* words, uint64_t n_words);
hd_status hd_read_basic_dag_file(const char* path, double bounds[6], uint32_t* levels, uint32_t* words,
                                 uint64_t capacity, uint64_t* n_words);
/* HashDAGFactory::save_dag_to_file / load_dag_from_file (hash_dag_factory.cpp:200-246): DAGInfo,
 * levels, firstNodeIndex, poolMaxSize, pageTableSize, poolTop, the pool pages, the page table.
 * The dump holds no bucket sizes or colours; loading walks the DAG from its root and builds the
 * context's table from it (the same DAG, re-interned). */
hd_status hd_save_hash_dag_file(hd_ctx* ctx, const char* path, const double bounds_min[3], const double bounds_max[3]);
hd_status hd_load_hash_dag_file(hd_ctx* ctx, const char* path, double bounds[6]);
/* BasicDAGFactory::save_compressed_colors_to_file / load_compressed_colors_from_file
 * (basic_dag.cpp:34-89): topLevels (uint32), then enclosedLeaves (uint64), weights (uint32),
 * blocks (uint64), macro blocks (uint64), each as a uint64 count and the data. Weight words are
 * given and returned in memory order and stored byte-swapped (ColorUtils::swap_byte_order,
 * color_utils.h:85-105). sizes[4] = counts of enclosedLeaves, weights, blocks, macro: on reading
 * with any array NULL the counts only, else the capacities in and the counts out. */
hd_status hd_write_compressed_colors_file(const char* path, uint32_t top_levels, const uint64_t* enclosed,
                                          uint64_t n_enclosed, const uint32_t* weights, uint64_t n_weights,
                                          const uint64_t* blocks, uint64_t n_blocks, const uint64_t* macro,
                                          uint64_t n_macro);
hd_status hd_read_compressed_colors_file(const char* path, uint32_t* top_levels, uint64_t sizes[4], uint64_t* enclosed,
                                         uint32_t* weights, uint64_t* blocks, uint64_t* macro);
/* The voxel-count headers the reference's compressed colours pair with
 * (BasicDAGColorsBase::get_leaves_count, basic_dag.h:60-72): out_words = words with every
 * interior header's upper 24 bits set to its subtree's voxel count, or, for levels < top_levels,
 * to an index into `enclosed` holding the count. *n_enclosed: capacity in (enclosed may be NULL
 * for the count only), count out. */
hd_status hd_basic_dag_counts(const uint32_t* words, uint64_t n_words, uint32_t levels, uint32_t top_levels,
                              uint32_t* out_words, uint64_t* enclosed, uint64_t* n_enclosed);
/* create_hash_dag_colors (hash_dag_factory.cpp:101-150) over a BasicDAG with such headers: the
 * colour tree (8 entries per node, pre-order; at level color_tree_levels - 1 every child slot
 * gets a leaf whose offset is the voxel count before it) as this library loads it
 * (hd_load_colors: handle 0 is "no leaf", so the builder's leaf k is handle k + 1, offsets[0] = 0),
 * and scene_words = the words with the headers above the colour tree reduced to the child mask.
 * *n_nodes / *n_offsets: capacities in (all three arrays NULL for the counts only), counts out. */
hd_status hd_colors_from_basic_dag(const uint32_t* words, uint64_t n_words, uint32_t levels, uint32_t color_tree_levels,
                                   uint32_t top_levels, const uint64_t* enclosed, uint64_t n_enclosed,
                                   uint32_t* scene_words, uint32_t* nodes, uint64_t* n_nodes, uint64_t* offsets,
                                   uint64_t* n_offsets);
/* Checkpoint of a context, no counterpart in the reference: the table image (with bucket sizes
 * and full nodes) and the colour state, restored exactly into a context of the same geometry.
 * The undo history is not saved and is cleared on load. */
hd_status hd_checkpoint_save(hd_ctx* ctx, const char* path);
hd_status hd_checkpoint_load(hd_ctx* ctx, const char* path);

/* ---- replays (ReplayManager, replay.h:12-635 / replay.cpp:1-101) ----
 * A replay is the reference's CSV action stream: camera moves, edits, undo / redo and frame
 * ends. hd_replay_parse reads it as ReplayManager::load_csv does (cells split at ',' and
 * lines at '\n', trailing empty cells dropped, numbers read as std::stof / std::stoi would)
 * and scales every edit by float(2^scene_depth) / float(2^replay_depth) (replay.cpp:5). */
#define HD_REPLAY_SET_LOCATION 0u
#define HD_REPLAY_SET_ROTATION 1u
#define HD_REPLAY_SET_TOOL 2u    /* SetToolParameters: interactive tool state, no effect on the DAG */
#define HD_REPLAY_EDIT 3u        /* EditSphere / EditCube / EditCopy / EditPaint / EditFill: `edit` */
#define HD_REPLAY_UNDO 4u
#define HD_REPLAY_REDO 5u
#define HD_REPLAY_END_FRAME 6u

typedef struct hd_replay hd_replay;
typedef struct hd_replay_action {
    uint32_t type;          /* HD_REPLAY_* */
    uint32_t frame;         /* frame the action belongs to (frames end at HD_REPLAY_END_FRAME) */
    double location[3];     /* SetLocation: camera position */
    double rotation[9];     /* SetRotation: camera rotation, rows */
    hd_edit edit;           /* edits: scaled to the scene depth */
} hd_replay_action;

/* ReplayManager::load_csv (replay.h:565-607); unknown actions or malformed cells are errors
 * (message: hd_last_error(NULL)). */
hd_status hd_replay_parse(const char* text, uint64_t length, uint32_t scene_depth, uint32_t replay_depth,
                          hd_replay** out);
/* the parsed actions (owned by the replay) and the frame count */
hd_status hd_replay_actions(const hd_replay* replay, const hd_replay_action** actions, uint64_t* n_actions,
                            uint32_t* n_frames);
void hd_replay_free(hd_replay* replay);

#ifdef __cplusplus
}
#endif

#endif /* HASHDAG_HIP_H */
