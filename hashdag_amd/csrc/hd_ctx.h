// Context object behind the C ABI (one per device-resident HashDAG).
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/hashdag_hip.h"
#include "hd_edit.h"
#include "hd_internal.h"

namespace hd {
// Scratch of one stream's persistent traces: tile queue heads, and per pixel tile the packed
// shadow rays (origin + pixel), their lit / shadowed colours and their count.
struct TraceScratch {
    hipStream_t stream = nullptr;
    uint32_t* queues = nullptr;
    uint4* rec = nullptr;
    uint2* shades = nullptr;
    uint32_t* cnt = nullptr;
    uint64_t tiles_cap = 0;
    // rays handed from the capped per-tile walk to the wavefront-per-ray search, and from it
    // back to the plain walk: pixel lists of pixels_cap entries, counts in long_count[0..1]
    uint32_t* long_items = nullptr;
    uint32_t* rest_items = nullptr;
    uint32_t* long_count = nullptr;
    // walks stopped at the per-tile cap and saved for the resume kernel (schedule 3)
    uint32_t* resume_items = nullptr;
    uint32_t* resume_state = nullptr;
    uint64_t pixels_cap = 0;
};

// One undo step (HashDAGUndoRedo::Frame, hash_dag.h:29-45): the roots before / after and the
// colour-leaf descriptors the pass replaced (old leaf data stays in the colour arena).
struct UndoFrame {
    uint32_t prev_root = 0, new_root = 0;
    std::vector<uint32_t> handles;
    std::vector<LeafDesc> prev, next;
};
}  // namespace hd

struct hd_ctx {
    hd::Geom g{};
    hd_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // hash table
    uint32_t* vpool = nullptr;          // total_virtual + kVpoolPad words
    uint32_t* page_table = nullptr;     // total_pages
    uint32_t* bucket_sizes = nullptr;   // total_buckets
    uint32_t* bucket_snapshot = nullptr;  // total_buckets: sizes before the running edit pass
    uint32_t* full_nodes = nullptr;     // levels
    uint32_t* counters = nullptr;       // [0] pool_top, [1] error flags, [2..] scratch counters
    // leaf level: one tag byte per 2-word slot of every bucket (bits 24-31 of the leaf's hash),
    // valid below each bucket's size; finds scan the tags before comparing leaves
    uint8_t* leaf_tags = nullptr;
    // edit pass measurement (hd_edit_count_bytes): device counters of the kernels, host sums of
    // the passes' own statistics
    unsigned long long* edit_bytes = nullptr;
    bool count_edit_bytes = false;
    // pass steps on levels with at most edit_block_max work items run together in one
    // single-workgroup launch; passes of at most edit_persist_max edits as one cooperative
    // launch (0: never)
    uint32_t edit_block_max = 0;
    uint32_t edit_persist_max = 0;
    uint32_t edit_persist_per_cu = 2;  // resident blocks per CU of that launch (at most what fits)
    uint64_t edit_bytes_host[3] = {};
    uint32_t root = 0;
    bool loaded = false;

    // colours
    uint32_t* cnodes = nullptr;
    uint64_t n_cnodes = 0;
    uint64_t* coffsets = nullptr;
    uint64_t n_coffsets = 0;
    hd::LeafDesc* cleaves = nullptr;
    uint64_t n_cleaves = 0;
    std::vector<hd::LeafDesc> h_leaves;  // host mirror of the leaf descriptors
    uint64_t* arena = nullptr;          // colour arena (64-bit words)
    uint64_t arena_top = 0;
    hd::LeafDesc main{};
    bool colors_loaded = false;

    // tracer
    uint32_t tile_w = 8, tile_h = 8;
    uint32_t fetch_mode = 0;  // tracer node fetch: 0 = pointer then header, 1 = whole node
    unsigned long long* byte_counters = nullptr;  // 3 counters for instrumented launches
    // tracer schedule: 0 = one wavefront per pixel tile, 1 = persistent wavefronts fed from
    // tile queues, refilled once refill_at lanes are idle
    uint32_t trace_persist = 0, refill_at = 16;
    // per-tile schedule: walks longer than long_steps steps (0: no cap) are finished by the
    // wavefront-per-ray search
    uint32_t long_steps = 0;
    // schedule 3: the per-tile walk stops at resume_steps and compacted wavefronts continue it
    uint32_t resume_steps = 0;
    uint32_t persist_blocks[4] = {};   // resident blocks per persistent kernel (0: not measured)
    // scratch of the persistent schedule, one set per stream traced on (traces on different
    // streams of one context may overlap)
    std::vector<hd::TraceScratch> trace_scratch;

    // edits
    uint32_t* d_root = nullptr;  // device copy of the root (edit passes chain through it)
    hd::EditScratch edit;
    std::vector<hd::UndoFrame> undo, redo;
    // GC scratch, kept between collections and grown when a call needs more (freed with the
    // context): slot k holds (buffer, bytes)
    std::vector<std::pair<void*, uint64_t>> gc_slots;
    // bumped whenever a device buffer or launch shape that a captured graph bakes in changes
    uint64_t ptr_gen = 0;
};

namespace hd {
constexpr uint64_t kVpoolPad = 64;  // words after the virtual space (vector reads past a node end)
constexpr uint32_t kArenaTopWord = 8;  // counters[8..9]: committed colour arena top (64-bit)

void edit_free(hd_ctx* ctx);
// releases the GC scratch kept in the context
void gc_scratch_free(hd_ctx* ctx);
// device address of the static edit scratch on ctx's device, looked up once per device
hd_status static_scratch_setup(hd_ctx* ctx);
// the persistent tracer's scratch for stream s, grown to hold a W x H frame (allocates: not
// under stream capture); trace_scratch_free releases every stream's set
hd_status trace_reserve(hd_ctx* ctx, hipStream_t s, uint32_t W, uint32_t H, TraceScratch** out = nullptr);
void trace_scratch_free(hd_ctx* ctx);
// recomputes the leaf tags of every leaf-level bucket from the table (after uploads and GC)
hd_status leaf_tags_rebuild(hd_ctx* ctx, hipStream_t s);
// bytes of the leaf tag array of a geometry
inline uint64_t leaf_tag_bytes(const Geom& g) {
    return (uint64_t)g.buckets_per_level(g.leaf_level()) * (g.bucket_size(g.leaf_level()) / 2);
}
hd_status fail(hd_ctx* ctx, hd_status code, const std::string& msg);
hd_status hip_fail(hd_ctx* ctx, hipError_t e, const char* what);
DevDag dev_dag(const hd_ctx* ctx);
DevColors dev_colors(const hd_ctx* ctx);
// physical pool pages 0 .. pool_top - 1 into `pool` (at most cap_words words), zeros after them
hd_status download_pool(hd_ctx* ctx, uint32_t* pool, uint64_t cap_words);
}  // namespace hd

#define HD_HIP(ctx, call)                                          \
    do {                                                           \
        hipError_t e_ = (call);                                    \
        if (e_ != hipSuccess) return hd::hip_fail(ctx, e_, #call); \
    } while (0)
