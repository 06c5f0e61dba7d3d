// Device edit pipeline: shared structures (see hd_edit.hip).
#pragma once
#include <cstddef>

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "hd_internal.h"

struct hd_ctx;

namespace hd {

constexpr uint32_t EDIT_SPHERE_ADD = 0, EDIT_SPHERE_CARVE = 1, EDIT_BOX_ADD = 2, EDIT_BOX_CARVE = 3,
                   EDIT_SPHERE_PAINT = 4, EDIT_COPY = 5, EDIT_FILL = 6;
// kind of the C ABI's decompressing copy (hd_edit.kind); its DevEdit has kind EDIT_COPY, dec = 1
constexpr uint32_t EDIT_COPY_DEC = 7;
// SphereNoiseEditor adding / carving (hd_edit.kind); both DevEdits have kind EDIT_NOISE (the
// buffer they read holds the voxels' values after the edit)
constexpr uint32_t EDIT_NOISE = 8, EDIT_NOISE_ADD = 8, EDIT_NOISE_CARVE = 9;
inline bool is_copy_kind(uint32_t kind) { return kind == EDIT_COPY || kind == EDIT_COPY_DEC; }
inline bool is_noise_kind(uint32_t kind) { return kind == EDIT_NOISE_ADD || kind == EDIT_NOISE_CARVE; }
// editors that read a voxel buffer built from the DAG before their pass (each starts a pass)
inline bool is_buffer_kind(uint32_t kind) { return kind == EDIT_FILL || is_noise_kind(kind); }

// error bits reported by a pass
constexpr uint32_t ERR_SCRATCH = 1, ERR_BUCKET = 2, ERR_POOL = 4, ERR_HEADER = 8, ERR_COLOR_SCRATCH = 16,
                   ERR_COLOR_ARENA = 32, ERR_COLOR_TREE = 64, ERR_UNDO = 128, ERR_EMPTY = 256,
                   ERR_LEAF_SIZE = 512, ERR_KIND = 1024;

constexpr uint32_t kMaxEditLevels = 17;
constexpr uint32_t kUndoCap = 1u << 20;

// An editor with its derived bounds (hash_dag_editors.h:57-127): box min/max = c -/+ r,
// r2 = r * r, `single` = rgb101210 colour of add editors (sphere red, box magenta).
// A copy is a box around dest + (center - source) whose voxels read source + (p - dest).
struct DevEdit {
    uint32_t kind;
    float c[3];
    float radius;
    float r2;
    float bmin[3], bmax[3];
    uint32_t single;
    float src[3], dst[3];
    uint32_t swirl;      // copy: turn the relative position about y by y / swirl_period
    float swirl_period;
    uint32_t pad;
    double inv[9];       // copy: inverse of the copy's transform (row-major, FP64)
    // fill and noise: per buffer voxel (x + size * (y + size * z), buffer origin src[]) bit 0 =
    // the voxel is set after the edit, bit 1 = the flood fill set it; colours of the fill's
    // bounding box (lo, dim) for the voxels it set
    const uint8_t* fill_state;
    const uint64_t* fill_colors;
    uint32_t fill_size;
    uint32_t fill_lo[3], fill_dim[3];
    // copy with decompression (kind COPY with dec = 1, hash_dag_editors.h:437-616): the source
    // box's float minimum center - radius, its truncation, and the side of its voxel buffer
    uint32_t dec;
    float cmin[3];
    uint32_t icmin[3];
    uint32_t dsize;
    uint32_t pad3;
};

// The leading 64 bytes of a DevEdit: what the sphere / box editors' tests and colours read,
// loaded with one wide load per edit.
struct alignas(16) EditHead {
    uint32_t kind;
    float c[3];
    float radius;
    float r2;
    float bmin[3], bmax[3];
    uint32_t single;
    float src[3];
};
static_assert(sizeof(EditHead) == 64 && sizeof(DevEdit) % 16 == 0, "edit head layout");
static_assert(offsetof(DevEdit, single) == offsetof(EditHead, single) && offsetof(DevEdit, src) == offsetof(EditHead, src),
              "edit head layout");

// Device buffers of one prepared fill edit (owned by the host call that runs its pass).
struct FillBuffers {
    uint8_t* state = nullptr;
    uint64_t* colors = nullptr;
    uint8_t* tables = nullptr;  // noise: the permutation tables
    void release() {
        hipFree(state);
        hipFree(colors);
        hipFree(tables);
        state = nullptr;
        colors = nullptr;
        tables = nullptr;
    }
    ~FillBuffers() { release(); }
};

// float -> uint32 as x86-64 code converts it (truncate to int64, keep the low 32 bits)
inline uint32_t f2u_host(float f) { return (uint32_t)(int64_t)f; }

// Work items of one level (structure of arrays, capacity `cap`; child arrays are 8 * cap).
struct Items {
    uint32_t *px, *py, *pz;   // path of the item's node
    uint32_t* old_;           // effective old node (may be a full node after a full cover)
    uint32_t* orig;           // the node the parent actually points to
    uint32_t* parent;
    uint8_t* slot;
    uint32_t *estart, *ecount;  // edit sub-list in elist[level]
    uint32_t* aux;              // levels < ctl: colour node; levels >= ctl: task (level-ctl item)
    uint32_t* handle;           // tasks: colour leaf handle before the pass
    uint32_t* newhandle;        // tasks: unique leaf handle written by the pass
    uint32_t *cmode, *cmodeval;  // old colours: 0 = old leaf, 1 = single colour cmodeval, 2 = none
    uint8_t* ckind;             // per child slot: keep / empty / full / item
    uint32_t *cval, *corig, *cold, *cnew, *ccolor, *cstart, *costart;
    uint32_t *nidx, *ncount;    // result: new node index and voxel count
    uint32_t *cont, *clen, *hash, *rep, *state;  // pending node content (9 words) for the table
    uint32_t *nstart, *ostart;  // colour starts relative to the task
    uint64_t* cbase;            // tasks: base of the task's colour array
};

struct EditPass {
    Geom g;
    uint32_t levels, ctl;
    uint32_t* vpool;
    uint32_t* page_table;
    uint32_t* bucket_sizes;
    const uint32_t* full_nodes;
    uint32_t* pool_top;
    uint32_t* errors;
    uint8_t* leaf_tags;  // tag byte per leaf slot (leaf level), kept by k_add, read by k_find
    // algorithmic bytes when measuring (else null): [0] node words read by the descent,
    // [1] bucket bytes read by finds
    unsigned long long* abytes;
    uint32_t* n_cnodes;
    uint32_t* n_leaves;
    uint32_t* undo_count;
    uint32_t pool_pages;
    uint32_t* count;   // items per level
    uint32_t* ecount;  // edit-list words per level
    Items items[kMaxEditLevels];
    uint32_t* elist[kMaxEditLevels];
    uint32_t cap[kMaxEditLevels], ecap[kMaxEditLevels];  // item / edit-list capacity per level
    uint32_t* hset;
    uint32_t hcap;
    const DevEdit* edits;
    uint32_t single_edit;
    uint32_t wide_levels;  // bit l: level l's expand runs a wavefront per child slot
    unsigned long long* stat_voxels;
    unsigned long long* stat_new_words;
    unsigned long long* color_top;
    unsigned long long* arena_top;
    unsigned long long* stat_colors;  // colours staged (color_top also counts the chunk padding)
    uint32_t* find_count;             // per level: representatives listed for k_find
    uint64_t* colors;
    uint64_t* cmarks;  // per 64-colour chunk of `colors`: kUniformChunk | colour, or 0 (explicit)
    uint64_t color_cap;
    uint32_t* cnodes;
    uint32_t cnodes_cap;
    const uint64_t* offsets;
    LeafDesc* leaves;
    uint32_t leaves_cap;
    LeafDesc main;
    uint64_t* arena;
    uint64_t arena_cap;
    uint32_t* undo_handles;
    LeafDesc* undo_descs;
    uint32_t undo_cap;
};

// Per-pass scratch. Item arrays and edit lists are sized per level (one allocation each) from
// what passes ask for; the uploaded edits and the undo log survive re-reserves.
struct EditScratch {
    uint32_t levels = 0, hcap = 0, edits_cap = 0, undo_cap = 0;
    uint32_t item_cap[kMaxEditLevels]{}, elist_cap[kMaxEditLevels]{};
    uint64_t color_cap = 0;
    Items items[kMaxEditLevels]{};
    uint32_t* elist[kMaxEditLevels]{};
    void* level_mem[kMaxEditLevels]{};
    uint32_t* hset = nullptr;
    uint64_t* colors = nullptr;
    uint64_t* cmarks = nullptr;
    uint32_t* dcount = nullptr;
    uint32_t* decount = nullptr;
    uint64_t* dstats = nullptr;
    uint32_t* undo_handles = nullptr;
    LeafDesc* undo_descs = nullptr;
    LeafDesc* undo_next = nullptr;
    DevEdit* edits_dev = nullptr;
    std::vector<void*> allocs;
};

// Lays out one level's item arrays and edit list from `mem` (nullptr: sizes only); returns
// the bytes they take (hd_edit.hip).
uint64_t carve_level(Items& I, uint32_t*& elist, bool leafLevel, uint64_t c, uint64_t e, char* mem);

}  // namespace hd
