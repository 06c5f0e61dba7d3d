// C ABI of the device edits: HashDAG::edit_threads for one edit per pass (plain) or a
// whole batch fused into one pass, edited-voxel counts, undo / redo over root history and
// colour-leaf descriptors (HashDAGUndoRedo, hash_dag.h:20-212).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "hd_ctx.h"

namespace hd {

hd_status edit_reserve(hd_ctx* ctx, const uint32_t* items, const uint32_t* elist, uint64_t colors);
void edit_free(hd_ctx* ctx);
hd_status enqueue_pass(hd_ctx* ctx, uint32_t first, uint32_t n, bool colors, hipStream_t s,
                       const uint32_t* dN = nullptr, bool special = true, const EditScratch* scratch = nullptr,
                       const DevEdit* hostEdits = nullptr);
hd_status edit_ensure_edits(hd_ctx* ctx, uint32_t n);
hd_status enqueue_frame_from_device(hd_ctx* ctx, const hd_ray_params* d_params, uint32_t W, uint32_t H,
                                    uint32_t* d_paths, uint32_t* d_colors, float bias, float fog, hipStream_t s);
hd_status edit_reserve_undo(hd_ctx* ctx, uint64_t entries);
hd_status edit_scatter_leaves(hd_ctx* ctx, const uint32_t* handles, const LeafDesc* descs, uint32_t n);
hd_status fill_prepare(hd_ctx* ctx, const hd_edit& e, DevEdit& d, FillBuffers& fb, hipStream_t s);
hd_status noise_prepare(hd_ctx* ctx, const hd_edit& e, DevEdit& d, FillBuffers& fb, hipStream_t s);
void edit_restore_leaves(hd_ctx* ctx, hipStream_t s, const EditScratch* scratch = nullptr);
hipError_t edit_gather_next(hd_ctx* ctx, uint32_t n, LeafDesc* host, hipStream_t s);

namespace {

// edited-voxel count of one editor: voxels of [0, E)^3 whose unit box the editor touches
// (the reference's "num voxels", hash_dag.h:374-378; touch = should_edit_impl on the voxel).
__device__ __forceinline__ bool touch_voxel(const DevEdit& e, float x, float y, float z) {
    const float x1 = x + 1.0f, y1 = y + 1.0f, z1 = z + 1.0f;
    if (e.kind == EDIT_BOX_ADD || e.kind == EDIT_BOX_CARVE)
        return !(e.bmin[0] >= x1 || e.bmin[1] >= y1 || e.bmin[2] >= z1 || e.bmax[0] <= x || e.bmax[1] <= y ||
                 e.bmax[2] <= z || x >= e.bmax[0] || y >= e.bmax[1] || z >= e.bmax[2] || x1 <= e.bmin[0] ||
                 y1 <= e.bmin[1] || z1 <= e.bmin[2]);
    float dist = e.r2;
    if (e.c[0] < x) dist -= (e.c[0] - x) * (e.c[0] - x);
    else if (e.c[0] > x1) dist -= (e.c[0] - x1) * (e.c[0] - x1);
    if (e.c[1] < y) dist -= (e.c[1] - y) * (e.c[1] - y);
    else if (e.c[1] > y1) dist -= (e.c[1] - y1) * (e.c[1] - y1);
    if (e.c[2] < z) dist -= (e.c[2] - z) * (e.c[2] - z);
    else if (e.c[2] > z1) dist -= (e.c[2] - z1) * (e.c[2] - z1);
    return dist > 0;
}

// One block per editor; threads stride over the (x, z) columns of its bounding square; the
// touched y values of a column form an interval (the predicate is monotone in |y - c.y|),
// found by two binary searches around the centre row.
__global__ void k_count_voxels(const DevEdit* __restrict__ edits, uint32_t n, const uint32_t* __restrict__ dN,
                               uint32_t E, unsigned long long* __restrict__ out) {
    if (dN) n = min(n, *dN);  // graph replays read the edit count from device memory
    for (uint32_t ei = blockIdx.x; ei < n; ei += gridDim.x) {
        const DevEdit e = edits[ei];
        if (e.kind == EDIT_COPY || e.kind == EDIT_FILL || e.kind == EDIT_NOISE) continue;  // counted on the host
        const float r = e.radius;
        const int64_t x0 = std::max<int64_t>(0, (int64_t)floorf(e.c[0] - r) - 1);
        const int64_t x1 = std::min<int64_t>((int64_t)E - 1, (int64_t)ceilf(e.c[0] + r) + 1);
        const int64_t z0 = std::max<int64_t>(0, (int64_t)floorf(e.c[2] - r) - 1);
        const int64_t z1 = std::min<int64_t>((int64_t)E - 1, (int64_t)ceilf(e.c[2] + r) + 1);
        const int64_t ylo = std::max<int64_t>(0, (int64_t)floorf(e.c[1] - r) - 1);
        const int64_t yhi = std::min<int64_t>((int64_t)E - 1, (int64_t)ceilf(e.c[1] + r) + 1);
        if (x1 < x0 || z1 < z0 || yhi < ylo) continue;
        const int64_t nx = x1 - x0 + 1, nz = z1 - z0 + 1;
        unsigned long long acc = 0;
        const int64_t ymid = std::min(yhi, std::max(ylo, (int64_t)floorf(e.c[1])));
        for (int64_t col = threadIdx.x; col < nx * nz; col += blockDim.x) {
            const float x = (float)(x0 + col % nx), z = (float)(z0 + col / nx);
            if (!touch_voxel(e, x, (float)ymid, z)) continue;
            // first touched y in [ylo, ymid] and last touched y in [ymid, yhi]
            int64_t a = ylo, b = ymid;
            while (a < b) {
                const int64_t m = (a + b) / 2;
                if (touch_voxel(e, x, (float)m, z)) b = m;
                else a = m + 1;
            }
            const int64_t first = a;
            a = ymid;
            b = yhi;
            while (a < b) {
                const int64_t m = (a + b + 1) / 2;
                if (touch_voxel(e, x, (float)m, z)) a = m;
                else b = m - 1;
            }
            acc += (unsigned long long)(a - first + 1);
        }
        atomicAdd(out, acc);
    }
}

__global__ void k_reset(uint32_t* counters, uint64_t* stats) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        counters[1] = 0;  // errors
        counters[4] = 0;  // undo log
        for (int i = 0; i < 32; ++i) stats[i] = 0;  // stats, and the per-level find-list counts
        stats[3] = *reinterpret_cast<const uint64_t*>(counters + kArenaTopWord);  // committed arena top
    }
}

// The copy's transform; all zeros stands for the identity (callers that zero the struct).
static void copy_transform(const hd_edit& e, double T[9]) {
    bool zero = true;
    for (int i = 0; i < 9; ++i) zero &= e.transform[i] == 0.0;
    for (int i = 0; i < 9; ++i) T[i] = zero ? (i % 4 == 0 ? 1.0 : 0.0) : e.transform[i];
}

DevEdit make_edit(const hd_edit& e) {
    DevEdit d{};
    d.kind = e.kind;
    for (int i = 0; i < 3; ++i) d.c[i] = e.center[i];
    d.radius = e.radius;
    d.r2 = e.radius * e.radius;
    float box[3] = {e.center[0], e.center[1], e.center[2]};
    float boxRadius = e.radius;
    if (e.kind == EDIT_COPY_DEC) {
        // BoxEditorBase(center - source + dest, radius); values read through a buffer of
        // 2 ceil(radius) voxels per axis from truncate(center - radius) (hash_dag_editors.h:455-466)
        d.kind = EDIT_COPY;
        d.dec = 1;
        for (int i = 0; i < 3; ++i) {
            box[i] = e.center[i] - e.source[i] + e.dest[i];
            d.cmin[i] = e.center[i] - e.radius;
            d.icmin[i] = f2u_host(d.cmin[i]);
            d.inv[4 * i] = 1.0;  // identity: the transform and swirl do not apply
        }
        d.dsize = 2u * (uint32_t)std::ceil(e.radius);
    } else if (e.kind == EDIT_COPY) {
        // BoxEditorBase(dest + float3(T * (center - source)), radius * max|T * (1,1,1)|)
        // (hash_dag_editors.h:350-351; Matrix3x3 in FP64, Matrix3x3.h:190-243, 399-406)
        double T[9];
        copy_transform(e, T);
        const double r[3] = {(double)(e.center[0] - e.source[0]), (double)(e.center[1] - e.source[1]),
                             (double)(e.center[2] - e.source[2])};
        double rowMax = 0;
        for (int i = 0; i < 3; ++i) {
            box[i] = e.dest[i] + (float)(T[3 * i] * r[0] + T[3 * i + 1] * r[1] + T[3 * i + 2] * r[2]);
            rowMax = std::max(rowMax, std::fabs(T[3 * i] * 1.0 + T[3 * i + 1] * 1.0 + T[3 * i + 2] * 1.0));
        }
        boxRadius = (float)((double)e.radius * rowMax);
        // Matrix3x3::Inverse: adjugate times 1 / Determinate
        const double a[9] = {T[8] * T[4] - T[7] * T[5], T[7] * T[2] - T[8] * T[1], T[5] * T[1] - T[4] * T[2],
                             T[6] * T[5] - T[8] * T[3], T[8] * T[0] - T[6] * T[2], T[3] * T[2] - T[5] * T[0],
                             T[7] * T[3] - T[6] * T[4], T[6] * T[1] - T[7] * T[0], T[4] * T[0] - T[3] * T[1]};
        const double det = T[0] * (T[8] * T[4] - T[7] * T[5]) - T[3] * (T[8] * T[1] - T[7] * T[2]) +
                           T[6] * (T[5] * T[1] - T[4] * T[2]);
        const double sc = 1 / det;
        for (int i = 0; i < 9; ++i) d.inv[i] = a[i] * sc;
        d.swirl = e.swirl ? 1u : 0u;
        d.swirl_period = e.swirl_period;
    }
    for (int i = 0; i < 3; ++i) {
        d.bmin[i] = box[i] - boxRadius;
        d.bmax[i] = box[i] + boxRadius;
        d.src[i] = e.source[i];
        d.dst[i] = e.dest[i];
    }
    auto c101210 = [](float r, float g, float b) {
        return (uint32_t)(r * 1023.0f) | ((uint32_t)(g * 4095.0f) << 10) | ((uint32_t)(b * 1023.0f) << 22);
    };
    d.single = (e.kind == EDIT_BOX_ADD || e.kind == EDIT_BOX_CARVE) ? c101210(1, 0, 1) : c101210(1, 0, 0);
    return d;
}

// "num voxels" of copies: the box volume as a 32-bit product (BoxEditorBase::edit_count_impl,
// hash_dag_editors.h:92-96; hash_dag.h:374-378 reports it instead of the voxel count)
uint64_t copy_voxels(const hd_edit* e, uint32_t n) {
    uint64_t sum = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (!is_copy_kind(e[i].kind)) continue;
        const DevEdit d = make_edit(e[i]);
        sum += (uint32_t)(uint32_t)(d.bmax[0] - d.bmin[0]) * (uint32_t)(d.bmax[1] - d.bmin[1]) *
               (uint32_t)(d.bmax[2] - d.bmin[2]);
    }
    return sum;
}

// "num voxels" of a fill in a pass with other editors: the voxels of the scene whose unit box
// overlaps the fill's bounds, the voxels should_edit lets through (hash_dag_editors.h:76-91)
uint64_t fill_voxels(const DevEdit& d, uint32_t E) {
    uint64_t n = 1;
    for (int a = 0; a < 3; ++a) {
        const int64_t lo = std::max<int64_t>(0, (int64_t)std::floor((double)d.bmin[a]));
        const int64_t hi = std::min<int64_t>((int64_t)E - 1, (int64_t)std::ceil((double)d.bmax[a]) - 1);
        if (hi < lo) return 0;
        n *= (uint64_t)(hi - lo + 1);
    }
    return n;
}

struct Readback {
    uint32_t counters[8];
    uint64_t stats[8];
    uint32_t root;
    uint32_t count_ctl;
};

}  // namespace

static hd_status ensure_color_tree(hd_ctx* ctx) {
    if (ctx->colors_loaded) return HD_OK;
    const uint32_t zeros[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint64_t off0 = 0;
    return hd_load_colors(ctx, zeros, 8, &off0, 1, nullptr, 0, nullptr, 0, nullptr, 0);
}

static hd_status grow_arena(hd_ctx* ctx, uint64_t need) {
    const uint64_t cap = ctx->cfg.color_arena_words;
    if (ctx->arena_top + need <= cap) return HD_OK;
    const uint64_t ncap = std::max<uint64_t>(cap * 2, ctx->arena_top + need + (1u << 24));
    uint64_t* na = nullptr;
    HD_HIP(ctx, hipMalloc(&na, ncap * 8));
    HD_HIP(ctx, hipMemsetAsync(na, 0, ncap * 8, ctx->stream));
    HD_HIP(ctx, hipMemcpyAsync(na, ctx->arena, ctx->arena_top * 8, hipMemcpyDeviceToDevice, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    hipFree(ctx->arena);
    ctx->arena = na;
    ctx->ptr_gen++;
    ctx->cfg.color_arena_words = ncap;
    return HD_OK;
}

// Colour tree entries / unique leaf slots: reallocate (copying what is in use) so that at
// least `nodes` entries and `leaves` slots fit.
static hd_status grow_color_tables(hd_ctx* ctx, uint64_t nodes, uint64_t leaves) {
    HD_HIP(ctx, hipDeviceSynchronize());
    ctx->ptr_gen++;
    if (nodes > ctx->cfg.color_nodes_cap) {
        const uint64_t cap = std::min<uint64_t>(std::max<uint64_t>(nodes + nodes / 2, 2ull * ctx->cfg.color_nodes_cap), 0xFFFFFFF0u);
        if (cap < nodes) return fail(ctx, HD_ERR_COLOR_FULL, "colour tree exceeds 2^32 entries");
        uint32_t* nn = nullptr;
        HD_HIP(ctx, hipMalloc(&nn, cap * 4));
        HD_HIP(ctx, hipMemset(nn, 0, cap * 4));
        HD_HIP(ctx, hipMemcpy(nn, ctx->cnodes, (uint64_t)ctx->cfg.color_nodes_cap * 4, hipMemcpyDeviceToDevice));
        hipFree(ctx->cnodes);
        ctx->cnodes = nn;
        ctx->cfg.color_nodes_cap = (uint32_t)cap;
    }
    if (leaves > ctx->cfg.color_leaves_cap) {
        const uint64_t cap = std::min<uint64_t>(std::max<uint64_t>(leaves + leaves / 2, 2ull * ctx->cfg.color_leaves_cap), 0x7FFFFFFFu);
        if (cap < leaves) return fail(ctx, HD_ERR_COLOR_FULL, "more than 2^31 colour leaves");
        LeafDesc* nl = nullptr;
        HD_HIP(ctx, hipMalloc(&nl, cap * sizeof(LeafDesc)));
        HD_HIP(ctx, hipMemcpy(nl, ctx->cleaves, (uint64_t)ctx->cfg.color_leaves_cap * sizeof(LeafDesc), hipMemcpyDeviceToDevice));
        hipFree(ctx->cleaves);
        ctx->cleaves = nl;
        ctx->cfg.color_leaves_cap = (uint32_t)cap;
    }
    return HD_OK;
}

// The undo frame of a committed pass: the touched slots with their previous and new leaves.
static hd_status capture_frame(hd_ctx* ctx, uint32_t nu, UndoFrame& frame, hipStream_t s) {
    const EditScratch& S = ctx->edit;
    frame.handles.resize(nu);
    frame.prev.resize(nu);
    frame.next.resize(nu);
    if (nu) {
        HD_HIP(ctx, hipMemcpyAsync(frame.handles.data(), S.undo_handles, nu * 4, hipMemcpyDeviceToHost, s));
        HD_HIP(ctx, hipMemcpyAsync(frame.prev.data(), S.undo_descs, nu * sizeof(LeafDesc), hipMemcpyDeviceToHost, s));
        HD_HIP(ctx, edit_gather_next(ctx, nu, frame.next.data(), s));
        HD_HIP(ctx, hipStreamSynchronize(s));
    }
    return HD_OK;
}

// Colour tree, default scratch and undo log of a context's first edit.
static hd_status edit_prepare(hd_ctx* ctx) {
    hd_status st = ensure_color_tree(ctx);
    if (st != HD_OK) return st;
    if (!ctx->edit.level_mem[0]) {
        uint32_t items[kMaxEditLevels], elist[kMaxEditLevels];
        for (uint32_t l = 0; l < kMaxEditLevels; ++l) {
            items[l] = 1u << 16;
            elist[l] = 1u << 18;
        }
        st = edit_reserve(ctx, items, elist, 1u << 22);
        if (st != HD_OK) return st;
    }
    return edit_reserve_undo(ctx, kUndoCap);
}

// `special`: the pass holds a copy or a fill editor (leaf kernels built with their code).
// `fixed`: run in this scratch and grow nothing (the colour arena, tables and undo log
// included); a pass that needs more fails with the previous state kept.
static hd_status run_pass(hd_ctx* ctx, uint32_t first, uint32_t n, hipStream_t s, Readback& rb, UndoFrame* frame,
                          bool special, const EditScratch* fixed = nullptr, const DevEdit* hostEdits = nullptr) {
    const bool colors = ctx->colors_loaded && ctx->g.leaf_level() > ctx->g.color_tree_levels;
    for (int attempt = 0; attempt < 40; ++attempt) {
        hd_status st = HD_OK;
        if (!fixed) st = grow_arena(ctx, 1u << 22);
        if (st != HD_OK) return st;
        const EditScratch& S = fixed ? *fixed : ctx->edit;
        const uint64_t bucketBytes = (uint64_t)ctx->g.total_buckets() * 4;
        HD_HIP(ctx, hipMemcpyAsync(ctx->bucket_snapshot, ctx->bucket_sizes, bucketBytes, hipMemcpyDeviceToDevice, s));
        k_reset<<<1, 1, 0, s>>>(ctx->counters, S.dstats);
        st = enqueue_pass(ctx, first, n, colors, s, nullptr, special, fixed, hostEdits);
        if (st != HD_OK) return st;
        if (n > 1) {
            k_count_voxels<<<std::min<uint32_t>(n, 4096), 256, 0, s>>>(S.edits_dev + first, n, nullptr,
                                                                      1u << ctx->g.levels,
                                                                      reinterpret_cast<unsigned long long*>(S.dstats + 0));
        }
        HD_HIP(ctx, hipMemcpyAsync(rb.counters, ctx->counters, sizeof(rb.counters), hipMemcpyDeviceToHost, s));
        HD_HIP(ctx, hipMemcpyAsync(rb.stats, S.dstats, sizeof(rb.stats), hipMemcpyDeviceToHost, s));
        HD_HIP(ctx, hipMemcpyAsync(&rb.root, ctx->d_root, 4, hipMemcpyDeviceToHost, s));
        HD_HIP(ctx, hipMemcpyAsync(&rb.count_ctl, S.dcount + ctx->g.color_tree_levels, 4, hipMemcpyDeviceToHost, s));
        HD_HIP(ctx, hipStreamSynchronize(s));
        const uint32_t err = rb.counters[1];
        if (err) {
            // The root was not switched. Leaf slots the pass already re-pointed get their
            // previous leaves back. The buckets get their sizes from before the pass: what it
            // appended is dropped (pages it mapped stay mapped, as after GC), so a retry or a
            // later find never meets a node of a failed pass.
            HD_HIP(ctx, hipMemcpyAsync(ctx->bucket_sizes, ctx->bucket_snapshot, bucketBytes, hipMemcpyDeviceToDevice, s));
            HD_HIP(ctx, hipStreamSynchronize(s));
            if (rb.counters[4]) {
                edit_restore_leaves(ctx, s, fixed);
                HD_HIP(ctx, hipStreamSynchronize(s));
            }
            if (rb.counters[0] > ctx->cfg.pool_pages) {
                // page allocations that failed still advanced the page counter: bring it back to
                // the budget, which every page below it is part of
                const uint32_t top = (uint32_t)ctx->cfg.pool_pages;
                HD_HIP(ctx, hipMemcpyAsync(ctx->counters, &top, 4, hipMemcpyHostToDevice, s));
                HD_HIP(ctx, hipStreamSynchronize(s));
            }
        }
        if (fixed) {
            if (err & (ERR_SCRATCH | ERR_COLOR_SCRATCH | ERR_UNDO))
                return fail(ctx, HD_ERR_OOM, "edit batch too large for the static scratch (previous state kept)");
            if (err & (ERR_COLOR_ARENA | ERR_COLOR_TREE))
                return fail(ctx, HD_ERR_COLOR_FULL, "colour arena or tables full; a pass without a fixed scratch grows them (previous state kept)");
        }
        if (err & (ERR_SCRATCH | ERR_COLOR_SCRATCH)) {
            // The device counters hold what the pass asked for. Levels below the first one that
            // overflowed only saw that level's capped items: scale their need by its shortfall.
            uint32_t cnt[64], ecnt[64];
            HD_HIP(ctx, hipMemcpy(cnt, S.dcount, sizeof(cnt), hipMemcpyDeviceToHost));
            HD_HIP(ctx, hipMemcpy(ecnt, S.decount, sizeof(ecnt), hipMemcpyDeviceToHost));
            const uint32_t L = ctx->g.leaf_level();
            uint32_t items[kMaxEditLevels] = {}, elist[kMaxEditLevels] = {};
            double scale = 1.0;
            for (uint32_t l = 0; l <= L; ++l) {
                const double ci = (double)cnt[l] * scale, ce = (double)ecnt[l] * scale;
                items[l] = S.item_cap[l];
                elist[l] = S.elist_cap[l];
                if (ci > S.item_cap[l]) items[l] = (uint32_t)std::min<double>(ci * 1.25 + 64, 1u << 30);
                if (ce > S.elist_cap[l]) elist[l] = (uint32_t)std::min<double>(ce * 1.25 + 64, 1u << 30);
                const double over = std::max((double)cnt[l] / std::max<uint32_t>(S.item_cap[l], 1),
                                             (double)ecnt[l] / std::max<uint32_t>(S.elist_cap[l], 1));
                if (over > 1.0) scale = std::max(scale, over);
            }
            const uint64_t cc = (err & ERR_COLOR_SCRATCH) ? std::max<uint64_t>(S.color_cap * 2, rb.stats[2] + rb.stats[2] / 4)
                                                          : S.color_cap;
            bool grows = cc > S.color_cap;
            for (uint32_t l = 0; l <= L; ++l) grows |= items[l] > S.item_cap[l] || elist[l] > S.elist_cap[l];
            if (!grows) return fail(ctx, HD_ERR_OOM, "edit scratch overflow that growth cannot fix");
            st = edit_reserve(ctx, items, elist, cc);
            if (st != HD_OK) return st;
            continue;
        }
        if (err & ERR_COLOR_ARENA) {
            st = grow_arena(ctx, std::max<uint64_t>(rb.stats[3] - ctx->arena_top, 1u << 22) * 2);
            if (st != HD_OK) return st;
            continue;
        }
        if (err & ERR_BUCKET) return fail(ctx, HD_ERR_BUCKET_FULL, "hash bucket full during edit (old root kept)");
        if (err & ERR_POOL) return fail(ctx, HD_ERR_POOL_FULL, "page pool exhausted during edit (old root kept)");
        if (err & ERR_HEADER) return fail(ctx, HD_ERR_HEADER_OVERFLOW, "voxel count >= 2^24 in a node header");
        if (err & ERR_LEAF_SIZE) return fail(ctx, HD_ERR_COLOR_FULL, "a colour subtree holds more than 160 macro blocks");
        if (err & ERR_COLOR_TREE) {
            // the allocation counters ran past the capacities by what the pass asked for
            st = grow_color_tables(ctx, (uint64_t)rb.counters[2] + 64, (uint64_t)rb.counters[3] + 64);
            if (st != HD_OK) return st;
            continue;
        }
        if (err & ERR_UNDO) {
            st = edit_reserve_undo(ctx, (uint64_t)rb.counters[4] + rb.counters[4] / 2);
            if (st != HD_OK) return st;
            continue;
        }
        if (err & ERR_EMPTY) return fail(ctx, HD_ERR_INVALID, "edit would leave an empty scene (old root kept)");
        if (err & ERR_KIND) return fail(ctx, HD_ERR_INVALID, "copy or fill editor in a pass built without them");
        if (ctx->count_edit_bytes) {
            ctx->edit_bytes_host[0] += rb.stats[1] * 4;                  // node words written
            ctx->edit_bytes_host[2] += (rb.stats[3] - ctx->arena_top) * 8;  // colour leaf words written
        }
        ctx->arena_top = rb.stats[3];
        ctx->n_cnodes = rb.counters[2];
        ctx->n_cleaves = rb.counters[3];
        if (frame) return capture_frame(ctx, rb.counters[4], *frame, s);
        return HD_OK;
    }
    return fail(ctx, HD_ERR_OOM, "edit scratch could not be grown enough");
}

}  // namespace hd

using namespace hd;

extern "C" {

hd_status hd_edit_batch(hd_ctx* ctx, const hd_edit* edits, uint32_t n, uint32_t fused, hd_edit_stats* stats,
                        void* stream) {
    if (!ctx || (!edits && n) ) return fail(ctx, HD_ERR_INVALID, "bad arguments");
    if (!ctx->loaded) return fail(ctx, HD_ERR_INVALID, "no DAG loaded");
    if (n == 0) {
        if (stats) *stats = hd_edit_stats{0, 0, 0, ctx->root, 0, 0};
        return HD_OK;
    }
    for (uint32_t i = 0; i < n; ++i)
        if (edits[i].kind > EDIT_NOISE_CARVE || !(edits[i].radius > 0.f) || !std::isfinite(edits[i].radius))
            return fail(ctx, HD_ERR_INVALID, "unsupported editor kind or radius");
    for (uint32_t i = 0; i < n; ++i) {
        if (edits[i].kind == EDIT_COPY_DEC && !(edits[i].radius < 65536.f))
            return fail(ctx, HD_ERR_INVALID, "copy radius too large for its voxel buffer");
        if (edits[i].kind != EDIT_COPY) continue;
        // Matrix3x3::IsInvertible (Matrix3x3.h:245-248)
        double T[9];
        copy_transform(edits[i], T);
        const double det = T[0] * (T[8] * T[4] - T[7] * T[5]) - T[3] * (T[8] * T[1] - T[7] * T[2]) +
                           T[6] * (T[5] * T[1] - T[4] * T[2]);
        if (!(std::fabs(det) > 0.00001)) return fail(ctx, HD_ERR_INVALID, "copy transform is not invertible");
        if (edits[i].swirl && !(std::isfinite(edits[i].swirl_period) && edits[i].swirl_period != 0.f))
            return fail(ctx, HD_ERR_INVALID, "copy swirl period must be finite and non-zero");
    }
    hd_status st = edit_prepare(ctx);
    if (st != HD_OK) return st;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // editors to the device
    std::vector<DevEdit> dev(n);
    for (uint32_t i = 0; i < n; ++i) dev[i] = make_edit(edits[i]);
    st = edit_ensure_edits(ctx, n);
    if (st != HD_OK) return st;
    if (n) HD_HIP(ctx, hipMemcpyAsync(ctx->edit.edits_dev, dev.data(), (uint64_t)n * sizeof(DevEdit), hipMemcpyHostToDevice, s));
    HD_HIP(ctx, hipMemcpyAsync(ctx->d_root, &ctx->root, 4, hipMemcpyHostToDevice, s));
    hd_edit_stats out{};
    // a fill's or a noise editor's buffers are built from the DAG as it is when its pass starts
    FillBuffers fb;
    auto prepare_fill = [&](uint32_t i) -> hd_status {
        if (!is_buffer_kind(edits[i].kind)) return HD_OK;
        hd_status r = edits[i].kind == EDIT_FILL ? fill_prepare(ctx, edits[i], dev[i], fb, s)
                                                 : noise_prepare(ctx, edits[i], dev[i], fb, s);
        if (r != HD_OK) return r;
        HD_HIP(ctx, hipMemcpyAsync(ctx->edit.edits_dev + i, &dev[i], sizeof(DevEdit), hipMemcpyHostToDevice, s));
        return HD_OK;
    };
    if (fused) {
        if (n > ctx->edit.elist_cap[0]) {
            uint32_t items[kMaxEditLevels] = {}, elist[kMaxEditLevels] = {};
            elist[0] = n;
            st = edit_reserve(ctx, items, elist, 0);
            if (st != HD_OK) return st;
        }
        // A copy or a fill reads the DAG as the edits before it left it, so each starts a new
        // pass; every other run of editors fuses into one pass.
        for (uint32_t first = 0; first < n;) {
            uint32_t last = first + 1;
            while (last < n && !is_copy_kind(edits[last].kind) && !is_buffer_kind(edits[last].kind)) ++last;
            UndoFrame frame;
            frame.prev_root = ctx->root;
            Readback rb{};
            st = prepare_fill(first);
            if (st != HD_OK) return st;
            st = run_pass(ctx, first, last - first, s, rb, &frame,
                          is_copy_kind(edits[first].kind) || is_buffer_kind(edits[first].kind), nullptr,
                          dev.data() + first);
            if (st != HD_OK) return st;
            ctx->root = rb.root;
            frame.new_root = rb.root;
            ctx->undo.push_back(std::move(frame));
            ctx->redo.clear();
            out.num_voxels += rb.stats[0] + copy_voxels(edits + first, last - first);
            if (is_buffer_kind(edits[first].kind) && last - first > 1)
                out.num_voxels += fill_voxels(dev[first], 1u << ctx->g.levels);
            out.new_words += rb.stats[1];
            out.leaves_rebuilt += rb.counters[4];
            out.colors_encoded += rb.stats[4];
            out.passes++;
            first = last;
        }
    } else {
        for (uint32_t i = 0; i < n; ++i) {
            UndoFrame frame;
            frame.prev_root = ctx->root;
            Readback rb{};
            st = prepare_fill(i);
            if (st != HD_OK) return st;
            st = run_pass(ctx, i, 1, s, rb, &frame, is_copy_kind(edits[i].kind) || is_buffer_kind(edits[i].kind),
                          nullptr, dev.data() + i);
            if (st != HD_OK) return st;
            ctx->root = rb.root;
            frame.new_root = rb.root;
            ctx->undo.push_back(std::move(frame));
            ctx->redo.clear();
            out.num_voxels += rb.stats[0] + copy_voxels(edits + i, 1);
            out.new_words += rb.stats[1];
            out.leaves_rebuilt += rb.counters[4];
            out.colors_encoded += rb.stats[4];
            out.passes++;
        }
    }
    out.root = ctx->root;
    if (stats) *stats = out;
    return HD_OK;
}

static hd_status apply_frame(hd_ctx* ctx, const UndoFrame& f, bool undo) {
    const hd_status st = edit_scatter_leaves(ctx, f.handles.data(), undo ? f.prev.data() : f.next.data(),
                                             (uint32_t)f.handles.size());
    if (st != HD_OK) return st;
    ctx->root = undo ? f.prev_root : f.new_root;
    HD_HIP(ctx, hipMemcpy(ctx->d_root, &ctx->root, 4, hipMemcpyHostToDevice));
    return HD_OK;
}

hd_status hd_undo(hd_ctx* ctx) {
    if (!ctx) return fail(ctx, HD_ERR_INVALID, "null context");
    if (ctx->undo.empty()) return HD_OK;
    UndoFrame f = std::move(ctx->undo.back());
    ctx->undo.pop_back();
    hd_status st = apply_frame(ctx, f, true);
    if (st != HD_OK) return st;
    ctx->redo.push_back(std::move(f));
    return HD_OK;
}

hd_status hd_redo(hd_ctx* ctx) {
    if (!ctx) return fail(ctx, HD_ERR_INVALID, "null context");
    if (ctx->redo.empty()) return HD_OK;
    UndoFrame f = std::move(ctx->redo.back());
    ctx->redo.pop_back();
    hd_status st = apply_frame(ctx, f, false);
    if (st != HD_OK) return st;
    ctx->undo.push_back(std::move(f));
    return HD_OK;
}

hd_status hd_undo_clear(hd_ctx* ctx) {
    if (!ctx) return fail(ctx, HD_ERR_INVALID, "null context");
    ctx->undo.clear();
    ctx->redo.clear();
    return HD_OK;
}

hd_status hd_undo_depth(hd_ctx* ctx, uint32_t* undo, uint32_t* redo) {
    if (!ctx) return fail(ctx, HD_ERR_INVALID, "null context");
    if (undo) *undo = (uint32_t)ctx->undo.size();
    if (redo) *redo = (uint32_t)ctx->redo.size();
    return HD_OK;
}

hd_status hd_set_edit_launch(hd_ctx* ctx, uint32_t workgroup_max_items, uint32_t persist_max_edits,
                             uint32_t blocks_per_cu) {
    if (!ctx) return fail(ctx, HD_ERR_INVALID, "null context");
    if (blocks_per_cu < 1 || blocks_per_cu > 8) return fail(ctx, HD_ERR_INVALID, "blocks_per_cu must be 1..8");
    ctx->edit_block_max = workgroup_max_items;
    ctx->edit_persist_max = persist_max_edits;
    ctx->edit_persist_per_cu = blocks_per_cu;
    ctx->ptr_gen++;  // a captured pass bakes its launch form in
    return HD_OK;
}

hd_status hd_autotune_edit_launch(hd_ctx* ctx, const hd_edit* sample, uint32_t n, uint32_t reps,
                                  uint32_t* workgroup_max_items, float ms_per_edit[5]) {
    static const uint32_t kCandidates[5] = {0u, 32u, 128u, 512u, 2048u};
    if (!ctx || !sample || !n) return fail(ctx, HD_ERR_INVALID, "bad arguments");
    for (uint32_t i = 0; i < n; ++i)
        if (sample[i].kind > EDIT_SPHERE_PAINT) return fail(ctx, HD_ERR_INVALID, "sample editors must be sphere / box / paint");
    const uint32_t keepBlock = ctx->edit_block_max, keepPersist = ctx->edit_persist_max;
    ctx->edit_persist_max = 0;
    double best = 1e300;
    uint32_t pick = keepBlock;
    for (uint32_t c = 0; c < 5; ++c) {
        ctx->edit_block_max = kCandidates[c];
        double total = 0;
        for (uint32_t r = 0; r <= std::max(reps, 1u); ++r) {  // round 0 warms up (scratch sizes)
            for (uint32_t i = 0; i < n; ++i) {
                const auto t0 = std::chrono::steady_clock::now();
                hd_status st = hd_edit_batch(ctx, sample + i, 1, 0, nullptr, nullptr);
                const auto t1 = std::chrono::steady_clock::now();
                if (st == HD_OK) st = hd_undo(ctx);
                if (st != HD_OK) {
                    ctx->edit_block_max = keepBlock;
                    ctx->edit_persist_max = keepPersist;
                    return st;
                }
                if (r > 0) total += std::chrono::duration<double, std::milli>(t1 - t0).count();
            }
        }
        const double per = total / ((double)n * std::max(reps, 1u));
        if (ms_per_edit) ms_per_edit[c] = (float)per;
        if (per < best) {
            best = per;
            pick = kCandidates[c];
        }
    }
    ctx->edit_block_max = pick;
    ctx->edit_persist_max = keepPersist;
    ctx->ptr_gen++;
    if (workgroup_max_items) *workgroup_max_items = pick;
    return HD_OK;
}

hd_status hd_get_edit_launch(hd_ctx* ctx, uint32_t* workgroup_max_items, uint32_t* persist_max_edits,
                             uint32_t* blocks_per_cu) {
    if (!ctx) return fail(ctx, HD_ERR_INVALID, "null context");
    if (workgroup_max_items) *workgroup_max_items = ctx->edit_block_max;
    if (persist_max_edits) *persist_max_edits = ctx->edit_persist_max;
    if (blocks_per_cu) *blocks_per_cu = ctx->edit_persist_per_cu;
    return HD_OK;
}

hd_status hd_edit_count_bytes(hd_ctx* ctx, uint32_t on, uint64_t out[5]) {
    if (!ctx) return fail(ctx, HD_ERR_INVALID, "null context");
    HD_HIP(ctx, hipDeviceSynchronize());
    if (on) {
        HD_HIP(ctx, hipMemset(ctx->edit_bytes, 0, 8 * sizeof(unsigned long long)));
        for (uint64_t& v : ctx->edit_bytes_host) v = 0;
        ctx->count_edit_bytes = true;
        ctx->ptr_gen++;  // captured passes bake the counter pointer in
        return HD_OK;
    }
    unsigned long long dev[3] = {0, 0, 0};
    HD_HIP(ctx, hipMemcpy(dev, ctx->edit_bytes, sizeof(dev), hipMemcpyDeviceToHost));
    if (ctx->count_edit_bytes) ctx->ptr_gen++;
    ctx->count_edit_bytes = false;
    if (out) {
        out[0] = dev[0];
        out[1] = dev[1];
        out[2] = ctx->edit_bytes_host[0];
        out[3] = dev[2];
        out[4] = ctx->edit_bytes_host[2];
    }
    return HD_OK;
}

hd_status hd_reserve_edit_scratch(hd_ctx* ctx, uint32_t items_per_level, uint32_t edit_list_words, uint64_t colors) {
    if (!ctx || !items_per_level) return fail(ctx, HD_ERR_INVALID, "bad arguments");
    uint32_t items[kMaxEditLevels], elist[kMaxEditLevels];
    for (uint32_t l = 0; l < kMaxEditLevels; ++l) {
        items[l] = items_per_level;
        elist[l] = edit_list_words;
    }
    return edit_reserve(ctx, items, elist, colors);
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// Edit passes that allocate nothing: a fused pass run in a file-scope device array instead of
// the context's growable scratch, for callers that cannot call hipMalloc (code under graph
// capture, serving loops, harnesses that forbid it). Each GPU holds its own copy of the array
// until the process ends; every call on a GPU uses that GPU's copy, so calls that overlap on
// one GPU race.
// ---------------------------------------------------------------------------------------------
namespace hd {
namespace {

constexpr uint64_t kStaticScratchBytes = 32ull << 20;
constexpr uint32_t kStaticItems = 4096;          // work items per level (level l: at most 8^l)
constexpr uint32_t kStaticElist = 32768;         // edit-list words per level
constexpr uint64_t kStaticColors = 1ull << 20;   // staged colours
constexpr uint32_t kStaticUndo = 1u << 14;       // leaf slots one pass may re-point
constexpr uint32_t kStaticEdits = HD_STATIC_MAX_EDITS;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

std::mutex g_scratch_mu;
// device address of `scratch` per device ordinal, from hipGetSymbolAddress (the array's name
// is a host-side stand-in in host code, never a device address)
char* g_scratch_ptr[kMaxDevices] = {};

// The scratch view of a pass at depth `levels` over the array at `base`; the bytes it takes.
uint64_t static_layout(EditScratch& S, uint32_t levels, char* base) {
    S = EditScratch{};
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) -> char* {
        char* p = base ? base + off : nullptr;
        off += (bytes + 255) & ~255ull;
        return p;
    };
    S.levels = levels;
    S.dcount = reinterpret_cast<uint32_t*>(take(64 * 4));
    S.decount = reinterpret_cast<uint32_t*>(take(64 * 4));
    S.dstats = reinterpret_cast<uint64_t*>(take(64 * 8));
    S.edits_dev = reinterpret_cast<DevEdit*>(take((uint64_t)kStaticEdits * sizeof(DevEdit)));
    S.edits_cap = kStaticEdits;
    S.undo_handles = reinterpret_cast<uint32_t*>(take((uint64_t)kStaticUndo * 4));
    S.undo_descs = reinterpret_cast<LeafDesc*>(take((uint64_t)kStaticUndo * sizeof(LeafDesc)));
    S.undo_next = reinterpret_cast<LeafDesc*>(take((uint64_t)kStaticUndo * sizeof(LeafDesc)));
    S.undo_cap = kStaticUndo;
    S.hcap = 2 * kStaticItems;  // a power of two at least twice any level's items
    S.hset = reinterpret_cast<uint32_t*>(take((uint64_t)S.hcap * 4));
    S.colors = reinterpret_cast<uint64_t*>(take(kStaticColors * 8));
    S.cmarks = reinterpret_cast<uint64_t*>(take(kStaticColors / 64 * 8));
    S.color_cap = kStaticColors;
    const uint32_t L = levels - 2;
    for (uint32_t l = 0; l <= L; ++l) {
        const uint64_t c = l < 4 ? std::min<uint64_t>(1ull << (3 * l), kStaticItems) : kStaticItems;
        const uint64_t bytes = carve_level(S.items[l], S.elist[l], l == L, c, kStaticElist, nullptr);
        char* mem = take(bytes);
        if (base) carve_level(S.items[l], S.elist[l], l == L, c, kStaticElist, mem);
        S.item_cap[l] = (uint32_t)c;
        S.elist_cap[l] = kStaticElist;
    }
    return off;
}

}  // namespace

// Looks up (once per device, with that device current) the device address of this library's
// static edit scratch on ctx's device.
hd_status static_scratch_setup(hd_ctx* ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return fail(ctx, HD_ERR_INVALID, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[ctx->device]) return HD_OK;
    int cur = -1;
    HD_HIP(ctx, hipGetDevice(&cur));
    if (cur != ctx->device) return fail(ctx, HD_ERR_INVALID, "the context's device is not the current device");
    void* ptr = nullptr;
    HD_HIP(ctx, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(ctx, HD_ERR_HIP, "static edit scratch has no device address");
    g_scratch_ptr[ctx->device] = static_cast<char*>(ptr);
    return HD_OK;
}

}  // namespace hd

extern "C" {

hd_status hd_edit_batch_static(hd_ctx* ctx, const hd_edit* edits, uint32_t n, hd_edit_stats* stats, void* stream) {
    if (!ctx || (!edits && n)) return fail(ctx, HD_ERR_INVALID, "bad arguments");
    if (!ctx->loaded) return fail(ctx, HD_ERR_INVALID, "no DAG loaded");
    if (n > kStaticEdits) return fail(ctx, HD_ERR_INVALID, "more edits than HD_STATIC_MAX_EDITS");
    if (ctx->g.levels > kMaxEditLevels) return fail(ctx, HD_ERR_INVALID, "levels above the static scratch's 17");
    for (uint32_t i = 0; i < n; ++i)
        if (edits[i].kind > EDIT_SPHERE_PAINT || !(edits[i].radius > 0.f) || !std::isfinite(edits[i].radius))
            return fail(ctx, HD_ERR_INVALID, "static passes take sphere / box / paint editors with a finite radius");
    if (!ctx->colors_loaded) return fail(ctx, HD_ERR_INVALID, "no colour tree (hd_load_colors, or one hd_edit_batch call, sets it up)");
    if (n == 0) {
        if (stats) *stats = hd_edit_stats{0, 0, 0, ctx->root, 0, 0};
        return HD_OK;
    }
    hd_status st = static_scratch_setup(ctx);
    if (st != HD_OK) return st;
    char* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        base = g_scratch_ptr[ctx->device];
    }
    EditScratch S;
    if (static_layout(S, ctx->g.levels, nullptr) > kStaticScratchBytes)
        return fail(ctx, HD_ERR_INVALID, "static edit scratch layout exceeds the array");
    static_layout(S, ctx->g.levels, base);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // editors to the device in chunks staged on the stack
    constexpr uint32_t kChunk = 32;
    DevEdit dev[kChunk];
    for (uint32_t i = 0; i < n; i += kChunk) {
        const uint32_t m = std::min(kChunk, n - i);
        for (uint32_t k = 0; k < m; ++k) dev[k] = make_edit(edits[i + k]);
        HD_HIP(ctx, hipMemcpyAsync(S.edits_dev + i, dev, (uint64_t)m * sizeof(DevEdit), hipMemcpyHostToDevice, s));
        HD_HIP(ctx, hipStreamSynchronize(s));
    }
    HD_HIP(ctx, hipMemcpyAsync(ctx->d_root, &ctx->root, 4, hipMemcpyHostToDevice, s));
    Readback rb{};
    st = run_pass(ctx, 0, n, s, rb, nullptr, false, &S);
    if (st != HD_OK) return st;
    ctx->root = rb.root;
    // the pass re-pointed leaf slots without an undo frame: earlier frames no longer apply
    ctx->undo.clear();
    ctx->redo.clear();
    hd_edit_stats out{};
    out.num_voxels = rb.stats[0];
    out.new_words = rb.stats[1];
    out.leaves_rebuilt = rb.counters[4];
    out.colors_encoded = rb.stats[4];
    out.passes = 1;
    out.root = ctx->root;
    if (stats) *stats = out;
    return HD_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// Frame graph: one fused edit pass and the trace of the resulting DAG, captured once into a HIP
// graph and replayed with new edits and a new camera (both staged in pinned memory and copied
// by the graph's own memcpy nodes). A replay whose pass overflows a capacity is rolled back
// and re-run eagerly (which grows the capacity); the graph is re-captured when any buffer it
// bakes in has been reallocated.
// ---------------------------------------------------------------------------------------------
struct hd_frame_graph {
    hd_ctx* ctx = nullptr;
    uint32_t max_edits = 0, width = 0, height = 0;
    uint32_t* paths = nullptr;
    uint32_t* colors = nullptr;
    float bias = 1.f, fog = 0.f;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t gen = ~0ull;
    DevEdit* h_edits = nullptr;      // pinned staging
    uint32_t* h_n = nullptr;
    hd_ray_params* h_params = nullptr;
    Readback* h_rb = nullptr;
    uint32_t* d_n = nullptr;         // device copies read by the replay
    hd_ray_params* d_params = nullptr;
};

static void graph_drop(hd_frame_graph* g) {
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    g->exec = nullptr;
    g->graph = nullptr;
}

static hd_status graph_capture(hd_frame_graph* g) {
    hd_ctx* ctx = g->ctx;
    graph_drop(g);
    EditScratch& S = ctx->edit;
    const bool colors = ctx->colors_loaded && ctx->g.leaf_level() > ctx->g.color_tree_levels;
    hipStream_t s = g->stream;
    const bool trace = g->width != 0;  // edit-only graphs hold no trace
    if (trace && (ctx->trace_persist || ctx->long_steps)) {
        const hd_status rs = trace_reserve(ctx, s, g->width, g->height);
        if (rs != HD_OK) return rs;
    }
    HD_HIP(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipMemcpyAsync(g->d_n, g->h_n, 4, hipMemcpyHostToDevice, s);
    hipMemcpyAsync(S.edits_dev, g->h_edits, (uint64_t)g->max_edits * sizeof(DevEdit), hipMemcpyHostToDevice, s);
    if (trace) hipMemcpyAsync(g->d_params, g->h_params, sizeof(hd_ray_params), hipMemcpyHostToDevice, s);
    hipMemcpyAsync(ctx->bucket_snapshot, ctx->bucket_sizes, (uint64_t)ctx->g.total_buckets() * 4, hipMemcpyDeviceToDevice, s);
    k_reset<<<1, 1, 0, s>>>(ctx->counters, S.dstats);
    hd_status st = enqueue_pass(ctx, 0, g->max_edits, colors, s, g->d_n, false);  // kinds checked at capture
    k_count_voxels<<<std::min<uint32_t>(g->max_edits, 4096), 256, 0, s>>>(
        S.edits_dev, g->max_edits, g->d_n, 1u << ctx->g.levels, reinterpret_cast<unsigned long long*>(S.dstats + 0));
    if (st == HD_OK && trace)
        st = enqueue_frame_from_device(ctx, g->d_params, g->width, g->height, g->paths, g->colors, g->bias, g->fog, s);
    hipMemcpyAsync(g->h_rb->counters, ctx->counters, sizeof(g->h_rb->counters), hipMemcpyDeviceToHost, s);
    hipMemcpyAsync(g->h_rb->stats, S.dstats, sizeof(g->h_rb->stats), hipMemcpyDeviceToHost, s);
    hipMemcpyAsync(&g->h_rb->root, ctx->d_root, 4, hipMemcpyDeviceToHost, s);
    const hipError_t e = hipStreamEndCapture(s, &g->graph);
    if (st != HD_OK) return st;
    if (e != hipSuccess) return hip_fail(ctx, e, "frame graph capture");
    HD_HIP(ctx, hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    g->gen = ctx->ptr_gen;
    return HD_OK;
}

extern "C" {

hd_status hd_frame_graph_create(hd_ctx* ctx, uint32_t max_edits, uint32_t width, uint32_t height, uint32_t* d_paths,
                                uint32_t* d_colors, float shadow_bias, float fog_density, hd_frame_graph** out) {
    const bool editOnly = !width && !height && !d_paths && !d_colors;
    if (!ctx || !out || !max_edits || (!editOnly && (!d_paths || !d_colors || !width || !height)))
        return fail(ctx, HD_ERR_INVALID, "bad arguments");
    if (!ctx->loaded || !ctx->colors_loaded) return fail(ctx, HD_ERR_INVALID, "no DAG or colours loaded");
    auto* g = new hd_frame_graph();
    g->ctx = ctx;
    g->max_edits = max_edits;
    g->width = width;
    g->height = height;
    g->paths = d_paths;
    g->colors = d_colors;
    g->bias = shadow_bias;
    g->fog = fog_density;
    hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc(&g->h_edits, (uint64_t)max_edits * sizeof(DevEdit));
    if (e == hipSuccess) e = hipHostMalloc(&g->h_n, 4);
    if (e == hipSuccess) e = hipHostMalloc(&g->h_params, sizeof(hd_ray_params));
    if (e == hipSuccess) e = hipHostMalloc(&g->h_rb, sizeof(Readback));
    if (e == hipSuccess) e = hipMalloc(&g->d_n, 4);
    if (e == hipSuccess) e = hipMalloc(&g->d_params, sizeof(hd_ray_params));
    if (e != hipSuccess) {
        hd_frame_graph_destroy(g);
        return hip_fail(ctx, e, "frame graph buffers");
    }
    std::memset(g->h_edits, 0, (uint64_t)max_edits * sizeof(DevEdit));
    *g->h_n = 0;
    hd_status st = edit_prepare(ctx);
    if (st == HD_OK) st = edit_ensure_edits(ctx, max_edits);
    if (st != HD_OK) {
        hd_frame_graph_destroy(g);
        return st;
    }
    *out = g;
    return HD_OK;
}

hd_status hd_frame_graph_destroy(hd_frame_graph* g) {
    if (!g) return HD_OK;
    if (g->stream) hipStreamSynchronize(g->stream);
    graph_drop(g);
    hipHostFree(g->h_edits);
    hipHostFree(g->h_n);
    hipHostFree(g->h_params);
    hipHostFree(g->h_rb);
    hipFree(g->d_n);
    hipFree(g->d_params);
    if (g->stream) hipStreamDestroy(g->stream);
    delete g;
    return HD_OK;
}

hd_status hd_frame_graph_launch(hd_frame_graph* g, const hd_edit* edits, uint32_t n, const hd_ray_params* params,
                                hd_edit_stats* stats, uint32_t* replayed) {
    if (!g || (!params && g->width) || (n && !edits)) return fail(g ? g->ctx : nullptr, HD_ERR_INVALID, "bad arguments");
    hd_ctx* ctx = g->ctx;
    if (!ctx->loaded) return fail(ctx, HD_ERR_INVALID, "no DAG loaded");
    if (n > g->max_edits) return fail(ctx, HD_ERR_INVALID, "more edits than the graph was built for");
    for (uint32_t i = 0; i < n; ++i)
        if (edits[i].kind > EDIT_SPHERE_PAINT || !(edits[i].radius > 0.f) || !std::isfinite(edits[i].radius))
            return fail(ctx, HD_ERR_INVALID, "unsupported editor kind or radius (copies and fills run through hd_edit_batch)");
    // the same free-space guarantees as an eager pass, then (re)capture if buffers moved
    hd_status st = grow_arena(ctx, 1u << 22);
    if (st == HD_OK) st = edit_ensure_edits(ctx, g->max_edits);
    if (st != HD_OK) return st;
    if (!g->exec || g->gen != ctx->ptr_gen) {
        st = graph_capture(g);
        if (st != HD_OK) return st;
    }
    for (uint32_t i = 0; i < n; ++i) g->h_edits[i] = make_edit(edits[i]);
    *g->h_n = n;
    if (params) *g->h_params = *params;
    HD_HIP(ctx, hipGraphLaunch(g->exec, g->stream));
    HD_HIP(ctx, hipStreamSynchronize(g->stream));
    const Readback& rb = *g->h_rb;
    if (rb.counters[1]) {
        // a capacity ran out: drop what the pass appended, roll the leaf slots back and run the
        // frame eagerly (which grows it)
        HD_HIP(ctx, hipMemcpyAsync(ctx->bucket_sizes, ctx->bucket_snapshot, (uint64_t)ctx->g.total_buckets() * 4,
                                   hipMemcpyDeviceToDevice, g->stream));
        HD_HIP(ctx, hipStreamSynchronize(g->stream));
        if (rb.counters[0] > ctx->cfg.pool_pages) {
            const uint32_t top = (uint32_t)ctx->cfg.pool_pages;
            HD_HIP(ctx, hipMemcpyAsync(ctx->counters, &top, 4, hipMemcpyHostToDevice, g->stream));
            HD_HIP(ctx, hipStreamSynchronize(g->stream));
        }
        if (rb.counters[4]) {
            edit_restore_leaves(ctx, g->stream);
            HD_HIP(ctx, hipStreamSynchronize(g->stream));
        }
        if (replayed) *replayed = 0;
        st = hd_edit_batch(ctx, edits, n, 1, stats, g->stream);
        if (!g->width) {
            if (st == HD_OK) HD_HIP(ctx, hipStreamSynchronize(g->stream));
            return st;
        }
        if (st == HD_OK) st = hd_trace_paths(ctx, params, g->width, g->height, g->paths, g->stream);
        if (st == HD_OK) st = hd_trace_colors(ctx, g->width, g->height, g->paths, g->colors, g->stream);
        if (st == HD_OK)
            st = hd_trace_shadows(ctx, params, g->width, g->height, g->paths, g->colors, g->bias, g->fog, g->stream);
        if (st == HD_OK) HD_HIP(ctx, hipStreamSynchronize(g->stream));
        return st;
    }
    // commit, as a successful eager pass does
    UndoFrame frame;
    frame.prev_root = ctx->root;
    ctx->root = rb.root;
    frame.new_root = rb.root;
    ctx->arena_top = rb.stats[3];
    ctx->n_cnodes = rb.counters[2];
    ctx->n_cleaves = rb.counters[3];
    st = capture_frame(ctx, rb.counters[4], frame, g->stream);
    if (st != HD_OK) return st;
    ctx->undo.push_back(std::move(frame));
    ctx->redo.clear();
    if (stats) *stats = hd_edit_stats{rb.stats[0], rb.stats[1], rb.counters[4], rb.root, 1u, rb.stats[4]};
    if (replayed) *replayed = 1;
    return HD_OK;
}

}  // extern "C"
