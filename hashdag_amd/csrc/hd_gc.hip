// Device garbage collection: HashDAG::remove_stale_nodes(maxLevel) (hash_dag.h:407-547).
//
//   mark     level by level from the root: every child pointer of a live node sets its bit in
//            a bitmap over the virtual address space; the first setter appends the child to
//            the next level's frontier (the reference's per-level unordered_sets)
//   rebuild  levels maxLevel .. 0: gather every live node with its children remapped through
//            the old -> new map of the level below, reset the level's bucket sizes, append the
//            full node first (ADD_FULL_NODES_FIRST), place every live node in its bucket with
//            the append rule (in the order of the ranks the nodes draw in their bucket), record
//            old -> new
//
// Pages stay mapped (poolTop never shrinks), as in the reference. The result is canonically
// identical to the DAG before GC; node order within a bucket is the rank order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "hd_ctx.h"
#include "hd_table_dev.h"

namespace hd {
namespace {

constexpr uint32_t kGcGrid = 2048;
constexpr uint32_t GC_ERR_CAP = 1, GC_ERR_BUCKET = 2, GC_ERR_POOL = 4, GC_ERR_MAP = 8;

__device__ __forceinline__ uint32_t child_off(uint32_t mask, uint32_t c) { return __popc(mask & ((1u << c) - 1u)) + 1u; }

__global__ void k_gc_fill(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

// words used per level (sum of the level's bucket sizes): blockIdx.y is the level
__global__ __launch_bounds__(256) void k_gc_level_used(Geom g, const uint32_t* __restrict__ bucket_sizes,
                                                       unsigned long long* used) {
    const uint32_t l = blockIdx.y;
    const uint32_t nb = g.buckets_per_level(l);
    const uint32_t* __restrict__ bs = bucket_sizes + g.bucket_global(l, 0);
    unsigned long long sum = 0;
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) sum += bs[b];
#pragma unroll
    for (uint32_t off = 32; off; off >>= 1) sum += __shfl_down(sum, off, 64);
    if ((threadIdx.x & 63u) == 0 && sum) atomicAdd(&used[l], sum);
}

// level 0 frontier = the (distinct) roots kept alive
__global__ void k_gc_seed(uint32_t* frontier0, uint32_t* counts, const uint32_t* roots, uint32_t n_roots,
                          uint32_t levels) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (uint32_t l = 0; l <= levels; ++l) counts[l] = 0;
        counts[0] = n_roots;
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_roots; i += gridDim.x * blockDim.x)
        frontier0[i] = roots[i];
}

// mark children of the live nodes of `level` (an interior level): one lane per node, its
// children's pointers, bitmap words and atomics issued together; a wavefront appends the
// children it marked first with one atomic (in node order, child order within a node)
__global__ __launch_bounds__(256) void k_gc_mark(const uint32_t* __restrict__ vpool, uint32_t* __restrict__ bitmap,
                                                 const uint32_t* __restrict__ frontier, uint32_t* __restrict__ next,
                                                 uint32_t* counts, uint32_t level, uint32_t capCur, uint32_t cap,
                                                 uint64_t numVirtual, uint32_t* errors) {
    __shared__ uint32_t sTot[256 / 64 + 1];
    const uint32_t n = min(counts[level], capCur);
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint32_t stride = gridDim.x * blockDim.x;
    // block-uniform loop: the next frontier is appended with one atomic per block
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < n; b0 += stride) {
        const uint32_t i = b0 + threadIdx.x;
        uint32_t kid[8], nk = 0, firstBits = 0;
        if (i < n) {
            const uint32_t va = frontier[i];
            nk = __popc(vpool[va] & 0xFFu);
#pragma unroll
            for (uint32_t k = 0; k < 8; ++k) kid[k] = k < nk ? vpool[va + 1 + k] : 0u;
            uint32_t w[8];
            // shared children are marked by many parents: read first, so only the first few
            // pay for the atomic on a hot bitmap word
#pragma unroll
            for (uint32_t k = 0; k < 8; ++k)
                w[k] = k < nk && kid[k] < numVirtual
                           ? __hip_atomic_load(&bitmap[kid[k] >> 5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                           : ~0u;
#pragma unroll
            for (uint32_t k = 0; k < 8; ++k) {
                if (k >= nk) continue;
                if (kid[k] >= numVirtual) {  // not an address of this table
                    atomicOr(errors, GC_ERR_MAP);
                    continue;
                }
                const uint32_t bit = 1u << (kid[k] & 31u);
                if (!(w[k] & bit) && !(atomicOr(&bitmap[kid[k] >> 5], bit) & bit)) firstBits |= 1u << k;
            }
        }
        // block scan of the marked counts, one atomic for the block
        const uint32_t cnt = __popc(firstBits);
        uint32_t incl = cnt;
#pragma unroll
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t t = __shfl_up(incl, off, 64);
            if (lane >= off) incl += t;
        }
        if (lane == 63) sTot[wv] = incl;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tot = 0;
            for (uint32_t w = 0; w < nw; ++w) {
                const uint32_t x = sTot[w];
                sTot[w] = tot;
                tot += x;
            }
            sTot[nw] = tot ? atomicAdd(&counts[level + 1], tot) : 0u;
        }
        __syncthreads();
        uint32_t base = sTot[nw] + sTot[wv] + incl - cnt;
        __syncthreads();  // sTot is rewritten by the next iteration
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) {
            if (!((firstBits >> k) & 1u)) continue;
            if (base >= cap) atomicOr(errors, GC_ERR_CAP);
            else next[base] = kid[k];
            ++base;
        }
    }
}

// gather live nodes of `level` with their children's new addresses (the forwarding array of
// the level below) into tmp (9 words per node), and each node's bucket
__global__ __launch_bounds__(256) void k_gc_gather(Geom g, const uint32_t* __restrict__ vpool,
                                                   const uint32_t* __restrict__ frontier, const uint32_t* counts,
                                                   uint32_t level, uint32_t maxLevel, uint32_t cap,
                                                   const uint32_t* __restrict__ fwd, uint32_t* __restrict__ tmp,
                                                   uint32_t* __restrict__ tlen, uint32_t* __restrict__ bkt) {
    const uint32_t n = min(counts[level], cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t va = frontier[i];
        uint32_t* out = tmp + (uint64_t)i * 9;
        if (level == g.leaf_level()) {
            const uint32_t w0 = vpool[va], w1 = vpool[va + 1];
            out[0] = w0;
            out[1] = w1;
            tlen[i] = 2;
            bkt[i] = g.bucket_from_hash(level, hash_leaf((uint64_t)w0 | ((uint64_t)w1 << 32)));
            continue;
        }
        const uint32_t hdr = vpool[va];
        const uint32_t nk = __popc(hdr & 0xFFu);
        uint32_t c[8];
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) c[k] = k < nk ? vpool[va + 1 + k] : 0u;
        if (level != maxLevel) {
#pragma unroll
            for (uint32_t k = 0; k < 8; ++k)
                if (k < nk) c[k] = fwd[c[k]];
        }
        uint32_t node[9];
        node[0] = hdr;
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) node[1 + k] = c[k];
        out[0] = hdr;
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k)
            if (k < nk) out[1 + k] = c[k];
        tlen[i] = nk + 1;
        bkt[i] = g.bucket_from_hash(level, hash_words(node, nk + 1));
    }
}

__global__ void k_gc_full(TableRef T, uint32_t level, uint32_t* full_nodes, uint32_t* errors) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const Geom& g = T.g;
    uint32_t node[9];
    uint32_t n, h;
    if (level == g.leaf_level()) {
        node[0] = node[1] = 0xFFFFFFFFu;
        n = 2;
        h = hash_leaf(~0ull);
    } else {
        const uint32_t count = level < g.color_tree_levels ? 0u : (1u << (3 * (g.levels - level)));
        node[0] = (count << 8) | 0xFFu;
        for (int c = 0; c < 8; ++c) node[1 + c] = full_nodes[level + 1];
        n = 9;
        h = hash_words(node, 9);
    }
    const uint32_t va = table_append(T, level, g.bucket_from_hash(level, h), node, n);
    if (va == kAppendBucketFull) atomicOr(errors, GC_ERR_BUCKET);
    else if (va == kAppendPoolFull) atomicOr(errors, GC_ERR_POOL);
    else full_nodes[level] = va;
}

// The live nodes of a level are placed bucket by bucket: each node takes a rank within its
// bucket (the bucket from k_gc_gather, the rank from k_gc_bucket), the buckets' ranges of a level-wide order follow from a scan of the
// bucket counts (k_gc_tile_sums, k_gc_tile_scan, k_gc_tile_apply, k_gc_order), one lane per bucket lays its nodes out in rank order
// with the append rule of table_append (a node that does not fit in the rest of its page starts
// the next page; k_gc_layout), and every node is then written where it was placed
// (k_gc_write). This is what appending the nodes one at a time in rank order gives, without a
// compare-and-swap loop on each bucket's size.
__global__ __launch_bounds__(256) void k_gc_bucket(uint32_t level, const uint32_t* counts, uint32_t cap,
                                                   const uint32_t* __restrict__ bkt, uint32_t* __restrict__ rank,
                                                   uint32_t* bcount) {
    const uint32_t n = min(counts[level], cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        rank[i] = atomicAdd(&bcount[bkt[i]], 1u);
}

// exclusive scan of the bucket counts of one level, in tiles of 1024 buckets: the tiles' sums
// (k_gc_tile_sums), their exclusive scan (k_gc_tile_scan, one block), then each tile scanned
// from its offset (k_gc_tile_apply)
constexpr uint32_t kScanTile = 1024;
__device__ __forceinline__ uint32_t block_incl_scan1024(uint32_t v, uint32_t* lds, uint32_t& total) {
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    if (lane == 63) lds[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t w = 0; w < kScanTile / 64; ++w) {
            const uint32_t x = lds[w];
            lds[w] = run;
            run += x;
        }
        lds[kScanTile / 64] = run;
    }
    __syncthreads();
    v += lds[wv];
    total = lds[kScanTile / 64];
    __syncthreads();  // lds is reused by the caller's next call
    return v;
}
__global__ __launch_bounds__(kScanTile) void k_gc_tile_sums(const uint32_t* __restrict__ bcount, uint32_t nb,
                                                           uint32_t* __restrict__ tileSum) {
    __shared__ uint32_t lds[kScanTile / 64 + 1];
    const uint32_t i = blockIdx.x * kScanTile + threadIdx.x;
    uint32_t total;
    block_incl_scan1024(i < nb ? bcount[i] : 0u, lds, total);
    if (threadIdx.x == 0) tileSum[blockIdx.x] = total;
}
__global__ __launch_bounds__(kScanTile) void k_gc_tile_scan(uint32_t* __restrict__ tileSum, uint32_t nTiles) {
    __shared__ uint32_t lds[kScanTile / 64 + 1];
    uint32_t carry = 0;
    for (uint32_t t0 = 0; t0 < nTiles; t0 += kScanTile) {  // block-uniform
        const uint32_t t = t0 + threadIdx.x;
        const uint32_t v = t < nTiles ? tileSum[t] : 0u;
        uint32_t total;
        const uint32_t incl = block_incl_scan1024(v, lds, total);
        if (t < nTiles) tileSum[t] = carry + incl - v;
        carry += total;
    }
}
__global__ __launch_bounds__(kScanTile) void k_gc_tile_apply(const uint32_t* __restrict__ bcount, uint32_t nb,
                                                            const uint32_t* __restrict__ tileSum,
                                                            uint32_t* __restrict__ bstart) {
    __shared__ uint32_t lds[kScanTile / 64 + 1];
    const uint32_t i = blockIdx.x * kScanTile + threadIdx.x;
    const uint32_t v = i < nb ? bcount[i] : 0u;
    uint32_t total;
    const uint32_t incl = block_incl_scan1024(v, lds, total);
    if (i < nb) bstart[i] = tileSum[blockIdx.x] + incl - v;
}

// node lengths in level-wide (bucket, rank) order
__global__ __launch_bounds__(256) void k_gc_order(uint32_t level, const uint32_t* counts, uint32_t cap,
                                                  const uint32_t* __restrict__ tlen, const uint32_t* __restrict__ bkt,
                                                  const uint32_t* __restrict__ rank, const uint32_t* __restrict__ bstart,
                                                  uint32_t* __restrict__ olen) {
    const uint32_t n = min(counts[level], cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        olen[bstart[bkt[i]] + rank[i]] = tlen[i];
}

// one lane per bucket: positions of its nodes in rank order, after what the bucket holds
// already (the level's full node); ~0 for nodes that no longer fit
__global__ __launch_bounds__(256) void k_gc_layout(Geom g, uint32_t level, uint32_t* bucket_sizes,
                                                   const uint32_t* __restrict__ bcount,
                                                   const uint32_t* __restrict__ bstart,
                                                   const uint32_t* __restrict__ olen, uint32_t* __restrict__ opos,
                                                   uint32_t* errors) {
    const uint32_t nb = g.buckets_per_level(level), ps = g.page_size, capB = g.bucket_size(level);
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        uint32_t* bsPtr = &bucket_sizes[g.bucket_global(level, b)];
        uint32_t cur = *bsPtr;
        const uint32_t k0 = bstart[b], k1 = k0 + bcount[b];
        bool full = false;
        for (uint32_t k = k0; k < k1; ++k) {
            const uint32_t len = olen[k];
            const uint32_t left = ps - cur % ps;
            const uint32_t start = (left != ps && left < len) ? cur + left : cur;
            if (full || start + len > capB) {
                full = true;
                opos[k] = ~0u;
                continue;
            }
            opos[k] = start;
            cur = start + len;
        }
        if (full) atomicOr(errors, GC_ERR_BUCKET);
        *bsPtr = cur;
    }
}

// every node written where it was placed; its new address into the forwarding array
__global__ __launch_bounds__(256) void k_gc_write(TableRef T, uint32_t level, const uint32_t* counts, uint32_t cap,
                                                  const uint32_t* __restrict__ frontier,
                                                  const uint32_t* __restrict__ tmp, const uint32_t* __restrict__ tlen,
                                                  const uint32_t* __restrict__ bkt, const uint32_t* __restrict__ rank,
                                                  const uint32_t* __restrict__ bstart,
                                                  const uint32_t* __restrict__ opos, uint32_t* __restrict__ fwd,
                                                  uint32_t* errors) {
    const uint32_t n = min(counts[level], cap);
    const uint32_t ps = T.g.page_size;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t b = bkt[i];
        const uint32_t start = opos[bstart[b] + rank[i]];
        if (start == ~0u) continue;  // the bucket overflowed (error set by the layout)
        const uint32_t va = T.g.make_ptr(level, b, start);
        if (start % ps == 0 && !table_map_page(T, va / ps)) {
            atomicOr(errors, GC_ERR_POOL);
            continue;
        }
        const uint32_t* node = tmp + (uint64_t)i * 9;
        const uint32_t len = tlen[i];
        for (uint32_t k = 0; k < len; ++k) T.vpool[va + k] = node[k];
        fwd[frontier[i]] = va;
    }
}

// every kept root -> its address after GC
__global__ void k_gc_roots(const uint32_t* __restrict__ fwd, uint32_t* roots, uint32_t n_roots, uint32_t* errors) {
    if (*errors) return;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_roots; i += gridDim.x * blockDim.x)
        roots[i] = fwd[roots[i]];
}

}  // namespace
}  // namespace hd

using namespace hd;

// Roots kept alive by a GC: the current root, the caller's extra roots and (keep_undo) the
// roots of every undo / redo frame, each once.
static std::vector<uint32_t> gc_root_set(const hd_ctx* ctx, const uint32_t* extra, uint32_t n_extra, bool keep_undo) {
    std::vector<uint32_t> r{ctx->root};
    for (uint32_t i = 0; i < n_extra; ++i) r.push_back(extra[i]);
    if (keep_undo)
        for (const auto* frames : {&ctx->undo, &ctx->redo})
            for (const UndoFrame& f : *frames) {
                r.push_back(f.prev_root);
                r.push_back(f.new_root);
            }
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    return r;
}

static uint32_t remap(const std::vector<uint32_t>& from, const std::vector<uint32_t>& to, uint32_t v) {
    return to[std::lower_bound(from.begin(), from.end(), v) - from.begin()];
}

// GC scratch slots (fixed indices; frontier l at kSlotFrontier + l)
enum : uint32_t {
    kSlotBitmap, kSlotCounts, kSlotErrors, kSlotRoots, kSlotTmp, kSlotTlen, kSlotBkt, kSlotRank, kSlotOlen,
    kSlotOpos, kSlotBcount, kSlotBstart, kSlotTileSum, kSlotFwd, kSlotUsed, kSlotFrontier
};

// slot k of the context's GC scratch with at least `bytes` (grown by a quarter beyond the
// request, so tables that grow between collections reallocate rarely)
static hipError_t gc_slot(hd_ctx* ctx, uint32_t k, uint64_t bytes, void** p) {
    if (ctx->gc_slots.size() <= k) ctx->gc_slots.resize(k + 1, {nullptr, 0});
    auto& sl = ctx->gc_slots[k];
    if (sl.second < bytes) {
        if (sl.first) {
            const hipError_t e = hipFree(sl.first);
            sl = {nullptr, 0};
            if (e != hipSuccess) return e;
        }
        const uint64_t want = std::max<uint64_t>(bytes + bytes / 4, 256);
        void* q = nullptr;
        const hipError_t e = hipMalloc(&q, want);
        if (e != hipSuccess) return e;
        sl = {q, want};
    }
    *p = sl.first;
    return hipSuccess;
}

// words used per level 0..levels-1 (one small readback in place of every bucket size)
static hd_status level_used(hd_ctx* ctx, hipStream_t s, std::vector<uint64_t>& used) {
    const Geom& g = ctx->g;
    unsigned long long* d = nullptr;
    hipError_t e = gc_slot(ctx, kSlotUsed, (uint64_t)g.levels * 8, (void**)&d);
    if (e != hipSuccess) return hip_fail(ctx, e, "GC scratch");
    k_gc_fill<<<1, 64, 0, s>>>((uint32_t*)d, (uint64_t)g.levels * 2, 0u);
    k_gc_level_used<<<dim3(64, g.levels), 256, 0, s>>>(g, ctx->bucket_sizes, d);
    used.assign(g.levels, 0);
    HD_HIP(ctx, hipMemcpyAsync(used.data(), d, (uint64_t)g.levels * 8, hipMemcpyDeviceToHost, s));
    HD_HIP(ctx, hipStreamSynchronize(s));
    return HD_OK;
}

static hd_status gc_run(hd_ctx* ctx, uint32_t max_level, uint32_t* extra, uint32_t n_extra, bool keep_undo,
                        uint64_t stats[4], void* stream) {
    if (!ctx || !ctx->loaded) return fail(ctx, HD_ERR_INVALID, "no DAG loaded");
    const Geom& g = ctx->g;
    const uint32_t L = g.leaf_level();
    if (max_level > L) return fail(ctx, HD_ERR_INVALID, "max_level above the leaf level");
    if (n_extra && !extra) return fail(ctx, HD_ERR_INVALID, "extra roots missing");
    const std::vector<uint32_t> roots = gc_root_set(ctx, extra, n_extra, keep_undo);
    for (uint32_t r : roots)
        if (r >= g.total_virtual()) return fail(ctx, HD_ERR_INVALID, "a root is not an address of the table");
    const uint32_t nRoots = (uint32_t)roots.size();
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // capacities: used words per level bound the number of nodes
    std::vector<uint64_t> used;
    {
        const hd_status st = level_used(ctx, s, used);
        if (st != HD_OK) return st;
    }
    // every node takes at least 2 words, so a level holds at most used / 2 nodes
    auto fcap = [&](uint32_t l) {
        return (uint32_t)std::max<uint64_t>(std::max<uint64_t>(used[l] / 2 + 16, 16), l == 0 ? nRoots : 0);
    };
    uint64_t capMax = 16;
    for (uint32_t l = 0; l <= max_level; ++l) capMax = std::max<uint64_t>(capMax, fcap(l));
    if (capMax >= (1u << 30)) return fail(ctx, HD_ERR_INVALID, "level too large for GC scratch");
    const uint32_t cap = (uint32_t)capMax;

    // scratch from the context's slots (kept between calls; every kernel is given its sizes)
    auto alloc = [&](uint32_t slot, void** p, uint64_t bytes) -> hipError_t { return gc_slot(ctx, slot, bytes, p); };
    auto release = [&]() { hipStreamSynchronize(s); };
    const uint64_t bitmapWords = (g.total_virtual() + 31) / 32;
    uint32_t *bitmap, *counts, *errors, *tmp, *tlen, *fwd, *droots;
    std::vector<uint32_t*> frontier(max_level + 1, nullptr);
    hipError_t e = alloc(kSlotBitmap, (void**)&bitmap, bitmapWords * 4);
    if (e == hipSuccess) e = alloc(kSlotCounts, (void**)&counts, 64 * 4);
    if (e == hipSuccess) e = alloc(kSlotErrors, (void**)&errors, 4);
    if (e == hipSuccess) e = alloc(kSlotRoots, (void**)&droots, (uint64_t)nRoots * 4);
    for (uint32_t l = 0; l <= max_level && e == hipSuccess; ++l)
        e = alloc(kSlotFrontier + l, (void**)&frontier[l], (uint64_t)fcap(l) * 4);
    if (e == hipSuccess) e = alloc(kSlotTmp, (void**)&tmp, (uint64_t)cap * 36);
    if (e == hipSuccess) e = alloc(kSlotTlen, (void**)&tlen, (uint64_t)cap * 4);
    // placement scratch: bucket and rank per node, lengths and positions in placement order,
    // count and start per bucket of a level
    uint32_t maxBuckets = 1;
    for (uint32_t l = 0; l <= max_level; ++l) maxBuckets = std::max(maxBuckets, g.buckets_per_level(l));
    uint32_t *bkt, *rank, *olen, *opos, *bcount, *bstart, *tileSum;
    if (e == hipSuccess) e = alloc(kSlotBkt, (void**)&bkt, (uint64_t)cap * 4);
    if (e == hipSuccess) e = alloc(kSlotRank, (void**)&rank, (uint64_t)cap * 4);
    if (e == hipSuccess) e = alloc(kSlotOlen, (void**)&olen, (uint64_t)cap * 4);
    if (e == hipSuccess) e = alloc(kSlotOpos, (void**)&opos, (uint64_t)cap * 4);
    if (e == hipSuccess) e = alloc(kSlotBcount, (void**)&bcount, (uint64_t)maxBuckets * 4);
    if (e == hipSuccess) e = alloc(kSlotBstart, (void**)&bstart, (uint64_t)maxBuckets * 4);
    if (e == hipSuccess) e = alloc(kSlotTileSum, (void**)&tileSum, (uint64_t)(maxBuckets / kScanTile + 1) * 4);
    // old address -> new address of every live node: one word per virtual address (written
    // for live nodes only, read only for them)
    if (e == hipSuccess) e = alloc(kSlotFwd, (void**)&fwd, (uint64_t)g.total_virtual() * 4);
    if (e != hipSuccess) {
        release();
        gc_scratch_free(ctx);  // give back what a smaller call could not reuse
        return hip_fail(ctx, e, "GC scratch");
    }
    k_gc_fill<<<1024, 256, 0, s>>>(bitmap, bitmapWords, 0u);
    k_gc_fill<<<1, 64, 0, s>>>(errors, 1, 0u);
    e = hipMemcpyAsync(droots, roots.data(), (uint64_t)nRoots * 4, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
        release();
        return hip_fail(ctx, e, "GC roots");
    }
    k_gc_seed<<<(nRoots + 255) / 256, 256, 0, s>>>(frontier[0], counts, droots, nRoots, g.levels);
    for (uint32_t l = 0; l < max_level; ++l)
        k_gc_mark<<<kGcGrid, 256, 0, s>>>(ctx->vpool, bitmap, frontier[l], frontier[l + 1], counts, l, fcap(l),
                                          fcap(l + 1), g.total_virtual(), errors);
    {  // nothing has been modified yet: stop here if the mark overflowed or met a bad pointer
        uint32_t markErr = 0;
        HD_HIP(ctx, hipMemcpyAsync(&markErr, errors, 4, hipMemcpyDeviceToHost, s));
        e = hipStreamSynchronize(s);
        if (e != hipSuccess || markErr) {
            release();
            if (e != hipSuccess) return hip_fail(ctx, e, "GC mark");
            return fail(ctx, HD_ERR_INVALID, "GC mark: a child pointer is outside the table or a level overflowed");
        }
    }
    TableRef T{g, ctx->vpool, ctx->page_table, ctx->bucket_sizes, ctx->counters + 0, ctx->cfg.pool_pages};
    for (uint32_t l = max_level; l != (uint32_t)-1; --l) {
        k_gc_gather<<<kGcGrid, 256, 0, s>>>(g, ctx->vpool, frontier[l], counts, l, max_level, cap, fwd, tmp, tlen, bkt);
        // reset the level's buckets, then the full node first, then the live nodes
        const uint32_t b0 = g.bucket_global(l, 0);
        k_gc_fill<<<256, 256, 0, s>>>(ctx->bucket_sizes + b0, g.buckets_per_level(l), 0u);
        k_gc_full<<<1, 1, 0, s>>>(T, l, ctx->full_nodes, errors);
        const uint32_t nb = g.buckets_per_level(l);
        k_gc_fill<<<256, 256, 0, s>>>(bcount, nb, 0u);
        k_gc_bucket<<<kGcGrid, 256, 0, s>>>(l, counts, cap, bkt, rank, bcount);
        const uint32_t nTiles = (nb + kScanTile - 1) / kScanTile;
        k_gc_tile_sums<<<nTiles, kScanTile, 0, s>>>(bcount, nb, tileSum);
        k_gc_tile_scan<<<1, kScanTile, 0, s>>>(tileSum, nTiles);
        k_gc_tile_apply<<<nTiles, kScanTile, 0, s>>>(bcount, nb, tileSum, bstart);
        k_gc_order<<<kGcGrid, 256, 0, s>>>(l, counts, cap, tlen, bkt, rank, bstart, olen);
        k_gc_layout<<<(nb + 255) / 256, 256, 0, s>>>(g, l, ctx->bucket_sizes, bcount, bstart, olen, opos, errors);
        k_gc_write<<<kGcGrid, 256, 0, s>>>(T, l, counts, cap, frontier[l], tmp, tlen, bkt, rank, bstart, opos, fwd,
                                           errors);
    }
    k_gc_roots<<<(nRoots + 255) / 256, 256, 0, s>>>(fwd, droots, nRoots, errors);
    uint32_t err = 0;
    std::vector<uint32_t> newRoots(nRoots);
    std::vector<uint32_t> cnt(max_level + 1);
    HD_HIP(ctx, hipMemcpyAsync(&err, errors, 4, hipMemcpyDeviceToHost, s));
    HD_HIP(ctx, hipMemcpyAsync(newRoots.data(), droots, (uint64_t)nRoots * 4, hipMemcpyDeviceToHost, s));
    HD_HIP(ctx, hipMemcpyAsync(cnt.data(), counts, cnt.size() * 4, hipMemcpyDeviceToHost, s));
    e = hipStreamSynchronize(s);
    release();
    if (e != hipSuccess || err) {
        // levels were rewritten in place up to the failure: the table no longer holds the DAG,
        // so every call on this context refuses until a DAG is loaded again
        ctx->loaded = false;
        ctx->undo.clear();
        ctx->redo.clear();
        if (e != hipSuccess) return hip_fail(ctx, e, "GC");
        return fail(ctx, err & GC_ERR_POOL ? HD_ERR_POOL_FULL : HD_ERR_BUCKET_FULL,
                    "GC failed (capacity); the table is no longer consistent: load a DAG again");
    }
    ctx->root = remap(roots, newRoots, ctx->root);
    HD_HIP(ctx, hipMemcpyAsync(ctx->d_root, &ctx->root, 4, hipMemcpyHostToDevice, s));
    {
        const hd_status ts = leaf_tags_rebuild(ctx, s);  // the leaf buckets were rewritten
        if (ts != HD_OK) return ts;
    }
    for (uint32_t i = 0; i < n_extra; ++i) extra[i] = remap(roots, newRoots, extra[i]);
    if (keep_undo) {
        for (auto* frames : {&ctx->undo, &ctx->redo})
            for (UndoFrame& f : *frames) {
                f.prev_root = remap(roots, newRoots, f.prev_root);
                f.new_root = remap(roots, newRoots, f.new_root);
            }
    } else {
        ctx->undo.clear();  // old roots point at collected nodes (engine.cpp:178-185 frees undo)
        ctx->redo.clear();
    }
    if (stats) {
        uint64_t live = 0;
        for (uint32_t v : cnt) live += v;
        uint64_t before = 0;
        for (uint32_t l = 0; l <= max_level; ++l) before += used[l];
        stats[0] = live;
        stats[1] = before;
        std::vector<uint64_t> used2;
        const hd_status st = level_used(ctx, s, used2);
        if (st != HD_OK) return st;
        uint64_t after = 0;
        for (uint32_t l = 0; l <= max_level; ++l) after += used2[l];
        stats[2] = after;
        stats[3] = 0;
    }
    return HD_OK;
}

namespace hd {
void gc_scratch_free(hd_ctx* ctx) {
    if (!ctx) return;
    for (auto& sl : ctx->gc_slots)
        if (sl.first) hipFree(sl.first);
    ctx->gc_slots.clear();
}
}  // namespace hd

extern "C" hd_status hd_gc(hd_ctx* ctx, uint32_t max_level, uint64_t stats[4], void* stream) {
    return gc_run(ctx, max_level, nullptr, 0, false, stats, stream);
}

extern "C" hd_status hd_gc_roots(hd_ctx* ctx, uint32_t max_level, uint32_t keep_undo, uint32_t* roots,
                                 uint32_t n_roots, uint64_t stats[4], void* stream) {
    return gc_run(ctx, max_level, roots, n_roots, keep_undo != 0, stats, stream);
}
