// Device edits: sphere / box add and carve, sphere paint, over the device-resident HashDAG.
//
// Semantics: HashDAG::edit_threads (hash_dag.h:255-405) with HashDagEdits::edit_threaded /
// edit_leaf (hash_dag_edits.h:106-643) and the editors of hash_dag_editors.h. The result is
// canonically identical to the reference's (same node set after flattening, same colours);
// physical addresses differ (the reference's are thread-order dependent too).
//
// One pass of the pipeline applies an ordered list of edits (one edit = the plain form; a
// whole batch = the fused form). Every work item carries the ordered sub-list of the edits
// that touch its node, so one traversal composes the edits exactly as applying them one by
// one would: per voxel the value/colour is folded through the list in order, and a node
// covered entirely by a full-add / full-carve keeps only the edits after the last cover.
//
//   expand   levels 0 .. L-1, top-down: per (item, child) filter the edit list, classify the
//            child (keep / empty / full / item), append child items (atomic slot allocation)
//   build    levels L .. 0, bottom-up: new leaf masks (one wavefront per leaf item, one lane
//            per voxel) and new interior nodes; unchanged nodes keep their index
//   dedup    identical new nodes of a level collapse to one representative (device hash set)
//   find     one wavefront per representative scans its bucket in 64-word strides
//   add      misses reserve space with a CAS on the bucket size (page padding rule of
//            hash_table.h:418-442) and write the node; new pages are numbered from poolTop
//   colours  per colour subtree (level color_tree_levels item): prefix sums of new / old voxel
//            counts, emission of every new colour (copied, filled, or folded per voxel), then
//            one workgroup re-encodes the subtree's leaf (block / weight / macro-block rules of
//            variable_weight_size_colors.h:557-721) into the colour arena
//
// Kernels read item counts from device memory (grid-stride loops), so a pass needs no host
// round trip and can be captured in a HIP graph.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <random>

#include <algorithm>
#include <cstring>
#include <vector>

#include "hd_ctx.h"
#include "hd_edit.h"

namespace hd {
namespace {

constexpr uint32_t KIND_KEEP = 0, KIND_EMPTY = 1, KIND_FULL = 2, KIND_ITEM = 3;
constexpr uint32_t kGrid = 2048;  // blocks of grid-stride kernels

// Wave-strided loops (one work item per wavefront): the index is made wave-uniform so the
// compiler keeps it in a scalar register and loads the item's fields through the scalar cache.
// The pass kernels' bodies take the grid as an argument (VGrid), so one persistent launch can run
// them in sequence with the same work division.
struct VGrid {
    uint32_t block, nblocks;
};
__device__ __forceinline__ uint32_t wave_index(VGrid G) {
    return __builtin_amdgcn_readfirstlane((G.block * blockDim.x + threadIdx.x) >> 6);
}
__device__ __forceinline__ uint32_t wave_count(VGrid G) { return (G.nblocks * blockDim.x) >> 6; }

// Wave-aggregated counters (every lane of the wavefront must call these): one atomic per
// wavefront instead of one per lane, so millions of items do not queue on one address.
__device__ __forceinline__ uint32_t wave_scan32(uint32_t v) {
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}
__device__ __forceinline__ uint64_t wave_scan64(uint64_t v) {
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t lo = __shfl_up((uint32_t)v, off, 64), hi = __shfl_up((uint32_t)(v >> 32), off, 64);
        if (lane >= off) v += ((uint64_t)hi << 32) | lo;
    }
    return v;
}
// exclusive offset of this lane's v in a block reserved from *ctr
__device__ __forceinline__ uint32_t wave_alloc32(uint32_t* ctr, uint32_t v) {
    const uint32_t incl = wave_scan32(v);
    const uint32_t total = __shfl(incl, 63, 64);
    uint32_t base = 0;
    if (total != 0 && (threadIdx.x & 63u) == 63u) base = atomicAdd(ctr, total);
    base = __shfl(base, 63, 64);
    return base + incl - v;
}
__device__ __forceinline__ uint64_t wave_alloc64(unsigned long long* ctr, uint64_t v) {
    const uint64_t incl = wave_scan64(v);
    const uint32_t lo = __shfl((uint32_t)incl, 63, 64), hi = __shfl((uint32_t)(incl >> 32), 63, 64);
    const uint64_t total = ((uint64_t)hi << 32) | lo;
    unsigned long long base = 0;
    if (total != 0 && (threadIdx.x & 63u) == 63u) base = atomicAdd(ctr, (unsigned long long)total);
    const uint64_t b = ((uint64_t)__shfl((uint32_t)(base >> 32), 63, 64) << 32) | __shfl((uint32_t)base, 63, 64);
    return b + incl - v;
}
__device__ __forceinline__ void wave_add64(unsigned long long* ctr, uint64_t v) { (void)wave_alloc64(ctr, v); }
// Block-wide wave_alloc32 (one atomic per block); block-uniform control flow, `lds` holds
// waves + 1 words.
__device__ __forceinline__ uint32_t block_alloc32(uint32_t* lds, uint32_t* ctr, uint32_t v) {
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint32_t incl = wave_scan32(v);
    if (lane == 63) lds[wv] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (uint32_t w = 0; w < nw; ++w) {
            const uint32_t x = lds[w];
            lds[w] = tot;
            tot += x;
        }
        lds[nw] = tot ? atomicAdd(ctr, tot) : 0u;
    }
    __syncthreads();
    const uint32_t out = lds[nw] + lds[wv] + incl - v;
    __syncthreads();
    return out;
}
// Block-wide form of wave_alloc32 for two counters: one atomic per counter per block (a counter
// every wavefront of a large grid hits is limited by how fast one address takes atomics). Every
// thread of the block calls it, in block-uniform control flow; `lds` holds 2 * (waves + 1) words.
__device__ __forceinline__ void block_alloc32x2(uint32_t* lds, uint32_t* ctrA, uint32_t a, uint32_t* ctrB, uint32_t b,
                                                uint32_t& outA, uint32_t& outB) {
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint32_t ia = wave_scan32(a), ib = wave_scan32(b);
    if (lane == 63) {
        lds[wv] = ia;
        lds[nw + 1 + wv] = ib;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ta = 0, tb = 0;
        for (uint32_t w = 0; w < nw; ++w) {
            const uint32_t xa = lds[w], xb = lds[nw + 1 + w];
            lds[w] = ta;  // exclusive prefix of the wavefront totals
            lds[nw + 1 + w] = tb;
            ta += xa;
            tb += xb;
        }
        lds[nw] = ta ? atomicAdd(ctrA, ta) : 0u;
        lds[2 * nw + 1] = tb ? atomicAdd(ctrB, tb) : 0u;
    }
    __syncthreads();
    outA = lds[nw] + lds[wv] + ia - a;
    outB = lds[2 * nw + 1] + lds[nw + 1 + wv] + ib - b;
    __syncthreads();  // the words are reused by the next call
}
__device__ __forceinline__ uint64_t shfl64(uint64_t v, uint32_t src) {
    return ((uint64_t)__shfl((uint32_t)(v >> 32), src, 64) << 32) | __shfl((uint32_t)v, src, 64);
}

// ---------------------------------------------------------------------------------------------
// Editors (hash_dag_editors.h:9-330). FP32, contraction off.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float sqf(float f) { return f * f; }
__device__ __forceinline__ float tmaxf2(float a, float b) { return a > b ? a : b; }

// hash_dag_editors.h:76-91 / 112-126
template <class E>
__device__ __forceinline__ bool e_touch(const E& e, float x0, float y0, float z0, float x1, float y1, float z1) {
    if (e.kind == EDIT_BOX_ADD || e.kind == EDIT_BOX_CARVE || e.kind == EDIT_COPY || e.kind == EDIT_FILL ||
        e.kind == EDIT_NOISE) {
        return !(e.bmin[0] >= x1 || e.bmin[1] >= y1 || e.bmin[2] >= z1 || e.bmax[0] <= x0 || e.bmax[1] <= y0 ||
                 e.bmax[2] <= z0 || x0 >= e.bmax[0] || y0 >= e.bmax[1] || z0 >= e.bmax[2] || x1 <= e.bmin[0] ||
                 y1 <= e.bmin[1] || z1 <= e.bmin[2]);
    }
    float dist = e.r2;
    if (e.c[0] < x0) dist -= sqf(e.c[0] - x0);
    else if (e.c[0] > x1) dist -= sqf(e.c[0] - x1);
    if (e.c[1] < y0) dist -= sqf(e.c[1] - y0);
    else if (e.c[1] > y1) dist -= sqf(e.c[1] - y1);
    if (e.c[2] < z0) dist -= sqf(e.c[2] - z0);
    else if (e.c[2] > z1) dist -= sqf(e.c[2] - z1);
    return dist > 0;
}
// hash_dag_editors.h:233-252 / 280-295: 1 = full (add covers), 2 = empty (carve covers), 0 = no
template <class E>
__device__ __forceinline__ uint32_t e_cover(const E& e, float x0, float y0, float z0, float x1, float y1,
                                            float z1) {
    if (e.kind == EDIT_BOX_ADD || e.kind == EDIT_BOX_CARVE) {
        const bool in = e.bmin[0] <= x0 && e.bmin[1] <= y0 && e.bmin[2] <= z0 && e.bmax[0] >= x1 && e.bmax[1] >= y1 &&
                        e.bmax[2] >= z1;
        return in ? (e.kind == EDIT_BOX_ADD ? 1u : 2u) : 0u;
    }
    if (e.kind == EDIT_SPHERE_ADD || e.kind == EDIT_SPHERE_CARVE) {
        const bool in = tmaxf2(sqf(x0 - e.c[0]), sqf(x1 - e.c[0])) + tmaxf2(sqf(y0 - e.c[1]), sqf(y1 - e.c[1])) +
                            tmaxf2(sqf(z0 - e.c[2]), sqf(z1 - e.c[2])) <
                        e.r2;
        return in ? (e.kind == EDIT_SPHERE_ADD ? 1u : 2u) : 0u;
    }
    return 0u;
}

// CompressedColor packed in 64 bits: colorBits | weight << 32 | bpw << 40
__device__ __forceinline__ uint64_t pack_color(uint32_t bits, uint32_t w, uint32_t bpw) {
    return (uint64_t)bits | ((uint64_t)w << 32) | ((uint64_t)bpw << 40);
}
__device__ __forceinline__ float clampf(float v, float a, float b) { return v < a ? a : (v > b ? b : v); }
__device__ __forceinline__ uint32_t to101210(float r, float g, float b) {
    r = clampf(r, 0.f, 1.f);
    g = clampf(g, 0.f, 1.f);
    b = clampf(b, 0.f, 1.f);
    return ((uint32_t)(r * 1023.0f)) | ((uint32_t)(g * 4095.0f) << 10) | ((uint32_t)(b * 1023.0f) << 22);
}
// CompressedColor::get_color (variable_weight_size_colors.h:144-154)
__device__ __forceinline__ void unpack_rgb(uint64_t c, float& r, float& g, float& b) {
    const uint32_t bits = (uint32_t)c, w = (uint32_t)(c >> 32) & 0xFFu, bpw = (uint32_t)(c >> 40) & 0xFFu;
    if (bpw == 0) {
        r = (float)(bits & 0x3FFu) / 1023.0f;
        g = (float)((bits >> 10) & 0xFFFu) / 4095.0f;
        b = (float)((bits >> 22) & 0x3FFu) / 1023.0f;
        return;
    }
    const float f = (float)w / (float)((1u << bpw) - 1u);
    const float omf = 1 - f;
    r = (float)(bits & 0x1Fu) / 31.0f * omf + (float)((bits >> 16) & 0x1Fu) / 31.0f * f;
    g = (float)((bits >> 5) & 0x3Fu) / 63.0f * omf + (float)((bits >> 21) & 0x3Fu) / 63.0f * f;
    b = (float)((bits >> 11) & 0x1Fu) / 31.0f * omf + (float)((bits >> 27) & 0x1Fu) / 31.0f * f;
}
// get_new_color of the editors (hash_dag_editors.h:264-269, 307-312, 323-329)
template <class E>
__device__ __forceinline__ uint64_t e_color(const E& e, float px, float py, float pz, uint64_t prev) {
    if (e.kind != EDIT_SPHERE_PAINT) return pack_color(e.single, 0, 0);
    float r, g, b;
    unpack_rgb(prev, r, g, b);
    const float dx = px - e.c[0], dy = py - e.c[1], dz = pz - e.c[2];
    const float len = __fsqrt_rn(dx * dx + dy * dy + dz * dz);
    const float k = clampf(1 - __fdiv_rn(len, e.radius), 0.f, .1f);
    const float omk = 1 - k;
    return pack_color(to101210(r * omk + 0.0f * k, g * omk + 0.0f * k, b * omk + 1.0f * k), 0, 0);
}
template <class E>
__device__ __forceinline__ bool e_value(const E& e, bool old) {
    return e.kind == EDIT_SPHERE_PAINT ? old : (e.kind == EDIT_SPHERE_ADD || e.kind == EDIT_BOX_ADD);
}

__device__ __forceinline__ EditHead load_head(const DevEdit* edits, uint32_t ei) {
    return *reinterpret_cast<const EditHead*>(edits + ei);
}

// ---------------------------------------------------------------------------------------------
// Colour leaf decode (variable_weight_size_colors.h:343-413) returning the packed colour
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t hdr_index(uint32_t h) { return h & 16383u; }
__device__ __forceinline__ uint32_t hdr_bpw(uint32_t h) { return (h >> 16) == 0xFFFFu ? 0u : ((h >> 14) & 3u) + 1u; }

__device__ uint64_t leaf_decode(const uint64_t* __restrict__ arena, const LeafDesc& d, uint64_t colorIndex) {
    if (d.offset != ~0ull) colorIndex += d.offset;
    const uint64_t* blocks = arena + d.blocks;
    const uint64_t* macro = arena + d.macro;
    const uint32_t local = (uint32_t)(colorIndex & 16383u);
    const uint64_t m = colorIndex >> 14;
    uint32_t lo = (uint32_t)macro[2 * m];
    uint32_t hi = 2 * (m + 1) < d.n_macro ? (uint32_t)macro[2 * (m + 1)] - 1u : (uint32_t)d.n_blocks - 1u;
    uint32_t pos = (lo + hi) / 2;
    uint64_t blk = blocks[pos];
    while (hdr_index((uint32_t)blk) != local && lo <= hi) {
        if (hdr_index((uint32_t)blk) > local) hi = pos - 1;
        else lo = pos + 1;
        pos = (lo + hi) / 2;
        blk = blocks[pos];
    }
    const uint32_t h = (uint32_t)blk, bits = (uint32_t)(blk >> 32), bpw = hdr_bpw(h);
    if (bpw == 0) return pack_color(bits, 0, 0);
    const uint64_t bit = macro[2 * m + 1] + (h >> 16) + (uint64_t)((local - hdr_index(h)) * bpw);
    const uint32_t* w = reinterpret_cast<const uint32_t*>(arena) + d.weights32;
    const uint64_t byte = bit >> 3;
    const uint32_t w0 = w[byte >> 2], sh = (uint32_t)(byte & 3u) * 8u;
    const uint32_t b0 = (w0 >> sh) & 0xFFu;
    const uint32_t b1 = (byte & 3u) == 3u ? (w[(byte >> 2) + 1] & 0xFFu) : ((w0 >> (sh + 8)) & 0xFFu);
    const uint32_t dst = (b0 << 8) | b1;
    const uint32_t wq = (dst >> (16u - bpw - (uint32_t)(bit & 7u))) & ((1u << bpw) - 1u);
    return pack_color(bits, wq, bpw);
}
// Runs of consecutive colours decoded by a wavefront: lane r (0 <= r < span) takes colour
// first + r. `pos` is a block whose start is at most `first` and after which no block starts
// before `first` (~0u: search it); the 64 block headers after it are read at once, the block
// starts among them that fall in [first, first + 64) are gathered into a bit mask, and lane r's
// block is pos plus the number of those starts at or before r: no search per colour. pos
// moves to the block of colour first + span - 1, the hint for a run that starts right after.
// Lanes outside the run (active false) or past the leaf's valid colours return 0. Called by
// every lane of the wavefront with the same first, span and pos.
__device__ uint64_t leaf_decode_run(const uint64_t* __restrict__ arena, const LeafDesc& d, uint64_t first,
                                    uint32_t span, int32_t r, bool active, uint32_t& pos) {
    const uint64_t limit = (d.n_macro / 2) * 16384ull;  // leaf_valid_index: 2 i / 16384 < n_macro
    const uint64_t x0 = first + (d.offset != ~0ull ? d.offset : 0ull);
    if (d.n_blocks == 0 || x0 >= limit) return 0;
    const uint64_t* blocks = arena + d.blocks;
    const uint64_t* macro = arena + d.macro;
    const uint64_t m0 = x0 >> 14;
    if (pos == ~0u) {  // the block holding colour x0 (the search of leaf_decode)
        const uint32_t local = (uint32_t)(x0 & 16383u);
        uint32_t lo = (uint32_t)macro[2 * m0];
        uint32_t hi = 2 * (m0 + 1) < d.n_macro ? (uint32_t)macro[2 * (m0 + 1)] - 1u : (uint32_t)d.n_blocks - 1u;
        uint32_t p = (lo + hi) / 2;
        uint64_t blk = blocks[p];
        while (hdr_index((uint32_t)blk) != local && lo <= hi) {
            if (hdr_index((uint32_t)blk) > local) hi = p - 1;
            else lo = p + 1;
            p = (lo + hi) / 2;
            blk = blocks[p];
        }
        pos = p;
    }
    const uint32_t lane = threadIdx.x & 63u;
    // block pos + 1 + lane: its macro block (a run of 64 colours crosses at most one macro
    // block start, and pos may end the macro block before x0's), its start, and its bit.
    // Blocks from macro block m0 + 2 on start past the run (a window can reach them when
    // macro block m0 + 1 holds few blocks) and are left out.
    const uint32_t mFirst = (uint32_t)macro[2 * m0];
    const uint32_t mNext = 2 * (m0 + 1) < d.n_macro ? (uint32_t)macro[2 * (m0 + 1)] : (uint32_t)d.n_blocks;
    const uint32_t mNext2 = 2 * (m0 + 2) < d.n_macro ? (uint32_t)macro[2 * (m0 + 2)] : (uint32_t)d.n_blocks;
    const uint64_t wb = (uint64_t)pos + 1 + lane;
    uint64_t wblk = 0, bit = 0;
    if (wb < mNext2) {
        wblk = blocks[wb];
        const uint64_t wm = wb < mFirst ? m0 - 1 : (wb < mNext ? m0 : m0 + 1);
        const uint64_t st = (wm << 14) + hdr_index((uint32_t)wblk);
        if (st >= x0 && st < x0 + 64) bit = 1ull << (st - x0);
    }
#pragma unroll
    for (uint32_t off = 32; off > 0; off >>= 1) bit |= shfl64(bit, lane ^ off);
    const uint64_t starts = bit;  // the same in every lane
    const uint64_t b0 = blocks[pos];
    const uint32_t rr = r < 0 ? 0u : (uint32_t)r;
    const uint32_t nb = (uint32_t)__popcll(starts & (rr >= 63 ? ~0ull : (2ull << rr) - 1ull));
    const uint64_t wsel = shfl64(wblk, nb == 0 ? 0u : nb - 1);  // every lane takes part
    const uint64_t blk = nb == 0 ? b0 : wsel;
    pos += (uint32_t)__popcll(starts & (span >= 64 ? ~0ull : (1ull << span) - 1ull));
    const uint64_t x = x0 + rr;
    if (!active || r < 0 || x >= limit) return 0;
    const uint32_t h = (uint32_t)blk, bits = (uint32_t)(blk >> 32), bpw = hdr_bpw(h);
    if (bpw == 0) return pack_color(bits, 0, 0);
    const uint32_t local = (uint32_t)(x & 16383u);
    const uint64_t m = x >> 14;
    const uint64_t wbit = macro[2 * m + 1] + (h >> 16) + (uint64_t)((local - hdr_index(h)) * bpw);
    const uint32_t* w = reinterpret_cast<const uint32_t*>(arena) + d.weights32;
    const uint64_t byte = wbit >> 3;
    const uint32_t w0 = w[byte >> 2], sh = (uint32_t)(byte & 3u) * 8u;
    const uint32_t c0 = (w0 >> sh) & 0xFFu;
    const uint32_t c1 = (byte & 3u) == 3u ? (w[(byte >> 2) + 1] & 0xFFu) : ((w0 >> (sh + 8)) & 0xFFu);
    const uint32_t dst = (c0 << 8) | c1;
    const uint32_t wq = (dst >> (16u - bpw - (uint32_t)(wbit & 7u))) & ((1u << bpw) - 1u);
    return pack_color(bits, wq, bpw);
}
// is_valid() && is_valid_index(i) (variable_weight_size_colors.h:247-262)
__device__ __forceinline__ bool leaf_valid_index(const LeafDesc& d, uint64_t i) {
    if (d.n_blocks == 0) return false;
    if (d.offset != ~0ull) i += d.offset;
    return 2 * i / 16384u < d.n_macro;
}

// ---------------------------------------------------------------------------------------------
// Copy editor reads (hash_dag_editors.h:364-434): the DAG as it was when the pass started
// ---------------------------------------------------------------------------------------------
// float -> uint32 as x86-64 code converts it (truncate to int64, keep the low 32 bits)
__device__ __forceinline__ uint32_t f2u_x86(float f) { return (uint32_t)(int64_t)f; }

// source voxel of a destination voxel corner (destination_to_source, hash_dag_editors.h:364-379):
// p - dest in FP32, swirled about y, through the inverse transform in FP64, plus source. The
// swirl's cos / sin are FP64 rounded to FP32 (the reference calls the host's cosf / sinf).
__device__ __forceinline__ void copy_source(const DevEdit& e, float px, float py, float pz, uint32_t& sx, uint32_t& sy,
                                            uint32_t& sz) {
    float rx = px - e.dst[0], ry = py - e.dst[1], rz = pz - e.dst[2];
    if (e.swirl) {
        const float angle = __fdiv_rn(ry, e.swirl_period);
        const float c = (float)cos((double)angle), s = (float)sin((double)angle);
        const float x = rx, z = rz;
        rx = x * c - z * s;
        rz = x * s + z * c;
    }
    const double vx = (double)rx, vy = (double)ry, vz = (double)rz;
    const float tx = (float)(e.inv[0] * vx + e.inv[1] * vy + e.inv[2] * vz);
    const float ty = (float)(e.inv[3] * vx + e.inv[4] * vy + e.inv[5] * vz);
    const float tz = (float)(e.inv[6] * vx + e.inv[7] * vy + e.inv[8] * vz);
    sx = f2u_x86(e.src[0] + tx);
    sy = f2u_x86(e.src[1] + ty);
    sz = f2u_x86(e.src[2] + tz);
}

// source voxel of a destination voxel corner for either copy: destination_to_source for the
// copy without decompression; for the decompressing copy the voxel buffer's cell
// truncate(position - dest + source - copyMin), read at intCopyMin + cell, and false outside
// the buffer (get_new_value / get_new_color keep the old voxel, hash_dag_editors.h:582-611)
__device__ __forceinline__ bool copy_source_ok(const DevEdit& e, float px, float py, float pz, uint32_t& sx,
                                               uint32_t& sy, uint32_t& sz) {
    if (!e.dec) {
        copy_source(e, px, py, pz, sx, sy, sz);
        return true;
    }
    const uint32_t x = f2u_x86(px - e.dst[0] + e.src[0] - e.cmin[0]);
    const uint32_t y = f2u_x86(py - e.dst[1] + e.src[1] - e.cmin[1]);
    const uint32_t z = f2u_x86(pz - e.dst[2] + e.src[2] - e.cmin[2]);
    if (!(x < e.dsize && y < e.dsize && z < e.dsize)) return false;
    sx = e.icmin[0] + x;
    sy = e.icmin[1] + y;
    sz = e.icmin[2] + z;
    return true;
}

// DAGUtils::is_empty (dag_utils.h:174-252) on the DAG before the pass, for the region
// [start, start + size) and a node level maxLevel >= 1: the recursion keeps the nodes above
// maxLevel whose inclusive box [min, min + 2^shift - 1] passes the overlap test (min < start +
// size and min + 2^shift - 1 > start per axis) and reports "not empty" as soon as it reaches a
// node of maxLevel. Every ancestor of an overlapping node overlaps too, so the region is empty
// exactly when no overlapping node of level maxLevel - 1 exists with a child. Per axis those
// nodes are a run of at most two indices, found directly; each candidate is one root-to-node
// walk.
__device__ bool source_region_empty(const uint32_t* __restrict__ vpool, uint32_t root, uint32_t levels,
                                    uint32_t maxLevel, uint32_t sx, uint32_t sy, uint32_t sz, uint32_t size) {
    const uint32_t k = maxLevel - 1, shift = levels - k;
    const uint32_t start[3] = {sx, sy, sz};
    uint32_t lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = (uint32_t)(((uint64_t)start[a] + 1) >> shift);
        const uint64_t h = ((uint64_t)start[a] + size - 1) >> shift;
        hi[a] = (uint32_t)min<uint64_t>(h, (1ull << k) - 1);
        if (lo[a] > hi[a]) return true;
    }
    for (uint32_t ax = lo[0]; ax <= hi[0]; ++ax)
        for (uint32_t ay = lo[1]; ay <= hi[1]; ++ay)
            for (uint32_t az = lo[2]; az <= hi[2]; ++az) {
                uint32_t va = root;
                bool exists = true;
                for (uint32_t level = 0; level < k && exists; ++level) {
                    const uint32_t b = k - level - 1;
                    const uint32_t c = (((ax >> b) & 1u) << 2) | (((ay >> b) & 1u) << 1) | ((az >> b) & 1u);
                    const uint32_t m = vpool[va] & 0xFFu;
                    if (!(m & (1u << c))) exists = false;
                    else va = vpool[va + __popc(m & ((1u << c) - 1u)) + 1u];
                }
                if (exists && (vpool[va] & 0xFFu) != 0u) return false;
            }
    return true;
}

// fill buffer voxel of a voxel corner: truncate(position - min) (hash_dag_editors.h:213-214);
// returns the state byte (0 outside the buffer)
__device__ __forceinline__ uint32_t fill_state_at(const DevEdit& e, float px, float py, float pz, uint64_t& idx) {
    const uint32_t x = f2u_x86(px - e.src[0]), y = f2u_x86(py - e.src[1]), z = f2u_x86(pz - e.src[2]);
    const uint32_t S = e.fill_size;
    if (x >= S || y >= S || z >= S) return 0u;
    idx = (uint64_t)x + (uint64_t)S * y + (uint64_t)S * S * z;
    return e.fill_state[idx];
}
// SphereNoiseEditor::get_new_color (hash_dag_editors.h:794-808): rgb565 min / max (precomputed
// in `single`), 3 bits of weight |p - c| / boxRadius - 0.3 (radius holds the box radius 1.5 r)
__device__ __forceinline__ uint64_t noise_color(const DevEdit& e, float px, float py, float pz) {
    const float dx = px - e.c[0], dy = py - e.c[1], dz = pz - e.c[2];
    const float len = __fsqrt_rn(dx * dx + dy * dy + dz * dz);
    const int32_t w = (int32_t)((__fdiv_rn(len, e.radius) - 0.3f) * 7.0f);
    return pack_color(e.single, (uint32_t)(w < 0 ? 0 : (w > 7 ? 7 : w)), 3u);
}
__device__ __forceinline__ uint64_t fill_color_at(const DevEdit& e, uint64_t idx) {
    const uint64_t S = e.fill_size;
    const uint32_t x = (uint32_t)(idx % S), y = (uint32_t)(idx / S % S), z = (uint32_t)(idx / (S * S));
    return e.fill_colors[(uint64_t)(x - e.fill_lo[0]) +
                         (uint64_t)e.fill_dim[0] * ((y - e.fill_lo[1]) + (uint64_t)e.fill_dim[1] * (z - e.fill_lo[2]))];
}

// 64-bit leaf mask of the leaf at leaf coordinates (lx, ly, lz) (voxel >> 2), 0 when absent
// (DAGUtils::get_value, dag_utils.h:138-172, stopped at the leaf)
__device__ uint64_t source_leaf(const uint32_t* __restrict__ vpool, uint32_t root, uint32_t levels, uint32_t lx,
                                uint32_t ly, uint32_t lz) {
    const uint32_t L = levels - 2;
    uint32_t va = root;
    for (uint32_t level = 0; level < L; ++level) {
        const uint32_t s = L - level - 1;  // bit of the leaf coordinates below this level
        const uint32_t c = (((lx >> s) & 1u) << 2) | (((ly >> s) & 1u) << 1) | ((lz >> s) & 1u);
        const uint32_t m = vpool[va] & 0xFFu;
        if (!(m & (1u << c))) return 0;
        va = vpool[va + __popc(m & ((1u << c) - 1u)) + 1u];
    }
    return *reinterpret_cast<const uint64_t*>(vpool + va);
}

__device__ __forceinline__ uint32_t leaf_bit(uint32_t x, uint32_t y, uint32_t z) {
    return ((x & 1u) ? 4u : 0u) | ((y & 1u) ? 2u : 0u) | ((z & 1u) ? 1u : 0u) | ((x & 2u) ? 32u : 0u) |
           ((y & 2u) ? 16u : 0u) | ((z & 2u) ? 8u : 0u);
}

__device__ __forceinline__ uint32_t child_off(uint32_t mask, uint32_t c) { return __popc(mask & ((1u << c) - 1u)) + 1u; }

__device__ __forceinline__ void set_error(uint32_t* flags, uint32_t bit) { atomicOr(flags, bit); }

// CopyEditorWithDecompression::should_edit's empty check (COPY_EMPTY_CHECKS,
// hash_dag_editors.h:555-577): a node above levels - 4 whose source region (its box moved by
// source - dest, when that lies at non-negative coordinates) is empty in the DAG before the
// pass is not edited
__device__ __forceinline__ bool copy_source_empty(const EditPass& P, uint32_t ei, float x0, float y0, float z0,
                                                  uint32_t level) {
    const DevEdit& e = P.edits[ei];
    if (!e.dec) return false;
    const float ox = x0 - e.dst[0] + e.src[0], oy = y0 - e.dst[1] + e.src[1], oz = z0 - e.dst[2] + e.src[2];
    if (!(ox >= 0 && oy >= 0 && oz >= 0)) return false;
    return source_region_empty(P.vpool, P.items[0].orig[0], P.levels, level, f2u_x86(ox), f2u_x86(oy), f2u_x86(oz),
                               1u << (P.levels - level));
}

// ---------------------------------------------------------------------------------------------
// expand: items of level l -> child classification and items of level l+1
// ---------------------------------------------------------------------------------------------
// does edit ei (head h) take part in the child box [x0, x1) at level l + 1: its box overlaps,
// and a decompressing copy's source region there is not empty
__device__ __forceinline__ bool expand_touch(const EditPass& P, const EditHead& h, uint32_t ei, uint32_t l, float x0,
                                             float y0, float z0, float x1, float y1, float z1) {
    if (!e_touch(h, x0, y0, z0, x1, y1, z1)) return false;
    return !(h.kind == EDIT_COPY && l + 1 + 4 < P.levels && copy_source_empty(P, ei, x0, y0, z0, l + 1));
}

// child slot c of item i: its position, box, and its node under the item's effective old node
// and under its original node
struct SlotRefs {
    uint32_t cx, cy, cz, childOld, childOrig;
    float x0, y0, z0, x1, y1, z1;
};
__device__ __forceinline__ SlotRefs slot_refs(const EditPass& P, const Items& I, uint32_t i, uint32_t c,
                                              uint32_t depth, uint64_t& nodeBytes) {
    SlotRefs r;
    r.cx = (I.px[i] << 1) | ((c >> 2) & 1u);
    r.cy = (I.py[i] << 1) | ((c >> 1) & 1u);
    r.cz = (I.pz[i] << 1) | (c & 1u);
    r.childOld = kInvalidNode;
    r.childOrig = kInvalidNode;
    const uint32_t old = I.old_[i], orig = I.orig[i];
    if (old != kInvalidNode) {
        const uint32_t cm = P.vpool[old] & 0xFFu;
        if (cm & (1u << c)) r.childOld = P.vpool[old + child_off(cm, c)];
        if (c == 0) nodeBytes += 4u * (1u + (uint32_t)__popc(cm));
    }
    if (orig == old) {
        r.childOrig = r.childOld;
    } else if (orig != kInvalidNode) {
        const uint32_t cm = P.vpool[orig] & 0xFFu;
        if (cm & (1u << c)) r.childOrig = P.vpool[orig + child_off(cm, c)];
        if (c == 0) nodeBytes += 4u * (1u + (uint32_t)__popc(cm));
    }
    const float side = (float)(1u << depth);
    r.x0 = (float)(r.cx << depth);
    r.y0 = (float)(r.cy << depth);
    r.z0 = (float)(r.cz << depth);
    r.x1 = r.x0 + side;
    r.y1 = r.y0 + side;
    r.z1 = r.z0 + side;
    return r;
}

// Classifies a child slot from its filtered edits (m edits touch it; the last full / empty
// cover among them is the lastCover-th, ~0 for none): keep, full, empty or a child item.
// Writes the slot's corig / cold / cval / cnew / ccolor; coveredVoxels is the voxel count of a
// full / empty early exit.
__device__ __forceinline__ uint32_t slot_classify(const EditPass& P, const Items& I, uint32_t l, uint32_t slot,
                                                  uint32_t depth, const SlotRefs& r, uint32_t m, uint32_t lastCover,
                                                  uint32_t coverKind, uint32_t coverColor,
                                                  unsigned long long& coveredVoxels) {
    const uint32_t L = P.levels - 2, ctl = P.ctl;
    // old voxel count of the effective old child (colour offsets, "keep" counts)
    uint32_t oldCount = 0;
    if (l + 1 >= ctl && r.childOld != kInvalidNode) {
        if (l + 1 == L) {
            const uint64_t lf = *reinterpret_cast<const uint64_t*>(P.vpool + r.childOld);
            oldCount = __popcll(lf);
        } else {
            oldCount = P.vpool[r.childOld] >> 8;
        }
    }
    I.corig[slot] = r.childOrig;
    I.cold[slot] = oldCount;
    const bool covered = lastCover != ~0u;
    uint32_t kind;
    if (m == 0) {
        kind = KIND_KEEP;
        I.cval[slot] = r.childOld;
        I.cnew[slot] = oldCount;
    } else {
        const uint32_t rest = covered ? m - lastCover - 1 : m;
        // a covered colour subtree root stays an item (its colour leaf is rebuilt: the
        // reference's task runs edit_leaf<color_tree_levels>, which takes the early exit)
        if (covered && rest == 0 && l + 1 > ctl) {
            kind = coverKind == 1 ? KIND_FULL : KIND_EMPTY;
            I.cval[slot] = coverKind == 1 ? P.full_nodes[l + 1] : kInvalidNode;
            I.cnew[slot] = coverKind == 1 ? (1u << (3 * depth)) : 0u;
        } else {
            kind = KIND_ITEM;
        }
        // numVoxels of a plain edit: a full / empty early exit counts the whole node
        if (covered && rest == 0) coveredVoxels = 1ull << (3 * depth);
        // colour of a full cover (add_large_single_color of the covering editor)
        if (covered) I.ccolor[slot] = coverColor;
    }
    return kind;
}

// child item ci of a slot, all but its edit list (rest words from eo)
__device__ __forceinline__ void slot_child_item(const EditPass& P, const Items& I, const Items& C, uint32_t l,
                                                uint32_t i, uint32_t c, uint32_t slot, uint32_t ci, uint32_t eo,
                                                uint32_t rest, const SlotRefs& r, uint32_t lastCover,
                                                uint32_t coverKind) {
    const uint32_t ctl = P.ctl;
    C.px[ci] = r.cx;
    C.py[ci] = r.cy;
    C.pz[ci] = r.cz;
    C.orig[ci] = r.childOrig;
    C.parent[ci] = i;
    C.slot[ci] = (uint8_t)c;
    C.estart[ci] = eo;
    C.ecount[ci] = rest;
    if (lastCover != ~0u) {
        C.old_[ci] = coverKind == 1 ? P.full_nodes[l + 1] : kInvalidNode;
        C.cmode[ci] = coverKind == 1 ? 1u : 2u;  // old colours: single colour / none
        C.cmodeval[ci] = I.ccolor[slot];
    } else {
        C.old_[ci] = r.childOld;
        C.cmode[ci] = I.cmode[i];
        C.cmodeval[ci] = I.cmodeval[i];
    }
    // colour tree: node of the child (levels < ctl), or the subtree's leaf handle / task
    const uint32_t cn = I.aux[i];
    if (l + 1 < ctl) {
        C.aux[ci] = (l == 0 || cn > 0) ? P.cnodes[cn + c] : 0u;
    } else if (l + 1 == ctl) {
        C.aux[ci] = ci;  // task id = own index at level ctl
        const uint32_t h = (l == 0 || cn > 0) ? P.cnodes[cn + c] : 0u;
        C.handle[ci] = h;
    } else {
        C.aux[ci] = cn;  // task of the ancestor at level ctl
    }
    I.cval[slot] = ci;
}

// ---------------------------------------------------------------------------------------------
// expand: items of level l -> child classification and items of level l+1
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void d_expand(const EditPass& P, uint32_t l, VGrid G) {
    __shared__ uint32_t sAlloc[2 * (256 / 64 + 1)];
    __shared__ uint32_t sErr;
    const uint32_t n = min(P.count[l], P.cap[l]);
    if (threadIdx.x == 0) sErr = *P.errors;
    __syncthreads();
    if (sErr) return;  // block-uniform: the loop below synchronises the block
    const Items I = P.items[l];
    const Items C = P.items[l + 1];
    const uint32_t D = P.levels, ctl = P.ctl;
    const uint32_t depth = D - (l + 1);  // child cube side = 2^depth
    const uint64_t total = (uint64_t)n * 8;
    uint64_t nodeBytes = 0;  // header + child pointers of the items' nodes (measurement)
    // block-uniform loop: child items and edit-list words are allocated with one atomic each
    // per block
    for (uint64_t b0 = (uint64_t)G.block * blockDim.x; b0 < total; b0 += (uint64_t)G.nblocks * blockDim.x) {
        const uint64_t t = b0 + threadIdx.x;
        const bool active = t < total;
        const uint32_t i = (uint32_t)(t >> 3), c = (uint32_t)(t & 7u);
        const uint32_t slot = i * 8 + c;
        uint32_t kind = KIND_KEEP, m = 0, lastCover = ~0u, coverKind = 0, es = 0, ec = 0;
        uint64_t touch = 0;
        SlotRefs r{};
        unsigned long long coveredVoxels = 0;
        if (active) {
            r = slot_refs(P, I, i, c, depth, nodeBytes);
            es = I.estart[i];
            ec = I.ecount[i];
            // filter the edit list, two edit heads in flight; remember the last cover at levels
            // >= ctl (and its colour) and which of the first 64 edits touch
            uint32_t coverColor = 0;
            for (uint32_t k = 0; k < ec; k += 2) {
                const bool two = k + 1 < ec;
                const uint32_t e0 = P.elist[l][es + k];
                const uint32_t e1 = two ? P.elist[l][es + k + 1] : e0;
                const EditHead h0 = load_head(P.edits, e0), h1 = load_head(P.edits, e1);
#pragma unroll
                for (uint32_t u = 0; u < 2; ++u) {
                    if (u == 1 && !two) break;
                    const EditHead& h = u ? h1 : h0;
                    if (!expand_touch(P, h, u ? e1 : e0, l, r.x0, r.y0, r.z0, r.x1, r.y1, r.z1)) continue;
                    if (k + u < 64) touch |= 1ull << (k + u);
                    if (l + 1 >= ctl) {
                        const uint32_t cv = e_cover(h, r.x0, r.y0, r.z0, r.x1, r.y1, r.z1);
                        if (cv) {
                            lastCover = m;
                            coverKind = cv;
                            coverColor = h.single;
                        }
                    }
                    m++;
                }
            }
            kind = slot_classify(P, I, l, slot, depth, r, m, lastCover, coverKind, coverColor, coveredVoxels);
        }
        const bool isItem = active && kind == KIND_ITEM;
        const uint32_t rest = isItem ? (lastCover == ~0u ? m : m - lastCover - 1) : 0u;
        if (P.single_edit) wave_add64(P.stat_voxels, coveredVoxels);
        uint32_t ci, eo;
        block_alloc32x2(sAlloc, &P.count[l + 1], isItem ? 1u : 0u, &P.ecount[l + 1], rest, ci, eo);
        if (!active) continue;
        if (isItem && (ci >= P.cap[l + 1] || eo + rest > P.ecap[l + 1])) {
            set_error(P.errors, ERR_SCRATCH);
            I.ckind[slot] = (uint8_t)kind;
            continue;
        }
        if (isItem) {
            // the child's edit list: the touching edits after the last cover, in list order
            uint32_t q = 0, w = 0;
            for (uint32_t k = 0; k < ec; ++k) {
                const uint32_t ei = P.elist[l][es + k];
                if (k < 64 ? !((touch >> k) & 1ull)
                           : !expand_touch(P, load_head(P.edits, ei), ei, l, r.x0, r.y0, r.z0, r.x1, r.y1, r.z1))
                    continue;
                if (lastCover == ~0u || q > lastCover) P.elist[l + 1][eo + w++] = ei;
                q++;
            }
            slot_child_item(P, I, C, l, i, c, slot, ci, eo, rest, r, lastCover, coverKind);
        }
        I.ckind[slot] = (uint8_t)kind;
    }
    if (P.abytes) wave_add64(P.abytes + 0, nodeBytes);
}

// expand with a wavefront per child slot, for levels whose items hold long edit lists (the top
// levels of a pass over many editors): the lanes test 64 edits of the list at once, two ballots
// give the touching edits and the covers, and the child's list is written in list order by a
// second sweep from the chunk of the last cover. Same classification and items as d_expand.
__device__ __forceinline__ void d_expand_wide(const EditPass& P, uint32_t l, VGrid G) {
    __shared__ uint32_t sErr;
    if (threadIdx.x == 0) sErr = *P.errors;
    __syncthreads();
    if (sErr) return;
    const uint32_t n = min(P.count[l], P.cap[l]);
    const Items I = P.items[l];
    const Items C = P.items[l + 1];
    const uint32_t ctl = P.ctl, depth = P.levels - (l + 1);
    const uint32_t lane = threadIdx.x & 63u, wpb = blockDim.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    uint64_t nodeBytes = 0;
    // wavefront-uniform loop
    for (uint64_t slot = (uint64_t)G.block * wpb + (threadIdx.x >> 6); slot < (uint64_t)n * 8;
         slot += (uint64_t)G.nblocks * wpb) {
        const uint32_t i = (uint32_t)(slot >> 3), c = (uint32_t)(slot & 7u);
        uint64_t nb = 0;
        const SlotRefs r = slot_refs(P, I, i, c, depth, nb);
        if (lane == 0) nodeBytes += nb;
        const uint32_t es = I.estart[i], ec = I.ecount[i];
        uint32_t m = 0, lastCover = ~0u, coverKind = 0, coverColor = 0, coverBase = 0, coverQ = 0;
        for (uint32_t b = 0; b < ec; b += 64) {
            const uint32_t k = b + lane;
            bool t = false;
            uint32_t cv = 0, col = 0;
            if (k < ec) {
                const uint32_t ei = P.elist[l][es + k];
                const EditHead h = load_head(P.edits, ei);
                t = expand_touch(P, h, ei, l, r.x0, r.y0, r.z0, r.x1, r.y1, r.z1);
                if (t && l + 1 >= ctl) {
                    cv = e_cover(h, r.x0, r.y0, r.z0, r.x1, r.y1, r.z1);
                    col = h.single;
                }
            }
            const uint64_t T = __ballot(t), V = __ballot(cv != 0);
            if (V) {
                const uint32_t hb = 63u - (uint32_t)__builtin_clzll(V);
                lastCover = m + (uint32_t)__popcll(T & ((1ull << hb) - 1ull));
                coverKind = __shfl(cv, hb, 64);
                coverColor = __shfl(col, hb, 64);
                coverBase = b;
                coverQ = m;
            }
            m += (uint32_t)__popcll(T);
        }
        unsigned long long coveredVoxels = 0;
        uint32_t kind = 0;
        if (lane == 0)
            kind = slot_classify(P, I, l, (uint32_t)slot, depth, r, m, lastCover, coverKind, coverColor, coveredVoxels);
        kind = __shfl(kind, 0, 64);
        if (P.single_edit && lane == 0 && coveredVoxels) atomicAdd(P.stat_voxels, coveredVoxels);
        if (kind == KIND_ITEM) {
            const uint32_t rest = lastCover == ~0u ? m : m - lastCover - 1;
            uint32_t ci = 0, eo = 0;
            if (lane == 0) {
                ci = atomicAdd(&P.count[l + 1], 1u);
                eo = rest ? atomicAdd(&P.ecount[l + 1], rest) : 0u;
            }
            ci = __shfl(ci, 0, 64);
            eo = __shfl(eo, 0, 64);
            if (ci >= P.cap[l + 1] || eo + rest > P.ecap[l + 1]) {
                if (lane == 0) {
                    set_error(P.errors, ERR_SCRATCH);
                    I.ckind[slot] = (uint8_t)kind;
                }
                continue;
            }
            // the touching edits after the last cover, in list order
            uint32_t q = coverQ, w = 0;
            for (uint32_t b = coverBase; b < ec; b += 64) {
                const uint32_t k = b + lane;
                bool t = false;
                uint32_t ei = 0;
                if (k < ec) {
                    ei = P.elist[l][es + k];
                    t = expand_touch(P, load_head(P.edits, ei), ei, l, r.x0, r.y0, r.z0, r.x1, r.y1, r.z1);
                }
                const uint64_t T = __ballot(t);
                const bool emit = t && (lastCover == ~0u || q + (uint32_t)__popcll(T & below) > lastCover);
                const uint64_t E = __ballot(emit);
                if (emit) P.elist[l + 1][eo + w + (uint32_t)__popcll(E & below)] = ei;
                w += (uint32_t)__popcll(E);
                q += (uint32_t)__popcll(T);
            }
            if (lane == 0) slot_child_item(P, I, C, l, i, c, (uint32_t)slot, ci, eo, rest, r, lastCover, coverKind);
        }
        if (lane == 0) I.ckind[slot] = (uint8_t)kind;
    }
    if (P.abytes) wave_add64(P.abytes + 0, nodeBytes);
}
__global__ __launch_bounds__(256) void k_expand(EditPass P, uint32_t l) {
    d_expand(P, l, VGrid{blockIdx.x, gridDim.x});
}
__global__ __launch_bounds__(256) void k_expand_wide(EditPass P, uint32_t l) {
    d_expand_wide(P, l, VGrid{blockIdx.x, gridDim.x});
}

// ---------------------------------------------------------------------------------------------
// leaf level: one thread per leaf item, its 64 voxels as bit masks (bit = child1 * 8 + child2).
// Thread-per-item keeps many items' loads in flight; the per-voxel editor tests are cheap.
// ---------------------------------------------------------------------------------------------
// Voxel bit of a leaf for the offsets (xi, yi, zi) in 0..3: child1 * 8 + child2, child1 from
// the offsets' high bits and child2 from their low bits (x, y, z from bit 2 down).
__device__ __forceinline__ constexpr uint32_t leaf_voxel_bit(uint32_t xi, uint32_t yi, uint32_t zi) {
    return ((((xi >> 1) << 2) | ((yi >> 1) << 1) | (zi >> 1)) << 3) | (((xi & 1u) << 2) | ((yi & 1u) << 1) | (zi & 1u));
}
// The voxels of the leaf at (bx, by, bz) * 4 that e_touch accepts. Both of its tests split by
// axis: the box test is an overlap test per axis, and the sphere test subtracts a term per axis
// from r² in x, y, z order, each term depending on that axis's coordinate only (a term that is
// not subtracted is 0, and dist - 0 is dist). So the x terms are taken once for the 4 x
// offsets, then the y terms for the 16 (x, y) pairs, then the z terms for the 64 voxels: the
// same FP32 operations per voxel in the same order as e_touch.
template <class E>
__device__ __forceinline__ uint64_t leaf_touch_mask(const E& e, uint32_t bx, uint32_t by, uint32_t bz) {
    const uint32_t base[3] = {bx, by, bz};
    if (e.kind == EDIT_BOX_ADD || e.kind == EDIT_BOX_CARVE || e.kind == EDIT_COPY || e.kind == EDIT_FILL ||
        e.kind == EDIT_NOISE) {
        uint32_t in[3];  // per axis: bit i = offset i overlaps
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            in[a] = 0;
#pragma unroll
            for (uint32_t i = 0; i < 4; ++i) {
                const float v0 = (float)(base[a] | i), v1 = v0 + 1.0f;
                const bool out = e.bmin[a] >= v1 || e.bmax[a] <= v0 || v0 >= e.bmax[a] || v1 <= e.bmin[a];
                in[a] |= out ? 0u : (1u << i);
            }
        }
        uint64_t tm = 0;
#pragma unroll
        for (uint32_t xi = 0; xi < 4; ++xi)
#pragma unroll
            for (uint32_t yi = 0; yi < 4; ++yi)
#pragma unroll
                for (uint32_t zi = 0; zi < 4; ++zi)
                    if (((in[0] >> xi) & (in[1] >> yi) & (in[2] >> zi)) & 1u) tm |= 1ull << leaf_voxel_bit(xi, yi, zi);
        return tm;
    }
    float term[3][4];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            const float v0 = (float)(base[a] | i), v1 = v0 + 1.0f;
            term[a][i] = e.c[a] < v0 ? sqf(e.c[a] - v0) : (e.c[a] > v1 ? sqf(e.c[a] - v1) : 0.0f);
        }
    uint64_t tm = 0;
#pragma unroll
    for (uint32_t xi = 0; xi < 4; ++xi) {
        const float dx = e.r2 - term[0][xi];
#pragma unroll
        for (uint32_t yi = 0; yi < 4; ++yi) {
            const float dxy = dx - term[1][yi];
#pragma unroll
            for (uint32_t zi = 0; zi < 4; ++zi)
                if (dxy - term[2][zi] > 0) tm |= 1ull << leaf_voxel_bit(xi, yi, zi);
        }
    }
    return tm;
}

// kSpecial: the pass may hold copy / fill editors (their code is built in only then)
template <bool kSpecial>
__device__ __forceinline__ void d_leaf(const EditPass& P, VGrid G) {
    const uint32_t L = P.levels - 2;
    const uint32_t n = min(P.count[L], P.cap[L]);
    if (*P.errors) return;
    const Items I = P.items[L];
    uint64_t leafBytes = 0;  // old and original leaves read (measurement)
    for (uint32_t t0 = G.block * blockDim.x + (threadIdx.x & ~63u); t0 < n; t0 += G.nblocks * blockDim.x) {
        const uint32_t i = t0 + (threadIdx.x & 63u);
        const bool active = i < n;
        uint64_t touched = 0;
        if (active) {
            const uint32_t bx = I.px[i] << 2, by = I.py[i] << 2, bz = I.pz[i] << 2;
            const uint32_t old = I.old_[i], orig = I.orig[i];
            leafBytes += (old != kInvalidNode ? 8u : 0u) + (orig != kInvalidNode && orig != old ? 8u : 0u);
            uint64_t v = old != kInvalidNode ? *reinterpret_cast<const uint64_t*>(P.vpool + old) : 0ull;
            const uint64_t origMask = orig != kInvalidNode ? *reinterpret_cast<const uint64_t*>(P.vpool + orig) : 0ull;
            const uint32_t es = I.estart[i], ec = I.ecount[i];
            uint32_t cx = ~0u, cy = ~0u, cz = ~0u;  // copies: last source leaf looked up
            uint64_t cachedLeaf = 0;
            for (uint32_t k = 0; k < ec; ++k) {
                const uint32_t ei = P.elist[L][es + k];
                const EditHead h = load_head(P.edits, ei);
                if (!kSpecial && (h.kind == EDIT_COPY || h.kind == EDIT_FILL || h.kind == EDIT_NOISE)) {
                    set_error(P.errors, ERR_KIND);
                    break;
                }
                const uint64_t tm = leaf_touch_mask(h, bx, by, bz);
                const DevEdit& e = P.edits[ei];
                if (kSpecial && h.kind == EDIT_COPY) {
                    // get_new_value: an empty touched voxel takes its source voxel's value
                    const uint32_t srcRoot = P.items[0].orig[0];
                    uint64_t cand = tm & ~v, add = 0;
                    while (cand) {
                        const uint32_t b = (uint32_t)__builtin_ctzll(cand);
                        cand &= cand - 1;
                        const uint32_t c1 = b >> 3, c2 = b & 7u;
                        const float fx = (float)(bx | (((c1 >> 2) & 1u) << 1) | ((c2 >> 2) & 1u));
                        const float fy = (float)(by | (((c1 >> 1) & 1u) << 1) | ((c2 >> 1) & 1u));
                        const float fz = (float)(bz | ((c1 & 1u) << 1) | (c2 & 1u));
                        uint32_t sx, sy, sz;
                        if (!copy_source_ok(e, fx, fy, fz, sx, sy, sz)) continue;
                        if ((sx >> 2) != cx || (sy >> 2) != cy || (sz >> 2) != cz) {
                            cx = sx >> 2;
                            cy = sy >> 2;
                            cz = sz >> 2;
                            cachedLeaf = source_leaf(P.vpool, srcRoot, P.levels, cx, cy, cz);
                        }
                        if ((cachedLeaf >> leaf_bit(sx, sy, sz)) & 1ull) add |= 1ull << b;
                    }
                    v |= add;
                    continue;  // a copy's "num voxels" is its box volume, counted on the host
                }
                if (kSpecial && (h.kind == EDIT_FILL || h.kind == EDIT_NOISE)) {
                    // BufferEditor::get_new_value: set where the buffer voxel is set
                    uint64_t cand = tm & ~v, add = 0;
                    while (cand) {
                        const uint32_t b = (uint32_t)__builtin_ctzll(cand);
                        cand &= cand - 1;
                        const uint32_t c1 = b >> 3, c2 = b & 7u;
                        const float fx = (float)(bx | (((c1 >> 2) & 1u) << 1) | ((c2 >> 2) & 1u));
                        const float fy = (float)(by | (((c1 >> 1) & 1u) << 1) | ((c2 >> 1) & 1u));
                        const float fz = (float)(bz | ((c1 & 1u) << 1) | (c2 & 1u));
                        uint64_t idx;
                        if (fill_state_at(e, fx, fy, fz, idx) & 1u) add |= 1ull << b;
                    }
                    v |= add;
                    touched |= tm;
                    continue;
                }
                if (h.kind == EDIT_SPHERE_ADD || h.kind == EDIT_BOX_ADD) v |= tm;
                else if (h.kind != EDIT_SPHERE_PAINT) v &= ~tm;
                touched |= tm;
            }
            I.ncount[i] = __popcll(v);
            if (v == origMask) {
                I.nidx[i] = orig;
                I.state[i] = 0;
            } else if (v == 0) {
                I.nidx[i] = kInvalidNode;
                I.state[i] = 0;
            } else {
                I.clen[i] = 2;
                I.cont[i * 9 + 0] = (uint32_t)v;
                I.cont[i * 9 + 1] = (uint32_t)(v >> 32);
                I.hash[i] = hash_leaf(v);
                I.state[i] = 1;
            }
        }
        if (P.single_edit) wave_add64(P.stat_voxels, (uint64_t)__popcll(touched));
    }
    if (P.abytes) wave_add64(P.abytes + 0, leafBytes);
}
template <bool kSpecial>
__global__ __launch_bounds__(256) void k_leaf(EditPass P) {
    d_leaf<kSpecial>(P, VGrid{blockIdx.x, gridDim.x});
}

// ---------------------------------------------------------------------------------------------
// build interior items of level l from their children's results
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void d_build(const EditPass& P, uint32_t l, VGrid G) {
    const uint32_t n = min(P.count[l], P.cap[l]);
    if (*P.errors) return;
    const Items I = P.items[l];
    const Items C = P.items[l + 1];
    const uint32_t D = P.levels, ctl = P.ctl;
    for (uint32_t i = G.block * blockDim.x + threadIdx.x; i < n; i += G.nblocks * blockDim.x) {
        uint32_t node[9];
        uint32_t nn = 1, mask = 0;
        uint64_t count = 0;
        bool changed = false;
        for (uint32_t c = 0; c < 8; ++c) {
            const uint32_t slot = i * 8 + c;
            const uint32_t kind = I.ckind[slot];
            uint32_t idx, cnt;
            if (kind == KIND_ITEM) {
                const uint32_t ci = I.cval[slot];
                idx = C.nidx[ci];
                cnt = C.ncount[ci];
                I.cnew[slot] = cnt;
            } else {
                idx = I.cval[slot];
                cnt = I.cnew[slot];
            }
            changed |= idx != I.corig[slot];
            if (idx != kInvalidNode) {
                mask |= 1u << c;
                node[nn++] = idx;
                count += cnt;
            }
        }
        (void)D;
        if (l >= ctl && count >= (1u << 24)) set_error(P.errors, ERR_HEADER);
        I.ncount[i] = (uint32_t)count;
        if (!changed) {
            I.nidx[i] = I.orig[i];
            I.state[i] = 0;
            continue;
        }
        if (!mask) {
            I.nidx[i] = kInvalidNode;
            I.state[i] = 0;
            continue;
        }
        node[0] = mask | (l >= ctl ? (uint32_t)(count << 8) : 0u);
        for (uint32_t k = 0; k < nn; ++k) I.cont[i * 9 + k] = node[k];
        I.clen[i] = nn;
        I.hash[i] = hash_words(node, nn);
        I.state[i] = 1;
    }
}
__global__ __launch_bounds__(256) void k_build(EditPass P, uint32_t l) {
    d_build(P, l, VGrid{blockIdx.x, gridDim.x});
}

// ---------------------------------------------------------------------------------------------
// dedup / find / add / resolve for the pending items of one level
// ---------------------------------------------------------------------------------------------
__global__ void k_fill(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

// Slots of the dedup hash set a level uses: the power of two at or above twice its item
// count (load factor at most 1/2), at most the whole set.
__device__ __forceinline__ uint32_t hset_slots(const EditPass& P, uint32_t l) {
    const uint32_t n = min(P.count[l], P.cap[l]);
    uint32_t m = 64;
    while (m < P.hcap && m < 2u * n) m <<= 1;
    return min(m, P.hcap);
}

// clears the slots of the dedup hash set that level l will use
__device__ __forceinline__ void d_hset_clear(const EditPass& P, uint32_t l, VGrid G) {
    const uint32_t m = hset_slots(P, l);
    for (uint32_t i = G.block * blockDim.x + threadIdx.x; i < m; i += G.nblocks * blockDim.x) P.hset[i] = 0u;
}
__global__ void k_hset_clear(EditPass P, uint32_t l) {
    d_hset_clear(P, l, VGrid{blockIdx.x, gridDim.x});
}

__device__ __forceinline__ bool same_content(const Items& I, uint32_t a, uint32_t b) {
    const uint32_t n = I.clen[a];
    if (n != I.clen[b]) return false;
    for (uint32_t k = 0; k < n; ++k)
        if (I.cont[a * 9 + k] != I.cont[b * 9 + k]) return false;
    return true;
}

// Representatives (the first item of each distinct content) are listed for k_find, in the
// level's nstart words (free until the colour phase), with one atomic per block.
__device__ __forceinline__ void d_dedup(const EditPass& P, uint32_t l, VGrid G) {
    __shared__ uint32_t sAlloc[256 / 64 + 1];
    __shared__ uint32_t sErr;
    const uint32_t n = min(P.count[l], P.cap[l]);
    if (threadIdx.x == 0) sErr = *P.errors;
    __syncthreads();
    if (sErr) return;  // block-uniform: the loop below synchronises the block
    const Items I = P.items[l];
    const uint32_t hmask = hset_slots(P, l) - 1;
    for (uint32_t b0 = G.block * blockDim.x; b0 < n; b0 += G.nblocks * blockDim.x) {
        const uint32_t i = b0 + threadIdx.x;
        bool isRep = false;
        if (i < n && I.state[i] == 1) {
            uint32_t s = I.hash[i] & hmask;
            for (;;) {
                // read first: many items can share one content (and so one slot), and a CAS per
                // item on one address would serialise them all
                uint32_t cur = __atomic_load_n(&P.hset[s], __ATOMIC_RELAXED);
                if (cur == 0) {
                    cur = atomicCAS(&P.hset[s], 0u, i + 1);
                    if (cur == 0) {
                        I.rep[i] = i;
                        isRep = true;
                        break;
                    }
                }
                const uint32_t j = cur - 1;
                if (I.hash[j] == I.hash[i] && same_content(I, i, j)) {
                    I.rep[i] = j;
                    break;
                }
                s = (s + 1) & hmask;
            }
        }
        const uint32_t at = block_alloc32(sAlloc, &P.find_count[l], isRep ? 1u : 0u);
        if (isRep) I.nstart[at] = i;
    }
}
__global__ __launch_bounds__(256) void k_dedup(EditPass P, uint32_t l) {
    d_dedup(P, l, VGrid{blockIdx.x, gridDim.x});
}

// One wavefront per representative: scan the bucket for an identical node (any word
// position inside the bucket's used size whose words equal the node and that does not
// cross a page boundary holds the same node content).
__device__ __forceinline__ void d_find(const EditPass& P, uint32_t l, VGrid G) {
    const uint32_t n = min(P.count[l], P.cap[l]);
    if (*P.errors) return;
    const Items I = P.items[l];
    const Geom g = P.g;
    const uint32_t lane = threadIdx.x & 63u;
    const bool leafLevel = l == g.leaf_level();
    // the representatives listed by k_dedup
    const uint32_t nReps = min(P.find_count[l], n);
    uint64_t scanBytes = 0;  // bucket bytes this lane read (measurement)
    for (uint32_t r = wave_index(G); r < nReps; r += wave_count(G)) {
        const uint32_t i = I.nstart[r];
        const uint32_t bucket = g.bucket_from_hash(l, I.hash[i]);
        const uint32_t base = g.make_ptr(l, bucket, 0);
        const uint32_t bs = P.bucket_sizes[g.bucket_global(l, bucket)];
        const uint32_t nw = I.clen[i];
        uint32_t found = ~0u;
        if (leafLevel && P.leaf_tags) {
            // slots (2 words each) whose tag byte matches, lowest first, confirmed on the leaf;
            // 4 tags per lane, 256 slots per wavefront step
            const uint64_t key = (uint64_t)I.cont[i * 9] | ((uint64_t)I.cont[i * 9 + 1] << 32);
            const uint32_t want = I.hash[i] >> 24, nslots = bs / 2;
            const uint8_t* tags = P.leaf_tags + (uint64_t)bucket * (g.bucket_size(l) / 2);
            for (uint32_t s0 = 0; s0 < nslots && found == ~0u; s0 += 256) {
                const uint32_t sl = s0 + 4 * lane;
                uint32_t hitSlot = ~0u;
                if (sl < nslots) {
                    const uint32_t four = *reinterpret_cast<const uint32_t*>(tags + sl);
                    scanBytes += min(4u, nslots - sl);
                    for (uint32_t k = 0; k < 4 && hitSlot == ~0u; ++k)
                        if (sl + k < nslots && ((four >> (8 * k)) & 0xFFu) == want) {
                            scanBytes += 8;
                            if (*reinterpret_cast<const uint64_t*>(P.vpool + base + 2 * (sl + k)) == key) hitSlot = sl + k;
                        }
                }
                const uint64_t b = __ballot(hitSlot != ~0u);
                if (b) found = 2u * (uint32_t)__shfl((int)hitSlot, __ffsll((long long)b) - 1, 64);
            }
        } else if (leafLevel) {
            const uint64_t key = (uint64_t)I.cont[i * 9] | ((uint64_t)I.cont[i * 9 + 1] << 32);
            for (uint32_t p0 = 0; p0 < bs && found == ~0u; p0 += 128) {
                const uint32_t pos = p0 + 2 * lane;
                if (pos + 2 <= bs) scanBytes += 8;
                const bool hit = pos + 2 <= bs && *reinterpret_cast<const uint64_t*>(P.vpool + base + pos) == key;
                const uint64_t b = __ballot(hit);
                if (b) found = p0 + 2 * (uint32_t)__ffsll((long long)b) - 2;
            }
        } else {
            // The reference walks node by node from each page start and compares only nodes
            // with start + n < min(bucketSize, pageEnd). The wave compares all positions at
            // once, then confirms the first candidate lies on that walk.
            const uint32_t h0 = I.cont[i * 9];
            const uint32_t ps = g.page_size;
            for (uint32_t p0 = 0; p0 + nw < bs && found == ~0u; p0 += 64) {
                const uint32_t pos = p0 + lane;
                const uint32_t pageEnd = min(bs, pos - pos % ps + ps);
                if (pos < bs) scanBytes += 4;
                bool hit = pos + nw < pageEnd && P.vpool[base + pos] == h0;
                if (hit)
                    for (uint32_t k = 1; k < nw; ++k) {
                        scanBytes += 4;
                        if (P.vpool[base + pos + k] != I.cont[i * 9 + k]) {
                            hit = false;
                            break;
                        }
                    }
                uint64_t b = __ballot(hit);
                while (b) {
                    const uint32_t c = p0 + (uint32_t)__ffsll((long long)b) - 1;  // wave-uniform
                    uint32_t w = c - c % ps;
                    while (w < c) w += 1u + __popc(P.vpool[base + w] & 0xFFu);  // advances >= 1 word
                    if (w == c) {
                        found = c;
                        break;
                    }
                    b &= b - 1;
                }
            }
        }
        if (lane == 0) {
            if (found != ~0u) {
                I.nidx[i] = base + found;
                I.state[i] = 2;  // resolved by find
            }
        }
    }
    if (P.abytes) wave_add64(P.abytes + 1, scanBytes);
}
__global__ __launch_bounds__(256) void k_find(EditPass P, uint32_t l) {
    d_find(P, l, VGrid{blockIdx.x, gridDim.x});
}

__device__ bool alloc_page(const EditPass& P, uint32_t page) {
    if (atomicCAS(&P.page_table[page], 0u, 0xFFFFFFFFu) != 0u) return true;  // already mapped (after GC)
    const uint32_t phys = atomicAdd(P.pool_top, 1u);
    if (phys >= P.pool_pages) {
        set_error(P.errors, ERR_POOL);
        atomicExch(&P.page_table[page], 0u);
        return false;
    }
    atomicExch(&P.page_table[page], phys);
    return true;
}

__device__ __forceinline__ void d_add(const EditPass& P, uint32_t l, VGrid G) {
    const uint32_t n = min(P.count[l], P.cap[l]);
    if (*P.errors) return;
    const Items I = P.items[l];
    const Geom g = P.g;
    const uint32_t ps = g.page_size;
    for (uint32_t t0 = G.block * blockDim.x + (threadIdx.x & ~63u); t0 < n; t0 += G.nblocks * blockDim.x) {
        const uint32_t i = t0 + (threadIdx.x & 63u);
        uint32_t added = 0;
        if (i < n && I.state[i] == 1 && I.rep[i] == i) {
            const uint32_t nw = I.clen[i];
            const uint32_t bucket = g.bucket_from_hash(l, I.hash[i]);
            uint32_t* bsPtr = &P.bucket_sizes[g.bucket_global(l, bucket)];
            const uint32_t cap = g.bucket_size(l);
            uint32_t start;
            bool ok = true;
            uint32_t cur = *reinterpret_cast<volatile uint32_t*>(bsPtr);
            for (;;) {
                const uint32_t left = ps - cur % ps;
                start = (left == ps || left < nw) ? (left == ps ? cur : cur + left) : cur;
                if (start + nw > cap) {
                    ok = false;
                    break;
                }
                const uint32_t prev = atomicCAS(bsPtr, cur, start + nw);
                if (prev == cur) break;
                cur = prev;
            }
            if (!ok) {
                set_error(P.errors, ERR_BUCKET);
            } else {
                const uint32_t va = g.make_ptr(l, bucket, start);
                if (start % ps != 0 || alloc_page(P, va / ps)) {
                    for (uint32_t k = 0; k < nw; ++k) P.vpool[va + k] = I.cont[i * 9 + k];
                    if (P.leaf_tags && l == g.leaf_level())
                        P.leaf_tags[(uint64_t)bucket * (cap / 2) + start / 2] = (uint8_t)(I.hash[i] >> 24);
                    I.nidx[i] = va;
                    added = nw;
                }
            }
        }
        wave_add64(P.stat_new_words, added);
    }
}
__global__ __launch_bounds__(256) void k_add(EditPass P, uint32_t l) {
    d_add(P, l, VGrid{blockIdx.x, gridDim.x});
}

__device__ __forceinline__ void d_resolve(const EditPass& P, uint32_t l, VGrid G) {
    const uint32_t n = min(P.count[l], P.cap[l]);
    if (*P.errors) return;
    const Items I = P.items[l];
    for (uint32_t i = G.block * blockDim.x + threadIdx.x; i < n; i += G.nblocks * blockDim.x)
        if (I.state[i] == 1 && I.rep[i] != i) I.nidx[i] = I.nidx[I.rep[i]];
}
__global__ __launch_bounds__(256) void k_resolve(EditPass P, uint32_t l) {
    d_resolve(P, l, VGrid{blockIdx.x, gridDim.x});
}

// ---------------------------------------------------------------------------------------------
// colours
// ---------------------------------------------------------------------------------------------
// new / old colour starts of the child items (relative to the task's colour array): one lane per
// child slot, the 8 slots of an item in 8 consecutive lanes (their starts by a scan over the 8)
__device__ __forceinline__ void d_color_starts(const EditPass& P, uint32_t l, VGrid G) {
    const uint32_t n = min(P.count[l], P.cap[l]);
    if (*P.errors) return;
    const Items I = P.items[l];
    const Items C = P.items[l + 1];
    const uint64_t total = (uint64_t)n * 8;
    // the loop bound is a multiple of 8 per step, so the 8 lanes of an item step together
    for (uint64_t t = (uint64_t)G.block * blockDim.x + threadIdx.x; t < ((total + 63) & ~63ull);
         t += (uint64_t)G.nblocks * blockDim.x) {
        const bool active = t < total;
        const uint32_t slot = (uint32_t)t, i = slot >> 3, c = slot & 7u;
        const uint32_t cn = active ? I.cnew[slot] : 0u, co = active ? I.cold[slot] : 0u;
        uint32_t sn = cn, so = co;  // inclusive scans over the item's 8 slots
#pragma unroll
        for (uint32_t d = 1; d < 8; d <<= 1) {
            const uint32_t xn = __shfl_up(sn, d, 8), xo = __shfl_up(so, d, 8);
            if (c >= d) {
                sn += xn;
                so += xo;
            }
        }
        if (!active) continue;
        uint32_t ns = 0, os = 0;
        if (l == P.ctl) {
            if (c == 0) {
                I.nstart[i] = 0;
                I.ostart[i] = 0;
            }
        } else {
            ns = I.nstart[i];
            os = I.ostart[i];
        }
        ns += sn - cn;
        os += so - co;
        if (I.ckind[slot] == KIND_ITEM) {
            const uint32_t ci = I.cval[slot];
            C.nstart[ci] = ns;
            C.ostart[ci] = os;
        }
        I.cstart[slot] = ns;
        I.costart[slot] = os;
    }
}
__global__ __launch_bounds__(256) void k_color_starts(EditPass P, uint32_t l) {
    d_color_starts(P, l, VGrid{blockIdx.x, gridDim.x});
}

// colour array of every task with colours: base offset in the colour scratch. Arrays start on
// a 64-colour chunk; the chunk markers of the wave's range are cleared here (the colour
// writers mark the chunks they fill with one colour).
__device__ __forceinline__ void d_color_alloc(const EditPass& P, VGrid G) {
    const uint32_t ctl = P.ctl;
    const uint32_t n = min(P.count[ctl], P.cap[ctl]);
    if (*P.errors) return;
    const Items T = P.items[ctl];
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t t0 = G.block * blockDim.x + (threadIdx.x & ~63u); t0 < n; t0 += G.nblocks * blockDim.x) {
        const uint32_t i = t0 + lane;
        const uint32_t cnt = i < n ? T.ncount[i] : 0u;
        const uint64_t acnt = ((uint64_t)cnt + 63u) & ~63ull;
        const uint64_t base = wave_alloc64(P.color_top, acnt);
        wave_add64(P.stat_colors, cnt);
        if (i < n) T.cbase[i] = base;
        // the wave's arrays are contiguous: [base of lane 0, end of lane 63)
        const uint64_t lo = shfl64(base, 0), hi = shfl64(base + acnt, 63);
        if (hi > P.color_cap) {
            if (lane == 0) set_error(P.errors, ERR_COLOR_SCRATCH);
            continue;
        }
        for (uint64_t k = lo / 64 + lane; k < hi / 64; k += 64) P.cmarks[k] = 0;
    }
}
__global__ __launch_bounds__(256) void k_color_alloc(EditPass P) {
    d_color_alloc(P, VGrid{blockIdx.x, gridDim.x});
}

__device__ __forceinline__ LeafDesc task_leaf(const EditPass& P, uint32_t handle) {
    LeafDesc d{};
    if (handle == 0) return d;  // no leaf: n_blocks = 0 (invalid)
    if (handle & 0x80000000u) return P.leaves[handle & 0x7FFFFFFFu];
    d = P.main;
    d.offset = P.offsets[handle];
    return d;
}

// DAGUtils::get_color (dag_utils.h:423-523) of voxel (x, y, z) in the DAG at `root`: the colour
// at the voxel's rank in its colour subtree, CompressedColor{} (0) when absent. The colour
// slots hold the pass's starting colours until k_encode writes the new leaves.
__device__ uint64_t source_color(const EditPass& P, uint32_t root, uint32_t x, uint32_t y, uint32_t z) {
    const uint32_t levels = P.levels, L = levels - 2, ctl = P.ctl;
    const uint32_t* __restrict__ vpool = P.vpool;
    uint64_t nof = 0;
    uint32_t colorNode = 0, handle = 0;
    bool haveLeaf = false;
    uint32_t va = root;
    for (uint32_t level = 1; level <= L; ++level) {
        const uint32_t m = vpool[va] & 0xFFu;
        const uint32_t s = levels - level;
        const uint32_t child = (((x >> s) & 1u) << 2) | (((y >> s) & 1u) << 1) | ((z >> s) & 1u);
        if (!(m & (1u << child))) return 0;
        if (level - 1 < ctl) {
            colorNode = P.cnodes[colorNode + child];
            if (level == ctl) {
                handle = colorNode;
                haveLeaf = true;
            } else if (!colorNode) {
                return 0;
            }
        }
        if (level == L) {
            for (uint32_t c = 0; c < child; ++c)
                if (m & (1u << c)) nof += __popcll(*reinterpret_cast<const uint64_t*>(vpool + vpool[va + child_off(m, c)]));
            const uint64_t lf = *reinterpret_cast<const uint64_t*>(vpool + vpool[va + child_off(m, child)]);
            nof += __popcll(lf & ((1ull << leaf_bit(x, y, z)) - 1ull));
            break;
        }
        if (level > ctl)
            for (uint32_t c = 0; c < child; ++c)
                if (m & (1u << c)) nof += vpool[vpool[va + child_off(m, c)]] >> 8;
        va = vpool[va + child_off(m, child)];
    }
    if (!haveLeaf) return 0;
    const LeafDesc d = task_leaf(P, handle);
    if (!leaf_valid_index(d, nof)) return 0;
    return leaf_decode(P.arena, d, nof);
}

__device__ __forceinline__ uint64_t old_color(const EditPass& P, uint32_t mode, uint32_t modeval, const LeafDesc& d,
                                              uint64_t index) {
    if (mode == 1) return pack_color(modeval, 0, 0);
    if (leaf_valid_index(d, index)) return leaf_decode(P.arena, d, index);
    return 0;  // CompressedColor{}
}

// A chunk of 64 staged colours that all share one colour with no weights is stored as a marker
// (kUniformChunk | colour bits) instead of the colours. Block starts compare the colour bits
// and the bits per weight (kColorKey) of neighbouring colours.
constexpr uint64_t kUniformChunk = 1ull << 63;
constexpr uint64_t kColorKey = 0x0000FF00FFFFFFFFull;

// keep / full slots of interior items at levels >= ctl: one wavefront per slot. Whole chunks
// of one colour become markers; the partial chunks at either end are written as colours.
__device__ __forceinline__ void d_color_copy(const EditPass& P, uint32_t l, VGrid G) {
    const uint32_t n = min(P.count[l], P.cap[l]);
    if (*P.errors) return;
    const Items I = P.items[l];
    const Items T = P.items[P.ctl];
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t w = wave_index(G); w < n * 8; w += wave_count(G)) {
        const uint32_t i = w >> 3, slot = w;
        const uint32_t kind = I.ckind[slot];
        if (kind != KIND_KEEP && kind != KIND_FULL) continue;
        const uint32_t cnt = I.cnew[slot];
        if (cnt == 0) continue;
        const uint32_t task = I.aux[i];
        const uint64_t cb = T.cbase[task];  // a multiple of 64
        uint64_t* out = P.colors + cb;
        uint64_t* marks = P.cmarks + cb / 64;
        const uint32_t cs = I.cstart[slot], ce = cs + cnt;
        if (kind == KIND_FULL) {
            const uint64_t col = pack_color(I.ccolor[slot], 0, 0);
            const uint32_t a = min((cs + 63u) & ~63u, ce), b = max(ce & ~63u, a);
            for (uint32_t k = cs + lane; k < a; k += 64) out[k] = col;
            for (uint32_t k = a / 64 + lane; k < b / 64; k += 64) marks[k] = kUniformChunk | col;
            for (uint32_t k = b + lane; k < ce; k += 64) out[k] = col;
        } else {
            const LeafDesc d = task_leaf(P, T.handle[task]);
            const uint32_t mode = I.cmode[i], mv = I.cmodeval[i];
            const uint64_t os = (uint64_t)I.costart[slot] - cs;
            uint32_t pos = ~0u;  // block of the colour before the chunk (decoded runs)
            for (uint32_t k0 = cs & ~63u; k0 < ce; k0 += 64) {
                const uint32_t k = k0 + lane;
                const bool in = k >= cs && k < ce;
                uint64_t col;
                if (mode == 1) {
                    col = in ? pack_color(mv, 0, 0) : 0ull;
                } else {  // the chunk's old colours decoded as one run
                    const uint32_t b0 = k0 < cs ? cs - k0 : 0u, e0 = min(ce - k0, 64u);
                    col = leaf_decode_run(P.arena, d, os + k0 + b0, e0 - b0, (int32_t)lane - (int32_t)b0, in, pos);
                }
                if (k0 >= cs && k0 + 64 <= ce) {
                    const uint64_t c0 = shfl64(col, 0);
                    if (__ballot(((col ^ c0) & kColorKey) != 0 || (col >> 40) != 0) == 0) {
                        if (lane == 0) marks[k0 / 64] = kUniformChunk | (uint32_t)c0;
                        continue;
                    }
                }
                if (in) out[k] = col;
            }
        }
    }
}
__global__ __launch_bounds__(256) void k_color_copy(EditPass P, uint32_t l) {
    d_color_copy(P, l, VGrid{blockIdx.x, gridDim.x});
}

// keep / full slots of the items at level L - 1 (their children are leaves: at most 64 colours
// a slot): one wavefront per item, the colours of all its keep / full slots laid out over the
// lanes (lane j of round r takes colour r * 64 + j of the item's slots in slot order).
__device__ __forceinline__ void d_color_copy_leaves(const EditPass& P, VGrid G) {
    const uint32_t l = P.levels - 3;
    const uint32_t n = min(P.count[l], P.cap[l]);
    if (*P.errors) return;
    const Items I = P.items[l];
    const Items T = P.items[P.ctl];
    const uint32_t lane = threadIdx.x & 63u, s = lane & 7u;
    for (uint32_t i = wave_index(G); i < n; i += wave_count(G)) {
        const uint32_t slot = i * 8 + s;
        const uint32_t kind = I.ckind[slot];
        const uint32_t cnt = (kind == KIND_KEEP || kind == KIND_FULL) ? I.cnew[slot] : 0u;
        // exclusive prefix of the slot counts over lanes 0..7 (lanes 8.. repeat them)
        uint32_t off = cnt;
#pragma unroll
        for (uint32_t d = 1; d < 8; d <<= 1) {
            const uint32_t t = __shfl_up(off, d, 8);
            if (s >= d) off += t;
        }
        const uint32_t total = __shfl(off, 7, 64);
        off -= cnt;
        if (total == 0) continue;
        const uint32_t cs = I.cstart[slot], cos = I.costart[slot], ccol = I.ccolor[slot];
        const uint32_t task = I.aux[i];
        const uint64_t cb = T.cbase[task];
        const LeafDesc d = task_leaf(P, T.handle[task]);
        const uint32_t mode = I.cmode[i], mv = I.cmodeval[i];
        for (uint32_t r = 0; r < total; r += 64) {
            const uint32_t idx = r + lane;
            // the slot holding colour idx: the last one starting at or before it
            uint32_t j = 0;
#pragma unroll
            for (uint32_t step = 4; step > 0; step >>= 1)
                if (__shfl(off, j + step, 64) <= idx) j += step;
            const uint32_t k = idx - __shfl(off, j, 64);
            const uint32_t kj = __shfl(kind, j, 64), csj = __shfl(cs, j, 64), cosj = __shfl(cos, j, 64);
            const uint32_t ccj = __shfl(ccol, j, 64);
            if (idx >= total) continue;
            uint64_t col;
            if (kj == KIND_FULL) col = pack_color(ccj, 0, 0);
            else if (mode == 1) col = pack_color(mv, 0, 0);
            else col = leaf_valid_index(d, (uint64_t)cosj + k) ? leaf_decode(P.arena, d, (uint64_t)cosj + k) : 0ull;
            P.colors[cb + csj + k] = col;
        }
    }
}
__global__ __launch_bounds__(256) void k_color_copy_leaves(EditPass P) {
    d_color_copy_leaves(P, VGrid{blockIdx.x, gridDim.x});
}

// leaf_decode of ascending colour indices of one leaf. An index in the macro block of the
// previous one walks forward from the previous block (blocks start at ascending indices, so
// the last block starting at or before the index is reached by that walk); another macro block
// starts a binary search for that last block, which is the block leaf_decode's search finds.
struct LeafCursor {
    const uint64_t* arena;
    LeafDesc d;
    uint64_t m = ~0ull, macroBit = 0, blk = 0;
    uint32_t pos = 0, hiM = 0;
    __device__ uint64_t get(uint64_t index) {
        if (!leaf_valid_index(d, index)) return 0;  // CompressedColor{}
        const uint64_t G = index + (d.offset != ~0ull ? d.offset : 0ull);
        const uint64_t* blocks = arena + d.blocks;
        const uint64_t* macro = arena + d.macro;
        const uint64_t mm = G >> 14;
        const uint32_t local = (uint32_t)(G & 16383u);
        if (mm != m) {
            m = mm;
            macroBit = macro[2 * mm + 1];
            uint32_t lo = (uint32_t)macro[2 * mm];
            hiM = 2 * (mm + 1) < d.n_macro ? (uint32_t)macro[2 * (mm + 1)] - 1u : (uint32_t)d.n_blocks - 1u;
            uint32_t hi = hiM;  // blocks[lo] starts at index 0 of the macro block
            while (lo < hi) {
                const uint32_t mid = lo + (hi - lo + 1) / 2;
                if (hdr_index((uint32_t)blocks[mid]) <= local) lo = mid;
                else hi = mid - 1;
            }
            pos = lo;
            blk = blocks[pos];
        } else {
            while (pos < hiM) {
                const uint64_t nb = blocks[pos + 1];
                if (hdr_index((uint32_t)nb) > local) break;
                ++pos;
                blk = nb;
            }
        }
        const uint32_t h = (uint32_t)blk, bits = (uint32_t)(blk >> 32), bpw = hdr_bpw(h);
        if (bpw == 0) return pack_color(bits, 0, 0);
        const uint64_t bit = macroBit + (h >> 16) + (uint64_t)((local - hdr_index(h)) * bpw);
        const uint32_t* w = reinterpret_cast<const uint32_t*>(arena) + d.weights32;
        const uint64_t byte = bit >> 3;
        const uint32_t w0 = w[byte >> 2], sh = (uint32_t)(byte & 3u) * 8u;
        const uint32_t b0 = (w0 >> sh) & 0xFFu;
        const uint32_t b1 = (byte & 3u) == 3u ? (w[(byte >> 2) + 1] & 0xFFu) : ((w0 >> (sh + 8)) & 0xFFu);
        const uint32_t dst = (b0 << 8) | b1;
        const uint32_t wq = (dst >> (16u - bpw - (uint32_t)(bit & 7u))) & ((1u << bpw) - 1u);
        return pack_color(bits, wq, bpw);
    }
};

// Leaf items of a pass without copy / fill editors, one thread per item. The voxels set at the
// end are those of the leaf node the item resolved to (k_leaf's masks). Without a paint editor
// in the list, a voxel that is set at the end took the colour of the last add editor that
// touched it (an add sets the voxel and its colour, a carve clears it, so the last editor
// touching a voxel that ends set is an add), else it kept its old colour: the editors are
// walked backwards, each add claiming the set voxels it touches that no later add claimed, and
// the unclaimed set voxels read their old colours in ascending order through a LeafCursor.
// With a paint editor (whose colour depends on the colour before it) every set voxel follows
// the edit list on its own, as a lane of k_color_leaf does.
__device__ __forceinline__ void d_color_leaf_items(const EditPass& P, VGrid G) {
    const uint32_t L = P.levels - 2;
    const uint32_t n = min(P.count[L], P.cap[L]);
    if (*P.errors) return;
    const Items I = P.items[L];
    const Items T = P.items[P.ctl];
    for (uint32_t i = G.block * blockDim.x + threadIdx.x; i < n; i += G.nblocks * blockDim.x) {
        const uint32_t es = I.estart[i], ec = I.ecount[i];
        const uint32_t bx = I.px[i] << 2, by = I.py[i] << 2, bz = I.pz[i] << 2;
        const uint32_t old = I.old_[i];
        const uint64_t oldMask = old != kInvalidNode ? *reinterpret_cast<const uint64_t*>(P.vpool + old) : 0ull;
        // the voxels k_leaf set: the leaf node the item resolved to
        const uint32_t leaf = I.nidx[i];
        const uint64_t v = leaf != kInvalidNode ? *reinterpret_cast<const uint64_t*>(P.vpool + leaf) : 0ull;
        bool paint = false;
        for (uint32_t k = 0; k < ec; ++k) {
            const uint32_t kind = P.edits[P.elist[L][es + k]].kind;
            if (kind == EDIT_COPY || kind == EDIT_FILL || kind == EDIT_NOISE) set_error(P.errors, ERR_KIND);
            paint |= kind == EDIT_SPHERE_PAINT;
        }
        if (!v) continue;
        const uint32_t task = I.aux[i];
        uint64_t* out = P.colors + T.cbase[task] + I.nstart[i];
        if (paint) {
            const uint32_t ostart = I.ostart[i], cmode = I.cmode[i], cmodeval = I.cmodeval[i];
            LeafCursor cur{P.arena, task_leaf(P, T.handle[task])};
            uint64_t rest = v;
            while (rest) {
                const uint32_t b = (uint32_t)__builtin_ctzll(rest);
                rest &= rest - 1;
                const uint32_t c1 = b >> 3, c2 = b & 7u;
                const float fx = (float)(bx | (((c1 >> 2) & 1u) << 1) | ((c2 >> 2) & 1u));
                const float fy = (float)(by | (((c1 >> 1) & 1u) << 1) | ((c2 >> 1) & 1u));
                const float fz = (float)(bz | ((c1 & 1u) << 1) | (c2 & 1u));
                const bool prev = (oldMask >> b) & 1ull;
                // oldColor() of a voxel that was empty is CompressedColor{} (hash_dag_edits.h:520-540);
                // read at most once per voxel, in ascending voxel order
                auto old_of_voxel = [&]() -> uint64_t {
                    if (!prev) return 0ull;
                    if (cmode == 1) return pack_color(cmodeval, 0, 0);
                    return cur.get(ostart + __popcll(oldMask & ((1ull << b) - 1ull)));
                };
                bool vb = prev, have = false;
                uint64_t col = 0;
                for (uint32_t k = 0; k < ec; ++k) {
                    const EditHead h = load_head(P.edits, P.elist[L][es + k]);
                    if (!e_touch(h, fx, fy, fz, fx + 1.0f, fy + 1.0f, fz + 1.0f)) continue;
                    const bool nv = e_value(h, vb);
                    if (nv) {
                        uint64_t pc = 0;
                        if (h.kind == EDIT_SPHERE_PAINT && vb) {
                            if (!have) col = old_of_voxel();
                            pc = col;
                        }
                        col = e_color(h, fx, fy, fz, pc);
                        have = true;
                    }
                    vb = nv;
                }
                if (!have) col = old_of_voxel();
                out[__popcll(v & ((1ull << b) - 1ull))] = col;
            }
            continue;
        }
        uint64_t claimed = 0;
        for (uint32_t k = ec; k-- > 0 && (v & ~claimed);) {
            const EditHead h = load_head(P.edits, P.elist[L][es + k]);
            if (h.kind != EDIT_SPHERE_ADD && h.kind != EDIT_BOX_ADD) continue;
            uint64_t fresh = leaf_touch_mask(h, bx, by, bz) & v & ~claimed;
            claimed |= fresh;
            const uint64_t col = pack_color(h.single, 0, 0);
            while (fresh) {
                const uint32_t b = (uint32_t)__builtin_ctzll(fresh);
                fresh &= fresh - 1;
                out[__popcll(v & ((1ull << b) - 1ull))] = col;
            }
        }
        uint64_t kept = v & ~claimed;  // set before the pass (an unclaimed set voxel was not touched by an add)
        if (!kept) continue;
        const uint32_t ostart = I.ostart[i];
        if (I.cmode[i] == 1) {
            const uint64_t col = pack_color(I.cmodeval[i], 0, 0);
            while (kept) {
                const uint32_t b = (uint32_t)__builtin_ctzll(kept);
                kept &= kept - 1;
                out[__popcll(v & ((1ull << b) - 1ull))] = col;
            }
            continue;
        }
        LeafCursor cur{P.arena, task_leaf(P, T.handle[task])};
        while (kept) {
            const uint32_t b = (uint32_t)__builtin_ctzll(kept);
            kept &= kept - 1;
            out[__popcll(v & ((1ull << b) - 1ull))] = cur.get(ostart + __popcll(oldMask & ((1ull << b) - 1ull)));
        }
    }
}
__global__ __launch_bounds__(256) void k_color_leaf_items(EditPass P) {
    d_color_leaf_items(P, VGrid{blockIdx.x, gridDim.x});
}

// A leaf item's fields that k_color_leaf reads first.
struct LeafMeta {
    uint32_t es, ec, px, py, pz, old, aux, cmode, cmodeval, ostart, nstart;
};
__device__ __forceinline__ LeafMeta load_meta(const Items& I, uint32_t i) {
    return LeafMeta{I.estart[i], I.ecount[i], I.px[i],    I.py[i],     I.pz[i],    I.old_[i],
                    I.aux[i],    I.cmode[i],  I.cmodeval[i], I.ostart[i], I.nstart[i]};
}
// Edit head of lane k of a wavefront (the leading words only: kind .. single).
__device__ __forceinline__ EditHead head_of_lane(const EditHead& mh, uint32_t k) {
    EditHead h;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&mh);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&h);
#pragma unroll
    for (int w = 0; w < 13; ++w) dst[w] = (uint32_t)__builtin_amdgcn_readlane((int)src[w], (int)k);
    return h;
}

// Leaf items of a pass with a copy or fill editor: one lane per voxel folds the voxel's value
// and colour through the edit list. A wavefront takes a run of consecutive items (neighbours
// share edits, so their edit heads stay cached), reads the next item's fields while it works
// on one, and loads an item's edit heads into the lanes at once (lane k: edit k), each edit
// then read out of its lane.
__device__ __forceinline__ void d_color_leaf(const EditPass& P, VGrid G) {
    const uint32_t L = P.levels - 2;
    const uint32_t n = min(P.count[L], P.cap[L]);
    if (*P.errors) return;
    const Items I = P.items[L];
    const Items T = P.items[P.ctl];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t run = max(1u, min(32u, n / (4u * wave_count(G))));
    for (uint32_t i0 = wave_index(G) * run; i0 < n; i0 += wave_count(G) * run) {
        const uint32_t iEnd = min(i0 + run, n);
        LeafMeta next = load_meta(I, i0);
        for (uint32_t i = i0; i < iEnd; ++i) {
            const LeafMeta m = next;
            if (i + 1 < iEnd) next = load_meta(I, i + 1);
            const uint32_t es = m.es, ec = m.ec;
            const uint32_t c1 = lane >> 3, c2 = lane & 7u;
            const uint32_t x = (m.px << 2) | (((c1 >> 2) & 1u) << 1) | ((c2 >> 2) & 1u);
            const uint32_t y = (m.py << 2) | (((c1 >> 1) & 1u) << 1) | ((c2 >> 1) & 1u);
            const uint32_t z = (m.pz << 2) | ((c1 & 1u) << 1) | (c2 & 1u);
            uint64_t oldMask = 0;
            if (m.old != kInvalidNode) oldMask = *reinterpret_cast<const uint64_t*>(P.vpool + m.old);
            const bool prev = (oldMask >> lane) & 1u;
            const uint32_t oldRank = __popcll(oldMask & ((1ull << lane) - 1ull));
            const uint32_t task = m.aux;
            const LeafDesc d = task_leaf(P, T.handle[task]);
            const uint32_t cmode = m.cmode, cmodeval = m.cmodeval;
            const uint64_t oldIndex = m.ostart + oldRank;
            // oldColor() of a voxel that was empty is CompressedColor{} (hash_dag_edits.h:520-540)
            auto voxel_old_color = [&]() -> uint64_t {
                if (!prev) return 0ull;
                if (cmode == 1) return pack_color(cmodeval, 0, 0);
                return leaf_valid_index(d, oldIndex) ? leaf_decode(P.arena, d, oldIndex) : 0ull;
            };
            bool v = prev;
            uint64_t col = 0;
            bool haveCol = false;
            const float fx = (float)x, fy = (float)y, fz = (float)z;
            auto apply = [&](const EditHead& h, uint32_t ei) {
                if (!e_touch(h, fx, fy, fz, fx + 1.0f, fy + 1.0f, fz + 1.0f)) return;
                if (h.kind == EDIT_FILL) {
                    // FillEditorColors::get_new_color: a voxel the flood fill set takes its colour
                    if (!v) {
                        const DevEdit& e = P.edits[ei];
                        uint64_t idx = 0;
                        const uint32_t st = fill_state_at(e, fx, fy, fz, idx);
                        if (st & 1u) {
                            if (st & 2u) {
                                col = fill_color_at(e, idx);
                                haveCol = true;
                            }
                            v = true;
                        }
                    }
                    return;
                }
                if (h.kind == EDIT_NOISE) {
                    // BufferEditor::get_new_value, SphereNoiseEditor::get_new_color: a voxel that
                    // becomes set takes the editor's colour
                    if (!v) {
                        uint64_t idx = 0;
                        if (fill_state_at(P.edits[ei], fx, fy, fz, idx) & 1u) {
                            col = noise_color(P.edits[ei], fx, fy, fz);
                            haveCol = true;
                            v = true;
                        }
                    }
                    return;
                }
                if (h.kind == EDIT_COPY) {
                    // get_new_value / get_new_color: an empty voxel takes its source's value and colour
                    if (!v) {
                        const DevEdit& e = P.edits[ei];
                        uint32_t sx, sy, sz;
                        if (!copy_source_ok(e, fx, fy, fz, sx, sy, sz)) return;
                        const uint32_t srcRoot = P.items[0].orig[0];
                        const uint64_t sl = source_leaf(P.vpool, srcRoot, P.levels, sx >> 2, sy >> 2, sz >> 2);
                        if ((sl >> leaf_bit(sx, sy, sz)) & 1ull) {
                            col = source_color(P, srcRoot, sx, sy, sz);
                            haveCol = true;
                            v = true;
                        }
                    }
                    return;
                }
                const bool nv = e_value(h, v);
                if (nv) {
                    uint64_t pc = 0;
                    if (h.kind == EDIT_SPHERE_PAINT && v) {
                        if (!haveCol) col = voxel_old_color();
                        pc = col;
                    }
                    col = e_color(h, fx, fy, fz, pc);
                    haveCol = true;
                }
                v = nv;
            };
            for (uint32_t k0 = 0; k0 < ec; k0 += 64) {
                const uint32_t myEdit = k0 + lane < ec ? P.elist[L][es + k0 + lane] : 0u;
                EditHead mh{};
                if (k0 + lane < ec) mh = load_head(P.edits, myEdit);
                const uint32_t kn = min(64u, ec - k0);
                for (uint32_t k = 0; k < kn; ++k)
                    apply(head_of_lane(mh, k), (uint32_t)__builtin_amdgcn_readlane((int)myEdit, (int)k));
            }
            const uint64_t newMask = __ballot(v);
            if (v) {
                if (!haveCol) col = voxel_old_color();
                const uint32_t newRank = __popcll(newMask & ((1ull << lane) - 1ull));
                P.colors[T.cbase[task] + m.nstart + newRank] = col;
            }
        }
    }
}
__global__ __launch_bounds__(256) void k_color_leaf(EditPass P) {
    d_color_leaf(P, VGrid{blockIdx.x, gridDim.x});
}

// colour tree: make sure every task with colours has colour-tree nodes down to its leaf slot
// and a unique leaf handle (hash_dag_edits.h:225-249); one thread, tasks in index order.
// One thread per task: walk the colour tree to the task's leaf slot, creating missing colour
// nodes (8 zeroed entries; installed with a CAS, a loser's block stays unused), give the task a
// unique leaf slot (a shared or empty slot gets a new one) and log the slot's previous leaf
// for undo. Colour-node entries only ever change from 0 to a fixed value, so a stale read is
// either that value or 0, and a 0 goes through the CAS. Leaf slots are per task (one task per
// level-ctl node position), so they are written without atomics.
__device__ __forceinline__ void d_color_tree(const EditPass& P, VGrid G) {
    if (*P.errors) return;
    const uint32_t ctl = P.ctl;
    const uint32_t n = min(P.count[ctl], P.cap[ctl]);
    const Items T = P.items[ctl];
    for (uint32_t t0 = G.block * blockDim.x + (threadIdx.x & ~63u); t0 < n; t0 += G.nblocks * blockDim.x) {
        const uint32_t i = t0 + (threadIdx.x & 63u);
        bool ok = i < n && T.ncount[i] != 0;
        if (i < n && !ok) T.newhandle[i] = 0;
        uint32_t idx = 0, c = 0, h = 0;
        if (ok) {
            const uint32_t x = T.px[i], y = T.py[i], z = T.pz[i];
            for (uint32_t lvl = 1; lvl < ctl; ++lvl) {
                const uint32_t s = ctl - lvl;
                const uint32_t cc = (((x >> s) & 1u) << 2) | (((y >> s) & 1u) << 1) | ((z >> s) & 1u);
                uint32_t e = __atomic_load_n(&P.cnodes[idx + cc], __ATOMIC_RELAXED);
                if (e == 0) {
                    const uint32_t ne = atomicAdd(P.n_cnodes, 8u);  // rare: new colour-tree nodes
                    if (ne + 8 > P.cnodes_cap) {
                        set_error(P.errors, ERR_COLOR_TREE);
                        ok = false;
                        break;
                    }
                    const uint32_t prev = atomicCAS(&P.cnodes[idx + cc], 0u, ne);
                    e = prev == 0 ? ne : prev;
                }
                idx = e;
            }
            c = ((x & 1u) << 2) | ((y & 1u) << 1) | (z & 1u);
            if (ok) h = __atomic_load_n(&P.cnodes[idx + c], __ATOMIC_RELAXED);
            if (!ok) T.newhandle[i] = 0;
        }
        const LeafDesc prevDesc = ok ? task_leaf(P, h) : LeafDesc{};
        const bool needSlot = ok && (h == 0 || !(h & 0x80000000u));
        const uint32_t li = wave_alloc32(P.n_leaves, needSlot ? 1u : 0u);
        if (needSlot && li >= P.leaves_cap) {
            set_error(P.errors, ERR_COLOR_TREE);
            T.newhandle[i] = 0;
            ok = false;
        }
        // every undo-log entry handed out below is written (a failed pass restores from it)
        const uint32_t u = wave_alloc32(P.undo_count, ok ? 1u : 0u);
        if (!ok) continue;
        if (needSlot) {
            h = li | 0x80000000u;
            // the slot keeps resolving to the previous colours until the pass commits
            P.leaves[li] = prevDesc;
            __atomic_store_n(&P.cnodes[idx + c], h, __ATOMIC_RELAXED);
        }
        T.newhandle[i] = h;
        // undo log: handle + the leaf it resolved to before this pass
        if (u < P.undo_cap) {
            P.undo_handles[u] = h;
            P.undo_descs[u] = prevDesc;
        } else {
            set_error(P.errors, ERR_UNDO);
        }
    }
}
__global__ __launch_bounds__(256) void k_color_tree(EditPass P) {
    d_color_tree(P, VGrid{blockIdx.x, gridDim.x});
}

// A failed pass: every slot in the undo log gets back the leaf it resolved to before.
__global__ __launch_bounds__(256) void k_restore_leaves(LeafDesc* leaves, const uint32_t* handles,
                                                        const LeafDesc* descs, const uint32_t* count, uint32_t cap) {
    const uint32_t n = min(*count, cap);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        leaves[handles[k] & 0x7FFFFFFFu] = descs[k];
}

// leaves[handles[k]] = descs[k] (undo / redo of a frame; a frame names each slot once).
__global__ __launch_bounds__(256) void k_scatter_leaves(LeafDesc* leaves, const uint32_t* handles,
                                                        const LeafDesc* descs, uint32_t n) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        leaves[handles[k] & 0x7FFFFFFFu] = descs[k];
}

// The leaves a pass wrote, in undo-log order (the redo half of an undo frame).
__global__ __launch_bounds__(256) void k_gather_leaves(const LeafDesc* leaves, const uint32_t* handles,
                                                       LeafDesc* out, uint32_t n) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        out[k] = leaves[handles[k] & 0x7FFFFFFFu];
}

// Colour leaf encoder (k_encode): per task, a counting walk over the colours (blocks and weight
// bits), then a writing walk in chunks of 64 with wave scans carried across chunks. Weight bits
// of a chunk are gathered in LDS and each weight word is written once (the word a chunk leaves
// partial is carried to the next chunk).
constexpr uint32_t kEncBlock = 256;
constexpr uint32_t kEncWaves = kEncBlock / 64;

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (uint32_t off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t lane) {
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t shfl_up64(uint64_t v) {
    return ((uint64_t)__shfl_up((uint32_t)(v >> 32), 1, 64) << 32) | __shfl_up((uint32_t)v, 1, 64);
}

// One task's staged colours, chunk by chunk, in colour order. A block starts where the colour
// index is a multiple of 16384 or the colour bits / bits per weight differ from the previous
// colour's (the same rule as the leaf builder). Uniform chunks (markers) hold no weights and
// can start a block only at their first colour, so a run of them is handled 64 chunks at a
// time. kWrite = false counts blocks and weight bits; kWrite = true writes blocks, macro-block
// records and weight words at the counted sizes. Weight bits of an explicit chunk are gathered
// in LDS and each weight word is written once (the word a chunk leaves partial is carried).
template <bool kWrite>
__device__ __forceinline__ void encode_colors(const uint64_t* __restrict__ cols, const uint64_t* __restrict__ marks,
                                              uint32_t cnt, uint32_t lane, uint32_t* buf, uint64_t* blocks,
                                              uint64_t* macro, uint32_t* weights, uint32_t& outBlocks,
                                              uint32_t& outBits) {
    uint32_t carryBlocks = 0, carryBits = 0, macroBit = 0, carryWord = 0, laneBits = 0;
    uint64_t prev = 0;  // the colour before the current chunk (unused at colour 0)
    const uint32_t nCh = (cnt + 63u) / 64u;
    for (uint32_t g0 = 0; g0 < nCh; g0 += 64) {
        const uint32_t gl = min(64u, nCh - g0);
        const uint64_t mk = lane < gl ? marks[g0 + lane] : 0ull;
        const uint64_t marked = __ballot((mk & kUniformChunk) != 0);
        uint32_t j = 0;
        while (j < gl) {
            if ((marked >> j) & 1ull) {
                // uniform chunks j .. e-1 (lane k: chunk g0 + k)
                const uint64_t open = ~marked >> j;
                const uint32_t e = min(open ? j + (uint32_t)__ffsll((long long)open) - 1u : 64u, gl);
                const uint64_t col = mk & 0xFFFFFFFFull;
                const uint32_t upc = __shfl_up((uint32_t)col, 1, 64);
                const bool in = lane >= j && lane < e;
                const uint32_t pos = (g0 + lane) * 64u;
                const bool mstart = in && (pos & 16383u) == 0;
                const bool st = mstart || (in && (lane == j ? ((prev ^ col) & kColorKey) != 0 : upc != (uint32_t)col));
                const uint64_t sb = __ballot(st);
                if constexpr (kWrite) {
                    const uint32_t bi = carryBlocks + (uint32_t)__popcll(sb & (~0ull >> (63u - lane)));
                    if (st) blocks[bi - 1] = (col << 32) | (0xFFFFu << 16) | (pos & 16383u);
                    if (mstart) {
                        macro[2 * (pos >> 14)] = bi - 1;
                        macro[2 * (pos >> 14) + 1] = carryBits;
                    }
                    if (__ballot(mstart)) macroBit = carryBits;
                }
                carryBlocks += (uint32_t)__popcll(sb);
                prev = __shfl((uint32_t)col, e - 1, 64);
                j = e;
                continue;
            }
            // explicit chunk
            const uint32_t c0 = (g0 + j) * 64u, i = c0 + lane;
            uint64_t c = 0;
            uint32_t bpw = 0;
            if (i < cnt) {
                c = cols[i];
                bpw = (uint32_t)(c >> 40) & 0xFFu;
            }
            const uint64_t up = shfl_up64(c);
            const uint64_t p = lane == 0 ? prev : up;
            const uint32_t start = i < cnt && ((i & 16383u) == 0 || ((p ^ c) & kColorKey) != 0);
            prev = shfl64(c, 63);
            ++j;
            if constexpr (!kWrite) {
                carryBlocks += (uint32_t)__popcll(__ballot(start));
                laneBits += bpw;
            } else {
                const uint32_t sA = wave_incl_scan(start, lane), sB = wave_incl_scan(bpw, lane);
                const uint32_t blockIdxIncl = carryBlocks + sA;  // block index + 1 if start
                const uint32_t bitOff = carryBits + sB - bpw;    // exclusive bit offset
                // only lane 0 can sit on a macro-block start (16384 is a multiple of 64)
                uint32_t myMacroBit = macroBit;
                if ((c0 & 16383u) == 0) {
                    myMacroBit = carryBits;
                    macroBit = carryBits;
                    if (lane == 0) {
                        macro[2 * (c0 >> 14)] = blockIdxIncl - 1;
                        macro[2 * (c0 >> 14) + 1] = bitOff;
                    }
                }
                if (i < cnt && start) {
                    const uint32_t rel = bitOff - myMacroBit;
                    const uint32_t hdr = bpw == 0 ? ((0xFFFFu << 16) | (i & 16383u))
                                                  : ((rel << 16) | ((bpw - 1) << 14) | (i & 16383u));
                    blocks[blockIdxIncl - 1] = ((uint64_t)(uint32_t)c << 32) | hdr;
                }
                // weight bits of the chunk: [B0, B1), words w0 .. ; MSB-first stream bit k ->
                // stored word k / 32, byte (k % 32) / 8, bit 7 - k % 8
                const uint32_t B0 = carryBits, B1 = carryBits + __shfl(sB, 63, 64);
                const uint32_t w0 = B0 >> 5;
                if (lane < 12) buf[lane] = lane == 0 ? carryWord : 0u;
                wave_sync();
                if (bpw) {
                    const uint32_t wvv = (uint32_t)(c >> 32) & 0xFFu;
                    for (uint32_t b = 0; b < bpw; ++b) {
                        if (!((wvv >> (bpw - 1 - b)) & 1u)) continue;
                        const uint32_t k = bitOff + b;
                        atomicOr(&buf[(k >> 5) - w0], 1u << (((k & 31u) >> 3) * 8u + (7u - (k & 7u))));
                    }
                }
                wave_sync();
                const uint32_t full = (B1 >> 5) - w0;  // complete words of the chunk
                if (lane < full) weights[w0 + lane] = buf[lane];
                carryWord = (B1 & 31u) ? buf[full] : 0u;
                wave_sync();
                carryBlocks = __shfl(blockIdxIncl, 63, 64);
                carryBits = B1;
            }
        }
    }
    outBlocks = carryBlocks;
    if constexpr (kWrite) {
        outBits = carryBits;
        if (lane == 0 && (carryBits & 31u)) weights[carryBits >> 5] = carryWord;
    } else {
        outBits = wave_sum(laneBits);
    }
}

// One wavefront per task: encode its colour array into a new unique leaf in the arena. The grid
// holds a wavefront per task (up to a cap, then a stride), so the hardware hands out tasks to
// free wavefronts as they finish: task sizes vary by orders of magnitude, and a fixed share of
// tasks per wavefront would leave the kernel's time to the wavefront with the largest share.
// Each task reserves its arena words with one atomic once its blocks and weight bits are
// counted.
__device__ __forceinline__ void d_encode(const EditPass& P, VGrid G) {
    const uint32_t ctl = P.ctl;
    const uint32_t n = min(P.count[ctl], P.cap[ctl]);
    if (*P.errors) return;
    const Items T = P.items[ctl];
    __shared__ uint32_t sBits[kEncWaves][12];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    uint32_t* buf = sBits[wv];
    uint64_t stagedBytes = 0;  // chunk markers and explicit colours (measurement)
    for (uint32_t t = G.block * kEncWaves + wv; t < n; t += G.nblocks * kEncWaves) {
        const uint32_t cnt = T.ncount[t];
        if (cnt == 0) continue;
        if ((cnt + 16383u) / 16384u > 160) {  // more than 160 * 16384 colours in one subtree
            if (lane == 0) set_error(P.errors, ERR_LEAF_SIZE);
            continue;
        }
        const uint64_t cb = T.cbase[t];
        if (P.abytes) {
            // each marker and explicit colour (8 B) is written once by the staging kernels and
            // read twice here (the counting walk, then the writing walk)
            const uint32_t nCh = (cnt + 63u) / 64u;
            for (uint32_t c = lane; c < nCh; c += 64) {
                const bool uniform = (P.cmarks[cb / 64 + c] & kUniformChunk) != 0;
                stagedBytes += 24u * (1u + (uniform ? 0u : min(64u, cnt - c * 64u)));
            }
        }
        uint32_t nBlocks, nBits;
        encode_colors<false>(P.colors + cb, P.cmarks + cb / 64, cnt, lane, buf, nullptr, nullptr, nullptr, nBlocks,
                             nBits);
        const uint32_t nMacro = (cnt + 16383u) / 16384u;
        const uint32_t nWeights = (nBits + 31) / 32;
        const uint64_t words = (uint64_t)nBlocks + 2ull * nMacro + (nWeights + 1) / 2 + 1;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(P.arena_top, (unsigned long long)words);
        base = shfl64(base, 0);
        if (base + words > P.arena_cap) {
            if (lane == 0) set_error(P.errors, ERR_COLOR_ARENA);
            continue;
        }
        uint64_t* blocks = P.arena + base;
        uint64_t* macro = blocks + nBlocks;
        uint32_t* weights = reinterpret_cast<uint32_t*>(macro + 2 * nMacro);
        uint32_t wb, wbits;
        encode_colors<true>(P.colors + cb, P.cmarks + cb / 64, cnt, lane, buf, blocks, macro, weights, wb, wbits);
        if (lane < 2) weights[nWeights + lane] = 0u;  // padding read by the 16-bit window
        if (lane == 0) {
            LeafDesc d;
            d.blocks = base;
            d.n_blocks = nBlocks;
            d.macro = base + nBlocks;
            d.n_macro = 2ull * nMacro;
            d.weights32 = 2ull * (base + nBlocks + 2ull * nMacro);
            d.n_weights = nWeights;
            d.offset = ~0ull;
            d.pad = 0;
            P.leaves[T.newhandle[t] & 0x7FFFFFFFu] = d;
        }
    }
    if (P.abytes) wave_add64(P.abytes + 2, stagedBytes);
}
__global__ __launch_bounds__(kEncBlock) void k_encode(EditPass P) {
    d_encode(P, VGrid{blockIdx.x, gridDim.x});
}

// ---- fill editor preparation (BufferEditor / FillEditorBase, hash_dag_editors.h:144-209, 629-671) ----
// value of every buffer voxel: DAGUtils::get_value at get_world_path (min + float(x), ...)
__global__ __launch_bounds__(256) void k_fill_load(const uint32_t* __restrict__ vpool, const uint32_t* d_root,
                                                   uint32_t levels, float mx, float my, float mz, uint32_t S,
                                                   uint8_t* __restrict__ out) {
    const uint64_t n = (uint64_t)S * S * S;
    const uint32_t root = *d_root;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)(i % S), y = (uint32_t)(i / S % S), z = (uint32_t)(i / ((uint64_t)S * S));
        const uint32_t wx = f2u_x86(mx + (float)x), wy = f2u_x86(my + (float)y), wz = f2u_x86(mz + (float)z);
        const uint64_t lf = source_leaf(vpool, root, levels, wx >> 2, wy >> 2, wz >> 2);
        out[i] = (uint8_t)((lf >> leaf_bit(wx, wy, wz)) & 1ull);
    }
}

// ---- noise editor preparation (SphereNoiseEditor, hash_dag_editors.h:762-792) ----
// FastNoise's 3D Perlin fractal with the editor's settings (FastNoise.cpp:195-205, 328-333,
// 760-797, 841-888): float, quintic interpolation, 3 octaves of lacunarity 2 and gain 0.5;
// tables = m_perm (512 bytes) then m_perm12 (512 bytes), staged in LDS
__device__ __forceinline__ int noise_floor(float f) { return f >= 0 ? (int)f : (int)f - 1; }
__device__ __forceinline__ float noise_lerp(float a, float b, float t) { return a + t * (b - a); }
__device__ __forceinline__ float noise_quintic(float t) { return t * t * t * (t * (t * 6 - 15) + 10); }
__device__ __forceinline__ float noise_grad(const uint8_t* perm, const uint8_t* perm12, uint32_t offset, int x, int y,
                                            int z, float dx, float dy, float dz) {
    // the 12 gradients (FastNoise.cpp:48-66): x = +-1 for the first 8, y = +-1 for 0-3 and 8-11,
    // z = +-1 from 4 on
    const uint32_t k = perm12[(x & 0xff) + perm[(y & 0xff) + perm[(z & 0xff) + offset]]];
    const float gx = k < 8 ? ((k & 1u) ? -1.f : 1.f) : 0.f;
    const float gy = k < 4 ? ((k & 2u) ? -1.f : 1.f) : (k < 8 ? 0.f : ((k & 1u) ? -1.f : 1.f));
    const float gz = k < 4 ? 0.f : ((k & 2u) ? -1.f : 1.f);
    return dx * gx + dy * gy + dz * gz;
}
__device__ float perlin_single(const uint8_t* perm, const uint8_t* perm12, uint32_t offset, float x, float y,
                               float z) {
    const int x0 = noise_floor(x), y0 = noise_floor(y), z0 = noise_floor(z);
    const int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
    const float xs = noise_quintic(x - (float)x0), ys = noise_quintic(y - (float)y0), zs = noise_quintic(z - (float)z0);
    const float xd0 = x - (float)x0, yd0 = y - (float)y0, zd0 = z - (float)z0;
    const float xd1 = xd0 - 1, yd1 = yd0 - 1, zd1 = zd0 - 1;
    const float xf00 = noise_lerp(noise_grad(perm, perm12, offset, x0, y0, z0, xd0, yd0, zd0),
                                  noise_grad(perm, perm12, offset, x1, y0, z0, xd1, yd0, zd0), xs);
    const float xf10 = noise_lerp(noise_grad(perm, perm12, offset, x0, y1, z0, xd0, yd1, zd0),
                                  noise_grad(perm, perm12, offset, x1, y1, z0, xd1, yd1, zd0), xs);
    const float xf01 = noise_lerp(noise_grad(perm, perm12, offset, x0, y0, z1, xd0, yd0, zd1),
                                  noise_grad(perm, perm12, offset, x1, y0, z1, xd1, yd0, zd1), xs);
    const float xf11 = noise_lerp(noise_grad(perm, perm12, offset, x0, y1, z1, xd0, yd1, zd1),
                                  noise_grad(perm, perm12, offset, x1, y1, z1, xd1, yd1, zd1), xs);
    return noise_lerp(noise_lerp(xf00, xf10, ys), noise_lerp(xf01, xf11, ys), zs);
}

// Buffer voxel values after the noise edit: set to `adding` where the distance to the buffer
// centre (uint3 difference, FP64) is below r * (1 + noise), else the DAG's value at the world
// voxel truncate(min + p) (BufferEditor::get_value, hash_dag_editors.h:185-195)
__global__ __launch_bounds__(256) void k_noise_state(const uint32_t* __restrict__ vpool, const uint32_t* d_root,
                                                     uint32_t levels, float mx, float my, float mz, uint32_t S,
                                                     float r, float frequency, float bounding,
                                                     const uint8_t* __restrict__ tables, uint32_t adding,
                                                     uint8_t* __restrict__ out) {
    __shared__ uint8_t sTab[1024];
    for (uint32_t i = threadIdx.x; i < 1024; i += blockDim.x) sTab[i] = tables[i];
    __syncthreads();
    const uint8_t* perm = sTab;
    const uint8_t* perm12 = sTab + 512;
    const uint64_t n = (uint64_t)S * S * S;
    const uint32_t root = *d_root, c0 = S / 2;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)(i % S), y = (uint32_t)(i / S % S), z = (uint32_t)(i / ((uint64_t)S * S));
        // GetPerlinFractal(float(x), float(y), float(z)) with FBM (FastNoise.cpp:760-797)
        float px = (float)x * frequency, py = (float)y * frequency, pz = (float)z * frequency;
        float sum = perlin_single(perm, perm12, perm[0], px, py, pz), amp = 1;
        for (uint32_t o = 1; o < 3; ++o) {
            px *= 2.f;
            py *= 2.f;
            pz *= 2.f;
            amp *= 0.5f;
            sum += perlin_single(perm, perm12, perm[o], px, py, pz) * amp;
        }
        const float noise = sum * bounding;
        const uint32_t dx = c0 - x, dy = c0 - y, dz = c0 - z;
        const double distance = __dsqrt_rn((double)(dx * dx + dy * dy + dz * dz));
        uint32_t v;
        if (distance < (double)r * (1. + (double)noise)) {
            v = adding;
        } else {
            const uint32_t wx = f2u_x86(mx + (float)x), wy = f2u_x86(my + (float)y), wz = f2u_x86(mz + (float)z);
            const uint64_t lf = source_leaf(vpool, root, levels, wx >> 2, wy >> 2, wz >> 2);
            v = (uint32_t)((lf >> leaf_bit(wx, wy, wz)) & 1ull);
        }
        out[i] = (uint8_t)v;
    }
}

// DAGUtils::get_color of the given voxels (the solid voxels the flood fill loaded)
__global__ __launch_bounds__(256) void k_fill_colors(EditPass P, const uint32_t* d_root, const uint32_t* __restrict__ xyz,
                                                     uint32_t n, uint64_t* __restrict__ out) {
    const uint32_t root = *d_root;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = source_color(P, root, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

__global__ void k_init_pass(EditPass P, uint32_t root, uint32_t first, uint32_t nEdits, const uint32_t* dN) {
    // level-0 item = the root, carrying the edits [first, first + nEdits)
    if (dN) nEdits = min(nEdits, *dN);
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nEdits) P.elist[0][t] = first + t;
    if (t == 0) {
        for (uint32_t l = 0; l < P.levels; ++l) {
            P.count[l] = 0;
            P.ecount[l] = 0;
        }
        P.count[0] = 1;
        P.ecount[0] = nEdits;
        const Items I = P.items[0];
        I.px[0] = I.py[0] = I.pz[0] = 0;
        I.old_[0] = root;
        I.orig[0] = root;
        I.parent[0] = 0;
        I.slot[0] = 0;
        I.estart[0] = 0;
        I.ecount[0] = nEdits;
        I.aux[0] = 0;  // colour node of the root
        I.cmode[0] = 0;
        I.cmodeval[0] = 0;
        *P.color_top = 0;
    }
}

__device__ __forceinline__ void d_finish(const EditPass& P, uint32_t* root_out, VGrid G) {
    if (G.block == 0 && threadIdx.x == 0 && *P.errors == 0) {
        const uint32_t r = P.items[0].nidx[0];
        if (r == kInvalidNode) {
            set_error(P.errors, ERR_EMPTY);
        } else {
            *root_out = r;
            // commit the colour arena top (a failed pass leaves the committed value)
            *reinterpret_cast<unsigned long long*>(P.pool_top + kArenaTopWord) = *P.arena_top;
        }
    }
}
__global__ void k_finish(EditPass P, uint32_t* root_out) {
    d_finish(P, root_out, VGrid{blockIdx.x, gridDim.x});
}

}  // namespace

// Host: scratch management and pass orchestration
// Frees the per-pass scratch; the uploaded edit array and the undo log survive.
static void scratch_free(hd_ctx* ctx) {
    EditScratch& S = ctx->edit;
    for (void* p : S.allocs) hipFree(p);
    hipFree(S.hset);
    hipFree(S.colors);
    hipFree(S.cmarks);
    for (void*& p : S.level_mem) {
        hipFree(p);
        p = nullptr;
    }
    DevEdit* edits = S.edits_dev;
    const uint32_t ecap = S.edits_cap;
    uint32_t* uh = S.undo_handles;
    LeafDesc *ud = S.undo_descs, *un = S.undo_next;
    const uint32_t ucap = S.undo_cap;
    S = EditScratch{};
    S.edits_dev = edits;
    S.edits_cap = ecap;
    S.undo_handles = uh;
    S.undo_descs = ud;
    S.undo_next = un;
    S.undo_cap = ucap;
}

// Lays the item arrays of one level (capacity c) and its edit list (e words) out from `mem`
// (nullptr: sizes them only) in 256-byte aligned fields; returns the bytes they take. The
// leaf level has no child-slot arrays.
uint64_t carve_level(Items& I, uint32_t*& elist, bool leafLevel, uint64_t c, uint64_t e, char* mem) {
    const uint64_t c8 = leafLevel ? 0 : 8 * c;
    struct Field {
        void** p;
        uint64_t bytes;
    };
    I = Items{};
    const Field f[] = {
        {(void**)&I.px, c * 4},       {(void**)&I.py, c * 4},      {(void**)&I.pz, c * 4},
        {(void**)&I.old_, c * 4},     {(void**)&I.orig, c * 4},    {(void**)&I.parent, c * 4},
        {(void**)&I.slot, c},         {(void**)&I.estart, c * 4},  {(void**)&I.ecount, c * 4},
        {(void**)&I.aux, c * 4},      {(void**)&I.handle, c * 4},  {(void**)&I.newhandle, c * 4},
        {(void**)&I.cmode, c * 4},    {(void**)&I.cmodeval, c * 4}, {(void**)&I.ckind, c8},
        {(void**)&I.cval, c8 * 4},    {(void**)&I.corig, c8 * 4},  {(void**)&I.cold, c8 * 4},
        {(void**)&I.cnew, c8 * 4},    {(void**)&I.ccolor, c8 * 4}, {(void**)&I.cstart, c8 * 4},
        {(void**)&I.costart, c8 * 4}, {(void**)&I.nidx, c * 4},    {(void**)&I.ncount, c * 4},
        {(void**)&I.cont, c * 36},    {(void**)&I.clen, c * 4},    {(void**)&I.hash, c * 4},
        {(void**)&I.rep, c * 4},      {(void**)&I.state, c * 4},   {(void**)&I.nstart, c * 4},
        {(void**)&I.ostart, c * 4},   {(void**)&I.cbase, c * 8},   {(void**)&elist, e * 4},
    };
    auto up = [](uint64_t v) { return (v + 255) & ~255ull; };
    uint64_t off = 0;
    for (const Field& x : f) {
        *x.p = (mem && x.bytes) ? mem + off : nullptr;
        off += up(x.bytes);
    }
    return std::max<uint64_t>(off, 256);
}

// One allocation per level holding its item arrays and edit list.
static hipError_t alloc_level(EditScratch& S, uint32_t l, bool leafLevel, uint64_t c, uint64_t e) {
    const uint64_t total = carve_level(S.items[l], S.elist[l], leafLevel, c, e, nullptr);
    char* mem = nullptr;
    const hipError_t err = hipMalloc(&mem, total);
    if (err != hipSuccess) return err;
    S.level_mem[l] = mem;
    carve_level(S.items[l], S.elist[l], leafLevel, c, e, mem);
    S.item_cap[l] = (uint32_t)c;
    S.elist_cap[l] = (uint32_t)e;
    return hipSuccess;
}

// Grows the scratch to at least items[l] / elist[l] per level (levels 0 .. leaf level) and
// `colors` staged colours. Levels that already fit keep their memory.
hd_status edit_reserve(hd_ctx* ctx, const uint32_t* items, const uint32_t* elist, uint64_t colors) {
    EditScratch& S = ctx->edit;
    const uint32_t levels = ctx->g.levels, L = ctx->g.leaf_level();
    if (S.levels != levels) {
        hipDeviceSynchronize();
        ctx->ptr_gen++;
        scratch_free(ctx);
        S.levels = levels;
        auto alloc = [&](void** p, uint64_t bytes) -> bool {
            if (hipMalloc(p, std::max<uint64_t>(bytes, 16)) != hipSuccess) return false;
            S.allocs.push_back(*p);
            return true;
        };
        if (!(alloc((void**)&S.dcount, 64 * 4) && alloc((void**)&S.decount, 64 * 4) && alloc((void**)&S.dstats, 64 * 8))) {
            scratch_free(ctx);
            return fail(ctx, HD_ERR_OOM, "edit scratch allocation failed");
        }
    }
    bool synced = false;
    auto sync_once = [&]() {
        if (!synced) hipDeviceSynchronize();  // no pass may still use memory being freed
        synced = true;
        ctx->ptr_gen++;
    };
    uint32_t maxItems = 0;
    for (uint32_t l = 0; l <= L; ++l) {
        // level l holds at most 8^l items
        const uint64_t bound = l < 10 ? (1ull << (3 * l)) : ~0ull;
        const uint64_t want = std::min<uint64_t>(std::max<uint64_t>(items[l], 1), bound);
        const uint64_t c = std::max<uint64_t>(want, S.item_cap[l]);
        const uint64_t e = std::max<uint64_t>(std::max<uint64_t>(elist[l], 16), S.elist_cap[l]);
        maxItems = std::max<uint32_t>(maxItems, (uint32_t)c);
        if (c == S.item_cap[l] && e == S.elist_cap[l] && S.level_mem[l]) continue;
        sync_once();
        hipFree(S.level_mem[l]);
        S.level_mem[l] = nullptr;
        S.item_cap[l] = S.elist_cap[l] = 0;
        const hipError_t err = alloc_level(S, l, l == L, c, e);
        if (err != hipSuccess) return hip_fail(ctx, err, "edit scratch (items)");
    }
    uint32_t hcap = 1;
    while (hcap < 2 * maxItems) hcap <<= 1;
    if (hcap > S.hcap) {
        sync_once();
        hipFree(S.hset);
        S.hset = nullptr;
        S.hcap = 0;
        const hipError_t err = hipMalloc(&S.hset, (uint64_t)hcap * 4);
        if (err != hipSuccess) return hip_fail(ctx, err, "edit scratch (hash set)");
        S.hcap = hcap;
    }
    if (colors > S.color_cap) {
        sync_once();
        hipFree(S.colors);
        hipFree(S.cmarks);
        S.colors = nullptr;
        S.cmarks = nullptr;
        S.color_cap = 0;
        colors = (colors + 63) & ~63ull;  // whole chunks, one marker each
        hipError_t err = hipMalloc(&S.colors, colors * 8);
        if (err == hipSuccess) err = hipMalloc(&S.cmarks, colors / 64 * 8);
        if (err != hipSuccess) return hip_fail(ctx, err, "edit scratch (colours)");
        S.color_cap = colors;
    }
    return HD_OK;
}

// Undo log of a pass (touched leaf slots): kept across scratch re-reserves, grown on demand.
hd_status edit_reserve_undo(hd_ctx* ctx, uint64_t entries) {
    EditScratch& S = ctx->edit;
    if (entries <= S.undo_cap) return HD_OK;
    const uint32_t cap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(entries, kUndoCap), 1u << 30);
    hipDeviceSynchronize();
    ctx->ptr_gen++;
    hipFree(S.undo_handles);
    hipFree(S.undo_descs);
    hipFree(S.undo_next);
    S.undo_handles = nullptr;
    S.undo_descs = S.undo_next = nullptr;
    S.undo_cap = 0;
    hipError_t e = hipMalloc(&S.undo_handles, (uint64_t)cap * 4);
    if (e == hipSuccess) e = hipMalloc(&S.undo_descs, (uint64_t)cap * sizeof(LeafDesc));
    if (e == hipSuccess) e = hipMalloc(&S.undo_next, (uint64_t)cap * sizeof(LeafDesc));
    if (e != hipSuccess) return hip_fail(ctx, e, "undo log");
    S.undo_cap = cap;
    return HD_OK;
}

// Device array for n edits (kept across scratch re-reserves; its contents stay valid).
hd_status edit_ensure_edits(hd_ctx* ctx, uint32_t n) {
    EditScratch& S = ctx->edit;
    if (n <= S.edits_cap) return HD_OK;
    const uint32_t cap = std::max<uint32_t>(n, 4096);
    DevEdit* ne = nullptr;
    const hipError_t e = hipMalloc(&ne, (uint64_t)cap * sizeof(DevEdit));
    if (e != hipSuccess) return hip_fail(ctx, e, "edit array");
    if (S.edits_dev) {
        hipDeviceSynchronize();
        hipFree(S.edits_dev);
    }
    ctx->ptr_gen++;
    S.edits_dev = ne;
    S.edits_cap = cap;
    return HD_OK;
}

void edit_free(hd_ctx* ctx) {
    EditScratch& S = ctx->edit;
    scratch_free(ctx);
    hipFree(S.edits_dev);
    hipFree(S.undo_handles);
    hipFree(S.undo_descs);
    hipFree(S.undo_next);
    S = EditScratch{};
}

namespace {
// tags of the leaf slots below each leaf-level bucket's size: one block per bucket at a time
__global__ __launch_bounds__(256) void k_leaf_tags(Geom g, const uint32_t* __restrict__ vpool,
                                                   const uint32_t* __restrict__ bucket_sizes, uint8_t* __restrict__ tags) {
    const uint32_t L = g.leaf_level(), nb = g.buckets_per_level(L), slots = g.bucket_size(L) / 2;
    for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const uint32_t n = min(bucket_sizes[g.bucket_global(L, b)], g.bucket_size(L)) / 2;
        const uint32_t base = g.make_ptr(L, b, 0);
        for (uint32_t s = threadIdx.x; s < n; s += blockDim.x) {
            const uint64_t leaf = *reinterpret_cast<const uint64_t*>(vpool + base + 2 * s);
            tags[(uint64_t)b * slots + s] = (uint8_t)(hash_leaf(leaf) >> 24);
        }
    }
}
}  // namespace

hd_status leaf_tags_rebuild(hd_ctx* ctx, hipStream_t s) {
    k_leaf_tags<<<4096, 256, 0, s>>>(ctx->g, ctx->vpool, ctx->bucket_sizes, ctx->leaf_tags);
    HD_HIP(ctx, hipGetLastError());
    HD_HIP(ctx, hipStreamSynchronize(s));
    return HD_OK;
}

namespace {
// A whole pass in one launch: the pass kernels' bodies in the order the level-synchronous form
// launches them, each over the same per-level item queues (the items k_expand appends for the
// next level, the representatives k_dedup lists for k_find), with a grid-wide barrier between
// steps. Launched cooperatively, so every block is resident and every barrier is reached by
// every block: the step sequence depends only on the kernel's arguments, and a body that finds
// an error flag set returns to the sequence, not out of the kernel.
// The steps of a pass, in order, as (step << 8 | level): expand per level top-down, the
// leaves, then per level bottom-up build / clear / dedup / find / add / resolve, the colour
// steps, and the commit. Every launch form runs this sequence.
enum PassStep : uint32_t {
    ST_EXPAND, ST_LEAF, ST_BUILD, ST_HCLEAR, ST_DEDUP, ST_FIND, ST_ADD, ST_RESOLVE, ST_CSTARTS, ST_CALLOC,
    ST_CTREE, ST_CCOPY, ST_CCOPY_LEAVES, ST_CLEAF, ST_ENCODE, ST_FINISH
};
constexpr uint32_t kMaxSteps = 8 * kMaxEditLevels + 8;
struct StepList {
    uint16_t step[kMaxSteps];
    uint32_t n;
};

static StepList pass_step_list(uint32_t levels, uint32_t ctl, bool colors) {
    StepList sl{};
    auto add = [&](uint32_t st, uint32_t l) { sl.step[sl.n++] = (uint16_t)(st << 8 | l); };
    const uint32_t L = levels - 2;
    for (uint32_t l = 0; l < L; ++l) add(ST_EXPAND, l);
    add(ST_LEAF, L);
    for (int l = (int)L; l >= 0; --l) {
        if (l < (int)L) add(ST_BUILD, (uint32_t)l);
        add(ST_HCLEAR, (uint32_t)l);
        add(ST_DEDUP, (uint32_t)l);
        add(ST_FIND, (uint32_t)l);
        add(ST_ADD, (uint32_t)l);
        add(ST_RESOLVE, (uint32_t)l);
    }
    if (colors) {
        for (uint32_t l = ctl; l < L; ++l) add(ST_CSTARTS, l);
        add(ST_CALLOC, ctl);
        add(ST_CTREE, ctl);
        for (uint32_t l = ctl; l < L; ++l) add(l + 1 == L ? ST_CCOPY_LEAVES : ST_CCOPY, l);
        add(ST_CLEAF, L);
        add(ST_ENCODE, ctl);
    }
    add(ST_FINISH, 0);
    return sl;
}

template <bool kSpecial>
__device__ __forceinline__ void run_step(const EditPass& P, uint32_t code, uint32_t* root_out, VGrid G) {
    const uint32_t l = code & 0xFFu;
    switch (code >> 8) {
        case ST_EXPAND:
            if ((P.wide_levels >> l) & 1u) d_expand_wide(P, l, G);
            else d_expand(P, l, G);
            break;
        case ST_LEAF: d_leaf<kSpecial>(P, G); break;
        case ST_BUILD: d_build(P, l, G); break;
        case ST_HCLEAR: d_hset_clear(P, l, G); break;
        case ST_DEDUP: d_dedup(P, l, G); break;
        case ST_FIND: d_find(P, l, G); break;
        case ST_ADD: d_add(P, l, G); break;
        case ST_RESOLVE: d_resolve(P, l, G); break;
        case ST_CSTARTS: d_color_starts(P, l, G); break;
        case ST_CALLOC: d_color_alloc(P, G); break;
        case ST_CTREE: d_color_tree(P, G); break;
        case ST_CCOPY: d_color_copy(P, l, G); break;
        case ST_CCOPY_LEAVES: d_color_copy_leaves(P, G); break;
        case ST_CLEAF:
            if constexpr (kSpecial) d_color_leaf(P, G);
            else d_color_leaf_items(P, G);
            break;
        case ST_ENCODE: d_encode(P, G); break;
        default: d_finish(P, root_out, G); break;
    }
}

// The whole pass in one cooperative launch: every block resident, a grid-wide barrier between
// steps; the step sequence depends only on the kernel's arguments, and a step that finds an
// error flag set returns to the sequence, not out of the kernel, so every block reaches every
// barrier.
template <bool kSpecial>
__global__ __launch_bounds__(256) void k_pass_persistent(EditPass P, StepList sl, uint32_t* root_out) {
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    const VGrid G{blockIdx.x, gridDim.x};
    for (uint32_t i = 0; i < sl.n; ++i) {
        run_step<kSpecial>(P, sl.step[i], root_out, G);
        if (i + 1 < sl.n) grid.sync();
    }
}

// Steps [first, last) of a pass in one workgroup, a block barrier between steps: the steps of
// levels with few items (all of a small pass) in one ordinary launch.
template <bool kSpecial>
__global__ __launch_bounds__(256) void k_pass_block(EditPass P, StepList sl, uint32_t first, uint32_t last,
                                                    uint32_t* root_out) {
    for (uint32_t i = first; i < last; ++i) {
        run_step<kSpecial>(P, sl.step[i], root_out, VGrid{0, 1});
        if (i + 1 < last) __syncthreads();
    }
}

template <bool kSpecial>
hd_status launch_persistent(hd_ctx* ctx, EditPass& P, StepList& sl, hipStream_t s) {
    auto* fn = k_pass_persistent<kSpecial>;
    int perCU = 0;
    HD_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fn, 256, 0));
    int dev = 0, cus = 0;
    HD_HIP(ctx, hipGetDevice(&dev));
    HD_HIP(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (perCU < 1 || cus < 1) return fail(ctx, HD_ERR_HIP, "persistent edit pass: no resident block fits");
    const uint32_t blocks = std::min<uint32_t>((uint32_t)perCU, ctx->edit_persist_per_cu) * (uint32_t)cus;
    uint32_t* root_out = ctx->d_root;
    void* args[] = {&P, &sl, &root_out};
    HD_HIP(ctx, hipLaunchCooperativeKernel(reinterpret_cast<const void*>(fn), dim3(blocks), dim3(256), args, 0, s));
    return HD_OK;
}
}  // namespace

static void fill_pass(hd_ctx* ctx, EditPass& P, const EditScratch& S) {
    P.g = ctx->g;
    P.levels = ctx->g.levels;
    P.ctl = ctx->g.color_tree_levels;
    P.vpool = ctx->vpool;
    P.page_table = ctx->page_table;
    P.bucket_sizes = ctx->bucket_sizes;
    P.full_nodes = ctx->full_nodes;
    P.pool_top = ctx->counters + 0;
    P.errors = ctx->counters + 1;
    P.leaf_tags = ctx->leaf_tags;
    P.abytes = ctx->count_edit_bytes ? ctx->edit_bytes : nullptr;
    P.n_cnodes = ctx->counters + 2;
    P.n_leaves = ctx->counters + 3;
    P.undo_count = ctx->counters + 4;
    P.pool_pages = ctx->cfg.pool_pages;
    P.count = S.dcount;
    P.ecount = S.decount;
    for (uint32_t l = 0; l < P.levels; ++l) {
        P.items[l] = S.items[l];
        P.elist[l] = S.elist[l];
    }
    for (uint32_t l = 0; l < kMaxEditLevels; ++l) {
        P.cap[l] = S.item_cap[l];
        P.ecap[l] = S.elist_cap[l];
    }
    P.hset = S.hset;
    P.hcap = S.hcap;
    P.edits = S.edits_dev;
    P.stat_voxels = reinterpret_cast<unsigned long long*>(S.dstats + 0);
    P.stat_new_words = reinterpret_cast<unsigned long long*>(S.dstats + 1);
    P.color_top = reinterpret_cast<unsigned long long*>(S.dstats + 2);
    P.arena_top = reinterpret_cast<unsigned long long*>(S.dstats + 3);
    P.stat_colors = reinterpret_cast<unsigned long long*>(S.dstats + 4);
    P.find_count = reinterpret_cast<uint32_t*>(S.dstats + 16);  // 32 words, zeroed with the stats
    P.colors = S.colors;
    P.cmarks = S.cmarks;
    P.color_cap = S.color_cap;
    P.cnodes = ctx->cnodes;
    P.cnodes_cap = ctx->cfg.color_nodes_cap;
    P.offsets = ctx->coffsets;
    P.leaves = ctx->cleaves;
    P.leaves_cap = ctx->cfg.color_leaves_cap;
    P.main = ctx->main;
    P.arena = ctx->arena;
    P.arena_cap = ctx->cfg.color_arena_words;
    P.undo_handles = S.undo_handles;
    P.undo_descs = S.undo_descs;
    P.undo_cap = S.undo_cap;
}

// Enqueues one pass over edits [first, first + n) of the uploaded edit array.
// dN (optional): the edit count is read from device memory (at most n), for graph replays.
// scratch (optional): the scratch the pass runs in (default: the context's own).
// Upper bound of the work items of each level a pass over these editors creates: an item is a
// node some editor's box overlaps, so per editor and level the nodes of that level its box
// overlaps, capped by the level's node count and item capacity.
static void level_bounds(const EditPass& P, const DevEdit* e, uint32_t n, uint64_t out[kMaxEditLevels]) {
    const uint32_t D = P.levels, L = D - 2;
    const double world = (double)(1ull << D);
    for (uint32_t l = 0; l <= L && l < kMaxEditLevels; ++l) {
        const double side = (double)(1ull << (D - l));
        const uint64_t perAxis = 1ull << l;
        uint64_t sum = 0;
        for (uint32_t i = 0; i < n && sum < P.cap[l]; ++i) {
            uint64_t cnt = 1;
            for (int a = 0; a < 3; ++a) {
                const double lo = std::max(0.0, (double)e[i].bmin[a]), hi = std::min(world, (double)e[i].bmax[a]);
                if (!(hi >= lo)) {
                    cnt = 0;
                    break;
                }
                const uint64_t c0 = (uint64_t)(lo / side), c1 = std::min<uint64_t>((uint64_t)(hi / side), perAxis - 1);
                cnt *= c1 >= c0 ? c1 - c0 + 1 : 1;
            }
            sum += cnt;
        }
        out[l] = std::min<uint64_t>(sum, P.cap[l]);
    }
}

constexpr uint32_t kWideEdits = 64, kWideItems = 512;

hd_status enqueue_pass(hd_ctx* ctx, uint32_t first, uint32_t n, bool colors, hipStream_t s, const uint32_t* dN,
                       bool special, const EditScratch* scratch, const DevEdit* hostEdits) {
    EditPass P{};
    fill_pass(ctx, P, scratch ? *scratch : ctx->edit);
    // grid per level: two blocks per item the level can hold (a wavefront per child slot is the
    // most any step uses), at most kGrid; kGrid everywhere without the host's copy of the edits
    uint32_t grid[kMaxEditLevels];
    uint32_t gridMax = 1;
    uint64_t bound[kMaxEditLevels];
    for (uint32_t l = 0; l < kMaxEditLevels; ++l) bound[l] = P.cap[l];
    if (hostEdits && !dN) level_bounds(P, hostEdits, n, bound);
    for (uint32_t l = 0; l < kMaxEditLevels; ++l) {
        grid[l] = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(2 * bound[l], 1), kGrid);
        gridMax = std::max(gridMax, grid[l]);
    }
    P.single_edit = (n == 1 && !dN) ? 1u : 0u;
    // a wavefront per child slot where a level's items hold long edit lists: passes over at
    // least kWideEdits editors, on levels with at most kWideItems items (the grid's two blocks
    // per item hold a wavefront per slot)
    P.wide_levels = 0;
    if (!dN && n >= kWideEdits)
        for (uint32_t l = 0; l < kMaxEditLevels && l < 32; ++l)
            if (bound[l] <= kWideItems) P.wide_levels |= 1u << l;
    const uint32_t D = P.levels, L = D - 2;
    k_init_pass<<<(std::max<uint32_t>(n, 1) + 255) / 256, 256, 0, s>>>(P, 0xFFFFFFFFu, first, n, dN);
    // the root index lives in ctx->d_root (device) so passes chain without a host round trip
    hipMemcpyAsync(&P.items[0].old_[0], ctx->d_root, 4, hipMemcpyDeviceToDevice, s);
    hipMemcpyAsync(&P.items[0].orig[0], ctx->d_root, 4, hipMemcpyDeviceToDevice, s);
    StepList sl = pass_step_list(D, P.ctl, colors);
    // one cooperative launch for passes of at most edit_persist_max edits (not under capture)
    if (!dN && n <= ctx->edit_persist_max)
        return special ? launch_persistent<true>(ctx, P, sl, s) : launch_persistent<false>(ctx, P, sl, s);
    // otherwise a pass with at most edit_block_max work items on every level (by the editors'
    // boxes) is one single-workgroup launch, and any other pass a launch per step over grids
    // sized for each level. (Running only a bigger pass's small levels in a workgroup was
    // slower: a workgroup's four wavefronts scan the buckets of a level's finds one after
    // another, where a launch gives each its own wavefront.)
    uint64_t largest = 0;
    for (uint32_t l = 0; l <= L && l < kMaxEditLevels; ++l) largest = std::max(largest, bound[l]);
    const bool smallPass = largest <= (uint64_t)ctx->edit_block_max;
    const auto small = [&](uint32_t) { return smallPass; };
    // a wavefront per task slot (the task count is on the device), at most 2^18 of them
    const bool bounded = hostEdits && !dN;  // bound[] from the editors' boxes
    const auto wideGrid = [](uint64_t blocks) {
        return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(blocks, 1u), 1u << 16);
    };
    const uint32_t encBlocks =
        (uint32_t)std::min<uint64_t>(std::max<uint64_t>((bound[P.ctl] + kEncWaves - 1) / kEncWaves, 1u), 1u << 16);
    for (uint32_t i = 0; i < sl.n;) {
        if (small(sl.step[i])) {
            uint32_t j = i + 1;
            while (j < sl.n && small(sl.step[j])) ++j;
            if (special) k_pass_block<true><<<1, 256, 0, s>>>(P, sl, i, j, ctx->d_root);
            else k_pass_block<false><<<1, 256, 0, s>>>(P, sl, i, j, ctx->d_root);
            i = j;
            continue;
        }
        const uint32_t l = sl.step[i] & 0xFFu, gl = grid[l];
        // Grids from the level's item bound by each kernel's unit of work (a thread, or a
        // wavefront, per item or per child slot), up to 2^16 blocks: free wavefronts then take
        // the next unit as the hardware hands out blocks, where a fixed stride of units per
        // thread leaves the launch to the one with the largest share (item costs vary by orders
        // of magnitude). Without the editors' boxes (graph replays) the grid is gl.
        const auto units = [&](uint64_t n, uint32_t perBlock) {
            return bounded ? wideGrid((n + perBlock - 1) / perBlock) : gl;
        };
        const uint64_t b = bound[l];
        switch (sl.step[i] >> 8) {
            case ST_EXPAND:  // wide: a wavefront per child slot; otherwise a thread per slot
                if ((P.wide_levels >> l) & 1u) k_expand_wide<<<gl, 256, 0, s>>>(P, l);
                else k_expand<<<units(8 * b, 256), 256, 0, s>>>(P, l);
                break;
            case ST_LEAF:  // a thread per item
                if (special) k_leaf<true><<<gl, 256, 0, s>>>(P);
                else k_leaf<false><<<units(b, 256), 256, 0, s>>>(P);
                break;
            case ST_BUILD: k_build<<<units(b, 256), 256, 0, s>>>(P, l); break;  // a thread per item
            case ST_HCLEAR: k_hset_clear<<<std::min<uint32_t>(gl, 256), 256, 0, s>>>(P, l); break;
            case ST_DEDUP: k_dedup<<<units(b, 256), 256, 0, s>>>(P, l); break;  // a thread per item
            case ST_FIND: k_find<<<units(b, 4), 256, 0, s>>>(P, l); break;      // a wavefront per representative
            case ST_ADD: k_add<<<units(b, 256), 256, 0, s>>>(P, l); break;      // a thread per item
            case ST_RESOLVE: k_resolve<<<gl, 256, 0, s>>>(P, l); break;
            case ST_CSTARTS: k_color_starts<<<units(8 * b, 256), 256, 0, s>>>(P, l); break;  // a thread per slot
            case ST_CALLOC: k_color_alloc<<<gl, 256, 0, s>>>(P); break;
            case ST_CTREE: k_color_tree<<<gridMax, 256, 0, s>>>(P); break;
            // colour copies: a wavefront per slot (per item at the leaf parents)
            case ST_CCOPY: k_color_copy<<<units(8 * b, 4), 256, 0, s>>>(P, l); break;
            case ST_CCOPY_LEAVES: k_color_copy_leaves<<<units(b, 4), 256, 0, s>>>(P); break;
            case ST_CLEAF:  // special: a wavefront per item; otherwise a thread per item
                if (special) k_color_leaf<<<gl, 256, 0, s>>>(P);
                else k_color_leaf_items<<<units(b, 256), 256, 0, s>>>(P);
                break;
            case ST_ENCODE: k_encode<<<encBlocks, kEncBlock, 0, s>>>(P); break;
            default: k_finish<<<1, 64, 0, s>>>(P, ctx->d_root); break;
        }
        ++i;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "edit pass launch");
    return HD_OK;
}

// After a failed pass: give every leaf slot the pass touched its previous leaf back.
void edit_restore_leaves(hd_ctx* ctx, hipStream_t s, const EditScratch* scratch) {
    const EditScratch& S = scratch ? *scratch : ctx->edit;
    k_restore_leaves<<<256, 256, 0, s>>>(ctx->cleaves, S.undo_handles, S.undo_descs, ctx->counters + 4, S.undo_cap);
}

// Undo / redo: write n host descriptors into their leaf slots (one upload, one scatter).
// FillEditorBase + FillEditorColors construction (hash_dag_editors.h:629-728) for one fill edit
// on the DAG at the context's current root: the buffer voxels' values come from the device,
// the flood fill (a stack walk, as the reference's) and the colour scan run on the host, and
// the result is uploaded for the pass. The buffers stay in `fb` until the pass has run.
hd_status fill_prepare(hd_ctx* ctx, const hd_edit& e, DevEdit& d, FillBuffers& fb, hipStream_t s) {
    const float fmin[3] = {e.center[0] - e.radius, e.center[1] - e.radius, e.center[2] - e.radius};
    const double ce = std::ceil((double)e.radius);
    if (!(ce >= 1.0) || ce > 512.0) return fail(ctx, HD_ERR_INVALID, "fill radius must be in (0, 512]");
    const uint32_t S = 2u * (uint32_t)ce;
    const uint64_t n = (uint64_t)S * S * S;
    enum : uint8_t { Unknown = 0, LoadedTrue = 1, LoadedFalse = 2, SetTrue = 3 };
    std::vector<uint8_t> solid(n);
    {
        uint8_t* dsolid = nullptr;
        HD_HIP(ctx, hipMalloc(&dsolid, n));
        k_fill_load<<<(uint32_t)std::min<uint64_t>((n + 255) / 256, 65536), 256, 0, s>>>(
            ctx->vpool, ctx->d_root, ctx->g.levels, fmin[0], fmin[1], fmin[2], S, dsolid);
        hipError_t err = hipGetLastError();
        if (err == hipSuccess) err = hipMemcpyAsync(solid.data(), dsolid, n, hipMemcpyDeviceToHost, s);
        if (err == hipSuccess) err = hipStreamSynchronize(s);
        hipFree(dsolid);
        if (err != hipSuccess) return hip_fail(ctx, err, "fill buffer load");
    }
    auto idx = [S](uint32_t x, uint32_t y, uint32_t z) { return (uint64_t)x + (uint64_t)S * y + (uint64_t)S * S * z; };
    std::vector<uint8_t> st(n, Unknown);
    auto get_value = [&](uint32_t x, uint32_t y, uint32_t z) {
        uint8_t& v = st[idx(x, y, z)];
        if (v == Unknown) v = solid[idx(x, y, z)] ? LoadedTrue : LoadedFalse;
        return v == LoadedTrue || v == SetTrue;
    };
    const uint32_t c0 = S / 2;
    uint32_t lo[3] = {c0, c0, c0}, hi[3] = {c0, c0, c0};
    std::vector<uint32_t> stack{c0, c0, c0};
    while (!stack.empty()) {
        const uint32_t z = stack.back();
        stack.pop_back();
        const uint32_t y = stack.back();
        stack.pop_back();
        const uint32_t x = stack.back();
        stack.pop_back();
        st[idx(x, y, z)] = SetTrue;
        const uint32_t q[3] = {x, y, z};
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], q[a]);
            hi[a] = std::max(hi[a], q[a]);
        }
        static const int dd[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
        for (const auto& o : dd) {
            const uint32_t nx = x + (uint32_t)o[0], ny = y + (uint32_t)o[1], nz = z + (uint32_t)o[2];
            if (nx < S && ny < S && nz < S && !get_value(nx, ny, nz)) {
                stack.push_back(nx);
                stack.push_back(ny);
                stack.push_back(nz);
            }
        }
    }
    // colours of the loaded solid voxels in the filled bounding box, in the scan order
    const uint32_t dim[3] = {hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
    std::vector<uint32_t> loaded;
    for (uint32_t x = lo[0]; x <= hi[0]; ++x)
        for (uint32_t y = lo[1]; y <= hi[1]; ++y)
            for (uint32_t z = lo[2]; z <= hi[2]; ++z)
                if (st[idx(x, y, z)] == LoadedTrue) {
                    loaded.push_back(f2u_host(fmin[0] + (float)x));
                    loaded.push_back(f2u_host(fmin[1] + (float)y));
                    loaded.push_back(f2u_host(fmin[2] + (float)z));
                }
    const uint32_t nl = (uint32_t)(loaded.size() / 3);
    std::vector<uint64_t> lcol(nl);
    if (nl && ctx->colors_loaded) {
        uint32_t* dxyz = nullptr;
        uint64_t* dcol = nullptr;
        hipError_t err = hipMalloc(&dxyz, loaded.size() * 4);
        if (err == hipSuccess) err = hipMalloc(&dcol, (uint64_t)nl * 8);
        if (err == hipSuccess) err = hipMemcpyAsync(dxyz, loaded.data(), loaded.size() * 4, hipMemcpyHostToDevice, s);
        if (err == hipSuccess) {
            EditPass P{};
            fill_pass(ctx, P, ctx->edit);
            k_fill_colors<<<std::min<uint32_t>((nl + 255) / 256, 4096), 256, 0, s>>>(P, ctx->d_root, dxyz, nl, dcol);
            err = hipGetLastError();
        }
        if (err == hipSuccess) err = hipMemcpyAsync(lcol.data(), dcol, (uint64_t)nl * 8, hipMemcpyDeviceToHost, s);
        if (err == hipSuccess) err = hipStreamSynchronize(s);
        hipFree(dxyz);
        hipFree(dcol);
        if (err != hipSuccess) return hip_fail(ctx, err, "fill colours");
    }
    const uint64_t nb = (uint64_t)dim[0] * dim[1] * dim[2];
    std::vector<uint64_t> colors(nb, 0);
    {
        uint64_t color = 0;  // CompressedColor{}
        uint32_t k = 0;
        for (uint32_t x = lo[0]; x <= hi[0]; ++x)
            for (uint32_t y = lo[1]; y <= hi[1]; ++y)
                for (uint32_t z = lo[2]; z <= hi[2]; ++z) {
                    const uint8_t v = st[idx(x, y, z)];
                    if (v == LoadedTrue) color = lcol[k++];
                    else if (v == SetTrue)
                        colors[(x - lo[0]) + (uint64_t)dim[0] * ((y - lo[1]) + (uint64_t)dim[1] * (z - lo[2]))] = color;
                }
    }
    // state for the pass: bit 0 = set after the edit (get_value loads unknown voxels from the
    // same DAG), bit 1 = set by the flood fill
    std::vector<uint8_t> state(n);
    for (uint64_t i = 0; i < n; ++i) state[i] = (uint8_t)((st[i] == SetTrue ? 3u : 0u) | (solid[i] ? 1u : 0u));
    fb.release();
    HD_HIP(ctx, hipMalloc(&fb.state, n));
    HD_HIP(ctx, hipMalloc(&fb.colors, nb * 8));
    HD_HIP(ctx, hipMemcpyAsync(fb.state, state.data(), n, hipMemcpyHostToDevice, s));
    HD_HIP(ctx, hipMemcpyAsync(fb.colors, colors.data(), nb * 8, hipMemcpyHostToDevice, s));
    HD_HIP(ctx, hipStreamSynchronize(s));
    d.fill_state = fb.state;
    d.fill_colors = fb.colors;
    d.fill_size = S;
    for (int a = 0; a < 3; ++a) {
        d.src[a] = fmin[a];
        d.fill_lo[a] = lo[a];
        d.fill_dim[a] = dim[a];
        // FillEditorBase: boundsMin / boundsMax from the min corners of the filled voxels
        d.bmin[a] = fmin[a] + (float)lo[a];
        d.bmax[a] = fmin[a] + (float)hi[a];
    }
    return HD_OK;
}

// SphereNoiseEditor (hash_dag_editors.h:762-809): the buffer of a BufferEditor of radius 1.5 r
// (2 ceil(1.5 r) voxels per axis from center - 1.5 r) holds each voxel's value after the edit;
// the permutation is FastNoise::SetSeed's (FastNoise.cpp:207-225) for the editor's seed.
hd_status noise_prepare(hd_ctx* ctx, const hd_edit& e, DevEdit& d, FillBuffers& fb, hipStream_t s) {
    const float r = e.radius, br = 1.5f * r;
    const float ce = std::ceil(br);
    if (!(ce >= 1.0f) || ce > 512.0f) return fail(ctx, HD_ERR_INVALID, "noise radius must be in (0, 341]");
    const uint32_t S = 2u * (uint32_t)ce;
    const uint64_t n = (uint64_t)S * S * S;
    const float fmin[3] = {e.center[0] - br, e.center[1] - br, e.center[2] - br};
    const int seed = std::abs(1923 + (int32_t)e.center[0] + 23 * (int32_t)e.center[1] + 2938 * (int32_t)e.center[2]);
    uint8_t tab[1024];
    {
        std::mt19937_64 gen((std::mt19937_64::result_type)(int64_t)seed);
        uint8_t* perm = tab;
        uint8_t* perm12 = tab + 512;
        for (int i = 0; i < 256; ++i) perm[i] = (uint8_t)i;
        for (int j = 0; j < 256; ++j) {
            const int k = (int)(gen() % (uint64_t)(256 - j)) + j;
            const uint8_t l = perm[j];
            perm[j] = perm[j + 256] = perm[k];
            perm[k] = l;
            perm12[j] = perm12[j + 256] = (uint8_t)(perm[j] % 12);
        }
    }
    // FastNoise::CalculateFractalBounding for 3 octaves of gain 0.5 (FastNoise.cpp:227-237)
    float amp = 0.5f, ampFractal = 1.0f;
    for (int i = 1; i < 3; ++i) {
        ampFractal += amp;
        amp *= 0.5f;
    }
    const float bounding = 1.0f / ampFractal;
    fb.release();
    HD_HIP(ctx, hipMalloc(&fb.state, n));
    HD_HIP(ctx, hipMalloc(&fb.tables, sizeof(tab)));
    HD_HIP(ctx, hipMemcpyAsync(fb.tables, tab, sizeof(tab), hipMemcpyHostToDevice, s));
    k_noise_state<<<(uint32_t)std::min<uint64_t>((n + 255) / 256, 65536), 256, 0, s>>>(
        ctx->vpool, ctx->d_root, ctx->g.levels, fmin[0], fmin[1], fmin[2], S, r, 1 / r, bounding, fb.tables,
        e.kind == EDIT_NOISE_ADD ? 1u : 0u, fb.state);
    HD_HIP(ctx, hipGetLastError());
    HD_HIP(ctx, hipStreamSynchronize(s));
    d.kind = EDIT_NOISE;
    d.radius = br;  // BoxEditorBase::radius, which get_new_color divides by
    d.fill_state = fb.state;
    d.fill_size = S;
    auto rgb565 = [](float rr, float gg, float bb) {
        return (uint32_t)(uint16_t)(((uint16_t)(rr * 31.0f) << 0) | ((uint16_t)(gg * 63.0f) << 5) | ((uint16_t)(bb * 31.0f) << 11));
    };
    d.single = rgb565(94 / 255.f, 30 / 255.f, 0 / 255.f) | (rgb565(125 / 255.f, 100 / 255.f, 89 / 255.f) << 16);
    for (int a = 0; a < 3; ++a) {
        d.c[a] = e.center[a];
        d.src[a] = fmin[a];
        d.bmin[a] = fmin[a];
        d.bmax[a] = e.center[a] + br;
    }
    return HD_OK;
}

hd_status edit_scatter_leaves(hd_ctx* ctx, const uint32_t* handles, const LeafDesc* descs, uint32_t n) {
    if (n == 0) return HD_OK;
    const EditScratch& S = ctx->edit;
    if (n > S.undo_cap) return fail(ctx, HD_ERR_INVALID, "undo frame larger than the undo log");
    HD_HIP(ctx, hipMemcpy(S.undo_handles, handles, (uint64_t)n * 4, hipMemcpyHostToDevice));
    HD_HIP(ctx, hipMemcpy(S.undo_next, descs, (uint64_t)n * sizeof(LeafDesc), hipMemcpyHostToDevice));
    k_scatter_leaves<<<std::min<uint32_t>((n + 255) / 256, 1024), 256>>>(ctx->cleaves, S.undo_handles, S.undo_next, n);
    HD_HIP(ctx, hipGetLastError());
    HD_HIP(ctx, hipDeviceSynchronize());
    return HD_OK;
}

// After a committed pass: copy the n leaf descriptors the pass wrote (undo-log order) to host.
hipError_t edit_gather_next(hd_ctx* ctx, uint32_t n, LeafDesc* host, hipStream_t s) {
    const EditScratch& S = ctx->edit;
    if (n == 0) return hipSuccess;
    n = std::min(n, S.undo_cap);
    k_gather_leaves<<<std::min<uint32_t>((n + 255) / 256, 1024), 256, 0, s>>>(ctx->cleaves, S.undo_handles, S.undo_next, n);
    return hipMemcpyAsync(host, S.undo_next, (uint64_t)n * sizeof(LeafDesc), hipMemcpyDeviceToHost, s);
}

}  // namespace hd
