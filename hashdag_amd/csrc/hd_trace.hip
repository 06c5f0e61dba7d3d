// Tracer kernels for gfx950: primary paths, colours and shadows over the device-resident
// HashDAG. Semantics follow Tracer::trace_paths / trace_colors / trace_shadows
// (tracer.cu:7-697) bit for bit: FP64 ray setup and FP32 slab math in the reference's
// operation order, compiled without FP contraction.
//
// Mapping: one lane per pixel; a 64-lane wavefront covers a tile_w x tile_h pixel tile
// (8 x 8 by default) so neighbouring rays share their upper-level nodes; a 256-thread block
// holds 2 x 2 wave tiles. The DFS stack (node index, child mask, visit mask per level) lives
// in LDS, laid out [level][thread] so a wavefront's accesses are bank-conflict free. Node
// words are read straight from the virtual-address pool (no page-table indirection).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "hd_ctx.h"

namespace hd {
namespace {

constexpr uint32_t kMaxLevels = 17;
constexpr uint32_t kBlock = 256;

struct RayP {
    double cam[3], rmin[3], ddx[3], ddy[3];
};

struct F3 {
    float x, y, z;
};

__device__ __forceinline__ uint32_t child_off(uint32_t mask, uint32_t child) {
    return __popc(mask & ((1u << child) - 1u)) + 1u;
}

// base_dag.h:16-58: bit i of the mask = byte i of the leaf is non-zero (the bytes' bits are
// folded onto their bit 0, then the four bit-0s of each word gathered)
__device__ __forceinline__ uint32_t byte_nonzero4(uint32_t w) {
    w |= w >> 4;
    w |= w >> 2;
    w |= w >> 1;
    w &= 0x01010101u;
    return (w | (w >> 7) | (w >> 14) | (w >> 21)) & 0xFu;
}
__device__ __forceinline__ uint32_t leaf_first_mask(uint64_t leaf) {
    return byte_nonzero4((uint32_t)leaf) | (byte_nonzero4((uint32_t)(leaf >> 32)) << 4);
}
__device__ __forceinline__ uint32_t leaf_second_mask(uint64_t leaf, uint32_t c) {
    const uint32_t w = (c & 4) ? (uint32_t)(leaf >> 32) : (uint32_t)leaf;
    return (w >> ((c & 3) * 8)) & 0xFFu;
}

// tracer.cu:7-17: children are visited as c = k ^ order for k = 0..7. Bit k of the permuted
// mask is bit k ^ order of the mask, so its lowest set bit is the first child in that order.
__device__ __forceinline__ uint32_t permute_order(uint32_t order, uint32_t mask) {
    uint32_t m = mask;
    if (order & 4) m = ((m & 0x0Fu) << 4) | ((m & 0xF0u) >> 4);
    if (order & 2) m = ((m & 0x33u) << 2) | ((m & 0xCCu) >> 2);
    if (order & 1) m = ((m & 0x55u) << 1) | ((m & 0xAAu) >> 1);
    return m;
}
__device__ __forceinline__ uint32_t next_child(uint32_t order, uint32_t mask) {
    return (uint32_t)__builtin_ctz(permute_order(order, mask)) ^ order;
}

__device__ __forceinline__ float tmax3(float x, float y, float z) {
    const float m = y > z ? y : z;
    return x > m ? x : m;
}
__device__ __forceinline__ float tmin3(float x, float y, float z) {
    const float m = y < z ? y : z;
    return x < m ? x : m;
}

// tracer.cu:19-136
template <bool isRoot>
__device__ __forceinline__ uint32_t intersection_mask(uint32_t levels, uint32_t level, uint32_t px, uint32_t py,
                                                      uint32_t pz, F3 o, F3 d, F3 inv) {
    const uint32_t shift = levels - level;
    const float radius = (float)(1u << (shift - 1));
    const float cx = radius + (float)(px << shift);
    const float cy = radius + (float)(py << shift);
    const float cz = radius + (float)(pz << shift);
    const float rx = cx - o.x, ry = cy - o.y, rz = cz - o.z;
    const float tx = rx * inv.x, ty = ry * inv.y, tz = rz * inv.z;
    const float sx = radius * fabsf(inv.x), sy = radius * fabsf(inv.y), sz = radius * fabsf(inv.z);
    const float tmin = fmaxf(tmax3(tx - sx, ty - sy, tz - sz), 0.0f);
    const float tmax = tmin3(tx + sx, ty + sy, tz + sz);
    if (isRoot && (tmin >= tmax)) return 0;
    uint32_t mask;
    {
        const float h = 0.5f * (tmin + tmax);
        const float qx = h * d.x, qy = h * d.y, qz = h * d.z;
        const uint32_t first = (qx >= rx ? 4u : 0u) + (qy >= ry ? 2u : 0u) + (qz >= rz ? 1u : 0u);
        mask = 1u << first;
    }
    // the three face terms are evaluated unconditionally and kept by predicate (same IEEE
    // operations as the branches of tracer.cu:70-130, without divergent control flow)
    const float eps = 1e-4f;
    {
        const float qy = tx * d.y, qz = tx * d.z;
        const uint32_t A = (qy >= ry - eps ? 0xCCu : 0u) | (qy <= ry + eps ? 0x33u : 0u);
        const uint32_t B = (qz >= rz - eps ? 0xAAu : 0u) | (qz <= rz + eps ? 0x55u : 0u);
        mask |= (tmin <= tx && tx <= tmax) ? (A & B) : 0u;
    }
    {
        const float qx = ty * d.x, qz = ty * d.z;
        const uint32_t C = (qx >= rx - eps ? 0xF0u : 0u) | (qx <= rx + eps ? 0x0Fu : 0u);
        const uint32_t D = (qz >= rz - eps ? 0xAAu : 0u) | (qz <= rz + eps ? 0x55u : 0u);
        mask |= (tmin <= ty && ty <= tmax) ? (C & D) : 0u;
    }
    {
        const float qx = tz * d.x, qy = tz * d.y;
        const uint32_t E = (qx >= rx - eps ? 0xF0u : 0u) | (qx <= rx + eps ? 0x0Fu : 0u);
        const uint32_t F = (qy >= ry - eps ? 0xCCu : 0u) | (qy <= ry + eps ? 0x33u : 0u);
        mask |= (tmin <= tz && tz <= tmax) ? (E & F) : 0u;
    }
    return mask;
}

// Node fetch strategies (picked per context by the autotuner):
//   kFetchPointer  a descent loads the child pointer, then the child's header (two dependent
//                  loads of one word each)
//   kFetchNode     a descent loads the child's whole node (header + up to 8 children) with
//                  4-byte-aligned vector loads, so the next descent needs no pointer load; the
//                  pool has padding words after the virtual space, so reading past the last
//                  node is safe
constexpr int kFetchPointer = 0, kFetchNode = 1;

typedef uint32_t u32x4_a4 __attribute__((ext_vector_type(4), aligned(4)));

__device__ __forceinline__ void load_node(const uint32_t* __restrict__ vpool, uint32_t index, uint32_t nd[9]) {
    const u32x4_a4 a = *reinterpret_cast<const u32x4_a4*>(vpool + index);
    const u32x4_a4 b = *reinterpret_cast<const u32x4_a4*>(vpool + index + 4);
    nd[0] = a.x;
    nd[1] = a.y;
    nd[2] = a.z;
    nd[3] = a.w;
    nd[4] = b.x;
    nd[5] = b.y;
    nd[6] = b.z;
    nd[7] = b.w;
    nd[8] = vpool[index + 8];
}

// nd[k] for a lane-varying k in 1..8, as a select tree over constant indices (a dynamic
// index would send the array to scratch memory)
__device__ __forceinline__ uint32_t pick(const uint32_t nd[9], uint32_t k) {
    const uint32_t a = (k & 1u) ? nd[1] : nd[2];
    const uint32_t b = (k & 1u) ? nd[3] : nd[4];
    const uint32_t c = (k & 1u) ? nd[5] : nd[6];
    const uint32_t d = (k & 1u) ? nd[7] : nd[8];
    const uint32_t ab = (k <= 2u) ? a : b;
    const uint32_t cd = (k <= 6u) ? c : d;
    return k <= 4u ? ab : cd;
}

// LDS stack of one block: node index per level 1 .. leafLevel-1 (level 0 is the root; at the
// two leaf levels the 64-bit leaf stays in registers), then (child mask | visit << 8) per level.
struct Stack {
    uint32_t* idx;
    uint16_t* mask;
    __device__ __forceinline__ uint32_t& at_idx(uint32_t level) { return idx[(level - 1) * kBlock + threadIdx.x]; }
    __device__ __forceinline__ uint16_t& at_mask(uint32_t level) { return mask[level * kBlock + threadIdx.x]; }
};

__host__ __device__ __forceinline__ uint32_t stack_bytes(uint32_t levels) {
    return ((levels - 3) * 4u + levels * 2u) * kBlock;
}

__device__ __forceinline__ Stack stack_of(uint32_t levels) {
    extern __shared__ uint32_t sStack[];
    Stack st;
    st.idx = sStack;
    st.mask = reinterpret_cast<uint16_t*>(sStack + (levels - 3) * kBlock);
    return st;
}

// DFS state of one ray: the current node, its child mask and the children left to visit,
// the voxel path, and (bit l of alive) which stack levels still hold children to visit.
struct Walk {
    uint32_t level, px, py, pz, index, childMask, visit, alive;
    uint64_t leaf;
    uint32_t nd[9];  // the current node's words (kFetchNode)
};

constexpr uint32_t kWalkOn = 0, kWalkHit = 1, kWalkMiss = 2;

template <bool kCount, int kFetch>
__device__ __forceinline__ void walk_begin(const DevDag& dag, Walk& w, F3 o, F3 d, F3 inv, uint64_t& bytes) {
    w.level = 0;
    w.px = w.py = w.pz = 0;
    w.index = dag.root;
    w.leaf = 0;
    w.alive = 0;
    if (kFetch == kFetchNode) {
        load_node(dag.vpool, w.index, w.nd);
        w.childMask = w.nd[0] & 0xFFu;
    } else {
        w.childMask = dag.vpool[w.index] & 0xFFu;
    }
    if (kCount) bytes += 4;
    w.visit = w.childMask & intersection_mask<true>(dag.levels, 0, 0, 0, 0, o, d, inv);
}

// One step of the DFS over the DAG (tracer.cu:166-249 ordered first hit; tracer.cu:458-542
// any hit): back up if the current node has nothing left to visit, then descend into the
// next child. Returns kWalkHit with the voxel path in (px, py, pz), kWalkMiss with (0, 0, 0),
// or kWalkOn.
template <bool kCount, int kFetch>
__device__ __forceinline__ uint32_t walk_step(const DevDag& dag, Walk& w, F3 o, F3 d, F3 inv, uint32_t order, Stack st,
                                              uint64_t& bytes) {
    const uint32_t* __restrict__ vpool = dag.vpool;
    const uint32_t levels = dag.levels, leafLevel = dag.leaf_level;
    if (!w.visit) {
        // back up to the deepest level with children left (the DFS backtrack)
        const uint32_t below = w.alive & ((1u << w.level) - 1u);
        if (!below) {
            w.px = w.py = w.pz = 0;
            return kWalkMiss;
        }
        const uint32_t newLevel = 31u - (uint32_t)__builtin_clz(below);
        const uint32_t m = st.at_mask(newLevel);
        w.childMask = m & 0xFFu;
        w.visit = m >> 8;
        const uint32_t up = w.level - newLevel;
        w.px >>= up;
        w.py >>= up;
        w.pz >>= up;
        w.level = newLevel;
        if (w.level < leafLevel) {
            w.index = w.level == 0 ? dag.root : st.at_idx(w.level);
            if (kFetch == kFetchNode) load_node(vpool, w.index, w.nd);
        }
    }
    // children front to back along the ray, also for the any-hit walk: its result does not
    // depend on the order (tracer.cu:493 takes the highest bit first), and front to back
    // reaches the nearest occluder of a shadowed ray first
    const uint32_t next = next_child(order, w.visit);
    w.visit &= ~(1u << next);
    w.px = (w.px << 1) | ((next >> 2) & 1u);
    w.py = (w.py << 1) | ((next >> 1) & 1u);
    w.pz = (w.pz << 1) | (next & 1u);
    // only entries with children left are ever read back
    if (w.visit) {
        if (w.level > 0 && w.level < leafLevel) st.at_idx(w.level) = w.index;
        st.at_mask(w.level) = (uint16_t)(w.childMask | (w.visit << 8));
        w.alive |= 1u << w.level;
    } else {
        w.alive &= ~(1u << w.level);
    }
    w.level++;
    if (w.level == levels) return kWalkHit;
    if (w.level < leafLevel) {
        if (kFetch == kFetchNode) {
            w.index = pick(w.nd, child_off(w.childMask, next));
            load_node(vpool, w.index, w.nd);
            w.childMask = w.nd[0] & 0xFFu;
        } else {
            w.index = vpool[w.index + child_off(w.childMask, next)];
            w.childMask = vpool[w.index] & 0xFFu;
        }
        if (kCount) bytes += 8;
    } else if (w.level == leafLevel) {
        const uint32_t addr = kFetch == kFetchNode ? pick(w.nd, child_off(w.childMask, next))
                                                   : vpool[w.index + child_off(w.childMask, next)];
        w.leaf = *reinterpret_cast<const uint64_t*>(vpool + addr);
        w.childMask = leaf_first_mask(w.leaf);
        if (kCount) bytes += 12;
    } else {
        w.childMask = leaf_second_mask(w.leaf, next);
    }
    w.visit = w.childMask & intersection_mask<false>(levels, w.level, w.px, w.py, w.pz, o, d, inv);
    return kWalkOn;
}

// The whole walk of one ray. Returns true on a hit; (px, py, pz) is the voxel path on a hit,
// (0,0,0) on a miss.
template <bool kCount, int kFetch>
__device__ __forceinline__ bool descend(const DevDag& dag, F3 o, F3 d, F3 inv, uint32_t order, uint32_t& px,
                                        uint32_t& py, uint32_t& pz, Stack st, uint64_t& bytes) {
    Walk w;
    walk_begin<kCount, kFetch>(dag, w, o, d, inv, bytes);
    uint32_t r;
    do {
        r = walk_step<kCount, kFetch>(dag, w, o, d, inv, order, st, bytes);
    } while (r == kWalkOn);
    px = w.px;
    py = w.py;
    pz = w.pz;
    return r == kWalkHit;
}

// descend() for at most `cap` steps: kWalkHit / kWalkMiss as descend, kWalkOn when the walk
// was stopped at the cap (the ray is then traced by the wavefront-per-ray search).
template <int kFetch>
__device__ __forceinline__ uint32_t descend_capped(const DevDag& dag, F3 o, F3 d, F3 inv, uint32_t order, uint32_t& px,
                                                   uint32_t& py, uint32_t& pz, Stack st, uint32_t cap, Walk& w) {
    uint64_t unused = 0;
    walk_begin<false, kFetch>(dag, w, o, d, inv, unused);
    uint32_t r, steps = 0;
    do {
        r = walk_step<false, kFetch>(dag, w, o, d, inv, order, st, unused);
    } while (r == kWalkOn && ++steps < cap);
    px = w.px;
    py = w.py;
    pz = w.pz;
    return r;
}

// Rays whose walk stopped at the cap: appended as pixel indices, one atomic per wavefront.
struct LongRays {
    uint32_t* items;  // pixel indices (x + y * W)
    uint32_t* count;
    uint32_t cap;     // walk steps before a ray is handed over; ~0u: never
    uint32_t* state;  // with a state: the walk of entry i is saved at state[word * stride + i]
    uint32_t stride;
};
// slot of this lane's entry (lanes with give = false get ~0u); every lane of the wavefront calls it
__device__ __forceinline__ uint32_t long_slot(bool give, const LongRays& L) {
    const uint64_t m = __ballot(give);
    if (!m) return ~0u;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t base = 0;
    if (lane == (uint32_t)__ffsll((long long)m) - 1u) base = atomicAdd(L.count, (uint32_t)__popcll(m));
    base = __shfl(base, __ffsll((long long)m) - 1, 64);
    return give ? base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull)) : ~0u;
}
__device__ __forceinline__ void long_append(bool give, uint32_t pix, const LongRays& L) {
    const uint32_t slot = long_slot(give, L);
    if (give) L.items[slot] = pix;
}

// A walk stopped at the cap, saved so another launch continues it (pointer fetch): the walk's
// registers, then the LDS stack's node indices (levels 1 .. leafLevel - 1) and child masks
// (two levels per word). Only levels below the walk's level with children left are read back.
constexpr uint32_t kWalkWords = 10;
constexpr uint32_t kStateWords = kWalkWords + (kMaxLevels - 3) + (kMaxLevels + 1) / 2;
__device__ __forceinline__ void save_walk(const Walk& w, Stack st, uint32_t levels, uint32_t leafLevel,
                                          uint32_t* __restrict__ state, uint32_t stride, uint32_t slot) {
    uint32_t* q = state + slot;
    q[0 * (uint64_t)stride] = w.level;
    q[1 * (uint64_t)stride] = w.px;
    q[2 * (uint64_t)stride] = w.py;
    q[3 * (uint64_t)stride] = w.pz;
    q[4 * (uint64_t)stride] = w.index;
    q[5 * (uint64_t)stride] = w.childMask;
    q[6 * (uint64_t)stride] = w.visit;
    q[7 * (uint64_t)stride] = w.alive;
    q[8 * (uint64_t)stride] = (uint32_t)w.leaf;
    q[9 * (uint64_t)stride] = (uint32_t)(w.leaf >> 32);
    for (uint32_t l = 1; l < leafLevel && l < w.level; ++l) q[(uint64_t)(kWalkWords + l - 1) * stride] = st.at_idx(l);
    const uint32_t m0 = kWalkWords + (kMaxLevels - 3);
    for (uint32_t l = 0; l < levels && l < w.level; l += 2) {
        const uint32_t hi = l + 1 < levels ? st.at_mask(l + 1) : 0u;
        q[(uint64_t)(m0 + l / 2) * stride] = st.at_mask(l) | (hi << 16);
    }
}
__device__ __forceinline__ void restore_walk(Walk& w, Stack st, uint32_t levels, uint32_t leafLevel,
                                             const uint32_t* __restrict__ state, uint32_t stride, uint32_t slot) {
    const uint32_t* q = state + slot;
    w.level = q[0 * (uint64_t)stride];
    w.px = q[1 * (uint64_t)stride];
    w.py = q[2 * (uint64_t)stride];
    w.pz = q[3 * (uint64_t)stride];
    w.index = q[4 * (uint64_t)stride];
    w.childMask = q[5 * (uint64_t)stride];
    w.visit = q[6 * (uint64_t)stride];
    w.alive = q[7 * (uint64_t)stride];
    w.leaf = (uint64_t)q[8 * (uint64_t)stride] | ((uint64_t)q[9 * (uint64_t)stride] << 32);
    for (uint32_t l = 1; l < leafLevel && l < w.level; ++l) st.at_idx(l) = q[(uint64_t)(kWalkWords + l - 1) * stride];
    const uint32_t m0 = kWalkWords + (kMaxLevels - 3);
    for (uint32_t l = 0; l < levels && l < w.level; l += 2) {
        const uint32_t v = q[(uint64_t)(m0 + l / 2) * stride];
        st.at_mask(l) = (uint16_t)v;
        if (l + 1 < levels) st.at_mask(l + 1) = (uint16_t)(v >> 16);
    }
}

__device__ __forceinline__ void dnormalize(double& x, double& y, double& z) {
    const double len = __dsqrt_rn(x * x + y * y + z * z);
    const double s = 1.0 / len;
    x = s * x;
    y = s * y;
    z = s * z;
}

__device__ __forceinline__ void pixel_of(uint32_t tileW, uint32_t tileH, uint32_t& x, uint32_t& y) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    x = blockIdx.x * (2 * tileW) + (wave & 1u) * tileW + lane % tileW;
    y = blockIdx.y * (2 * tileH) + (wave >> 1) * tileH + lane / tileW;
}

template <bool kCount>
__device__ __forceinline__ void flush_count(uint64_t v, unsigned long long* counter) {
    if (!kCount) return;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor((unsigned long long)v, off, 64);
    if ((threadIdx.x & 63u) == 0) atomicAdd(counter, (unsigned long long)v);
}

// Primary ray of pixel (x, y) (tracer.cu:160-180): FP64 direction, FP32 origin, direction and
// inverse direction, child order from the direction's signs.
__device__ __forceinline__ void primary_ray(const RayP& p, uint32_t x, uint32_t y, F3& o, F3& d, F3& inv,
                                            uint32_t& order) {
    double dx = p.rmin[0] + (double)x * p.ddx[0] + (double)y * p.ddy[0] - p.cam[0];
    double dy = p.rmin[1] + (double)x * p.ddx[1] + (double)y * p.ddy[1] - p.cam[1];
    double dz = p.rmin[2] + (double)x * p.ddx[2] + (double)y * p.ddy[2] - p.cam[2];
    dnormalize(dx, dy, dz);
    o = F3{(float)p.cam[0], (float)p.cam[1], (float)p.cam[2]};
    d = F3{(float)dx, (float)dy, (float)dz};
    inv = F3{(float)(1.0 / (double)d.x), (float)(1.0 / (double)d.y), (float)(1.0 / (double)d.z)};
    order = (d.x < 0.f ? 4u : 0u) + (d.y < 0.f ? 2u : 0u) + (d.z < 0.f ? 1u : 0u);
}

// tracer.cu:145-252
// L.cap < ~0u: walks stopped at L.cap steps leave their pixel to the wavefront-per-ray search
// (appended to L) and write nothing.
template <bool kCount, int kFetch>
__global__ __launch_bounds__(kBlock) void k_trace_paths(DevDag dag, RayP p, const RayP* __restrict__ pmem, uint32_t W, uint32_t H,
                                                        uint4* __restrict__ out, uint32_t tileW, uint32_t tileH,
                                                        unsigned long long* counter, LongRays L) {
    const Stack st = stack_of(dag.levels);
    if (pmem) p = *pmem;
    if (dag.root_ptr) dag.root = *dag.root_ptr;
    uint32_t x, y;
    pixel_of(tileW, tileH, x, y);
    uint64_t bytes = 0;
    bool give = false;
    Walk w;
    if (x < W && y < H) {
        F3 o, d, inv;
        uint32_t order;
        primary_ray(p, x, y, o, d, inv, order);
        uint32_t vx, vy, vz;
        if (kCount || L.cap == ~0u) {
            descend<kCount, kFetch>(dag, o, d, inv, order, vx, vy, vz, st, bytes);
        } else {
            give = descend_capped<kFetch>(dag, o, d, inv, order, vx, vy, vz, st, L.cap, w) == kWalkOn;
        }
        if (!give) out[(uint64_t)(H - 1 - y) * W + x] = make_uint4(vx, vy, vz, 0u);
        if (kCount) bytes += 16;
    }
    if (!kCount && L.cap != ~0u) {
        const uint32_t slot = long_slot(give, L);
        if (give) {
            L.items[slot] = y * W + x;
            if (L.state) save_walk(w, st, dag.levels, dag.leaf_level, L.state, L.stride, slot);
        }
    }
    flush_count<kCount>(bytes, counter);
}

// ---- colours ----

__device__ __forceinline__ float clamp01(float v) { return v < 0.f ? 0.f : (v > 1.f ? 1.f : v); }

// color_utils.h:28-37
__device__ __forceinline__ uint32_t to_rgb888(float r, float g, float b) {
    r = clamp01(r);
    g = clamp01(g);
    b = clamp01(b);
    return ((uint32_t)(r * 255.0f)) | ((uint32_t)(g * 255.0f) << 8) | ((uint32_t)(b * 255.0f) << 16) | 0xff000000u;
}

struct Leaf {
    const uint64_t* blocks;
    const uint64_t* macro;
    const uint32_t* weights;
    uint64_t n_blocks, n_macro, offset;
};

__device__ __forceinline__ uint32_t hdr_index(uint32_t h) { return h & 16383u; }
__device__ __forceinline__ uint32_t hdr_bpw(uint32_t h) { return (h >> 16) == 0xFFFFu ? 0u : ((h >> 14) & 3u) + 1u; }

// color_utils.h:112-148 (big-endian 16-bit window over the byte-swapped weight words)
template <bool kCount>
__device__ __forceinline__ uint32_t extract_bits(uint32_t bits, const uint32_t* __restrict__ w, uint64_t bitPtr,
                                                 uint64_t& bytes) {
    const uint64_t byte = bitPtr >> 3;
    const uint32_t w0 = w[byte >> 2];
    const uint32_t sh = (uint32_t)(byte & 3u) * 8u;
    uint32_t b0 = (w0 >> sh) & 0xFFu, b1;
    if ((byte & 3u) == 3u) {
        b1 = w[(byte >> 2) + 1] & 0xFFu;
        if (kCount) bytes += 4;
    } else {
        b1 = (w0 >> (sh + 8)) & 0xFFu;
    }
    if (kCount) bytes += 4;
    const uint32_t dst = (b0 << 8) | b1;
    const uint32_t shift = 16u - bits - (uint32_t)(bitPtr & 7u);
    return (dst >> shift) & ((1u << bits) - 1u);
}

// variable_weight_size_colors.h:343-413 + CompressedColor::get_color (:144-154)
template <bool kCount>
__device__ __forceinline__ void leaf_color(const Leaf& L, uint64_t colorIndex, float& r, float& g, float& b,
                                           uint64_t& bytes) {
    if (L.offset != ~0ull) colorIndex += L.offset;
    const uint32_t local = (uint32_t)(colorIndex & 16383u);
    const uint64_t m = colorIndex >> 14;
    uint32_t lo = (uint32_t)L.macro[2 * m];
    uint32_t hi = 2 * (m + 1) < L.n_macro ? (uint32_t)L.macro[2 * (m + 1)] - 1u : (uint32_t)L.n_blocks - 1u;
    if (kCount) bytes += 16;
    uint32_t pos = (lo + hi) / 2;
    uint64_t blk = L.blocks[pos];
    if (kCount) bytes += 8;
    while (hdr_index((uint32_t)blk) != local && lo <= hi) {
        if (hdr_index((uint32_t)blk) > local) hi = pos - 1;
        else lo = pos + 1;
        pos = (lo + hi) / 2;
        blk = L.blocks[pos];
        if (kCount) bytes += 8;
    }
    const uint32_t h = (uint32_t)blk, bits = (uint32_t)(blk >> 32);
    const uint32_t bpw = hdr_bpw(h);
    if (bpw == 0) {
        r = (float)((bits >> 0) & 0x3FFu) / 1023.0f;
        g = (float)((bits >> 10) & 0xFFFu) / 4095.0f;
        b = (float)((bits >> 22) & 0x3FFu) / 1023.0f;
        return;
    }
    const uint64_t macroWeight = L.macro[2 * m + 1];
    if (kCount) bytes += 8;
    const uint64_t bit = macroWeight + (h >> 16) + (uint64_t)((local - hdr_index(h)) * bpw);
    const uint32_t wq = extract_bits<kCount>(bpw, L.weights, bit, bytes);
    const float f = (float)wq / (float)((1u << bpw) - 1u);
    const float r0 = (float)(bits & 0x1Fu) / 31.0f, g0 = (float)((bits >> 5) & 0x3Fu) / 63.0f,
                b0 = (float)((bits >> 11) & 0x1Fu) / 31.0f;
    const float r1 = (float)((bits >> 16) & 0x1Fu) / 31.0f, g1 = (float)((bits >> 21) & 0x3Fu) / 63.0f,
                b1 = (float)((bits >> 27) & 0x1Fu) / 31.0f;
    const float omf = 1 - f;
    r = r0 * omf + r1 * f;
    g = g0 * omf + g1 * f;
    b = b0 * omf + b1 * f;
}

__device__ __forceinline__ Leaf make_leaf(const DevColors& c, const LeafDesc& d, uint64_t offset) {
    Leaf L;
    L.blocks = c.arena + d.blocks;
    L.macro = c.arena + d.macro;
    L.weights = reinterpret_cast<const uint32_t*>(c.arena) + d.weights32;
    L.n_blocks = d.n_blocks;
    L.n_macro = d.n_macro;
    L.offset = offset;
    return L;
}

// tracer.cu:254-451 (debugColors = None, TOOL_OVERLAY = 0)
template <bool kCount>
__global__ __launch_bounds__(kBlock) void k_trace_colors(DevDag dag, DevColors colors, uint32_t W, uint32_t H,
                                                         const uint4* __restrict__ paths, uint32_t* __restrict__ out,
                                                         uint32_t tileW, uint32_t tileH, unsigned long long* counter) {
    if (dag.root_ptr) dag.root = *dag.root_ptr;
    uint32_t x, y;
    pixel_of(tileW, tileH, x, y);
    uint64_t bytes = 0;
    if (x < W && y < H) {
        const uint64_t pix = (uint64_t)y * W + x;
        const uint4 pp = paths[pix];
        if (kCount) bytes += 20;  // path read + colour write
        uint32_t result;
        const uint32_t* __restrict__ vpool = dag.vpool;
        if (pp.x == 0 && pp.y == 0 && pp.z == 0) {
            result = to_rgb888(187.0f / 255.f, 242.0f / 255.f, 250.0f / 255.f);
        } else {
            const uint32_t cb = (pp.x ^ pp.y ^ pp.z) & 1u;
            const uint32_t invalid = to_rgb888(1.0f, (float)cb, 1.f - (float)cb);
            const uint32_t ctl = dag.color_tree_levels, L = dag.leaf_level, levels = dag.levels;
            uint64_t nof = 0;
            uint32_t colorNode = 0;
            bool haveLeaf = false;
            Leaf leaf{};
            uint32_t level = 0;
            uint32_t nodeIndex = dag.root;
            result = 0;
            bool done = false;
            while (level < L) {
                level++;
                const uint32_t node = vpool[nodeIndex];
                const uint32_t mask = node & 0xFFu;
                const uint32_t s = levels - level;
                const uint32_t child = (((pp.x >> s) & 1u) << 2) | (((pp.y >> s) & 1u) << 1) | ((pp.z >> s) & 1u);
                if (kCount) bytes += 4;
                if (!(mask & (1u << child))) {
                    result = 0xFF00FFu;
                    done = true;
                    break;
                }
                if (level - 1 < ctl) {
                    colorNode = colors.nodes[colorNode + child];
                    if (kCount) bytes += 4;
                    if (level == ctl) {
                        if (colorNode & 0x80000000u) {
                            const LeafDesc dsc = colors.leaves[colorNode & 0x7FFFFFFFu];
                            leaf = make_leaf(colors, dsc, dsc.offset);
                            haveLeaf = dsc.n_blocks > 0;
                        } else {
                            leaf = make_leaf(colors, colors.main, colors.offsets[colorNode]);
                            haveLeaf = colors.main.n_blocks > 0;
                        }
                        if (kCount) bytes += 8;
                    } else if (!colorNode) {
                        result = invalid;
                        done = true;
                        break;
                    }
                }
                if (level == L) {
                    for (uint32_t c = 0; c < child; ++c) {
                        if (mask & (1u << c)) {
                            const uint32_t ci = vpool[nodeIndex + child_off(mask, c)];
                            nof += __popcll(*reinterpret_cast<const uint64_t*>(vpool + ci));
                            if (kCount) bytes += 12;
                        }
                    }
                    const uint32_t ci = vpool[nodeIndex + child_off(mask, child)];
                    const uint64_t lf = *reinterpret_cast<const uint64_t*>(vpool + ci);
                    if (kCount) bytes += 12;
                    const uint32_t bit = ((pp.x & 1u) ? 4u : 0u) | ((pp.y & 1u) ? 2u : 0u) | ((pp.z & 1u) ? 1u : 0u) |
                                         ((pp.x & 2u) ? 32u : 0u) | ((pp.y & 2u) ? 16u : 0u) | ((pp.z & 2u) ? 8u : 0u);
                    nof += __popcll(lf & ((1ull << bit) - 1ull));
                    break;
                } else {
                    if (level > ctl) {
                        for (uint32_t c = 0; c < child; ++c) {
                            if (mask & (1u << c)) {
                                nof += vpool[vpool[nodeIndex + child_off(mask, c)]] >> 8;
                                if (kCount) bytes += 8;
                            }
                        }
                    }
                    nodeIndex = vpool[nodeIndex + child_off(mask, child)];
                    if (kCount) bytes += 4;
                }
            }
            if (!done) {
                // is_valid / is_valid_index (variable_weight_size_colors.h:247-262)
                if (!haveLeaf || !(2 * (nof + (leaf.offset != ~0ull ? leaf.offset : 0ull)) / 16384u < leaf.n_macro)) {
                    result = invalid;
                } else {
                    float r, g, b;
                    leaf_color<kCount>(leaf, nof, r, g, b, bytes);
                    result = to_rgb888(r, g, b);
                }
            }
        }
        out[pix] = result;
    }
    flush_count<kCount>(bytes, counter);
}

// ---- shadows ----

__device__ __forceinline__ void sun_direction(float& x, float& y, float& z) {
    // normalize(make_float3(0.3f, 1.f, 0.5f)) with (1 / length) * v (cuda_math.h:51-52)
    const float len = __fsqrt_rn(0.3f * 0.3f + 1.f * 1.f + 0.5f * 0.5f);
    const float s = __fdiv_rn(1.0f, len);
    x = s * 0.3f;
    y = s * 1.f;
    z = s * 0.5f;
}

// Camera ray direction of pixel (x, yy) in FP64 (tracer.cu:630-636).
__device__ __forceinline__ void camera_dir(const RayP& p, uint32_t x, uint32_t yy, double& cx, double& cy, double& cz) {
    cx = p.rmin[0] + (double)x * p.ddx[0] + (double)yy * p.ddy[0] - p.cam[0];
    cy = p.rmin[1] + (double)x * p.ddx[1] + (double)yy * p.ddy[1] - p.cam[1];
    cz = p.rmin[2] + (double)x * p.ddx[2] + (double)yy * p.ddy[2] - p.cam[2];
    dnormalize(cx, cy, cz);
}

// Shadow ray origin: the exact FP64 hit point on the voxel's box (ray_box_intersection,
// tracer.cu:574-587) moved by bias along the sun direction (tracer.cu:640-655).
__device__ __forceinline__ F3 shadow_origin(const RayP& p, uint4 pp, double cx, double cy, double cz, float bias,
                                            F3 sun) {
    const double bx = (double)(float)pp.x, by = (double)(float)pp.y, bz = (double)(float)pp.z;
    const double t0x = (bx - p.cam[0]) / cx, t0y = (by - p.cam[1]) / cy, t0z = (bz - p.cam[2]) / cz;
    const double t1x = (bx + 1 - p.cam[0]) / cx, t1y = (by + 1 - p.cam[1]) / cy, t1z = (bz + 1 - p.cam[2]) / cz;
    const double mx = t0x < t1x ? t0x : t1x, my = t0y < t1y ? t0y : t1y, mz = t0z < t1z ? t0z : t1z;
    const double maxmin = fmax(fmax(mx, my), mz);
    const float hx = (float)(p.cam[0] + cx * maxmin);
    const float hy = (float)(p.cam[1] + cy * maxmin);
    const float hz = (float)(p.cam[2] + cz * maxmin);
    return F3{hx + bias * sun.x, hy + bias * sun.y, hz + bias * sun.z};
}

// Distance from the camera to the hit voxel's corner and the direction to it (tracer.cu:656-660).
__device__ __forceinline__ void hit_distance(const RayP& p, uint4 pp, double& distance, double& nx, double& ny,
                                             double& nz) {
    const double bx = (double)(float)pp.x, by = (double)(float)pp.y, bz = (double)(float)pp.z;
    const double vxd = bx - p.cam[0], vyd = by - p.cam[1], vzd = bz - p.cam[2];
    distance = __dsqrt_rn(vxd * vxd + vyd * vyd + vzd * vzd);
    nx = vxd / distance;
    ny = vyd / distance;
    nz = vzd / distance;
}

// Shading (tracer.cu:604-619) and applyFog (tracer.cu:551-572).
__device__ __forceinline__ uint32_t shade(uint32_t colorIn, float light, double distance, double nx, double ny,
                                          double nz, float fog, F3 sun) {
    const float k = clamp01(0.5f + light);
    float r = (float)((colorIn >> 0) & 0xFFu) / 255.0f * k;
    float g = (float)((colorIn >> 8) & 0xFFu) / 255.0f * k;
    float b = (float)((colorIn >> 16) & 0xFFu) / 255.0f * k;
    const float fd = fog * 0.00001f;
    const double fogAmount = 1.0 - exp(-distance * (double)fd);
    const double sunAmount = (double)1.01f * fmax(nx * (double)sun.x + ny * (double)sun.y + nz * (double)sun.z, 0.0);
    const float sa = (float)pow(sunAmount, 30.0);
    const float oms = 1 - sa;
    const float fr = 187.0f / 255.f * oms + 1.0f * sa;
    const float fgc = 242.0f / 255.f * oms + 1.0f * sa;
    const float fb = 250.0f / 255.f * oms + 1.0f * sa;
    const float fa = clamp01((float)fogAmount);
    const float omf = 1 - fa;
    r = r * omf + fr * fa;
    g = g * omf + fgc * fa;
    b = b * omf + fb * fa;
    return to_rgb888(r, g, b);
}

// shade() without fog: the colour scaled by clamp(0.5 + light, 0, 1).
__device__ __forceinline__ uint32_t shade_clear(uint32_t colorIn, float light) {
    const float k = clamp01(0.5f + light);
    const float r = (float)((colorIn >> 0) & 0xFFu) / 255.0f * k;
    const float g = (float)((colorIn >> 8) & 0xFFu) / 255.0f * k;
    const float b = (float)((colorIn >> 16) & 0xFFu) / 255.0f * k;
    return to_rgb888(r, g, b);
}

__device__ __forceinline__ bool null_path(uint4 pp) { return pp.x == 0 && pp.y == 0 && pp.z == 0; }

// tracer.cu:589-697
template <bool kCount, int kFetch>
__global__ __launch_bounds__(kBlock) void k_trace_shadows(DevDag dag, RayP p, const RayP* __restrict__ pmem, uint32_t W, uint32_t H,
                                                          const uint4* __restrict__ paths, uint32_t* __restrict__ colors,
                                                          float bias, float fog, uint32_t tileW, uint32_t tileH,
                                                          unsigned long long* counter, LongRays L) {
    const Stack st = stack_of(dag.levels);
    if (pmem) p = *pmem;
    if (dag.root_ptr) dag.root = *dag.root_ptr;
    uint32_t x, y;
    pixel_of(tileW, tileH, x, y);
    uint64_t bytes = 0;
    bool give = false;
    Walk w;
    if (x < W && y < H) {
        const uint64_t pix = (uint64_t)y * W + x;
        const uint4 pp = paths[pix];
        const uint32_t colorIn = colors[pix];
        if (kCount) bytes += 24;
        F3 sun;
        sun_direction(sun.x, sun.y, sun.z);
        // With zero fog density applyFog leaves the colour as it is (fogAmount = 1 - exp(-0) = 0,
        // so r * 1 + fogColor * 0 == r), and the FP64 fog terms are skipped.
        const bool noFog = fog == 0.f;
        if (null_path(pp)) {  // length(rayOrigin) == 0 (tracer.cu:644)
            if (noFog) {
                colors[pix] = shade_clear(colorIn, 1.f);
            } else {
                double cx, cy, cz;
                camera_dir(p, x, H - 1 - y, cx, cy, cz);
                colors[pix] = shade(colorIn, 1.f, 1e9, cx, cy, cz, fog, sun);
            }
        } else {
            double cx, cy, cz;
            camera_dir(p, x, H - 1 - y, cx, cy, cz);
            const F3 o = shadow_origin(p, pp, cx, cy, cz, bias, sun);
            const F3 inv{(float)(1.0 / (double)sun.x), (float)(1.0 / (double)sun.y), (float)(1.0 / (double)sun.z)};
            const uint32_t order = (sun.x < 0.f ? 4u : 0u) + (sun.y < 0.f ? 2u : 0u) + (sun.z < 0.f ? 1u : 0u);
            uint32_t vx, vy, vz;
            bool shadowed;
            if (kCount || L.cap == ~0u) {
                shadowed = descend<kCount, kFetch>(dag, o, sun, inv, order, vx, vy, vz, st, bytes);
            } else {
                const uint32_t r = descend_capped<kFetch>(dag, o, sun, inv, order, vx, vy, vz, st, L.cap, w);
                give = r == kWalkOn;
                shadowed = r == kWalkHit;
            }
            if (give) {
            } else if (noFog) {
                colors[pix] = shade_clear(colorIn, shadowed ? 0.f : 1.f);
            } else {
                double distance, nx, ny, nz;
                hit_distance(p, pp, distance, nx, ny, nz);
                colors[pix] = shade(colorIn, shadowed ? 0.f : 1.f, distance, nx, ny, nz, fog, sun);
            }
        }
    }
    if (!kCount && L.cap != ~0u) {
        const uint32_t slot = long_slot(give, L);
        if (give) {
            L.items[slot] = y * W + x;
            if (L.state) save_walk(w, st, dag.levels, dag.leaf_level, L.state, L.stride, slot);
        }
    }
    flush_count<kCount>(bytes, counter);
}

// ---- persistent wavefronts fed from tile queues ----
//
// The per-tile launch keeps a wavefront alive until its slowest ray ends, so lanes whose rays
// ended (or, for shadows, that hold sky pixels) sit idle. Here every wavefront stays resident
// and takes 64-pixel tiles (the context's tile shape) from one of kQueues queues, each a
// contiguous run of the image's tiles; once `refillAt` or more of its lanes are idle it hands
// them the untaken pixels of its current tile, then of the next one. The queue of a wavefront
// is picked from the XCD its block runs on, so one XCD's L2 serves one region of the image; a
// wavefront whose queue is empty moves on to the next queue. Each ray's walk is the per-tile
// kernel's walk, so every pixel's result is the same.
constexpr uint32_t kQueues = 64;
constexpr uint32_t kQueueStride = 32;  // words between queue heads (one 128-byte line each)

struct TileQ {
    uint32_t* heads;  // kQueues * kQueueStride words, zeroed before each launch
    uint32_t ntiles, tilesX;
};

__device__ __forceinline__ uint32_t queue_begin(uint32_t q, uint32_t ntiles) {
    return (uint32_t)(((uint64_t)ntiles * q) / kQueues);
}

// Next tile for this wavefront (wave-uniform), moving from queue q to the following ones as they
// run dry; ~0u once all kQueues queues have been found dry.
__device__ __forceinline__ uint32_t grab_tile(const TileQ& Q, uint32_t& q, uint32_t& tried) {
    while (tried < kQueues) {
        const uint32_t n = queue_begin(q + 1, Q.ntiles) - queue_begin(q, Q.ntiles);
        uint32_t* h = Q.heads + q * kQueueStride;
        const uint32_t seen = __builtin_amdgcn_readfirstlane(__hip_atomic_load(h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (seen < n) {
            uint32_t t = 0;
            if ((threadIdx.x & 63u) == 0) t = atomicAdd(h, 1u);
            t = __builtin_amdgcn_readfirstlane(t);
            if (t < n) return queue_begin(q, Q.ntiles) + t;
        }
        q = q + 1 == kQueues ? 0 : q + 1;
        ++tried;
    }
    return ~0u;
}

// position of the n-th (from 0) set bit of m (n < popcount(m))
__device__ __forceinline__ uint32_t nth_set_bit(uint64_t m, uint32_t n) {
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t width = 32; width >= 1; width >>= 1) {
        const uint32_t c = (uint32_t)__popcll(m & ((1ull << width) - 1ull));
        if (n >= c) {
            n -= c;
            m >>= width;
            pos += width;
        }
    }
    return pos;
}

__global__ void k_queue_reset(uint32_t* heads) {
    for (uint32_t i = threadIdx.x; i < kQueues * kQueueStride; i += blockDim.x) heads[i] = 0;
}

// Shadow rays of a frame for the persistent walk (tracer.cu:589-697 up to the walk): per pixel
// tile, the hit pixels' shadow origins and pixel indices packed to the front of the tile's 64
// records (count in cnt), with the colour each gets lit and in shadow; sky pixels (null paths,
// tracer.cu:644) are shaded here. One wavefront per tile, tiles numbered ty * tilesX + tx.
template <bool kFog>
__global__ __launch_bounds__(kBlock) void k_shadow_setup(RayP p, const RayP* __restrict__ pmem, uint32_t W, uint32_t H,
                                                         const uint4* __restrict__ paths, uint32_t* __restrict__ colors,
                                                         float bias, float fog, uint32_t tileW, uint32_t tileH,
                                                         uint32_t tilesX, uint32_t tilesY, uint4* __restrict__ rec,
                                                         uint2* __restrict__ shades, uint32_t* __restrict__ cnt) {
    if (pmem) p = *pmem;
    uint32_t x, y;
    pixel_of(tileW, tileH, x, y);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t tx = blockIdx.x * 2u + (wave & 1u), ty = blockIdx.y * 2u + (wave >> 1);
    const uint32_t t = ty * tilesX + tx;
    bool valid = false;
    uint4 r{};
    uint2 sh{};
    if (x < W && y < H) {
        const uint32_t pix = y * W + x;
        const uint4 pp = paths[pix];
        const uint32_t colorIn = colors[pix];
        F3 sun;
        sun_direction(sun.x, sun.y, sun.z);
        double cx, cy, cz;
        if (null_path(pp)) {
            if (kFog) {
                camera_dir(p, x, H - 1 - y, cx, cy, cz);
                colors[pix] = shade(colorIn, 1.f, 1e9, cx, cy, cz, fog, sun);
            } else {
                colors[pix] = shade_clear(colorIn, 1.f);
            }
        } else {
            camera_dir(p, x, H - 1 - y, cx, cy, cz);
            const F3 o = shadow_origin(p, pp, cx, cy, cz, bias, sun);
            r = make_uint4(__float_as_uint(o.x), __float_as_uint(o.y), __float_as_uint(o.z), pix);
            if (kFog) {
                double distance, nx, ny, nz;
                hit_distance(p, pp, distance, nx, ny, nz);
                sh = make_uint2(shade(colorIn, 1.f, distance, nx, ny, nz, fog, sun),
                                shade(colorIn, 0.f, distance, nx, ny, nz, fog, sun));
            } else {
                sh = make_uint2(shade_clear(colorIn, 1.f), shade_clear(colorIn, 0.f));
            }
            valid = true;
        }
    }
    const uint64_t m = __ballot(valid);
    if (valid) {
        const uint32_t k = t * 64u + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        rec[k] = r;
        shades[k] = sh;
    }
    if (lane == 0 && tx < tilesX && ty < tilesY) cnt[t] = (uint32_t)__popcll(m);
}

// kShadow = false: tracer.cu:145-252 per pixel of each tile (paths out). kShadow = true: the
// any-hit walk of tracer.cu:453-543 for the records k_shadow_setup packed per tile, then the
// pixel's lit or shadowed colour.
template <bool kShadow, int kFetch>
__global__ __launch_bounds__(kBlock) void k_trace_persist(DevDag dag, RayP p, const RayP* __restrict__ pmem, uint32_t W,
                                                          uint32_t H, uint4* __restrict__ paths,
                                                          uint32_t* __restrict__ colors, const uint4* __restrict__ rec,
                                                          const uint2* __restrict__ shades,
                                                          const uint32_t* __restrict__ cnt, uint32_t tileW,
                                                          uint32_t tileH, TileQ Q, uint32_t refillAt) {
    const Stack st = stack_of(dag.levels);
    if (!kShadow && pmem) p = *pmem;
    if (dag.root_ptr) dag.root = *dag.root_ptr;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    // blocks go to the XCDs round robin: the 8 queues of XCD blockIdx % 8 cover one eighth of
    // the image
    uint32_t q = (blockIdx.x % 8u) * 8u + ((blockIdx.x / 8u) * 4u + wave) % 8u;
    uint32_t tried = 0;
    bool exhausted = false;
    uint64_t cand = 0;      // lanes holding an untaken candidate of the current tile
    uint32_t candWork = 0;  // this lane's candidate: pixel (paths) or record index (shadows)
    bool active = false;
    uint32_t pix = 0;
    uint2 sh{};
    Walk w;
    F3 o{}, d{}, inv{};
    uint32_t order = 0;
    if (kShadow) {
        sun_direction(d.x, d.y, d.z);
        inv = F3{(float)(1.0 / (double)d.x), (float)(1.0 / (double)d.y), (float)(1.0 / (double)d.z)};
        order = (d.x < 0.f ? 4u : 0u) + (d.y < 0.f ? 2u : 0u) + (d.z < 0.f ? 1u : 0u);
    }
    uint64_t unused = 0;
    for (;;) {
        uint64_t idle = __ballot(!active);
        if (!exhausted && (uint32_t)__popcll(idle) >= refillAt) {
            while (idle) {
                if (!cand) {
                    const uint32_t t = grab_tile(Q, q, tried);
                    if (t == ~0u) {
                        exhausted = true;
                        break;
                    }
                    bool ok;
                    if (kShadow) {
                        const uint32_t n = __builtin_amdgcn_readfirstlane(cnt[t]);
                        ok = lane < n;
                        candWork = t * 64u + lane;
                    } else {
                        const uint32_t x = (t % Q.tilesX) * tileW + lane % tileW;
                        const uint32_t y = (t / Q.tilesX) * tileH + lane / tileW;
                        ok = x < W && y < H;
                        candWork = y * W + x;
                    }
                    cand = __ballot(ok);
                    continue;
                }
                // the idle lanes, in lane order, take the candidates in lane order
                const uint32_t k = min((uint32_t)__popcll(idle), (uint32_t)__popcll(cand));
                const uint32_t cr = (uint32_t)__popcll(cand & below), ir = (uint32_t)__popcll(idle & below);
                const bool give = ((cand >> lane) & 1ull) && cr < k;
                const bool take = ((idle >> lane) & 1ull) && ir < k;
                // every lane reads a lane (its own when not taking): the ir-th candidate's
                const uint32_t src = take ? nth_set_bit(cand, ir) : lane;
                const uint32_t work = (uint32_t)__shfl((int)candWork, (int)src, 64);
                cand &= ~__ballot(give);
                idle &= ~__ballot(take);
                if (take) {
                    if (kShadow) {
                        const uint4 r = rec[work];
                        sh = shades[work];
                        o = F3{__uint_as_float(r.x), __uint_as_float(r.y), __uint_as_float(r.z)};
                        pix = r.w;
                    } else {
                        pix = work;
                        primary_ray(p, pix % W, pix / W, o, d, inv, order);
                    }
                    walk_begin<false, kFetch>(dag, w, o, d, inv, unused);
                    active = true;
                }
            }
        }
        if (!__ballot(active)) {
            if (exhausted) break;
            continue;
        }
        if (active) {
            const uint32_t r = walk_step<false, kFetch>(dag, w, o, d, inv, order, st, unused);
            if (r != kWalkOn) {
                active = false;
                if (kShadow) {
                    colors[pix] = r == kWalkHit ? sh.y : sh.x;
                } else {
                    const uint32_t x = pix % W, y = pix / W;
                    paths[(uint64_t)(H - 1 - y) * W + x] = make_uint4(w.px, w.py, w.pz, 0u);
                }
            }
        }
    }
}

// ---- wavefront per ray: ordered beam search for the rays the per-tile walk handed over ----
//
// A few rays per frame take hundreds of DFS steps (rays grazing the terrain), and the wavefront
// holding one lives that long: the per-tile walks stop at a step cap and hand those rays here.
// The 64 lanes search one ray together. The frontier holds the next (at most) 64 unexpanded
// nodes of the ray's DFS in the reference's order, one per lane; each round every lane expands
// its node into the children the node's intersection mask lets through, in the order
// child ^ order, and the concatenation in lane order is again in DFS order (the subtrees of
// frontier nodes are disjoint and ordered). The first 64 form the next frontier, the rest go
// onto a per-wavefront ring whose top is the smallest; a short frontier takes entries back
// from the top. The first hit is found when the head of the frontier is a voxel (nothing
// before it in DFS order is left); an any-hit search stops at any voxel. When the ring would
// overflow it drops its largest entries and a search that then ends without a hit is retried
// by the plain walk (k_trace_rest); so is one that runs out of rounds.
constexpr uint32_t kBeamRing = 512;  // ring entries per wavefront (a power of two)
constexpr uint32_t kBeamRounds = 4096;

// entry = node index (or leaf address), a = px | (py low 15 bits) << 17, b = py >> 15 | pz << 2
// | level << 19 (paths have at most 17 bits per axis)
__device__ __forceinline__ void beam_unpack(uint32_t a, uint32_t b, uint32_t& level, uint32_t& px, uint32_t& py,
                                            uint32_t& pz) {
    px = a & 0x1FFFFu;
    py = (a >> 17) | ((b & 3u) << 15);
    pz = (b >> 2) & 0x1FFFFu;
    level = b >> 19;
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct BeamLds {
    uint32_t* f;  // 3 x 64: frontier staging (index, a, b)
    uint32_t* r;  // 3 x kBeamRing: ring (index, a, b)
};

// Returns 1 on a hit (voxel path in hx, hy, hz), 0 on a miss, 2 when the walk must be retried.
// Wave-uniform: every lane calls it with the same ray.
template <bool kAnyHit>
__device__ uint32_t beam_search(const DevDag& dag, F3 o, F3 d, F3 inv, uint32_t order, BeamLds lds, uint32_t& hx,
                                uint32_t& hy, uint32_t& hz) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t levels = dag.levels, L = dag.leaf_level;
    const uint32_t* __restrict__ vpool = dag.vpool;
    constexpr uint32_t kMask = kBeamRing - 1u;
    bool valid = lane == 0;
    uint32_t level = 0, px = 0, py = 0, pz = 0, index = dag.root;
    uint32_t top = 0, bottom = 0;  // ring holds [bottom, top); top - 1 is the smallest entry
    bool truncated = false;
    for (uint32_t round = 0; round < kBeamRounds; ++round) {
        if (!__ballot(valid)) return truncated ? 2u : 0u;
        if (kAnyHit) {
            const uint64_t hm = __ballot(valid && level == levels);
            if (hm) {
                const int src = __ffsll((long long)hm) - 1;
                hx = (uint32_t)__shfl((int)px, src, 64);
                hy = (uint32_t)__shfl((int)py, src, 64);
                hz = (uint32_t)__shfl((int)pz, src, 64);
                return 1u;
            }
        } else if ((uint32_t)__shfl((int)level, 0, 64) == levels) {  // lane 0 holds the head
            hx = (uint32_t)__shfl((int)px, 0, 64);
            hy = (uint32_t)__shfl((int)py, 0, 64);
            hz = (uint32_t)__shfl((int)pz, 0, 64);
            return 1u;
        }
        // expand: the children the node's mask lets through, or the voxel entry itself
        uint32_t cm = 0, visit = 0, cnt = 0;
        uint32_t nd[9];
        if (valid && level < levels) {
            if (level < L) {
                load_node(vpool, index, nd);
                cm = nd[0] & 0xFFu;
            } else {
                const uint64_t leaf = *reinterpret_cast<const uint64_t*>(vpool + index);
                cm = level == L ? leaf_first_mask(leaf)
                                : leaf_second_mask(leaf, ((px & 1u) << 2) | ((py & 1u) << 1) | (pz & 1u));
            }
            const uint32_t im = level == 0 ? intersection_mask<true>(levels, 0, 0, 0, 0, o, d, inv)
                                           : intersection_mask<false>(levels, level, px, py, pz, o, d, inv);
            visit = cm & im;
            cnt = (uint32_t)__popc(visit);
        } else if (valid) {
            cnt = 1;
        }
        uint32_t incl = cnt;
#pragma unroll
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t t = (uint32_t)__shfl_up((int)incl, off, 64);
            if (lane >= off) incl += t;
        }
        const uint32_t T = (uint32_t)__shfl((int)incl, 63, 64);
        const uint32_t spilled = T > 64u ? T - 64u : 0u;
        if (top - bottom + spilled > kBeamRing) {
            bottom = top + spilled - kBeamRing;  // the largest entries make room
            truncated = true;
        }
        uint32_t pos = incl - cnt;
        auto put = [&](uint32_t idx, uint32_t lv, uint32_t x, uint32_t y, uint32_t z) {
            const uint32_t a = x | (y << 17), b = (y >> 15) | (z << 2) | (lv << 19);
            if (pos < 64u) {
                lds.f[pos] = idx;
                lds.f[64 + pos] = a;
                lds.f[128 + pos] = b;
            } else {
                const uint32_t q = (top + (T - 1u - pos)) & kMask;
                lds.r[q] = idx;
                lds.r[kBeamRing + q] = a;
                lds.r[2 * kBeamRing + q] = b;
            }
            ++pos;
        };
        if (valid && level == levels) {
            put(index, level, px, py, pz);
        } else if (valid) {
            uint32_t v = permute_order(order, visit);
            while (v) {
                const uint32_t c = (uint32_t)__builtin_ctz(v) ^ order;
                v &= v - 1u;
                const uint32_t ci = level < L ? pick(nd, child_off(cm, c)) : index;
                put(ci, level + 1, (px << 1) | ((c >> 2) & 1u), (py << 1) | ((c >> 1) & 1u), (pz << 1) | (c & 1u));
            }
        }
        top += spilled;
        wave_lds_sync();
        const uint32_t n = min(T, 64u);
        valid = lane < n;
        uint32_t ea = 0, eb = 0;
        if (valid) {
            index = lds.f[lane];
            ea = lds.f[64 + lane];
            eb = lds.f[128 + lane];
        }
        const uint32_t m = min(64u - n, top - bottom);
        if (lane >= n && lane < n + m) {
            const uint32_t q = (top - 1u - (lane - n)) & kMask;
            index = lds.r[q];
            ea = lds.r[kBeamRing + q];
            eb = lds.r[2 * kBeamRing + q];
            valid = true;
        }
        top -= m;
        if (valid) beam_unpack(ea, eb, level, px, py, pz);
        wave_lds_sync();
    }
    return 2u;
}

// The rays the per-tile walk handed over (L), one wavefront each; rays the search could not
// finish go to `rest`.
template <bool kShadow>
__global__ __launch_bounds__(kBlock) void k_trace_beam(DevDag dag, RayP p, const RayP* __restrict__ pmem, uint32_t W,
                                                       uint32_t H, uint4* __restrict__ paths,
                                                       uint32_t* __restrict__ colors, float bias, float fog,
                                                       LongRays L, LongRays rest) {
    __shared__ uint32_t sF[kBlock / 64][3 * 64];
    __shared__ uint32_t sR[kBlock / 64][3 * kBeamRing];
    if (pmem) p = *pmem;
    if (dag.root_ptr) dag.root = *dag.root_ptr;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const BeamLds lds{sF[wave], sR[wave]};
    const uint32_t n = *L.count;
    for (uint32_t r = blockIdx.x * (kBlock / 64) + wave; r < n; r += gridDim.x * (kBlock / 64)) {
        const uint32_t pix = __builtin_amdgcn_readfirstlane(L.items[r]);
        const uint32_t x = pix % W, y = pix / W;
        uint32_t hx = 0, hy = 0, hz = 0, res;
        F3 sun{};
        if (kShadow) {
            sun_direction(sun.x, sun.y, sun.z);
            const uint4 pp = paths[pix];
            double cx, cy, cz;
            camera_dir(p, x, H - 1 - y, cx, cy, cz);
            const F3 o = shadow_origin(p, pp, cx, cy, cz, bias, sun);
            const F3 inv{(float)(1.0 / (double)sun.x), (float)(1.0 / (double)sun.y), (float)(1.0 / (double)sun.z)};
            const uint32_t order = (sun.x < 0.f ? 4u : 0u) + (sun.y < 0.f ? 2u : 0u) + (sun.z < 0.f ? 1u : 0u);
            res = beam_search<true>(dag, o, sun, inv, order, lds, hx, hy, hz);
        } else {
            F3 o, d, inv;
            uint32_t order;
            primary_ray(p, x, y, o, d, inv, order);
            res = beam_search<false>(dag, o, d, inv, order, lds, hx, hy, hz);
        }
        if (lane != 0) continue;
        if (res == 2u) {
            rest.items[atomicAdd(rest.count, 1u)] = pix;
        } else if (kShadow) {
            const uint32_t colorIn = colors[pix];
            const float light = res == 1u ? 0.f : 1.f;
            if (fog == 0.f) {
                colors[pix] = shade_clear(colorIn, light);
            } else {
                const uint4 pp = paths[pix];
                double distance, nx, ny, nz;
                hit_distance(p, pp, distance, nx, ny, nz);
                colors[pix] = shade(colorIn, light, distance, nx, ny, nz, fog, sun);
            }
        } else {
            paths[(uint64_t)(H - 1 - y) * W + x] = res == 1u ? make_uint4(hx, hy, hz, 0u) : make_uint4(0u, 0u, 0u, 0u);
        }
    }
}

// The rays the beam search gave back: the plain walk without a cap, one lane each.
template <bool kShadow, int kFetch>
__global__ __launch_bounds__(kBlock) void k_trace_rest(DevDag dag, RayP p, const RayP* __restrict__ pmem, uint32_t W,
                                                       uint32_t H, uint4* __restrict__ paths,
                                                       uint32_t* __restrict__ colors, float bias, float fog,
                                                       LongRays rest) {
    const Stack st = stack_of(dag.levels);
    if (pmem) p = *pmem;
    if (dag.root_ptr) dag.root = *dag.root_ptr;
    const uint32_t n = *rest.count;
    uint64_t unused = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t pix = rest.items[i], x = pix % W, y = pix / W;
        uint32_t vx, vy, vz;
        if (kShadow) {
            F3 sun;
            sun_direction(sun.x, sun.y, sun.z);
            const uint4 pp = paths[pix];
            double cx, cy, cz;
            camera_dir(p, x, H - 1 - y, cx, cy, cz);
            const F3 o = shadow_origin(p, pp, cx, cy, cz, bias, sun);
            const F3 inv{(float)(1.0 / (double)sun.x), (float)(1.0 / (double)sun.y), (float)(1.0 / (double)sun.z)};
            const uint32_t order = (sun.x < 0.f ? 4u : 0u) + (sun.y < 0.f ? 2u : 0u) + (sun.z < 0.f ? 1u : 0u);
            const bool shadowed = descend<false, kFetch>(dag, o, sun, inv, order, vx, vy, vz, st, unused);
            const uint32_t colorIn = colors[pix];
            if (fog == 0.f) {
                colors[pix] = shade_clear(colorIn, shadowed ? 0.f : 1.f);
            } else {
                double distance, nx, ny, nz;
                hit_distance(p, pp, distance, nx, ny, nz);
                colors[pix] = shade(colorIn, shadowed ? 0.f : 1.f, distance, nx, ny, nz, fog, sun);
            }
        } else {
            F3 o, d, inv;
            uint32_t order;
            primary_ray(p, x, y, o, d, inv, order);
            descend<false, kFetch>(dag, o, d, inv, order, vx, vy, vz, st, unused);
            paths[(uint64_t)(H - 1 - y) * W + x] = make_uint4(vx, vy, vz, 0u);
        }
    }
}

// Walks stopped at the per-tile cap, continued by wavefronts over the compacted list (64
// consecutive entries per wavefront, so each wavefront holds only long walks): each lane
// restores its walk and steps on for at most `steps` steps; walks still unfinished are handed
// to the wavefront-per-ray search (list B). Pointer fetch only.
template <bool kShadow>
__global__ __launch_bounds__(kBlock) void k_trace_resume(DevDag dag, RayP p, const RayP* __restrict__ pmem, uint32_t W,
                                                         uint32_t H, uint4* __restrict__ paths,
                                                         uint32_t* __restrict__ colors, float bias, float fog,
                                                         LongRays R, LongRays B, uint32_t steps) {
    const Stack st = stack_of(dag.levels);
    if (pmem) p = *pmem;
    if (dag.root_ptr) dag.root = *dag.root_ptr;
    const uint32_t n = *R.count;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, waves = (gridDim.x * blockDim.x) >> 6;
    uint64_t unused = 0;
    for (uint32_t base = wave * 64u; base < n; base += waves * 64u) {
        const uint32_t i = base + lane;
        bool give = false;
        uint32_t pix = 0;
        if (i < n) {
            pix = R.items[i];
            const uint32_t x = pix % W, y = pix / W;
            F3 o, d, inv;
            uint32_t order;
            uint4 pp{};
            if (kShadow) {
                sun_direction(d.x, d.y, d.z);
                pp = paths[pix];
                double cx, cy, cz;
                camera_dir(p, x, H - 1 - y, cx, cy, cz);
                o = shadow_origin(p, pp, cx, cy, cz, bias, d);
                inv = F3{(float)(1.0 / (double)d.x), (float)(1.0 / (double)d.y), (float)(1.0 / (double)d.z)};
                order = (d.x < 0.f ? 4u : 0u) + (d.y < 0.f ? 2u : 0u) + (d.z < 0.f ? 1u : 0u);
            } else {
                primary_ray(p, x, y, o, d, inv, order);
            }
            Walk w;
            restore_walk(w, st, dag.levels, dag.leaf_level, R.state, R.stride, i);
            uint32_t r = kWalkOn, k = 0;
            while (r == kWalkOn && k++ < steps) r = walk_step<false, kFetchPointer>(dag, w, o, d, inv, order, st, unused);
            give = r == kWalkOn;
            if (!give) {
                if (kShadow) {
                    const uint32_t colorIn = colors[pix];
                    const bool shadowed = r == kWalkHit;
                    if (fog == 0.f) {
                        colors[pix] = shade_clear(colorIn, shadowed ? 0.f : 1.f);
                    } else {
                        double distance, nx, ny, nz;
                        hit_distance(p, pp, distance, nx, ny, nz);
                        colors[pix] = shade(colorIn, shadowed ? 0.f : 1.f, distance, nx, ny, nz, fog, d);
                    }
                } else {
                    paths[(uint64_t)(H - 1 - y) * W + x] = make_uint4(w.px, w.py, w.pz, 0u);
                }
            }
        }
        long_append(give, pix, B);
    }
}

__global__ void k_resume_reset(uint32_t* c) {
    if (threadIdx.x < 3) c[threadIdx.x] = 0;
}

inline RayP to_rayp(const hd_ray_params* p) {
    RayP r;
    for (int i = 0; i < 3; ++i) {
        r.cam[i] = p->camera_position[i];
        r.rmin[i] = p->ray_min[i];
        r.ddx[i] = p->ray_ddx[i];
        r.ddy[i] = p->ray_ddy[i];
    }
    return r;
}

inline dim3 grid_of(const hd_ctx* ctx, uint32_t W, uint32_t H) {
    return dim3((W + 2 * ctx->tile_w - 1) / (2 * ctx->tile_w), (H + 2 * ctx->tile_h - 1) / (2 * ctx->tile_h));
}

// The caller's stream as given; NULL is the HIP default stream (torch's default stream).
inline hipStream_t stream_of(hd_ctx*, void* s) { return reinterpret_cast<hipStream_t>(s); }

// Persistent schedule: the kernel of (shadow, fetch) with as many blocks as stay resident; shadows
// first pack their rays per tile (k_shadow_setup).
template <bool kShadow>
hd_status launch_persist(hd_ctx* ctx, const DevDag& dag, const RayP& p, const RayP* pm, uint32_t W, uint32_t H,
                         uint4* paths, uint32_t* colors, float bias, float fog, hipStream_t s) {
    const auto k = ctx->fetch_mode == kFetchNode ? k_trace_persist<kShadow, kFetchNode>
                                                 : k_trace_persist<kShadow, kFetchPointer>;
    const uint32_t smem = stack_bytes(ctx->g.levels);
    uint32_t& blocks = ctx->persist_blocks[(kShadow ? 2 : 0) + (ctx->fetch_mode == kFetchNode ? 1 : 0)];
    if (!blocks) {
        int perCU = 0;
        HD_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, reinterpret_cast<const void*>(k), kBlock, smem));
        hipDeviceProp_t prop;
        HD_HIP(ctx, hipGetDeviceProperties(&prop, ctx->device));
        blocks = (uint32_t)std::max(1, perCU) * (uint32_t)std::max(1, prop.multiProcessorCount);
    }
    TraceScratch* ts = nullptr;
    const hd_status rs = trace_reserve(ctx, s, W, H, &ts);
    if (rs != HD_OK) return rs;
    TileQ Q;
    Q.heads = ts->queues;
    Q.tilesX = (W + ctx->tile_w - 1) / ctx->tile_w;
    const uint32_t tilesY = (H + ctx->tile_h - 1) / ctx->tile_h;
    Q.ntiles = Q.tilesX * tilesY;
    uint4* rec = nullptr;
    uint2* shades = nullptr;
    uint32_t* cnt = nullptr;
    if (kShadow) {
        rec = ts->rec;
        shades = ts->shades;
        cnt = ts->cnt;
        const auto setup = fog == 0.f ? k_shadow_setup<false> : k_shadow_setup<true>;
        setup<<<grid_of(ctx, W, H), kBlock, 0, s>>>(p, pm, W, H, paths, colors, bias, fog, ctx->tile_w, ctx->tile_h,
                                                    Q.tilesX, tilesY, rec, shades, cnt);
        HD_HIP(ctx, hipGetLastError());
    }
    k_queue_reset<<<1, 256, 0, s>>>(Q.heads);
    k<<<blocks, kBlock, smem, s>>>(dag, p, pm, W, H, paths, colors, rec, shades, cnt, ctx->tile_w, ctx->tile_h, Q,
                                   std::max<uint32_t>(1, ctx->refill_at));
    HD_HIP(ctx, hipGetLastError());
    return HD_OK;
}

// Primary paths with the fetch strategy and schedule of the context; pm (device) overrides p
// when given.
// The handed-over rays of a capped per-tile walk: the wavefront-per-ray search, then the plain
// walk for the few it gives back.
template <bool kShadow>
hd_status launch_long(hd_ctx* ctx, const DevDag& dag, const RayP& p, const RayP* pm, uint32_t W, uint32_t H,
                      uint4* paths, uint32_t* colors, float bias, float fog, const LongRays& L, const LongRays& rest,
                      hipStream_t s) {
    k_trace_beam<kShadow><<<2048, kBlock, 0, s>>>(dag, p, pm, W, H, paths, colors, bias, fog, L, rest);
    HD_HIP(ctx, hipGetLastError());
    const auto k = ctx->fetch_mode == kFetchNode ? k_trace_rest<kShadow, kFetchNode> : k_trace_rest<kShadow, kFetchPointer>;
    k<<<64, kBlock, stack_bytes(ctx->g.levels), s>>>(dag, p, pm, W, H, paths, colors, bias, fog, rest);
    HD_HIP(ctx, hipGetLastError());
    return HD_OK;
}

// The per-tile walk's hand-over lists (cap ~0u when the context runs walks to the end): with a
// resume step (schedule 3, pointer fetch) the per-tile walk stops at resume_steps and saves its
// walk in R; the resume kernel continues to long_steps and hands what is left to L; the
// wavefront-per-ray search takes L and leaves what it could not hold in `rest`.
hd_status long_lists(hd_ctx* ctx, uint32_t W, uint32_t H, hipStream_t s, LongRays& L, LongRays& rest, LongRays& R) {
    L = LongRays{nullptr, nullptr, ~0u, nullptr, 0};
    rest = L;
    R = L;
    if (!ctx->long_steps) return HD_OK;
    TraceScratch* ts = nullptr;
    const hd_status rs = trace_reserve(ctx, s, W, H, &ts);
    if (rs != HD_OK) return rs;
    L = LongRays{ts->long_items, ts->long_count, ctx->long_steps, nullptr, 0};
    rest = LongRays{ts->rest_items, ts->long_count + 1, ~0u, nullptr, 0};
    if (ctx->resume_steps && ctx->fetch_mode == kFetchPointer && ts->resume_state)
        R = LongRays{ts->resume_items, ts->long_count + 2, ctx->resume_steps, ts->resume_state,
                     (uint32_t)ts->pixels_cap};
    k_resume_reset<<<1, 64, 0, s>>>(ts->long_count);
    HD_HIP(ctx, hipGetLastError());
    return HD_OK;
}

// Continues the walks saved in R (schedule 3) before the hand-over to the beam search.
template <bool kShadow>
hd_status launch_resume(hd_ctx* ctx, const DevDag& dag, const RayP& p, const RayP* pm, uint32_t W, uint32_t H,
                        uint4* paths, uint32_t* colors, float bias, float fog, const LongRays& R, const LongRays& L,
                        hipStream_t s) {
    k_trace_resume<kShadow><<<4096, kBlock, stack_bytes(ctx->g.levels), s>>>(dag, p, pm, W, H, paths, colors, bias, fog,
                                                                            R, L, ctx->long_steps - ctx->resume_steps);
    HD_HIP(ctx, hipGetLastError());
    return HD_OK;
}

hd_status launch_paths(hd_ctx* ctx, const DevDag& dag, const RayP& p, const RayP* pm, uint32_t W, uint32_t H,
                       uint4* out, hipStream_t s) {
    if (ctx->trace_persist) return launch_persist<false>(ctx, dag, p, pm, W, H, out, nullptr, 0.f, 0.f, s);
    LongRays L, rest, R;
    hd_status st = long_lists(ctx, W, H, s, L, rest, R);
    if (st != HD_OK) return st;
    const auto k = ctx->fetch_mode == kFetchNode ? k_trace_paths<false, kFetchNode> : k_trace_paths<false, kFetchPointer>;
    k<<<grid_of(ctx, W, H), kBlock, stack_bytes(ctx->g.levels), s>>>(dag, p, pm, W, H, out, ctx->tile_w, ctx->tile_h,
                                                                     nullptr, R.state ? R : L);
    HD_HIP(ctx, hipGetLastError());
    if (R.state) {
        st = launch_resume<false>(ctx, dag, p, pm, W, H, out, nullptr, 0.f, 0.f, R, L, s);
        if (st != HD_OK) return st;
    }
    if (L.cap != ~0u) return launch_long<false>(ctx, dag, p, pm, W, H, out, nullptr, 0.f, 0.f, L, rest, s);
    return HD_OK;
}

hd_status launch_shadows(hd_ctx* ctx, const DevDag& dag, const RayP& p, const RayP* pm, uint32_t W, uint32_t H,
                         const uint4* paths, uint32_t* colors, float bias, float fog, hipStream_t s) {
    if (ctx->trace_persist)
        return launch_persist<true>(ctx, dag, p, pm, W, H, const_cast<uint4*>(paths), colors, bias, fog, s);
    LongRays L, rest, R;
    hd_status st = long_lists(ctx, W, H, s, L, rest, R);
    if (st != HD_OK) return st;
    const auto k = ctx->fetch_mode == kFetchNode ? k_trace_shadows<false, kFetchNode> : k_trace_shadows<false, kFetchPointer>;
    k<<<grid_of(ctx, W, H), kBlock, stack_bytes(ctx->g.levels), s>>>(dag, p, pm, W, H, paths, colors, bias, fog,
                                                                     ctx->tile_w, ctx->tile_h, nullptr, R.state ? R : L);
    HD_HIP(ctx, hipGetLastError());
    if (R.state) {
        st = launch_resume<true>(ctx, dag, p, pm, W, H, const_cast<uint4*>(paths), colors, bias, fog, R, L, s);
        if (st != HD_OK) return st;
    }
    if (L.cap != ~0u)
        return launch_long<true>(ctx, dag, p, pm, W, H, const_cast<uint4*>(paths), colors, bias, fog, L, rest, s);
    return HD_OK;
}

}  // namespace
}  // namespace hd

namespace hd {
hd_status trace_reserve(hd_ctx* ctx, hipStream_t s, uint32_t W, uint32_t H, TraceScratch** out) {
    // tiles of the frame in any tile shape of 64 pixels: at most ceil(W / w) * ceil(H / h)
    uint64_t tiles = 0;
    for (uint32_t w = 1; w <= 64; w *= 2)
        tiles = std::max<uint64_t>(tiles, (uint64_t)((W + w - 1) / w) * ((H + 64 / w - 1) / (64 / w)));
    const uint64_t pixels = (uint64_t)W * H;
    TraceScratch* ts = nullptr;
    for (TraceScratch& t : ctx->trace_scratch)
        if (t.stream == s) ts = &t;
    const bool resume = ctx->resume_steps != 0;
    if (ts && ts->tiles_cap >= tiles && ts->pixels_cap >= pixels && (!resume || ts->resume_state)) {
        if (out) *out = ts;
        return HD_OK;
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HD_HIP(ctx, hipStreamIsCapturing(s, &cap));
    if (cap != hipStreamCaptureStatusNone)
        return fail(ctx, HD_ERR_INVALID, "persistent trace scratch must be reserved before stream capture");
    if (!ts) {
        ctx->trace_scratch.push_back(TraceScratch{});
        ts = &ctx->trace_scratch.back();
        ts->stream = s;
        HD_HIP(ctx, hipMalloc(&ts->queues, 64 * 32 * 4));
        HD_HIP(ctx, hipMalloc(&ts->long_count, 64));
    } else {
        HD_HIP(ctx, hipStreamSynchronize(s));
        hipFree(ts->rec);
        hipFree(ts->shades);
        hipFree(ts->cnt);
        hipFree(ts->long_items);
        hipFree(ts->rest_items);
        hipFree(ts->resume_items);
        hipFree(ts->resume_state);
        ts->rec = nullptr;
        ts->shades = nullptr;
        ts->cnt = nullptr;
        ts->long_items = ts->rest_items = ts->resume_items = ts->resume_state = nullptr;
        ts->tiles_cap = ts->pixels_cap = 0;
    }
    HD_HIP(ctx, hipMalloc(&ts->rec, tiles * 64 * sizeof(uint4)));
    HD_HIP(ctx, hipMalloc(&ts->shades, tiles * 64 * sizeof(uint2)));
    HD_HIP(ctx, hipMalloc(&ts->cnt, tiles * sizeof(uint32_t)));
    HD_HIP(ctx, hipMalloc(&ts->long_items, pixels * sizeof(uint32_t)));
    HD_HIP(ctx, hipMalloc(&ts->rest_items, pixels * sizeof(uint32_t)));
    if (resume) {
        if (pixels * kStateWords >= (1ull << 32)) return fail(ctx, HD_ERR_INVALID, "frame too large for resumed walks");
        HD_HIP(ctx, hipMalloc(&ts->resume_items, pixels * sizeof(uint32_t)));
        HD_HIP(ctx, hipMalloc(&ts->resume_state, pixels * kStateWords * sizeof(uint32_t)));
    }
    ts->tiles_cap = tiles;
    ts->pixels_cap = pixels;
    ctx->ptr_gen++;
    if (out) *out = ts;
    return HD_OK;
}

void trace_scratch_free(hd_ctx* ctx) {
    for (TraceScratch& t : ctx->trace_scratch) {
        hipFree(t.queues);
        hipFree(t.rec);
        hipFree(t.shades);
        hipFree(t.cnt);
        hipFree(t.long_items);
        hipFree(t.rest_items);
        hipFree(t.resume_items);
        hipFree(t.resume_state);
        hipFree(t.long_count);
    }
    ctx->trace_scratch.clear();
}

// The three tracer kernels with the ray parameters and the root read from device memory, so a
// captured graph traces whatever camera and DAG state the replay finds there.
hd_status enqueue_frame_from_device(hd_ctx* ctx, const hd_ray_params* d_params, uint32_t W, uint32_t H,
                                    uint32_t* d_paths, uint32_t* d_colors, float bias, float fog, hipStream_t s) {
    static_assert(sizeof(RayP) == sizeof(hd_ray_params), "ray parameter layouts differ");
    DevDag dag = dev_dag(ctx);
    dag.root_ptr = ctx->d_root;
    const RayP* pm = reinterpret_cast<const RayP*>(d_params);
    const RayP none{};
    hd_status st = launch_paths(ctx, dag, none, pm, W, H, reinterpret_cast<uint4*>(d_paths), s);
    if (st != HD_OK) return st;
    k_trace_colors<false><<<grid_of(ctx, W, H), kBlock, 0, s>>>(dag, dev_colors(ctx), W, H,
                                                                reinterpret_cast<const uint4*>(d_paths), d_colors,
                                                                ctx->tile_w, ctx->tile_h, nullptr);
    HD_HIP(ctx, hipGetLastError());
    return launch_shadows(ctx, dag, none, pm, W, H, reinterpret_cast<const uint4*>(d_paths), d_colors, bias, fog, s);
}
}  // namespace hd

using namespace hd;

extern "C" {

hd_status hd_set_trace_tile(hd_ctx* ctx, uint32_t tw, uint32_t th) {
    if (!ctx) return fail(ctx, HD_ERR_INVALID, "null context");
    if (tw == 0 && th == 0) {
        tw = 8;
        th = 8;
    }
    if (tw * th != 64) return fail(ctx, HD_ERR_INVALID, "tile must cover 64 pixels (one wavefront)");
    ctx->tile_w = tw;
    ctx->tile_h = th;
    ctx->ptr_gen++;
    return HD_OK;
}

hd_status hd_trace_paths(hd_ctx* ctx, const hd_ray_params* params, uint32_t W, uint32_t H, uint32_t* d_paths,
                         void* stream) {
    if (!ctx || !params || !d_paths || W == 0 || H == 0) return fail(ctx, HD_ERR_INVALID, "bad arguments");
    if (!ctx->loaded) return fail(ctx, HD_ERR_INVALID, "no DAG loaded");
    if (ctx->g.levels > kMaxLevels || ctx->g.levels < 3) return fail(ctx, HD_ERR_INVALID, "levels outside 3..17");
    return launch_paths(ctx, dev_dag(ctx), to_rayp(params), nullptr, W, H, reinterpret_cast<uint4*>(d_paths),
                        stream_of(ctx, stream));
}

hd_status hd_trace_colors(hd_ctx* ctx, uint32_t W, uint32_t H, const uint32_t* d_paths, uint32_t* d_colors,
                          void* stream) {
    if (!ctx || !d_paths || !d_colors || W == 0 || H == 0) return fail(ctx, HD_ERR_INVALID, "bad arguments");
    if (!ctx->loaded || !ctx->colors_loaded) return fail(ctx, HD_ERR_INVALID, "no DAG or colours loaded");
    k_trace_colors<false><<<grid_of(ctx, W, H), kBlock, 0, stream_of(ctx, stream)>>>(
        dev_dag(ctx), dev_colors(ctx), W, H, reinterpret_cast<const uint4*>(d_paths), d_colors, ctx->tile_w,
        ctx->tile_h, nullptr);
    HD_HIP(ctx, hipGetLastError());
    return HD_OK;
}

hd_status hd_trace_shadows(hd_ctx* ctx, const hd_ray_params* params, uint32_t W, uint32_t H, const uint32_t* d_paths,
                           uint32_t* d_colors, float bias, float fog, void* stream) {
    if (!ctx || !params || !d_paths || !d_colors || W == 0 || H == 0) return fail(ctx, HD_ERR_INVALID, "bad arguments");
    if (!ctx->loaded) return fail(ctx, HD_ERR_INVALID, "no DAG loaded");
    if (ctx->g.levels > kMaxLevels || ctx->g.levels < 3) return fail(ctx, HD_ERR_INVALID, "levels outside 3..17");
    return launch_shadows(ctx, dev_dag(ctx), to_rayp(params), nullptr, W, H, reinterpret_cast<const uint4*>(d_paths),
                          d_colors, bias, fog, stream_of(ctx, stream));
}

hd_status hd_set_trace_fetch(hd_ctx* ctx, uint32_t fetch_mode) {
    if (!ctx) return fail(ctx, HD_ERR_INVALID, "null context");
    if (fetch_mode > 1) return fail(ctx, HD_ERR_INVALID, "fetch mode is 0 or 1");
    ctx->fetch_mode = fetch_mode;
    ctx->ptr_gen++;
    return HD_OK;
}

hd_status hd_set_trace_schedule(hd_ctx* ctx, uint32_t schedule, uint32_t param) {
    if (!ctx) return fail(ctx, HD_ERR_INVALID, "null context");
    const uint32_t c1 = param & 0xFFFFu, c2 = param >> 16;
    if (schedule > 3 || (schedule == 1 && (param == 0 || param > 64)) || (schedule == 2 && param == 0) ||
        (schedule == 3 && (c1 == 0 || c2 <= c1)))
        return fail(ctx, HD_ERR_INVALID,
                    "schedule is 0 (per tile), 1 (persistent, refill at 1..64 idle lanes), 2 (per tile, walks "
                    "beyond param steps searched a wavefront per ray) or 3 (per tile to param & 0xFFFF steps, "
                    "resumed by compacted wavefronts to param >> 16 steps, the rest searched per ray)");
    ctx->trace_persist = schedule == 1 ? 1u : 0u;
    if (schedule == 1) ctx->refill_at = param;
    ctx->long_steps = schedule == 2 ? param : (schedule == 3 ? c2 : 0u);
    ctx->resume_steps = schedule == 3 ? c1 : 0u;
    ctx->ptr_gen++;
    return HD_OK;
}

hd_status hd_get_trace_schedule(hd_ctx* ctx, uint32_t out[2]) {
    if (!ctx || !out) return fail(ctx, HD_ERR_INVALID, "bad arguments");
    out[0] = ctx->trace_persist ? 1u : (ctx->resume_steps ? 3u : (ctx->long_steps ? 2u : 0u));
    out[1] = ctx->trace_persist ? ctx->refill_at
                                : (ctx->resume_steps ? (ctx->resume_steps | (ctx->long_steps << 16)) : ctx->long_steps);
    return HD_OK;
}

hd_status hd_get_trace_launch(hd_ctx* ctx, uint32_t out[3]) {
    if (!ctx || !out) return fail(ctx, HD_ERR_INVALID, "bad arguments");
    out[0] = ctx->tile_w;
    out[1] = ctx->tile_h;
    out[2] = ctx->fetch_mode;
    return HD_OK;
}

hd_status hd_autotune_trace(hd_ctx* ctx, const hd_ray_params* params, uint32_t n_params, uint32_t W, uint32_t H,
                            uint32_t* d_paths, uint32_t* d_colors, uint32_t reps, uint32_t* best_w, uint32_t* best_h,
                            float* best_ms) {
    if (!ctx || !params || !n_params || !d_paths || !d_colors || W == 0 || H == 0)
        return fail(ctx, HD_ERR_INVALID, "bad arguments");
    if (!ctx->loaded || !ctx->colors_loaded) return fail(ctx, HD_ERR_INVALID, "no DAG or colours loaded");
    static const uint32_t shapes[][2] = {{8, 8}, {16, 4}, {4, 16}, {32, 2}, {2, 32}, {64, 1}};
    static const uint32_t fetches[] = {kFetchPointer, kFetchNode};
    // schedule: one wavefront per tile; persistent wavefronts refilled at 8 / 24 idle lanes; one
    // wavefront per tile with walks beyond 96 / 128 / 192 steps searched a wavefront per ray
    // (three trials with resumed walks: per tile to 24 / 32 / 48 steps, resumed to 128 / 128 / 160)
    static const uint32_t schedules[][2] = {{0, 0},   {1, 8},   {1, 24},  {2, 96},           {2, 128},
                                            {2, 192}, {3, 24 | (128u << 16)}, {3, 32 | (128u << 16)},
                                            {3, 48 | (160u << 16)}};
    constexpr uint32_t kSchedules = sizeof(schedules) / sizeof(schedules[0]);
    constexpr uint32_t kTrials = 6 * 2 * kSchedules;  // tile shape x fetch strategy x schedule
    hipEvent_t e0, e1;
    HD_HIP(ctx, hipEventCreate(&e0));
    HD_HIP(ctx, hipEventCreate(&e1));
    float bestMs = 1e30f;
    uint32_t bw = ctx->tile_w, bh = ctx->tile_h, bf = ctx->fetch_mode, bp = ctx->trace_persist, br = ctx->refill_at,
             bl = ctx->long_steps, bs = ctx->resume_steps;
    reps = std::max<uint32_t>(reps, 1);
    for (uint32_t trial = 0; trial < kTrials; ++trial) {
        const auto& sh = shapes[trial % 6];
        ctx->tile_w = sh[0];
        ctx->tile_h = sh[1];
        ctx->fetch_mode = fetches[(trial / 6) % 2];
        const uint32_t* sch = schedules[trial / 12];
        if (sch[0] == 3 && ctx->fetch_mode != kFetchPointer) continue;  // resumed walks: pointer fetch only
        ctx->trace_persist = sch[0] == 1 ? 1u : 0u;
        if (sch[0] == 1) ctx->refill_at = sch[1];
        ctx->long_steps = sch[0] == 2 ? sch[1] : (sch[0] == 3 ? sch[1] >> 16 : 0u);
        ctx->resume_steps = sch[0] == 3 ? (sch[1] & 0xFFFFu) : 0u;
        float total = 0;
        for (uint32_t rep = 0; rep <= reps; ++rep) {  // rep 0 warms up
            HD_HIP(ctx, hipEventRecord(e0, ctx->stream));
            for (uint32_t k = 0; k < n_params; ++k) {
                hd_status st = hd_trace_paths(ctx, &params[k], W, H, d_paths, ctx->stream);
                if (st == HD_OK) st = hd_trace_colors(ctx, W, H, d_paths, d_colors, ctx->stream);
                if (st == HD_OK) st = hd_trace_shadows(ctx, &params[k], W, H, d_paths, d_colors, 1.0f, 0.0f, ctx->stream);
                if (st != HD_OK) {
                    hipEventDestroy(e0);
                    hipEventDestroy(e1);
                    return st;
                }
            }
            HD_HIP(ctx, hipEventRecord(e1, ctx->stream));
            HD_HIP(ctx, hipEventSynchronize(e1));
            float ms = 0;
            HD_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
            if (rep) total += ms;
        }
        total /= (float)reps;
        if (total < bestMs) {
            bestMs = total;
            bw = sh[0];
            bh = sh[1];
            bf = ctx->fetch_mode;
            bp = ctx->trace_persist;
            br = ctx->refill_at;
            bl = ctx->long_steps;
            bs = ctx->resume_steps;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    ctx->tile_w = bw;
    ctx->tile_h = bh;
    ctx->fetch_mode = bf;
    ctx->trace_persist = bp;
    ctx->refill_at = br;
    ctx->long_steps = bl;
    ctx->resume_steps = bs;
    ctx->ptr_gen++;
    if (best_w) *best_w = bw;
    if (best_h) *best_h = bh;
    if (best_ms) *best_ms = bestMs / (float)n_params;
    return HD_OK;
}

hd_status hd_trace_count_bytes(hd_ctx* ctx, const hd_ray_params* params, uint32_t W, uint32_t H, uint32_t* d_paths,
                               uint32_t* d_colors, float bias, float fog, uint64_t bytes[3]) {
    if (!ctx || !params || !d_paths || !d_colors || !bytes) return fail(ctx, HD_ERR_INVALID, "bad arguments");
    if (!ctx->loaded) return fail(ctx, HD_ERR_INVALID, "no DAG loaded");
    if (ctx->g.levels > kMaxLevels || ctx->g.levels < 3) return fail(ctx, HD_ERR_INVALID, "levels outside 3..17");
    const dim3 grid = grid_of(ctx, W, H);
    HD_HIP(ctx, hipMemsetAsync(ctx->byte_counters, 0, 3 * sizeof(unsigned long long), ctx->stream));
    k_trace_paths<true, kFetchPointer><<<grid, kBlock, stack_bytes(ctx->g.levels), ctx->stream>>>(dev_dag(ctx), to_rayp(params), nullptr, W, H,
                                                          reinterpret_cast<uint4*>(d_paths), ctx->tile_w, ctx->tile_h,
                                                          ctx->byte_counters + 0, LongRays{nullptr, nullptr, ~0u, nullptr, 0});
    HD_HIP(ctx, hipGetLastError());
    if (ctx->colors_loaded) {
        k_trace_colors<true><<<grid, kBlock, 0, ctx->stream>>>(dev_dag(ctx), dev_colors(ctx), W, H,
                                                               reinterpret_cast<const uint4*>(d_paths), d_colors,
                                                               ctx->tile_w, ctx->tile_h, ctx->byte_counters + 1);
        HD_HIP(ctx, hipGetLastError());
    }
    k_trace_shadows<true, kFetchPointer><<<grid, kBlock, stack_bytes(ctx->g.levels), ctx->stream>>>(dev_dag(ctx), to_rayp(params), nullptr, W, H,
                                                            reinterpret_cast<const uint4*>(d_paths), d_colors, bias, fog,
                                                            ctx->tile_w, ctx->tile_h, ctx->byte_counters + 2,
                                                            LongRays{nullptr, nullptr, ~0u, nullptr, 0});
    HD_HIP(ctx, hipGetLastError());
    unsigned long long b[3];
    HD_HIP(ctx, hipMemcpyAsync(b, ctx->byte_counters, sizeof(b), hipMemcpyDeviceToHost, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < 3; ++i) bytes[i] = b[i];
    return HD_OK;
}

}  // extern "C"
