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
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "hd_ctx.h"
#include "hd_edit.h"

namespace hd {
namespace {

constexpr uint32_t KIND_KEEP = 0, KIND_EMPTY = 1, KIND_FULL = 2, KIND_ITEM = 3;
constexpr uint32_t kGrid = 2048;  // blocks of grid-stride kernels

// ---------------------------------------------------------------------------------------------
// Editors (hash_dag_editors.h:9-330). FP32, contraction off.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float sqf(float f) { return f * f; }
__device__ __forceinline__ float tmaxf2(float a, float b) { return a > b ? a : b; }

// hash_dag_editors.h:76-91 / 112-126
__device__ __forceinline__ bool e_touch(const DevEdit& e, float x0, float y0, float z0, float x1, float y1, float z1) {
    if (e.kind == EDIT_BOX_ADD || e.kind == EDIT_BOX_CARVE) {
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
__device__ __forceinline__ uint32_t e_cover(const DevEdit& e, float x0, float y0, float z0, float x1, float y1,
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
__device__ __forceinline__ uint64_t e_color(const DevEdit& e, float px, float py, float pz, uint64_t prev) {
    if (e.kind != EDIT_SPHERE_PAINT) return pack_color(e.single, 0, 0);
    float r, g, b;
    unpack_rgb(prev, r, g, b);
    const float dx = px - e.c[0], dy = py - e.c[1], dz = pz - e.c[2];
    const float len = __fsqrt_rn(dx * dx + dy * dy + dz * dz);
    const float k = clampf(1 - __fdiv_rn(len, e.radius), 0.f, .1f);
    const float omk = 1 - k;
    return pack_color(to101210(r * omk + 0.0f * k, g * omk + 0.0f * k, b * omk + 1.0f * k), 0, 0);
}
__device__ __forceinline__ bool e_value(const DevEdit& e, bool old) {
    return e.kind == EDIT_SPHERE_PAINT ? old : (e.kind == EDIT_SPHERE_ADD || e.kind == EDIT_BOX_ADD);
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
// is_valid() && is_valid_index(i) (variable_weight_size_colors.h:247-262)
__device__ __forceinline__ bool leaf_valid_index(const LeafDesc& d, uint64_t i) {
    if (d.n_blocks == 0) return false;
    if (d.offset != ~0ull) i += d.offset;
    return 2 * i / 16384u < d.n_macro;
}

__device__ __forceinline__ uint32_t child_off(uint32_t mask, uint32_t c) { return __popc(mask & ((1u << c) - 1u)) + 1u; }

__device__ __forceinline__ void set_error(uint32_t* flags, uint32_t bit) { atomicOr(flags, bit); }

// ---------------------------------------------------------------------------------------------
// expand: items of level l -> child classification and items of level l+1
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_expand(EditPass P, uint32_t l) {
    const uint32_t n = min(P.count[l], P.cap);
    if (*P.errors) return;
    const Items I = P.items[l];
    const Items C = P.items[l + 1];
    const uint32_t D = P.levels, L = D - 2, ctl = P.ctl;
    const uint32_t depth = D - (l + 1);  // child cube side = 2^depth
    const float side = (float)(1u << depth);
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (uint64_t)n * 8; t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = (uint32_t)(t >> 3), c = (uint32_t)(t & 7u);
        const uint32_t cx = (I.px[i] << 1) | ((c >> 2) & 1u), cy = (I.py[i] << 1) | ((c >> 1) & 1u),
                       cz = (I.pz[i] << 1) | (c & 1u);
        // child of the effective old node, and of the original node
        uint32_t childOld = kInvalidNode, childOrig = kInvalidNode;
        const uint32_t old = I.old_[i], orig = I.orig[i];
        if (old != kInvalidNode) {
            const uint32_t m = P.vpool[old] & 0xFFu;
            if (m & (1u << c)) childOld = P.vpool[old + child_off(m, c)];
        }
        if (orig == old) {
            childOrig = childOld;
        } else if (orig != kInvalidNode) {
            const uint32_t m = P.vpool[orig] & 0xFFu;
            if (m & (1u << c)) childOrig = P.vpool[orig + child_off(m, c)];
        }
        const float x0 = (float)(cx << depth), y0 = (float)(cy << depth), z0 = (float)(cz << depth);
        const float x1 = x0 + side, y1 = y0 + side, z1 = z0 + side;
        const uint32_t es = I.estart[i], ec = I.ecount[i];
        // filter the edit list; remember the last cover at levels >= ctl
        uint32_t m = 0, lastCover = ~0u, coverKind = 0;
        for (uint32_t k = 0; k < ec; ++k) {
            const DevEdit& e = P.edits[P.elist[l][es + k]];
            if (!e_touch(e, x0, y0, z0, x1, y1, z1)) continue;
            if (l + 1 >= ctl) {
                const uint32_t cv = e_cover(e, x0, y0, z0, x1, y1, z1);
                if (cv) {
                    lastCover = m;
                    coverKind = cv;
                }
            }
            m++;
        }
        // old voxel count of the effective old child (colour offsets, "keep" counts)
        uint32_t oldCount = 0;
        if (l + 1 >= ctl && childOld != kInvalidNode) {
            if (l + 1 == L) {
                const uint64_t lf = *reinterpret_cast<const uint64_t*>(P.vpool + childOld);
                oldCount = __popcll(lf);
            } else {
                oldCount = P.vpool[childOld] >> 8;
            }
        }
        const uint32_t slot = i * 8 + c;
        I.corig[slot] = childOrig;
        I.cold[slot] = oldCount;
        uint32_t kind;
        const bool covered = lastCover != ~0u;
        if (m == 0) {
            kind = KIND_KEEP;
            I.cval[slot] = childOld;
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
            if (covered && rest == 0 && P.single_edit)
                atomicAdd(P.stat_voxels, (unsigned long long)(1ull << (3 * depth)));
            if (covered) {
                // colour of a full cover (add_large_single_color of the covering editor)
                uint32_t q = 0;
                for (uint32_t k = 0; k < ec; ++k) {
                    const DevEdit& e = P.edits[P.elist[l][es + k]];
                    if (!e_touch(e, x0, y0, z0, x1, y1, z1)) continue;
                    if (q == lastCover) {
                        I.ccolor[slot] = e.single;
                        break;
                    }
                    q++;
                }
            }
        }
        if (kind == KIND_ITEM) {
            const uint32_t ci = atomicAdd(&P.count[l + 1], 1u);
            const uint32_t rest = lastCover == ~0u ? m : m - lastCover - 1;
            const uint32_t eo = atomicAdd(&P.ecount[l + 1], rest);
            if (ci >= P.cap || eo + rest > P.ecap) {
                set_error(P.errors, ERR_SCRATCH);
                I.ckind[slot] = (uint8_t)kind;
                continue;
            }
            C.px[ci] = cx;
            C.py[ci] = cy;
            C.pz[ci] = cz;
            C.orig[ci] = childOrig;
            C.parent[ci] = i;
            C.slot[ci] = (uint8_t)c;
            C.estart[ci] = eo;
            C.ecount[ci] = rest;
            uint32_t q = 0, w = 0;
            for (uint32_t k = 0; k < ec; ++k) {
                const uint32_t ei = P.elist[l][es + k];
                const DevEdit& e = P.edits[ei];
                if (!e_touch(e, x0, y0, z0, x1, y1, z1)) continue;
                if (lastCover == ~0u || q > lastCover) P.elist[l + 1][eo + w++] = ei;
                q++;
            }
            if (lastCover != ~0u) {
                C.old_[ci] = coverKind == 1 ? P.full_nodes[l + 1] : kInvalidNode;
                C.cmode[ci] = coverKind == 1 ? 1u : 2u;  // old colours: single colour / none
                C.cmodeval[ci] = I.ccolor[slot];
            } else {
                C.old_[ci] = childOld;
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
        I.ckind[slot] = (uint8_t)kind;
    }
}

// ---------------------------------------------------------------------------------------------
// leaf level: one wavefront per leaf item, one lane per voxel (bit = child1 * 8 + child2)
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_leaf(EditPass P) {
    const uint32_t L = P.levels - 2;
    const uint32_t n = min(P.count[L], P.cap);
    if (*P.errors) return;
    const Items I = P.items[L];
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n; i += (gridDim.x * blockDim.x) >> 6) {
        const uint32_t c1 = lane >> 3, c2 = lane & 7u;
        const uint32_t x = (I.px[i] << 2) | (((c1 >> 2) & 1u) << 1) | ((c2 >> 2) & 1u);
        const uint32_t y = (I.py[i] << 2) | (((c1 >> 1) & 1u) << 1) | ((c2 >> 1) & 1u);
        const uint32_t z = (I.pz[i] << 2) | ((c1 & 1u) << 1) | (c2 & 1u);
        uint64_t oldMask = 0, origMask = 0;
        if (I.old_[i] != kInvalidNode) oldMask = *reinterpret_cast<const uint64_t*>(P.vpool + I.old_[i]);
        if (I.orig[i] != kInvalidNode) origMask = *reinterpret_cast<const uint64_t*>(P.vpool + I.orig[i]);
        bool v = (oldMask >> lane) & 1u;
        const float fx = (float)x, fy = (float)y, fz = (float)z;
        uint32_t touched = 0;
        for (uint32_t k = 0; k < I.ecount[i]; ++k) {
            const DevEdit& e = P.edits[P.elist[L][I.estart[i] + k]];
            if (e_touch(e, fx, fy, fz, fx + 1.0f, fy + 1.0f, fz + 1.0f)) {
                v = e_value(e, v);
                touched = 1;
            }
        }
        const uint64_t newMask = __ballot(v);
        if (P.single_edit) {
            const uint64_t t = __ballot(touched);
            if (lane == 0 && t) atomicAdd(P.stat_voxels, (unsigned long long)__popcll(t));
        }
        if (lane == 0) {
            I.ncount[i] = __popcll(newMask);
            if (newMask == origMask) {
                I.nidx[i] = I.orig[i];
                I.state[i] = 0;
            } else if (newMask == 0) {
                I.nidx[i] = kInvalidNode;
                I.state[i] = 0;
            } else {
                I.clen[i] = 2;
                I.cont[i * 9 + 0] = (uint32_t)newMask;
                I.cont[i * 9 + 1] = (uint32_t)(newMask >> 32);
                I.hash[i] = hash_leaf(newMask);
                I.state[i] = 1;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// build interior items of level l from their children's results
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_build(EditPass P, uint32_t l) {
    const uint32_t n = min(P.count[l], P.cap);
    if (*P.errors) return;
    const Items I = P.items[l];
    const Items C = P.items[l + 1];
    const uint32_t D = P.levels, ctl = P.ctl;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
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

// ---------------------------------------------------------------------------------------------
// dedup / find / add / resolve for the pending items of one level
// ---------------------------------------------------------------------------------------------
__global__ void k_fill(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

__device__ __forceinline__ bool same_content(const Items& I, uint32_t a, uint32_t b) {
    const uint32_t n = I.clen[a];
    if (n != I.clen[b]) return false;
    for (uint32_t k = 0; k < n; ++k)
        if (I.cont[a * 9 + k] != I.cont[b * 9 + k]) return false;
    return true;
}

__global__ __launch_bounds__(256) void k_dedup(EditPass P, uint32_t l) {
    const uint32_t n = min(P.count[l], P.cap);
    if (*P.errors) return;
    const Items I = P.items[l];
    const uint32_t hmask = P.hcap - 1;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (I.state[i] != 1) continue;
        uint32_t s = I.hash[i] & hmask;
        for (;;) {
            const uint32_t cur = atomicCAS(&P.hset[s], 0u, i + 1);
            if (cur == 0) {
                I.rep[i] = i;
                break;
            }
            const uint32_t j = cur - 1;
            if (I.hash[j] == I.hash[i] && same_content(I, i, j)) {
                I.rep[i] = j;
                break;
            }
            s = (s + 1) & hmask;
        }
    }
}

// One wavefront per representative: scan the bucket for an identical node (any word
// position inside the bucket's used size whose words equal the node and that does not
// cross a page boundary holds the same node content).
__global__ __launch_bounds__(256) void k_find(EditPass P, uint32_t l) {
    const uint32_t n = min(P.count[l], P.cap);
    if (*P.errors) return;
    const Items I = P.items[l];
    const Geom g = P.g;
    const uint32_t lane = threadIdx.x & 63u;
    const bool leafLevel = l == g.leaf_level();
    for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n; i += (gridDim.x * blockDim.x) >> 6) {
        if (I.state[i] != 1 || I.rep[i] != i) continue;
        const uint32_t bucket = g.bucket_from_hash(l, I.hash[i]);
        const uint32_t base = g.make_ptr(l, bucket, 0);
        const uint32_t bs = P.bucket_sizes[g.bucket_global(l, bucket)];
        const uint32_t nw = I.clen[i];
        uint32_t found = ~0u;
        if (leafLevel) {
            const uint64_t key = (uint64_t)I.cont[i * 9] | ((uint64_t)I.cont[i * 9 + 1] << 32);
            for (uint32_t p0 = 0; p0 < bs && found == ~0u; p0 += 128) {
                const uint32_t pos = p0 + 2 * lane;
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
                bool hit = pos + nw < pageEnd && P.vpool[base + pos] == h0;
                if (hit)
                    for (uint32_t k = 1; k < nw; ++k)
                        if (P.vpool[base + pos + k] != I.cont[i * 9 + k]) {
                            hit = false;
                            break;
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
}

__device__ bool alloc_page(EditPass& P, uint32_t page) {
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

__global__ __launch_bounds__(256) void k_add(EditPass P, uint32_t l) {
    const uint32_t n = min(P.count[l], P.cap);
    if (*P.errors) return;
    const Items I = P.items[l];
    const Geom g = P.g;
    const uint32_t ps = g.page_size;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (I.state[i] != 1 || I.rep[i] != i) continue;
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
            continue;
        }
        const uint32_t va = g.make_ptr(l, bucket, start);
        if (start % ps == 0 && !alloc_page(P, va / ps)) continue;
        for (uint32_t k = 0; k < nw; ++k) P.vpool[va + k] = I.cont[i * 9 + k];
        I.nidx[i] = va;
        atomicAdd(P.stat_new_words, (unsigned long long)nw);
    }
}

__global__ __launch_bounds__(256) void k_resolve(EditPass P, uint32_t l) {
    const uint32_t n = min(P.count[l], P.cap);
    if (*P.errors) return;
    const Items I = P.items[l];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (I.state[i] == 1 && I.rep[i] != i) I.nidx[i] = I.nidx[I.rep[i]];
}

// ---------------------------------------------------------------------------------------------
// colours
// ---------------------------------------------------------------------------------------------
// new / old colour starts of the child items (relative to the task's colour array)
__global__ __launch_bounds__(256) void k_color_starts(EditPass P, uint32_t l) {
    const uint32_t n = min(P.count[l], P.cap);
    if (*P.errors) return;
    const Items I = P.items[l];
    const Items C = P.items[l + 1];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (l == P.ctl) {
            I.nstart[i] = 0;
            I.ostart[i] = 0;
        }
        uint32_t ns = I.nstart[i], os = I.ostart[i];
        for (uint32_t c = 0; c < 8; ++c) {
            const uint32_t slot = i * 8 + c;
            if (I.ckind[slot] == KIND_ITEM) {
                const uint32_t ci = I.cval[slot];
                C.nstart[ci] = ns;
                C.ostart[ci] = os;
            }
            I.cstart[slot] = ns;
            I.costart[slot] = os;
            ns += I.cnew[slot];
            os += I.cold[slot];
        }
    }
}

// colour array of every task with colours: base offset in the colour scratch
__global__ void k_color_alloc(EditPass P) {
    const uint32_t ctl = P.ctl;
    const uint32_t n = min(P.count[ctl], P.cap);
    if (*P.errors) return;
    const Items T = P.items[ctl];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t cnt = T.ncount[i];
        uint64_t base = 0;
        if (cnt) {
            base = atomicAdd(P.color_top, (unsigned long long)cnt);
            if (base + cnt > P.color_cap) set_error(P.errors, ERR_COLOR_SCRATCH);
        }
        T.cbase[i] = base;
    }
}

__device__ __forceinline__ LeafDesc task_leaf(const EditPass& P, uint32_t handle) {
    LeafDesc d{};
    if (handle == 0) return d;  // no leaf: n_blocks = 0 (invalid)
    if (handle & 0x80000000u) return P.leaves[handle & 0x7FFFFFFFu];
    d = P.main;
    d.offset = P.offsets[handle];
    return d;
}

__device__ __forceinline__ uint64_t old_color(const EditPass& P, uint32_t mode, uint32_t modeval, const LeafDesc& d,
                                              uint64_t index) {
    if (mode == 1) return pack_color(modeval, 0, 0);
    if (leaf_valid_index(d, index)) return leaf_decode(P.arena, d, index);
    return 0;  // CompressedColor{}
}

// keep / full slots of interior items at levels >= ctl: one wavefront per slot
__global__ __launch_bounds__(256) void k_color_copy(EditPass P, uint32_t l) {
    const uint32_t n = min(P.count[l], P.cap);
    if (*P.errors) return;
    const Items I = P.items[l];
    const Items T = P.items[P.ctl];
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < n * 8; w += (gridDim.x * blockDim.x) >> 6) {
        const uint32_t i = w >> 3, slot = w;
        const uint32_t kind = I.ckind[slot];
        if (kind != KIND_KEEP && kind != KIND_FULL) continue;
        const uint32_t cnt = I.cnew[slot];
        if (cnt == 0) continue;
        const uint32_t task = I.aux[i];
        uint64_t* out = P.colors + T.cbase[task] + I.cstart[slot];
        if (kind == KIND_FULL) {
            const uint64_t col = pack_color(I.ccolor[slot], 0, 0);
            for (uint32_t k = lane; k < cnt; k += 64) out[k] = col;
        } else {
            const LeafDesc d = task_leaf(P, T.handle[task]);
            const uint32_t mode = I.cmode[i], mv = I.cmodeval[i];
            const uint64_t os = I.costart[slot];
            for (uint32_t k = lane; k < cnt; k += 64) out[k] = old_color(P, mode, mv, d, os + k);
        }
    }
}

// leaf items: fold each voxel's value and colour through the edit list
__global__ __launch_bounds__(256) void k_color_leaf(EditPass P) {
    const uint32_t L = P.levels - 2;
    const uint32_t n = min(P.count[L], P.cap);
    if (*P.errors) return;
    const Items I = P.items[L];
    const Items T = P.items[P.ctl];
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n; i += (gridDim.x * blockDim.x) >> 6) {
        const uint32_t c1 = lane >> 3, c2 = lane & 7u;
        const uint32_t x = (I.px[i] << 2) | (((c1 >> 2) & 1u) << 1) | ((c2 >> 2) & 1u);
        const uint32_t y = (I.py[i] << 2) | (((c1 >> 1) & 1u) << 1) | ((c2 >> 1) & 1u);
        const uint32_t z = (I.pz[i] << 2) | ((c1 & 1u) << 1) | (c2 & 1u);
        uint64_t oldMask = 0;
        if (I.old_[i] != kInvalidNode) oldMask = *reinterpret_cast<const uint64_t*>(P.vpool + I.old_[i]);
        const bool prev = (oldMask >> lane) & 1u;
        const uint32_t oldRank = __popcll(oldMask & ((1ull << lane) - 1ull));
        const uint32_t task = I.aux[i];
        const LeafDesc d = task_leaf(P, T.handle[task]);
        bool v = prev;
        uint64_t col = 0;
        bool haveCol = false;
        const float fx = (float)x, fy = (float)y, fz = (float)z;
        for (uint32_t k = 0; k < I.ecount[i]; ++k) {
            const DevEdit& e = P.edits[P.elist[L][I.estart[i] + k]];
            if (!e_touch(e, fx, fy, fz, fx + 1.0f, fy + 1.0f, fz + 1.0f)) continue;
            const bool nv = e_value(e, v);
            if (nv) {
                uint64_t pc = 0;
                if (e.kind == EDIT_SPHERE_PAINT && v) {
                    if (!haveCol) col = old_color(P, I.cmode[i], I.cmodeval[i], d, I.ostart[i] + oldRank);
                    pc = col;
                }
                col = e_color(e, fx, fy, fz, pc);
                haveCol = true;
            }
            v = nv;
        }
        const uint64_t newMask = __ballot(v);
        if (v) {
            if (!haveCol) col = old_color(P, I.cmode[i], I.cmodeval[i], d, I.ostart[i] + oldRank);
            const uint32_t newRank = __popcll(newMask & ((1ull << lane) - 1ull));
            P.colors[T.cbase[task] + I.nstart[i] + newRank] = col;
        }
    }
}

// colour tree: make sure every task with colours has colour-tree nodes down to its leaf slot
// and a unique leaf handle (hash_dag_edits.h:225-249); one thread, tasks in index order.
__global__ void k_color_tree(EditPass P) {
    if (blockIdx.x != 0 || threadIdx.x != 0 || *P.errors) return;
    const uint32_t ctl = P.ctl, D = P.levels;
    const uint32_t n = min(P.count[ctl], P.cap);
    const Items T = P.items[ctl];
    for (uint32_t i = 0; i < n; ++i) {
        if (T.ncount[i] == 0) {
            T.newhandle[i] = 0;
            continue;
        }
        uint32_t idx = 0;
        const uint32_t x = T.px[i], y = T.py[i], z = T.pz[i];
        bool ok = true;
        for (uint32_t lvl = 1; lvl < ctl; ++lvl) {
            const uint32_t s = ctl - lvl;
            const uint32_t c = (((x >> s) & 1u) << 2) | (((y >> s) & 1u) << 1) | ((z >> s) & 1u);
            uint32_t e = P.cnodes[idx + c];
            if (e == 0) {
                e = *P.n_cnodes;
                if (e + 8 > P.cnodes_cap) {
                    set_error(P.errors, ERR_COLOR_TREE);
                    ok = false;
                    break;
                }
                *P.n_cnodes = e + 8;
                P.cnodes[idx + c] = e;
            }
            idx = e;
        }
        if (!ok) break;
        const uint32_t c = ((x & 1u) << 2) | ((y & 1u) << 1) | (z & 1u);
        uint32_t h = P.cnodes[idx + c];
        const LeafDesc prevDesc = task_leaf(P, h);
        if (h == 0 || !(h & 0x80000000u)) {
            const uint32_t li = *P.n_leaves;
            if (li >= P.leaves_cap) {
                set_error(P.errors, ERR_COLOR_TREE);
                break;
            }
            *P.n_leaves = li + 1;
            h = li | 0x80000000u;
            P.cnodes[idx + c] = h;
        }
        T.newhandle[i] = h;
        // undo log: handle + the leaf it resolved to before this pass
        const uint32_t u = *P.undo_count;
        if (u < P.undo_cap) {
            P.undo_handles[u] = h;
            P.undo_descs[u] = prevDesc;
            *P.undo_count = u + 1;
        } else {
            set_error(P.errors, ERR_UNDO);
        }
        (void)D;
    }
}

// One workgroup per task: encode its colour array into a new unique leaf in the arena.
constexpr uint32_t kEncBlock = 256;

__global__ __launch_bounds__(kEncBlock) void k_encode(EditPass P) {
    const uint32_t ctl = P.ctl;
    const uint32_t n = min(P.count[ctl], P.cap);
    if (*P.errors) return;
    const Items T = P.items[ctl];
    __shared__ uint32_t sScanA[kEncBlock], sScanB[kEncBlock];
    __shared__ uint32_t sMacroBlock[160], sMacroBit[160];
    __shared__ uint64_t sBase;
    __shared__ uint32_t sCarryBlocks, sCarryBits;
    const uint32_t tid = threadIdx.x;
    for (uint32_t t = blockIdx.x; t < n; t += gridDim.x) {
        const uint32_t cnt = T.ncount[t];
        if (cnt == 0) continue;
        const uint64_t* cols = P.colors + T.cbase[t];
        const uint32_t nMacro = (cnt + 16383u) / 16384u;
        if (nMacro > 160) {
            if (tid == 0) set_error(P.errors, ERR_COLOR_SCRATCH);
            continue;
        }
        // pass 1: number of blocks and weight bits
        uint32_t myBlocks = 0, myBits = 0;
        for (uint32_t i = tid; i < cnt; i += kEncBlock) {
            const uint64_t c = cols[i];
            const uint32_t bpw = (uint32_t)(c >> 40) & 0xFFu;
            bool start = (i % 16384u) == 0;
            if (!start) {
                const uint64_t p = cols[i - 1];
                start = (uint32_t)p != (uint32_t)c || ((p >> 40) & 0xFFu) != bpw;
            }
            myBlocks += start;
            myBits += bpw;
        }
        sScanA[tid] = myBlocks;
        sScanB[tid] = myBits;
        __syncthreads();
        for (uint32_t s = kEncBlock / 2; s > 0; s >>= 1) {
            if (tid < s) {
                sScanA[tid] += sScanA[tid + s];
                sScanB[tid] += sScanB[tid + s];
            }
            __syncthreads();
        }
        const uint32_t nBlocks = sScanA[0], nBits = sScanB[0];
        const uint32_t nWeights = (nBits + 31) / 32;
        const uint64_t words = (uint64_t)nBlocks + 2ull * nMacro + (nWeights + 1) / 2 + 1;
        if (tid == 0) {
            sBase = atomicAdd(P.arena_top, (unsigned long long)words);
            sCarryBlocks = 0;
            sCarryBits = 0;
        }
        __syncthreads();
        const uint64_t base = sBase;
        if (base + words > P.arena_cap) {
            if (tid == 0) set_error(P.errors, ERR_COLOR_ARENA);
            __syncthreads();
            continue;
        }
        uint64_t* blocks = P.arena + base;
        uint64_t* macro = blocks + nBlocks;
        uint32_t* weights = reinterpret_cast<uint32_t*>(macro + 2 * nMacro);
        for (uint32_t k = tid; k < nWeights + 2; k += kEncBlock) weights[k] = 0u;
        __syncthreads();
        // pass 2: chunks of kEncBlock colours in order, exclusive scans carried across chunks
        for (uint32_t c0 = 0; c0 < cnt; c0 += kEncBlock) {
            const uint32_t i = c0 + tid;
            uint64_t c = 0;
            uint32_t bpw = 0, start = 0;
            if (i < cnt) {
                c = cols[i];
                bpw = (uint32_t)(c >> 40) & 0xFFu;
                start = (i % 16384u) == 0;
                if (!start) {
                    const uint64_t p = cols[i - 1];
                    start = (uint32_t)p != (uint32_t)c || ((p >> 40) & 0xFFu) != bpw;
                }
            }
            sScanA[tid] = start;
            sScanB[tid] = bpw;
            __syncthreads();
            // Hillis-Steele inclusive scan
            for (uint32_t off = 1; off < kEncBlock; off <<= 1) {
                const uint32_t a = tid >= off ? sScanA[tid - off] : 0u;
                const uint32_t b = tid >= off ? sScanB[tid - off] : 0u;
                __syncthreads();
                sScanA[tid] += a;
                sScanB[tid] += b;
                __syncthreads();
            }
            const uint32_t blockIdxIncl = sCarryBlocks + sScanA[tid];  // block index + 1 if start
            const uint32_t bitOff = sCarryBits + sScanB[tid] - bpw;    // exclusive bit offset
            if (i < cnt && (i % 16384u) == 0) {
                sMacroBlock[i / 16384u] = blockIdxIncl - 1;
                sMacroBit[i / 16384u] = bitOff;
            }
            __syncthreads();
            if (i < cnt) {
                const uint32_t m = i / 16384u;
                if (start) {
                    const uint32_t rel = bitOff - sMacroBit[m];
                    const uint32_t hdr = bpw == 0 ? ((0xFFFFu << 16) | (i % 16384u))
                                                  : ((rel << 16) | ((bpw - 1) << 14) | (i % 16384u));
                    blocks[blockIdxIncl - 1] = ((uint64_t)(uint32_t)c << 32) | hdr;
                }
                if (bpw) {
                    // MSB-first stream bit k -> stored word k / 32, byte (k % 32) / 8, bit 7 - k % 8
                    const uint32_t wv = (uint32_t)(c >> 32) & 0xFFu;
                    for (uint32_t b = 0; b < bpw; ++b) {
                        if (!((wv >> (bpw - 1 - b)) & 1u)) continue;
                        const uint32_t k = bitOff + b;
                        atomicOr(&weights[k >> 5], 1u << (((k & 31u) >> 3) * 8u + (7u - (k & 7u))));
                    }
                }
            }
            __syncthreads();
            if (tid == kEncBlock - 1) {
                sCarryBlocks = blockIdxIncl;
                sCarryBits = bitOff + bpw;
            }
            __syncthreads();
        }
        for (uint32_t m = tid; m < nMacro; m += kEncBlock) {
            macro[2 * m] = sMacroBlock[m];
            macro[2 * m + 1] = sMacroBit[m];
        }
        if (tid == 0) {
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
        __syncthreads();
    }
}

__global__ void k_init_pass(EditPass P, uint32_t root, uint32_t first, uint32_t nEdits) {
    // level-0 item = the root, carrying the edits [first, first + nEdits)
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

__global__ void k_finish(EditPass P, uint32_t* root_out) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && *P.errors == 0) {
        const uint32_t r = P.items[0].nidx[0];
        if (r == kInvalidNode) set_error(P.errors, ERR_EMPTY);
        else *root_out = r;
    }
}

}  // namespace

// Host: scratch management and pass orchestration
hd_status edit_reserve(hd_ctx* ctx, uint32_t item_cap, uint32_t elist_cap, uint64_t color_cap) {
    EditScratch& S = ctx->edit;
    const uint32_t levels = ctx->g.levels;
    if (S.item_cap >= item_cap && S.elist_cap >= elist_cap && S.color_cap >= color_cap && S.levels == levels) return HD_OK;
    edit_free(ctx);
    S.levels = levels;
    S.item_cap = item_cap;
    S.elist_cap = elist_cap;
    S.color_cap = color_cap;
    uint32_t hcap = 1;
    while (hcap < 2 * item_cap) hcap <<= 1;
    S.hcap = hcap;
    auto alloc = [&](void** p, uint64_t bytes) -> bool {
        if (hipMalloc(p, std::max<uint64_t>(bytes, 16)) != hipSuccess) return false;
        S.allocs.push_back(*p);
        return true;
    };
    const uint64_t c = item_cap, c8 = 8ull * item_cap;
    for (uint32_t l = 0; l < levels; ++l) {
        Items& I = S.items[l];
        bool ok = alloc((void**)&I.px, c * 4) && alloc((void**)&I.py, c * 4) && alloc((void**)&I.pz, c * 4) &&
                  alloc((void**)&I.old_, c * 4) && alloc((void**)&I.orig, c * 4) && alloc((void**)&I.parent, c * 4) &&
                  alloc((void**)&I.slot, c) && alloc((void**)&I.estart, c * 4) && alloc((void**)&I.ecount, c * 4) &&
                  alloc((void**)&I.aux, c * 4) && alloc((void**)&I.handle, c * 4) &&
                  alloc((void**)&I.newhandle, c * 4) && alloc((void**)&I.cmode, c * 4) &&
                  alloc((void**)&I.cmodeval, c * 4) && alloc((void**)&I.ckind, c8) && alloc((void**)&I.cval, c8 * 4) &&
                  alloc((void**)&I.corig, c8 * 4) && alloc((void**)&I.cold, c8 * 4) &&
                  alloc((void**)&I.cnew, c8 * 4) && alloc((void**)&I.ccolor, c8 * 4) &&
                  alloc((void**)&I.cstart, c8 * 4) && alloc((void**)&I.costart, c8 * 4) &&
                  alloc((void**)&I.nidx, c * 4) && alloc((void**)&I.ncount, c * 4) && alloc((void**)&I.cont, c * 36) &&
                  alloc((void**)&I.clen, c * 4) && alloc((void**)&I.hash, c * 4) && alloc((void**)&I.rep, c * 4) &&
                  alloc((void**)&I.state, c * 4) && alloc((void**)&I.nstart, c * 4) &&
                  alloc((void**)&I.ostart, c * 4) && alloc((void**)&I.cbase, c * 8) &&
                  alloc((void**)&S.elist[l], (uint64_t)elist_cap * 4);
        if (!ok) {
            edit_free(ctx);
            return fail(ctx, HD_ERR_OOM, "edit scratch allocation failed");
        }
    }
    bool ok = alloc((void**)&S.hset, (uint64_t)hcap * 4) && alloc((void**)&S.colors, color_cap * 8) &&
              alloc((void**)&S.dcount, 64 * 4) && alloc((void**)&S.decount, 64 * 4) && alloc((void**)&S.dstats, 64 * 8) &&
              alloc((void**)&S.undo_handles, (uint64_t)kUndoCap * 4) &&
              alloc((void**)&S.undo_descs, (uint64_t)kUndoCap * sizeof(LeafDesc)) && alloc((void**)&S.edits_dev, 4096 * sizeof(DevEdit));
    if (!ok) {
        edit_free(ctx);
        return fail(ctx, HD_ERR_OOM, "edit scratch allocation failed");
    }
    S.edits_cap = 4096;
    return HD_OK;
}

void edit_free(hd_ctx* ctx) {
    EditScratch& S = ctx->edit;
    for (void* p : S.allocs) hipFree(p);
    S.allocs.clear();
    S = EditScratch{};
}

static void fill_pass(hd_ctx* ctx, EditPass& P) {
    EditScratch& S = ctx->edit;
    P.g = ctx->g;
    P.levels = ctx->g.levels;
    P.ctl = ctx->g.color_tree_levels;
    P.vpool = ctx->vpool;
    P.page_table = ctx->page_table;
    P.bucket_sizes = ctx->bucket_sizes;
    P.full_nodes = ctx->full_nodes;
    P.pool_top = ctx->counters + 0;
    P.errors = ctx->counters + 1;
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
    P.cap = S.item_cap;
    P.ecap = S.elist_cap;
    P.hset = S.hset;
    P.hcap = S.hcap;
    P.edits = S.edits_dev;
    P.stat_voxels = reinterpret_cast<unsigned long long*>(S.dstats + 0);
    P.stat_new_words = reinterpret_cast<unsigned long long*>(S.dstats + 1);
    P.color_top = reinterpret_cast<unsigned long long*>(S.dstats + 2);
    P.arena_top = reinterpret_cast<unsigned long long*>(S.dstats + 3);
    P.colors = S.colors;
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
    P.undo_cap = kUndoCap;
}

// Enqueues one pass over edits [first, first + n) of the uploaded edit array.
hd_status enqueue_pass(hd_ctx* ctx, uint32_t first, uint32_t n, bool colors, hipStream_t s) {
    EditPass P{};
    fill_pass(ctx, P);
    P.single_edit = n == 1 ? 1u : 0u;
    const uint32_t D = P.levels, L = D - 2;
    k_init_pass<<<(n + 255) / 256, 256, 0, s>>>(P, 0xFFFFFFFFu, first, n);
    // the root index lives in ctx->d_root (device) so passes chain without a host round trip
    hipMemcpyAsync(&P.items[0].old_[0], ctx->d_root, 4, hipMemcpyDeviceToDevice, s);
    hipMemcpyAsync(&P.items[0].orig[0], ctx->d_root, 4, hipMemcpyDeviceToDevice, s);
    for (uint32_t l = 0; l < L; ++l) k_expand<<<kGrid, 256, 0, s>>>(P, l);
    k_leaf<<<kGrid, 256, 0, s>>>(P);
    for (int l = (int)L; l >= 0; --l) {
        if (l < (int)L) k_build<<<kGrid, 256, 0, s>>>(P, (uint32_t)l);
        k_fill<<<256, 256, 0, s>>>(P.hset, P.hcap, 0u);
        k_dedup<<<kGrid, 256, 0, s>>>(P, (uint32_t)l);
        k_find<<<kGrid, 256, 0, s>>>(P, (uint32_t)l);
        k_add<<<kGrid, 256, 0, s>>>(P, (uint32_t)l);
        k_resolve<<<kGrid, 256, 0, s>>>(P, (uint32_t)l);
    }
    if (colors) {
        for (uint32_t l = P.ctl; l < L; ++l) k_color_starts<<<kGrid, 256, 0, s>>>(P, l);
        k_color_alloc<<<kGrid, 256, 0, s>>>(P);
        k_color_tree<<<1, 64, 0, s>>>(P);
        for (uint32_t l = P.ctl; l < L; ++l) k_color_copy<<<kGrid, 256, 0, s>>>(P, l);
        k_color_leaf<<<kGrid, 256, 0, s>>>(P);
        k_encode<<<kGrid, kEncBlock, 0, s>>>(P);
    }
    k_finish<<<1, 64, 0, s>>>(P, ctx->d_root);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "edit pass launch");
    return HD_OK;
}

}  // namespace hd
