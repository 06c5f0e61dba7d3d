// Oracle: host restatement of Tracer::trace_paths / trace_colors / trace_shadows and the
// camera ray setup. TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
//
// The reference kernels are device code; where device overload resolution differs from host
// code it is followed: max(float,float) / max(double,double) in tracer.cu resolve to fmaxf /
// fmax under nvcc, while the vector max/min of cuda_math.h are ternary chains.
#include "oracle_core.hpp"

#include <cmath>

namespace orc {

// dag_tracer.cu:71-113 (all FP64; camera_view.h:8-27: right = -row0, up = row1, forward = row2)
RayParams get_trace_params(const double position[3], const double rot[9], const double bmin[3], const double bmax[3],
                           u32 levels, u32 W, u32 H) {
    const D3 pos{position[0], position[1], position[2]};
    const D3 fwd{rot[6], rot[7], rot[8]};
    const D3 up{rot[3], rot[4], rot[5]};
    const D3 right{-rot[0], -rot[1], -rot[2]};
    const D3 boundsMin{bmin[0], bmin[1], bmin[2]};
    const D3 boundsMax{bmax[0], bmax[1], bmax[2]};
    const double fov = 60.0 / 2.0 * (double(M_PI) / 180.);
    const double aspect = double(W) / double(H);
    const D3 X = right * std::sin(fov) * aspect;
    const D3 Y = up * std::sin(fov);
    const D3 Z = fwd * std::cos(fov);
    const D3 bl = pos + Z - Y - X;
    const D3 br = pos + Z - Y + X;
    const D3 tl = pos + Z + Y - X;
    const D3 translation{-boundsMin.x, -boundsMin.y, -boundsMin.z};
    const double side = double(1 << levels);
    const D3 scale = D3{side, side, side} / (boundsMax - boundsMin);
    const D3 fPos = (pos + translation) * scale;
    const D3 fBl = (bl + translation) * scale;
    const D3 fTl = (tl + translation) * scale;
    const D3 fBr = (br + translation) * scale;
    RayParams p;
    p.cameraPosition = fPos;
    p.rayMin = fBl;
    p.rayDDx = (fBr - fBl) * (1.0 / W);
    p.rayDDy = (fTl - fBl) * (1.0 / H);
    return p;
}

// tracer.cu:19-136
template <bool isRoot>
static u8 intersection_mask(u32 levels, u32 level, const Path& path, F3 o, F3 d, F3 inv) {
    const u32 shift = levels - level;
    const float radius = (float)(1u << (shift - 1));
    const F3 center = f3(radius) + path.as_position(shift);
    const F3 rel = center - o;
    const F3 tmid = rel * inv;
    float tmin, tmax;
    {
        const F3 slab = radius * F3{std::fabs(inv.x), std::fabs(inv.y), std::fabs(inv.z)};
        const F3 pmin = tmid - slab;
        tmin = std::fmax(tmax3(pmin), 0.0f);
        const F3 pmax = tmid + slab;
        tmax = tmin3(pmax);
    }
    if (isRoot && (tmin >= tmax)) return 0;
    u8 mask = 0;
    {
        const F3 por = (0.5f * (tmin + tmax)) * d;
        const u8 first = (u8)((por.x >= rel.x ? 4 : 0) + (por.y >= rel.y ? 2 : 0) + (por.z >= rel.z ? 1 : 0));
        mask |= (u8)(1u << first);
    }
    const float eps = 1e-4f;
    if (tmin <= tmid.x && tmid.x <= tmax) {
        const F3 por = tmid.x * d;
        u8 A = 0, B = 0;
        if (por.y >= rel.y - eps) A |= 0xCC;
        if (por.y <= rel.y + eps) A |= 0x33;
        if (por.z >= rel.z - eps) B |= 0xAA;
        if (por.z <= rel.z + eps) B |= 0x55;
        mask |= (u8)(A & B);
    }
    if (tmin <= tmid.y && tmid.y <= tmax) {
        const F3 por = tmid.y * d;
        u8 C = 0, D = 0;
        if (por.x >= rel.x - eps) C |= 0xF0;
        if (por.x <= rel.x + eps) C |= 0x0F;
        if (por.z >= rel.z - eps) D |= 0xAA;
        if (por.z <= rel.z + eps) D |= 0x55;
        mask |= (u8)(C & D);
    }
    if (tmin <= tmid.z && tmid.z <= tmax) {
        const F3 por = tmid.z * d;
        u8 E = 0, F = 0;
        if (por.x >= rel.x - eps) E |= 0xF0;
        if (por.x <= rel.x + eps) E |= 0x0F;
        if (por.y >= rel.y - eps) F |= 0xCC;
        if (por.y <= rel.y + eps) F |= 0x33;
        mask |= (u8)(E & F);
    }
    return mask;
}

// tracer.cu:7-17
static u8 next_child(u8 order, u8 mask) {
    for (u8 c = 0; c < 8; ++c) {
        const u8 ci = c ^ order;
        if (mask & (1u << ci)) return ci;
    }
    return 0;
}

struct StackEntry { u32 index; u8 childMask; u8 visitMask; };

// Algorithmic HBM bytes (page-table words excluded: the build's device pool is addressed
// by virtual word index): node header 4 B, child pointer 4 B, leaf 8 B.
// Child order of the any-hit walk. The reference visits the highest set bit first
// (tracer.cu:493); the device visits children front to back along the ray. The result of an
// any-hit walk does not depend on the order (it is whether any voxel is reached through the
// intersection masks, which depend only on the node and the ray), but the bytes it reads do:
// the byte-counter comparison switches this to the device's order.
static bool g_any_hit_front_to_back = false;

template <bool anyHit>
static bool descend(const DAG& dag, F3 o, F3 d, F3 inv, u8 order, Path& path, u64* bytes) {
    const Geometry& g = dag.g();
    const u32 levels = g.levels, leafLevel = g.leaf_level();
    u32 level = 0;
    StackEntry stack[32];
    StackEntry cache;
    u64 leaf = 0;
    cache.index = dag.root;
    cache.childMask = child_mask(dag.node(cache.index));
    if (bytes) *bytes += 4;
    cache.visitMask = cache.childMask & intersection_mask<true>(levels, 0, path, o, d, inv);
    for (;;) {
        {
            u32 newLevel = level;
            while (newLevel > 0 && !cache.visitMask) {
                newLevel--;
                cache = stack[newLevel];
            }
            if (newLevel == 0 && !cache.visitMask) {
                path = Path(0, 0, 0);
                return false;
            }
            path.ascend(level - newLevel);
            level = newLevel;
        }
        // tracer.cu:201 (ordered) / :493 (highest set bit first, 31 - clz)
        const u8 next = anyHit && !g_any_hit_front_to_back ? (u8)(31 - __builtin_clz((u32)cache.visitMask))
                                                            : next_child(order, cache.visitMask);
        cache.visitMask &= (u8)~(1u << next);
        path.descend(next);
        stack[level] = cache;
        level++;
        if (level == levels) return true;
        if (level < leafLevel) {
            cache.index = dag.child_index(cache.index, cache.childMask, next);
            cache.childMask = child_mask(dag.node(cache.index));
            if (bytes) *bytes += 8;
            cache.visitMask = cache.childMask & intersection_mask<false>(levels, level, path, o, d, inv);
        } else {
            u8 cm;
            if (level == leafLevel) {
                const u32 addr = dag.child_index(cache.index, cache.childMask, next);
                leaf = dag.leaf(addr);
                if (bytes) *bytes += 12;
                cm = leaf_first_mask(leaf);
            } else {
                cm = leaf_second_mask(leaf, next);
            }
            cache.childMask = cm;
            cache.visitMask = cache.childMask & intersection_mask<false>(levels, level, path, o, d, inv);
        }
    }
}

// tracer.cu:145-252 for one pixel; returns the stored path (0,0,0 on miss).
U3 trace_path(const DAG& dag, const RayParams& p, u32 px, u32 py, u64* bytes) {
    const F3 o = to_f3(p.cameraPosition);
    const D3 dd = normalized(p.rayMin + (double)px * p.rayDDx + (double)py * p.rayDDy - p.cameraPosition);
    const F3 d = to_f3(dd);
    const F3 inv = to_f3(D3{1. / (double)d.x, 1. / (double)d.y, 1. / (double)d.z});
    const u8 order = (u8)((d.x < 0.f ? 4 : 0) + (d.y < 0.f ? 2 : 0) + (d.z < 0.f ? 1 : 0));
    Path path(0, 0, 0);
    descend<false>(dag, o, d, inv, order, path, bytes);
    if (bytes) *bytes += 16;  // output
    return path.p;
}

// tracer.cu:453-543
bool intersect_any(const DAG& dag, F3 o, F3 d, u64* bytes) {
    const F3 inv = to_f3(D3{1. / (double)d.x, 1. / (double)d.y, 1. / (double)d.z});
    Path path(0, 0, 0);
    const u8 order = (u8)((d.x < 0.f ? 4 : 0) + (d.y < 0.f ? 2 : 0) + (d.z < 0.f ? 1 : 0));
    return descend<true>(dag, o, d, inv, order, path, bytes);
}

void set_any_hit_front_to_back(bool on) { g_any_hit_front_to_back = on; }

// tracer.cu:254-451 for one pixel (debugColors = None, TOOL_OVERLAY = 0)
u32 trace_color(const DAG& dag, const HashDAGColors& colors, U3 pp) {
    const Geometry& g = dag.g();
    const Path path(pp.x, pp.y, pp.z);
    if (path.is_null()) return float3_to_rgb888(F3{187, 242, 250} / 255.f);
    auto invalid = [&]() -> u32 {
        const u32 b = (path.p.x ^ path.p.y ^ path.p.z) & 0x1;
        return float3_to_rgb888(F3{1, (float)b, 1.f - (float)b});
    };
    u64 nof = 0;
    u32 colorNode = 0;
    LeafView leaf;  // default: invalid
    u32 level = 0;
    u32 nodeIndex = dag.root;
    while (level < g.leaf_level()) {
        level++;
        const u32 node = dag.node(nodeIndex);
        const u8 mask = child_mask(node);
        const u8 child = path.child_index(level, g.levels);
        if (!(mask & (1 << child))) return 0xFF00FF;
        if (level - 1 < g.color_tree_levels) {
            colorNode = colors.nodes[colorNode + child];
            if (level == g.color_tree_levels) {
                leaf = colors.get_leaf(colorNode);
            } else if (!colorNode) {
                return invalid();
            }
        }
        if (level == g.leaf_level()) {
            for (u8 c = 0; c < child; ++c)
                if (mask & (1u << c)) nof += popcll(dag.leaf(dag.child_index(nodeIndex, mask, c)));
            const u64 lf = dag.leaf(dag.child_index(nodeIndex, mask, child));
            const u32 bit = ((path.p.x & 1) ? 4 : 0) | ((path.p.y & 1) ? 2 : 0) | ((path.p.z & 1) ? 1 : 0) |
                            ((path.p.x & 2) ? 32 : 0) | ((path.p.y & 2) ? 16 : 0) | ((path.p.z & 2) ? 8 : 0);
            nof += popcll(lf & ((1ull << bit) - 1));
            break;
        } else {
            if (level > g.color_tree_levels) {
                for (u8 c = 0; c < child; ++c)
                    if (mask & (1u << c)) nof += dag.node(dag.child_index(nodeIndex, mask, c)) >> 8;
            }
            nodeIndex = dag.child_index(nodeIndex, mask, child);
        }
    }
    if (!leaf.is_valid() || !leaf.is_valid_index(nof)) return invalid();
    return float3_to_rgb888(leaf.get_color(nof).get_color());
}

// tracer.cu:546-549
F3 sun_direction() { return normalizef(F3{0.3f, 1.f, 0.5f}); }

static F3 lerp3(F3 a, F3 b, float f) { return a * (1 - f) + b * f; }

// tracer.cu:551-572 (the FOG #else branch)
static F3 apply_fog(F3 rgb, double distance, D3 rayDir, float fogDensity) {
    fogDensity *= 0.00001f;
    const double fogAmount = 1.0 - std::exp(-distance * fogDensity);
    const double sunAmount = 1.01f * std::fmax(dotd(rayDir, to_d3(sun_direction())), 0.0);
    const F3 fogColor = lerp3(F3{187, 242, 250} / 255.f, f3(1.0f), (float)std::pow(sunAmount, 30.0));
    return lerp3(rgb, fogColor, tclamp((float)fogAmount, 0.f, 1.f));
}

// tracer.cu:574-587
static D3 ray_box_intersection(D3 orig, D3 dir, D3 bmin, D3 bmax) {
    const D3 tmin = (bmin - orig) / dir;
    const D3 tmax = (bmax - orig) / dir;
    const D3 rmin{tmin.x < tmax.x ? tmin.x : tmax.x, tmin.y < tmax.y ? tmin.y : tmax.y, tmin.z < tmax.z ? tmin.z : tmax.z};
    const double maxmin = std::fmax(std::fmax(rmin.x, rmin.y), rmin.z);
    return orig + dir * maxmin;
}

// tracer.cu:589-697 for one pixel (EXACT_SHADOWS = 1, PER_VOXEL_FACE_SHADING = 0).
// colorIn is the colours image value at this pixel (read-modify-write).
u32 shade_shadow(const DAG& dag, const RayParams& p, u32 W, u32 H, u32 px, u32 py, U3 pp, u32 colorIn, float bias,
                 float fog, u64* bytes) {
    (void)W;
    auto shade = [&](float light, double distance, D3 dir) -> u32 {
        F3 c = rgb888_to_float3(colorIn);
        c = c * tclamp(0.5f + light, 0.f, 1.f);
        c = apply_fog(c, distance, dir, fog);
        return float3_to_rgb888(c);
    };
    const F3 origin{(float)pp.x, (float)pp.y, (float)pp.z};
    const D3 camDir = normalized(p.rayMin + (double)px * p.rayDDx + (double)(H - 1 - py) * p.rayDDy - p.cameraPosition);
    const D3 od = to_d3(origin);
    const D3 hit = ray_box_intersection(p.cameraPosition, camDir, od, od + 1.0);
    const F3 start = to_f3(hit);
    if (lengthf(origin) == 0.0f) return shade(1, 1e9, camDir);
    const F3 dir = sun_direction();
    const bool shadowed = intersect_any(dag, start + bias * dir, dir, bytes);
    const D3 v = od - p.cameraPosition;
    const double distance = lengthd(v);
    const D3 nv = v / distance;
    return shadowed ? shade(0, distance, nv) : shade(1, distance, nv);
}

}  // namespace orc
