// Oracle: editors (hash_dag_editors.h) and the two-pass edit recursion
// (hash_dag_edits.h:106-643, hash_dag.h:255-405), single-threaded and deterministic.
// TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
#include "oracle_core.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <exception>
#include <mutex>
#include <thread>
#include <cmath>
#include <map>
#include <stdexcept>
#include <tuple>

namespace orc {

namespace {

static float sq(float f) { return f * f; }
static float tmaxf(float a, float b) { return a > b ? a : b; }  // host ::max (cuda_math.h:7-8)

// float -> uint32 as x86-64 code converts it (truncate to int64, keep the low 32 bits): what
// the reference's uint32(p.x) yields for the source positions of a copy
static u32 to_u32(float f) { return (u32)(int64_t)f; }

// DAGUtils::get_value (dag_utils.h:138-172) from an explicit root
static bool value_at(const HashTable& t, u32 root, Path path) {
    const Geometry& g = t.g;
    u32 va = root;
    for (u32 level = 0; level < g.levels; ++level) {
        if (level < g.leaf_level()) {
            const u32 hdr = *t.ptr(va);
            const u8 mask = child_mask(hdr);
            const u8 c = path.child_index(level + 1, g.levels);
            if (!(mask & (1u << c))) return false;
            va = *t.ptr(va + child_offset(mask, c));
        } else {
            const u32* p = t.ptr(va);
            const u64 leaf = (u64)p[0] | ((u64)p[1] << 32);
            const u32 bit = ((path.p.x & 1) ? 4 : 0) | ((path.p.y & 1) ? 2 : 0) | ((path.p.z & 1) ? 1 : 0) |
                            ((path.p.x & 2) ? 32 : 0) | ((path.p.y & 2) ? 16 : 0) | ((path.p.z & 2) ? 8 : 0);
            return (leaf >> bit) & 1;
        }
    }
    return true;
}

// DAGUtils::get_color (dag_utils.h:423-523) from an explicit root: colour index = voxels
// before the path's voxel in its level-color_tree_levels subtree, looked up in that subtree's
// leaf; an absent voxel or colour gives CompressedColor{}
static CompressedColor color_at(const HashTable& t, u32 root, const HashDAGColors& colors, Path path) {
    const Geometry& g = t.g;
    const u32 ctl = g.color_tree_levels, L = g.leaf_level();
    u64 nof = 0;
    u32 colorNode = 0;
    LeafView leaf{};
    bool haveLeaf = false;
    u32 level = 0;
    u32 va = root;
    while (level < L) {
        level++;
        const u32 hdr = *t.ptr(va);
        const u8 mask = child_mask(hdr);
        const u8 child = path.child_index(level, g.levels);
        if (!(mask & (1u << child))) return CompressedColor{};
        if (level - 1 < ctl) {
            colorNode = colors.nodes[colorNode + child];
            if (level == ctl) {
                leaf = colors.get_leaf(colorNode);
                haveLeaf = true;
            } else if (!colorNode) {
                return CompressedColor{};
            }
        }
        if (level == L) {
            for (u8 c = 0; c < child; ++c)
                if (mask & (1u << c)) {
                    const u32* p = t.ptr(*t.ptr(va + child_offset(mask, c)));
                    nof += popcll((u64)p[0] | ((u64)p[1] << 32));
                }
            const u32* p = t.ptr(*t.ptr(va + child_offset(mask, child)));
            const u64 lf = (u64)p[0] | ((u64)p[1] << 32);
            const u32 bit = ((path.p.x & 1) ? 4 : 0) | ((path.p.y & 1) ? 2 : 0) | ((path.p.z & 1) ? 1 : 0) |
                            ((path.p.x & 2) ? 32 : 0) | ((path.p.y & 2) ? 16 : 0) | ((path.p.z & 2) ? 8 : 0);
            nof += popcll(lf & ((1ull << bit) - 1));
            break;
        }
        if (level > ctl)
            for (u8 c = 0; c < child; ++c)
                if (mask & (1u << c)) nof += *t.ptr(*t.ptr(va + child_offset(mask, c))) >> 8;
        va = *t.ptr(va + child_offset(mask, child));
    }
    if (!haveLeaf || !leaf.is_valid() || !leaf.is_valid_index(nof)) return CompressedColor{};
    return leaf.get_color(nof);
}

struct Editor {
    Edit e;
    F3 bmin{}, bmax{};  // BoxEditorBase (hash_dag_editors.h:57-74)
    float r2 = 0;       // SphereEditorBase (hash_dag_editors.h:99-110)
    // copy (CopyEditorWithoutDecompression<true>, hash_dag_editors.h:332-435): the DAG before
    // the edit and the colours it reads from, and the inverse of the copy's transform
    const HashTable* srcTable = nullptr;
    u32 srcRoot = 0;
    const HashDAGColors* srcColors = nullptr;
    double inv[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    // fill (FillEditorColors over BufferEditor, hash_dag_editors.h:129-222, 624-768): the buffer
    // of voxel states around the centre and the colours of the voxels the flood fill set
    enum : u8 { Unknown = 0, LoadedTrue = 1, LoadedFalse = 2, SetTrue = 3, SetFalse = 4 };
    mutable std::vector<u8> values;
    std::vector<CompressedColor> fillColors;
    F3 fmin{};
    u32 fsize = 0;
    // copy with decompression (hash_dag_editors.h:437-616): the source box's float and integer
    // minimum and the side of its voxel buffer
    F3 copyMin{};
    U3 intCopyMin{0, 0, 0};
    u32 dsize = 0;

    explicit Editor(const Edit& ed) : e(ed) {
        F3 c = e.center;
        float radius = e.radius;
        if (e.kind == COPY) {
            // BoxEditorBase(dest + float3(T * (center - source)), radius * max|T * (1,1,1)|)
            // (hash_dag_editors.h:350-351), Matrix3x3 in FP64 (Matrix3x3.h:190-243, 399-406)
            const double* T = e.transform;
            const F3 rel = e.center - e.source;
            const double r[3] = {(double)rel.x, (double)rel.y, (double)rel.z};
            c = e.dest + F3{(float)(T[0] * r[0] + T[1] * r[1] + T[2] * r[2]),
                            (float)(T[3] * r[0] + T[4] * r[1] + T[5] * r[2]),
                            (float)(T[6] * r[0] + T[7] * r[1] + T[8] * r[2])};
            const double rx = std::fabs(T[0] * 1.0 + T[1] * 1.0 + T[2] * 1.0);
            const double ry = std::fabs(T[3] * 1.0 + T[4] * 1.0 + T[5] * 1.0);
            const double rz = std::fabs(T[6] * 1.0 + T[7] * 1.0 + T[8] * 1.0);
            radius = (float)((double)e.radius * std::max(rx, std::max(ry, rz)));
            // Matrix3x3::Inverse: adjugate times 1 / Determinate
            const double a[9] = {T[8] * T[4] - T[7] * T[5], T[7] * T[2] - T[8] * T[1], T[5] * T[1] - T[4] * T[2],
                                 T[6] * T[5] - T[8] * T[3], T[8] * T[0] - T[6] * T[2], T[3] * T[2] - T[5] * T[0],
                                 T[7] * T[3] - T[6] * T[4], T[6] * T[1] - T[7] * T[0], T[4] * T[0] - T[3] * T[1]};
            const double v1 = T[0] * (T[8] * T[4] - T[7] * T[5]);
            const double v2 = T[3] * (T[8] * T[1] - T[7] * T[2]);
            const double v3 = T[6] * (T[5] * T[1] - T[4] * T[2]);
            const double s = 1 / (v1 - v2 + v3);
            for (int i = 0; i < 9; ++i) inv[i] = a[i] * s;
        }
        if (e.kind == NOISE_ADD || e.kind == NOISE_CARVE) radius = 1.5f * e.radius;  // BufferEditor(center, 1.5 r)
        if (e.kind == COPY_DEC) {
            // BoxEditorBase(center - source + dest, radius); a buffer of 2 ceil(radius) voxels
            // per axis from truncate(center - radius) (hash_dag_editors.h:455-466)
            c = e.center - e.source + e.dest;
            copyMin = e.center - f3(e.radius);
            intCopyMin = U3{to_u32(copyMin.x), to_u32(copyMin.y), to_u32(copyMin.z)};
            dsize = 2 * (u32)std::ceil(e.radius);
        }
        bmin = c - f3(radius);
        bmax = c + f3(radius);
        boxRadius = radius;
        r2 = e.radius * e.radius;
    }
    float boxRadius = 0;  // BoxEditorBase::radius
    bool is_noise() const { return e.kind == NOISE_ADD || e.kind == NOISE_CARVE; }
    bool is_box() const {
        return e.kind == BOX_ADD || e.kind == BOX_CARVE || e.kind == COPY || e.kind == FILL || e.kind == COPY_DEC ||
               is_noise();
    }

    // ---- CopyEditorWithDecompression ----
    // buffer voxel of a destination voxel corner: truncate(position - dest + source - copyMin)
    // (hash_dag_editors.h:588, 601); false outside the buffer
    bool dec_voxel(F3 pos, Path& src) const {
        const F3 q = pos - e.dest + e.source - copyMin;
        const u32 x = to_u32(q.x), y = to_u32(q.y), z = to_u32(q.z);
        if (!(x < dsize && y < dsize && z < dsize)) return false;
        src = Path(intCopyMin.x + x, intCopyMin.y + y, intCopyMin.z + z);
        return true;
    }
    // DAGUtils::is_empty (dag_utils.h:174-252) on the DAG before the edit: no node of maxLevel
    // overlaps [start, start + size) in the reference's inclusive-max test
    bool src_empty_rec(u32 level, u32 va, Path path, u32 maxLevel, U3 start, U3 size) const {
        if (level == maxLevel) return false;
        const Geometry& g = srcTable->g;
        const u32 shift = g.levels - level;
        const U3 mn{path.p.x << shift, path.p.y << shift, path.p.z << shift};
        const U3 mx{mn.x + ((1u << shift) - 1), mn.y + ((1u << shift) - 1), mn.z + ((1u << shift) - 1)};
        const U3 inMax{start.x + size.x, start.y + size.y, start.z + size.z};
        const bool overlaps = !(mn.x >= inMax.x || mn.y >= inMax.y || mn.z >= inMax.z || mx.x <= start.x ||
                                mx.y <= start.y || mx.z <= start.z || start.x >= mx.x || start.y >= mx.y ||
                                start.z >= mx.z || inMax.x <= mn.x || inMax.y <= mn.y || inMax.z <= mn.z);
        if (level < g.leaf_level() && !overlaps) return true;
        const u32 hdr = *srcTable->ptr(va);
        const u8 m = child_mask(hdr);
        for (u8 c = 0; c < 8; ++c) {
            if (!(m & (1u << c))) continue;
            Path np = path;
            np.descend(c);
            if (!src_empty_rec(level + 1, *srcTable->ptr(va + child_offset(m, c)), np, maxLevel, start, size))
                return false;
        }
        return true;
    }

    // ---- BufferEditor (hash_dag_editors.h:144-209) ----
    bool fvalid(u32 x, u32 y, u32 z) const { return x < fsize && y < fsize && z < fsize; }
    u64 findex(u32 x, u32 y, u32 z) const { return (u64)x + (u64)fsize * y + (u64)fsize * fsize * z; }
    Path world_path(u32 x, u32 y, u32 z) const {
        return Path(to_u32(fmin.x + (float)x), to_u32(fmin.y + (float)y), to_u32(fmin.z + (float)z));
    }
    // get_value: loads an unknown voxel from the DAG before the edit
    bool fget(u32 x, u32 y, u32 z) const {
        u8& v = values[findex(x, y, z)];
        if (v == Unknown) {
            const Path p = world_path(x, y, z);
            v = value_at(*srcTable, srcRoot, p) ? LoadedTrue : LoadedFalse;
        }
        return v == LoadedTrue || v == SetTrue;
    }
    // FillEditorBase (hash_dag_editors.h:629-671) then FillEditorColors (:688-728)
    void build_fill() {
        fmin = e.center - f3(e.radius);
        fsize = 2 * (u32)std::ceil(e.radius);
        values.assign((u64)fsize * fsize * fsize, Unknown);
        fillColors.assign(values.size(), CompressedColor{});
        const u32 c0 = fsize / 2;
        u32 mn[3] = {c0, c0, c0}, mx[3] = {c0, c0, c0};
        std::vector<std::array<u32, 3>> stack{{c0, c0, c0}};
        while (!stack.empty()) {
            const auto q = stack.back();
            stack.pop_back();
            values[findex(q[0], q[1], q[2])] = SetTrue;
            for (int a = 0; a < 3; ++a) {
                mn[a] = std::min(mn[a], q[a]);
                mx[a] = std::max(mx[a], q[a]);
            }
            static const int d[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
            for (const auto& o : d) {
                const u32 n0 = q[0] + (u32)o[0], n1 = q[1] + (u32)o[1], n2 = q[2] + (u32)o[2];
                if (fvalid(n0, n1, n2) && !fget(n0, n1, n2)) stack.push_back({n0, n1, n2});
            }
        }
        bmin = fmin + F3{(float)mn[0], (float)mn[1], (float)mn[2]};
        bmax = fmin + F3{(float)mx[0], (float)mx[1], (float)mx[2]};
        CompressedColor color{};
        for (u32 x = mn[0]; x <= mx[0]; ++x)
            for (u32 y = mn[1]; y <= mx[1]; ++y)
                for (u32 z = mn[2]; z <= mx[2]; ++z) {
                    const u64 i = findex(x, y, z);
                    if (values[i] == LoadedTrue) color = color_at(*srcTable, srcRoot, *srcColors, world_path(x, y, z));
                    else if (values[i] == SetTrue) fillColors[i] = color;
                }
    }
    // SphereNoiseEditor (hash_dag_editors.h:762-792): the buffer of a BufferEditor of radius
    // 1.5 r; voxels whose distance to the buffer centre (integer, FP64) is below
    // r * (1 + Perlin fractal) are set to isAdding, the others are loaded from the DAG on use
    void build_noise() {
        const float br = boxRadius;
        fmin = e.center - f3(br);
        fsize = 2 * (u32)std::ceil(br);
        values.assign((u64)fsize * fsize * fsize, Unknown);
        PerlinFbm noise;
        noise.set_seed(std::abs(1923 + (int32_t)e.center.x + 23 * (int32_t)e.center.y + 2938 * (int32_t)e.center.z));
        noise.frequency = 1 / e.radius;
        noise.set_octaves(3);
        const u32 c0 = fsize / 2;
        const u8 set = e.kind == NOISE_ADD ? SetTrue : SetFalse;
        for (u32 x = 0; x < fsize; ++x)
            for (u32 y = 0; y < fsize; ++y)
                for (u32 z = 0; z < fsize; ++z) {
                    // length(editCenter - position) on uint3: 32-bit differences and dot product
                    const u32 dx = c0 - x, dy = c0 - y, dz = c0 - z;
                    const double distance = std::sqrt((double)(u32)(dx * dx + dy * dy + dz * dz));
                    if (distance < e.radius * (1. + (double)noise.fractal((float)x, (float)y, (float)z)))
                        values[findex(x, y, z)] = set;
                }
    }
    // SphereNoiseEditor::get_new_color (hash_dag_editors.h:794-808)
    CompressedColor noise_color(F3 pos) const {
        CompressedColor c;
        c.bitsPerWeight = 3;
        const u16 mn = float3_to_rgb565(F3{94, 30, 0} / 255.f), mx = float3_to_rgb565(F3{125, 100, 89} / 255.f);
        c.colorBits = (u32)mn | ((u32)mx << 16);
        const F3 d = pos - e.center;
        const float len = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
        const int32_t w = (int32_t)((len / boxRadius - 0.3f) * (float)((1 << 3) - 1));
        c.weight = (u8)(w < 0 ? 0 : (w > 7 ? 7 : w));
        return c;
    }
    // destination_to_source (hash_dag_editors.h:364-379): position - dest, swirled about the y
    // axis (host cosf / sinf), through the inverse transform in FP64, plus source
    Path source_of(F3 pos) const {
        F3 rel = pos - e.dest;
        if (e.swirl) {
            const float angle = rel.y / e.swirlPeriod;
            const float x = rel.x, z = rel.z;
            rel.x = x * std::cos(angle) - z * std::sin(angle);
            rel.z = x * std::sin(angle) + z * std::cos(angle);
        }
        const double v[3] = {(double)rel.x, (double)rel.y, (double)rel.z};
        rel = F3{(float)(inv[0] * v[0] + inv[1] * v[1] + inv[2] * v[2]),
                 (float)(inv[3] * v[0] + inv[4] * v[1] + inv[5] * v[2]),
                 (float)(inv[6] * v[0] + inv[7] * v[1] + inv[8] * v[2])};
        const F3 p = e.source + rel;
        return Path(to_u32(p.x), to_u32(p.y), to_u32(p.z));
    }
    // BoxEditorBase::edit_count_impl (hash_dag_editors.h:92-96), 32-bit product
    u32 edit_count() const {
        const F3 ext = bmax - bmin;
        return (u32)ext.x * (u32)ext.y * (u32)ext.z;
    }
    bool adding() const { return e.kind == SPHERE_ADD || e.kind == BOX_ADD; }

    // hash_dag_editors.h:112-126 / 76-91
    bool should_edit_impl(F3 mn, F3 mx) const {
        if (is_box()) {
            return !(bmin.x >= mx.x || bmin.y >= mx.y || bmin.z >= mx.z || bmax.x <= mn.x || bmax.y <= mn.y ||
                     bmax.z <= mn.z || mn.x >= bmax.x || mn.y >= bmax.y || mn.z >= bmax.z || mx.x <= bmin.x ||
                     mx.y <= bmin.y || mx.z <= bmin.z);
        }
        const F3 c = e.center;
        float dist = r2;
        if (c.x < mn.x) dist -= sq(c.x - mn.x);
        else if (c.x > mx.x) dist -= sq(c.x - mx.x);
        if (c.y < mn.y) dist -= sq(c.y - mn.y);
        else if (c.y > mx.y) dist -= sq(c.y - mx.y);
        if (c.z < mn.z) dist -= sq(c.z - mn.z);
        else if (c.z > mx.z) dist -= sq(c.z - mx.z);
        return dist > 0;
    }
    bool contained_box(F3 mn, F3 mx) const {  // hash_dag_editors.h:233-252
        return bmin.x <= mn.x && bmin.y <= mn.y && bmin.z <= mn.z && bmax.x >= mx.x && bmax.y >= mx.y && bmax.z >= mx.z;
    }
    bool inside_sphere(F3 mn, F3 mx) const {  // hash_dag_editors.h:280-295
        const F3 c = e.center;
        return tmaxf(sq(mn.x - c.x), sq(mx.x - c.x)) + tmaxf(sq(mn.y - c.y), sq(mx.y - c.y)) +
                   tmaxf(sq(mn.z - c.z), sq(mx.z - c.z)) <
               r2;
    }
    bool is_full_impl(F3 mn, F3 mx) const {
        if (e.kind == BOX_ADD) return contained_box(mn, mx);
        if (e.kind == SPHERE_ADD) return inside_sphere(mn, mx);
        return false;
    }
    bool is_empty_impl(F3 mn, F3 mx) const {
        if (e.kind == BOX_CARVE) return contained_box(mn, mx);
        if (e.kind == SPHERE_CARVE) return inside_sphere(mn, mx);
        return false;
    }
    // Editor::should_edit/is_full/is_empty (hash_dag_editors.h:12-27); the decompressing copy
    // also skips nodes above levels - 4 whose source region is empty (COPY_EMPTY_CHECKS,
    // hash_dag_editors.h:561-582)
    bool should_edit(const Path& p, u32 depth) const {
        const F3 s = p.as_position(depth);
        if (!should_edit_impl(s, s + f3((float)(1u << depth)))) return false;
        if (e.kind == COPY_DEC) {
            const u32 levels = srcTable->g.levels, level = levels - depth;
            if (level < levels - 4) {
                const F3 off = s - e.dest + e.source;
                if (off.x >= 0 && off.y >= 0 && off.z >= 0) {
                    const U3 start{to_u32(off.x), to_u32(off.y), to_u32(off.z)};
                    const u32 n = 1u << depth;
                    if (src_empty_rec(0, srcRoot, Path(0, 0, 0), level, start, U3{n, n, n})) return false;
                }
            }
        }
        return true;
    }
    bool is_full(const Path& p, u32 depth) const {
        const F3 s = p.as_position(depth);
        return is_full_impl(s, s + f3((float)(1u << depth)));
    }
    bool is_empty(const Path& p, u32 depth) const {
        const F3 s = p.as_position(depth);
        return is_empty_impl(s, s + f3((float)(1u << depth)));
    }
    F3 single_color() const { return is_box() ? F3{1, 0, 1} : F3{1, 0, 0}; }
    // get_new_value (hash_dag_editors.h:263, 306, 322, 416-424)
    bool new_value(bool old, F3 pos) const {
        if (e.kind == COPY) return old ? old : value_at(*srcTable, srcRoot, source_of(pos));
        if (e.kind == COPY_DEC) {  // hash_dag_editors.h:582-592
            Path src;
            if (old || !dec_voxel(pos, src)) return old;
            return value_at(*srcTable, srcRoot, src);
        }
        if (e.kind == FILL || is_noise()) {  // BufferEditor::get_new_value (hash_dag_editors.h:211-220)
            const F3 fp = pos - fmin;
            const u32 x = to_u32(fp.x), y = to_u32(fp.y), z = to_u32(fp.z);
            if (!fvalid(x, y, z) || !fget(x, y, z)) return old;
            return true;
        }
        return e.kind == SPHERE_PAINT ? old : adding();
    }
    // hash_dag_editors.h:264-269, 307-312, 323-329, 425-434
    template <typename Fn>
    CompressedColor new_color(F3 pos, Fn oldColor, bool oldValue, bool newValue) const {
        CompressedColor c;
        if (e.kind == COPY) {
            if (oldValue == newValue || !newValue) return oldColor();
            return color_at(*srcTable, srcRoot, *srcColors, source_of(pos));
        }
        if (e.kind == COPY_DEC) {  // hash_dag_editors.h:593-611 (colours of the DAG before the edit)
            Path src;
            if (oldValue == newValue || !newValue || !dec_voxel(pos, src)) return oldColor();
            return color_at(*srcTable, srcRoot, *srcColors, src);
        }
        if (is_noise()) return oldValue == newValue ? oldColor() : noise_color(pos);
        if (e.kind == FILL) {  // FillEditorColors::get_new_color (hash_dag_editors.h:737-754)
            if (newValue && !oldValue) {
                const F3 fp = pos - fmin;
                const u32 x = to_u32(fp.x), y = to_u32(fp.y), z = to_u32(fp.z);
                if (fvalid(x, y, z) && values[findex(x, y, z)] == SetTrue) return fillColors[findex(x, y, z)];
            }
            return oldColor();
        }
        if (e.kind == SPHERE_PAINT) {
            const CompressedColor prev = oldColor();
            const float f = tclamp(1 - lengthf(pos - e.center) / e.radius, 0.f, .1f);
            const F3 a = prev.get_color();
            const F3 b{0, 0, 1};
            c.set_single_color(a * (1 - f) + b * f);
        } else {
            c.set_single_color(single_color());
        }
        return c;
    }
};

struct LeafResult { u32 nodeIndex; u32 count; };

struct LeafCtx {
    HashTable& t;
    LeafView leaf;
    ColorLeafBuilder* builder;  // null when colours are disabled
    u64 oldLeavesCount = 0;
    EditStats& stats;
};

// hash_dag_edits.h:335-643
LeafResult edit_leaf(u32 level, u32 nodeIndex, Path path, const Editor& ed, LeafCtx& ctx) {
    HashTable& t = ctx.t;
    const Geometry& g = t.g;
    const u32 levels = g.levels, leafLevel = g.leaf_level();
    ctx.stats.numNodes++;
    auto voxelCount = [&]() -> u32 {
        const u32* p = t.ptr(nodeIndex);
        return level < leafLevel ? (p[0] >> 8) : popc(p[0]) + popc(p[1]);
    };
    const u32 fullCount = 1u << (3 * (levels - level));
    if (!ed.should_edit(path, levels - level)) {
        u32 count = 0;
        if (nodeIndex != C_InvalidNodeIndex) {
            count = voxelCount();
            if (ctx.builder) {
                if (!ctx.leaf.is_valid()) throw std::runtime_error("oracle: copy from missing colour leaf");
                u64 start = ctx.oldLeavesCount;
                if (ctx.leaf.is_shared()) start += ctx.leaf.offset;
                ctx.leaf.copy_colors(*ctx.builder, start, count);
                ctx.oldLeavesCount += count;
            }
        }
        return {nodeIndex, count};
    }
    if (ed.is_empty(path, levels - level)) {
        if (ctx.builder && nodeIndex != C_InvalidNodeIndex) ctx.oldLeavesCount += voxelCount();
        ctx.stats.numVoxels += fullCount;
        return {C_InvalidNodeIndex, 0};
    }
    if (ed.is_full(path, levels - level)) {
        if (ctx.builder) {
            if (nodeIndex != C_InvalidNodeIndex) ctx.oldLeavesCount += voxelCount();
            ctx.builder->add_large_single_color(ed.single_color(), fullCount);
        }
        ctx.stats.numVoxels += fullCount;
        return {t.full_nodes[level], fullCount};
    }
    if (level == leafLevel) {
        u64 mask = 0;
        if (nodeIndex != C_InvalidNodeIndex) {
            const u32* p = t.ptr(nodeIndex);
            mask = (u64)p[0] | ((u64)p[1] << 32);
        }
        const u64 initial = mask;
        for (u8 c1 = 0; c1 < 8; ++c1) {
            Path p1 = path;
            p1.descend(c1);
            for (u8 c2 = 0; c2 < 8; ++c2) {
                Path p2 = p1;
                p2.descend(c2);
                const F3 pos = p2.as_position(0);
                const u64 bit = 1ull << (c1 * 8u + c2);
                const bool shouldEdit = ed.should_edit_impl(pos, pos + 1.0f);
                const bool prev = (mask & bit) != 0;
                const bool nv = shouldEdit ? ed.new_value(prev, pos) : prev;
                ctx.stats.numVoxels += shouldEdit;
                if (ctx.builder) {
                    if (nv) {
                        auto prevColor = [&]() -> CompressedColor {
                            if (prev && ctx.leaf.is_valid() && ctx.leaf.is_valid_index(ctx.oldLeavesCount))
                                return ctx.leaf.get_color(ctx.oldLeavesCount);
                            return CompressedColor{};
                        };
                        const CompressedColor col = shouldEdit ? ed.new_color(pos, prevColor, prev, nv) : prevColor();
                        ctx.builder->add(col);
                    }
                    ctx.oldLeavesCount += prev;
                }
                if (prev != nv) {
                    if (nv) mask |= bit;
                    else mask &= ~bit;
                }
            }
        }
        if (mask == initial) return {nodeIndex, popcll(mask)};
        if (mask == 0) return {C_InvalidNodeIndex, 0};
        return {t.find_or_add_leaf(mask), popcll(mask)};
    }
    const u32* np = nullptr;
    u8 cmask = 0;
    if (nodeIndex != C_InvalidNodeIndex) {
        np = t.ptr(nodeIndex);
        cmask = child_mask(np[0]);
    }
    u32 off = 1;
    u32 newChildren[8];
    u32 newCount = 0;
    bool changed = false;
    for (u8 c = 0; c < 8; ++c) {
        Path cp = path;
        cp.descend(c);
        const u32 childIndex = (cmask & (1u << c)) ? np[off++] : C_InvalidNodeIndex;
        const LeafResult r = edit_leaf(level + 1, childIndex, cp, ed, ctx);
        np = nodeIndex != C_InvalidNodeIndex ? t.ptr(nodeIndex) : nullptr;  // table storage is stable
        if (r.nodeIndex != C_InvalidNodeIndex) newCount += r.count;
        changed |= childIndex != r.nodeIndex;
        newChildren[c] = r.nodeIndex;
    }
    if (newCount >= (1u << 24)) throw std::runtime_error("oracle: voxel count overflows the node header");
    if (!changed) return {nodeIndex, newCount};
    u32 node[9];
    u32 n = 0;
    node[n++] = newCount << 8;
    for (u8 c = 0; c < 8; ++c) {
        if (newChildren[c] != C_InvalidNodeIndex) {
            node[n++] = newChildren[c];
            node[0] |= 1u << c;
        }
    }
    if (child_mask(node[0]) == 0) return {C_InvalidNodeIndex, 0};
    return {t.find_or_add_interior(level, n, node), newCount};
}

struct Task {
    u32 childNodeIndex;
    Path path;
    LeafView existing;
    std::unique_ptr<ColorLeafBuilder> builder;
    EditStats stats;
    u32 finalIndex = 0;
};

struct ColorToBuild {
    std::unique_ptr<ColorLeafBuilder> builder;
    LeafView oldLeaf;
    u32 newLeafIndex;
};

struct EditCtx {
    HashTable& t;
    HashDAGColors* colors;
    const Editor& ed;
    std::map<std::tuple<u32, u32, u32>, std::unique_ptr<Task>> tasks;
    std::vector<Task*> taskOrder;
    std::vector<ColorToBuild> toBuild;
    EditStats stats;
};

struct EditIndex { u32 nodeIndex; u32 colorIndex; };

// hash_dag_edits.h:106-333
EditIndex edit_threaded(bool queue, u32 level, u32 nodeIndex, u32 colorNodeIndex, Path path, EditCtx& ctx) {
    HashTable& t = ctx.t;
    const Geometry& g = t.g;
    const u32 levels = g.levels, ctl = g.color_tree_levels;
    ctx.stats.numNodes++;
    if (!ctx.ed.should_edit(path, levels - level)) return {queue ? C_InvalidNodeIndex : nodeIndex, C_IgnoreColorIndex};
    auto colorNode = [&](u32 idx) -> u32 { return ctx.colors ? ctx.colors->nodes[idx] : 0u; };
    auto allocateColorNode = [&]() {
        colorNodeIndex = ctx.colors->allocate_interior_node();
        for (int i = 0; i < 7; ++i) ctx.colors->allocate_interior_node();
    };
    u8 cmask = 0;
    if (nodeIndex != C_InvalidNodeIndex) cmask = child_mask(*t.ptr(nodeIndex));
    u32 off = 1;
    u32 newChildren[8];
    bool changed = false;
    for (u8 c = 0; c < 8; ++c) {
        Path np = path;
        np.descend(c);
        const u32 childNodeIndex = (cmask & (1u << c)) ? t.ptr(nodeIndex)[off++] : C_InvalidNodeIndex;
        u32 newChild;
        u32 childColor = 0;
        if (level == 0 || colorNodeIndex > 0) childColor = colorNode(colorNodeIndex + c);
        if (level == ctl - 1) {
            if (ctx.ed.should_edit(np, levels - level - 1)) {
                LeafView existing;
                u32 existingIndex = (u32)-1;
                if (childColor > 0 && ctx.colors) {
                    existing = ctx.colors->get_leaf(childColor);
                    existingIndex = childColor;
                }
                const auto key = std::make_tuple(np.p.x, np.p.y, np.p.z);
                if (queue) {
                    auto task = std::make_unique<Task>();
                    task->childNodeIndex = childNodeIndex;
                    task->path = np;
                    task->existing = existing;
                    if (ctx.colors && g.leaf_level() > ctl) task->builder = std::make_unique<ColorLeafBuilder>();
                    ctx.taskOrder.push_back(task.get());
                    ctx.tasks[key] = std::move(task);
                    newChild = C_InvalidNodeIndex;
                } else {
                    Task& task = *ctx.tasks.at(key);
                    newChild = task.finalIndex;
                    if (task.builder && task.builder->colorIndex > 0) {
                        if (existingIndex == (u32)-1 || !(existingIndex & 0x80000000u)) {
                            if (colorNodeIndex == 0) allocateColorNode();
                            childColor = ctx.colors->allocate_leaf();
                            ctx.colors->nodes[colorNodeIndex + c] = childColor;
                            existingIndex = childColor;
                        }
                        ctx.toBuild.push_back({std::move(task.builder), existing, existingIndex});
                    }
                }
            } else {
                newChild = queue ? C_InvalidNodeIndex : childNodeIndex;
            }
        } else {
            const EditIndex r = edit_threaded(queue, level + 1, childNodeIndex, childColor, np, ctx);
            newChild = r.nodeIndex;
            if (!queue && ctx.colors && r.colorIndex != childColor && r.colorIndex != C_IgnoreColorIndex &&
                r.colorIndex != C_RemoveColorIndex) {
                if (level > 0 && colorNodeIndex == 0) allocateColorNode();
                ctx.colors->nodes[colorNodeIndex + c] = r.colorIndex;
            }
        }
        if (!queue) {
            changed |= childNodeIndex != newChild;
            newChildren[c] = newChild;
        }
    }
    if (queue) return {C_InvalidNodeIndex, C_IgnoreColorIndex};
    if (!changed) return {nodeIndex, colorNodeIndex};
    u32 node[9];
    u32 n = 0;
    node[n++] = 0;
    for (u8 c = 0; c < 8; ++c) {
        if (newChildren[c] != C_InvalidNodeIndex) {
            node[n++] = newChildren[c];
            node[0] |= 1u << c;
        }
    }
    if (child_mask(node[0]) == 0) return {C_InvalidNodeIndex, C_IgnoreColorIndex};
    return {t.find_or_add_interior(level, n, node), colorNodeIndex};
}

}  // namespace

CompressedColor get_color_at(const DAG& dag, const HashDAGColors& colors, Path path, bool* ok) {
    const bool present = value_at(dag.t, dag.root, path);
    if (ok) *ok = present;
    return color_at(dag.t, dag.root, colors, path);
}

// hash_dag.h:255-405 (tasks run in creation order; undo frame kept when requested)
EditStats apply_edit(DAG& dag, HashDAGColors* colors, const Edit& e, UndoFrame* undo, int threads) {
    Editor ed(e);
    ed.srcTable = &dag.t;  // the copy reads the DAG as it was before this edit
    ed.srcRoot = dag.root;
    ed.srcColors = colors;
    if ((e.kind == COPY || e.kind == COPY_DEC || e.kind == FILL) && !colors)
        throw std::runtime_error("oracle: copy and fill need colours");
    if (e.kind == FILL) ed.build_fill();
    if (ed.is_noise()) ed.build_noise();
    EditCtx ctx{dag.t, colors, ed, {}, {}, {}, {}};
    if (undo) {
        undo->root = dag.root;
        undo->leaves.clear();
    }
    edit_threaded(true, 0, dag.root, 0, Path(0, 0, 0), ctx);
    // the tasks (hash_dag.h:330-356): in creation order, or taken by `threads` threads from a
    // shared index as the reference's thread pool takes them (NUM_THREADS, typedefs.h:54-56)
    auto run_task = [&](Task* task) {
        LeafCtx lc{dag.t, task->existing, task->builder.get(), 0, task->stats};
        task->finalIndex = edit_leaf(dag.g().color_tree_levels, task->childNodeIndex, task->path, ed, lc).nodeIndex;
    };
    if (threads <= 1 || ctx.taskOrder.size() < 2) {
        for (Task* task : ctx.taskOrder) run_task(task);
    } else {
        std::vector<std::mutex> locks(4096);
        dag.t.locks = &locks;
        std::atomic<size_t> next{0};
        std::exception_ptr failure;
        std::mutex failMu;
        std::vector<std::thread> pool;
        for (int k = 0; k < threads; ++k)
            pool.emplace_back([&]() {
                for (size_t i; (i = next.fetch_add(1)) < ctx.taskOrder.size();) {
                    try {
                        run_task(ctx.taskOrder[i]);
                    } catch (...) {
                        std::lock_guard<std::mutex> g(failMu);
                        failure = std::current_exception();
                        next = ctx.taskOrder.size();
                    }
                }
            });
        for (auto& th : pool) th.join();
        dag.t.locks = nullptr;
        if (failure) std::rethrow_exception(failure);
    }
    for (Task* task : ctx.taskOrder) {
        ctx.stats.numNodes += task->stats.numNodes;
        ctx.stats.numVoxels += task->stats.numVoxels;
    }
    dag.root = edit_threaded(false, 0, dag.root, 0, Path(0, 0, 0), ctx).nodeIndex;
    ctx.stats.leavesRebuilt = ctx.toBuild.size();
    if (e.kind == COPY || e.kind == COPY_DEC) ctx.stats.numVoxels = ed.edit_count();  // "num voxels" (hash_dag.h:374-378)
    for (auto& tb : ctx.toBuild) {
        ColorLeaf& ref = colors->leaves[tb.newLeafIndex & 0x7FFFFFFFu];
        // The undo frame keeps the leaf the subtree used before the edit: the shared main
        // leaf (with its offset) when the slot was freshly allocated, else the slot itself.
        ColorLeaf previous = ref;
        if (tb.oldLeaf.d == &colors->mainLeaf) {
            previous = colors->mainLeaf;
            previous.offset = tb.oldLeaf.offset;
        }
        tb.builder->build(ref);
        if (undo) undo->leaves.push_back({tb.newLeafIndex, {std::move(previous), ref}});
    }
    return ctx.stats;
}

}  // namespace orc
