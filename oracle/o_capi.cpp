// Oracle C API implementation. TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
#include "oracle_capi.h"

#include <algorithm>
#include <atomic>
#include <exception>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "oracle_core.hpp"

using namespace orc;

static thread_local std::string g_err;

#define GUARD_BEGIN try {
#define GUARD_END(fail)          \
    }                            \
    catch (const std::exception& ex) { \
        g_err = ex.what();       \
        return fail;             \
    }

static Geometry to_geom(const orc_geometry* g) {
    Geometry r;
    r.levels = g->levels;
    r.page_size = g->page_size;
    r.num_top_levels = g->num_top_levels;
    r.bucket_bits_top = g->bucket_bits_top;
    r.bucket_bits_low = g->bucket_bits_low;
    r.bucket_size_top = g->bucket_size_top;
    r.bucket_size_low = g->bucket_size_low;
    r.color_tree_levels = g->color_tree_levels;
    return r;
}

static RayParams to_params(const double p[12]) {
    RayParams r;
    r.cameraPosition = {p[0], p[1], p[2]};
    r.rayMin = {p[3], p[4], p[5]};
    r.rayDDx = {p[6], p[7], p[8]};
    r.rayDDy = {p[9], p[10], p[11]};
    return r;
}

template <typename Fn>
static void parallel_rows(uint32_t y0, uint32_t y1, int threads, Fn fn) {
    if (threads <= 1) {
        for (uint32_t y = y0; y < y1; ++y) fn(y);
        return;
    }
    std::vector<std::thread> pool;
    std::atomic<uint32_t> next{y0};
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&]() {
            for (;;) {
                const uint32_t y = next.fetch_add(1);
                if (y >= y1) return;
                fn(y);
            }
        });
    for (auto& th : pool) th.join();
}

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

uint32_t orc_murmurhash32(uint32_t h) { return murmurhash32(h); }
uint64_t orc_murmurhash64(uint64_t h) { return murmurhash64(h); }
uint32_t orc_murmurhash32xN(const uint32_t* w, uint64_t n, uint32_t seed) { return murmurhash32xN(w, n, seed); }

void* orc_dag_from_basic(const orc_geometry* g, uint32_t pool_pages, const uint32_t* data, uint64_t n) {
    GUARD_BEGIN
    auto* dag = new DAG();
    try {
        load_from_basic_dag(*dag, to_geom(g), pool_pages, data, n);
    } catch (...) {
        delete dag;
        throw;
    }
    return dag;
    GUARD_END(nullptr)
}

void* orc_dag_from_image(const orc_geometry* g, uint32_t pool_pages, const uint32_t* pool, const uint32_t* pt,
                         const uint32_t* bs, const uint32_t* full, uint32_t pool_top, uint32_t root) {
    GUARD_BEGIN
    auto* dag = new DAG();
    dag->t.create(to_geom(g), pool_pages);
    std::copy(pool, pool + dag->t.pool.size(), dag->t.pool.begin());
    std::copy(pt, pt + dag->t.page_table.size(), dag->t.page_table.begin());
    std::copy(bs, bs + dag->t.bucket_sizes.size(), dag->t.bucket_sizes.begin());
    std::copy(full, full + dag->t.full_nodes.size(), dag->t.full_nodes.begin());
    dag->t.pool_top = pool_top;
    dag->root = root;
    return dag;
    GUARD_END(nullptr)
}

void orc_dag_free(void* dag) { delete static_cast<DAG*>(dag); }
uint32_t orc_dag_root(void* dag) { return static_cast<DAG*>(dag)->root; }
uint32_t orc_dag_pool_top(void* dag) { return static_cast<DAG*>(dag)->t.pool_top; }

int orc_dag_export(void* d, uint32_t* pool, uint32_t* pt, uint32_t* bs, uint32_t* full) {
    auto* dag = static_cast<DAG*>(d);
    if (pool) std::copy(dag->t.pool.begin(), dag->t.pool.end(), pool);
    if (pt) std::copy(dag->t.page_table.begin(), dag->t.page_table.end(), pt);
    if (bs) std::copy(dag->t.bucket_sizes.begin(), dag->t.bucket_sizes.end(), bs);
    if (full) std::copy(dag->t.full_nodes.begin(), dag->t.full_nodes.end(), full);
    return 0;
}

uint64_t orc_canonical(void* d, uint32_t* out, uint64_t cap) {
    GUARD_BEGIN
    const std::vector<uint32_t> c = canonicalize(*static_cast<DAG*>(d));
    if (out && cap >= c.size()) std::copy(c.begin(), c.end(), out);
    return c.size();
    GUARD_END(0)
}

int orc_get_value(void* d, uint32_t x, uint32_t y, uint32_t z) { return get_value(*static_cast<DAG*>(d), Path(x, y, z)); }

// Dense occupancy by recursive expansion (levels <= 9 in practice).
static void dense_rec(const DAG& dag, u32 level, u32 va, u32 x, u32 y, u32 z, uint8_t* out) {
    const Geometry& g = dag.g();
    const u32 E = 1u << g.levels;
    if (level == g.leaf_level()) {
        const u64 leaf = dag.leaf(va);
        for (u32 b = 0; b < 64; ++b) {
            if (!((leaf >> b) & 1)) continue;
            const u32 c1 = b >> 3, c2 = b & 7;
            const u32 vx = x * 4 + ((c1 >> 2) & 1) * 2 + ((c2 >> 2) & 1);
            const u32 vy = y * 4 + ((c1 >> 1) & 1) * 2 + ((c2 >> 1) & 1);
            const u32 vz = z * 4 + (c1 & 1) * 2 + (c2 & 1);
            out[((u64)vx * E + vy) * E + vz] = 1;
        }
        return;
    }
    const u32 hdr = dag.node(va);
    const u8 m = child_mask(hdr);
    for (u8 c = 0; c < 8; ++c)
        if (m & (1u << c))
            dense_rec(dag, level + 1, dag.child_index(va, m, c), x * 2 + ((c >> 2) & 1), y * 2 + ((c >> 1) & 1),
                      z * 2 + (c & 1), out);
}

int orc_dense(void* d, uint8_t* out) {
    GUARD_BEGIN
    const DAG& dag = *static_cast<DAG*>(d);
    const u64 E = 1ull << dag.g().levels;
    std::fill(out, out + E * E * E, 0);
    dense_rec(dag, 0, dag.root, 0, 0, 0, out);
    return 0;
    GUARD_END(-1)
}

void* orc_colors_new(const uint32_t* nodes, uint64_t n_nodes, const uint64_t* offsets, uint64_t n_offsets,
                     const uint64_t* blocks, uint64_t n_blocks, const uint32_t* weights, uint64_t n_weights,
                     const uint64_t* macro, uint64_t n_macro) {
    auto* c = new HashDAGColors();
    c->nodes.assign(nodes, nodes + n_nodes);
    c->offsets.assign(offsets, offsets + n_offsets);
    c->mainLeaf.offset = ColorLeaf::kUnique;
    c->mainLeaf.blocks.assign(blocks, blocks + n_blocks);
    c->mainLeaf.weights.assign(weights, weights + n_weights);
    c->mainLeaf.macro.assign(macro, macro + n_macro);
    c->mainLeaf.valid = n_blocks > 0;
    return c;
}

uint32_t orc_colors_add_unique(void* cp, const uint64_t* blocks, uint64_t n_blocks, const uint32_t* weights,
                               uint64_t n_weights, const uint64_t* macro, uint64_t n_macro) {
    auto* c = static_cast<HashDAGColors*>(cp);
    const uint32_t h = c->allocate_leaf();
    ColorLeaf& l = c->leaves[h & 0x7FFFFFFFu];
    l.blocks.assign(blocks, blocks + n_blocks);
    l.weights.assign(weights, weights + n_weights);
    l.macro.assign(macro, macro + n_macro);
    l.valid = n_blocks > 0;
    return h;
}

uint32_t orc_colors_add_shared(void* cp, uint64_t offset) {
    auto* c = static_cast<HashDAGColors*>(cp);
    const uint32_t h = c->allocate_leaf();
    ColorLeaf& l = c->leaves[h & 0x7FFFFFFFu];
    l = c->mainLeaf;
    l.offset = offset;
    return h;
}

void orc_colors_free(void* c) { delete static_cast<HashDAGColors*>(c); }
uint64_t orc_colors_num_nodes(void* c) { return static_cast<HashDAGColors*>(c)->nodes.size(); }
uint64_t orc_colors_num_leaves(void* c) { return static_cast<HashDAGColors*>(c)->leaves.size(); }
int orc_colors_nodes(void* c, uint32_t* out) {
    auto& n = static_cast<HashDAGColors*>(c)->nodes;
    std::copy(n.begin(), n.end(), out);
    return 0;
}
int orc_colors_leaf_sizes(void* c, uint32_t i, uint64_t sizes[3]) {
    auto& l = static_cast<HashDAGColors*>(c)->leaves.at(i);
    sizes[0] = l.blocks.size();
    sizes[1] = l.weights.size();
    sizes[2] = l.macro.size();
    return l.offset == ColorLeaf::kUnique ? 0 : 1;
}
int orc_colors_leaf_copy(void* c, uint32_t i, uint64_t* blocks, uint32_t* weights, uint64_t* macro) {
    auto& l = static_cast<HashDAGColors*>(c)->leaves.at(i);
    std::copy(l.blocks.begin(), l.blocks.end(), blocks);
    std::copy(l.weights.begin(), l.weights.end(), weights);
    std::copy(l.macro.begin(), l.macro.end(), macro);
    return 0;
}

// Walk the colour tree down to the colour-leaf level in DFS order and decode every subtree.
static void color_canon_rec(const DAG& dag, const HashDAGColors& colors, u32 level, u32 va, u32 colorNode, Path path,
                            std::vector<u32>& out) {
    const Geometry& g = dag.g();
    const u32 ctl = g.color_tree_levels;
    const u32 hdr = dag.node(va);
    const u8 m = child_mask(hdr);
    for (u8 c = 0; c < 8; ++c) {
        if (!(m & (1u << c))) continue;
        Path cp = path;
        cp.descend(c);
        const u32 child = dag.child_index(va, m, c);
        const u32 cn = colors.nodes[colorNode + c];
        if (level + 1 == ctl) {
            const LeafView leaf = colors.get_leaf(cn);
            const u32 count = dag.node(child) >> 8;
            out.push_back(cp.p.x);
            out.push_back(cp.p.y);
            out.push_back(cp.p.z);
            out.push_back(count);
            for (u32 i = 0; i < count; ++i) {
                if (!leaf.is_valid() || !leaf.is_valid_index(i)) {
                    out.push_back(0xFFFFFFFFu);
                    out.push_back(0xFFFFFFFFu);
                    continue;
                }
                const CompressedColor col = leaf.get_color(i);
                out.push_back(col.colorBits);
                out.push_back((u32)col.weight | ((u32)col.bitsPerWeight << 8));
            }
        } else {
            if (cn == 0 && level > 0) continue;  // missing colour node
            color_canon_rec(dag, colors, level + 1, child, cn, cp, out);
        }
    }
}

// The raw words of every level-(color_tree_levels) subtree's colour leaf in DFS order:
// [x, y, z, kind] then, for a unique leaf (kind 1), n_blocks, n_weights, n_macro and the
// blocks / weights / macro words as stored; for a shared leaf (kind 2) its offset (lo, hi);
// kind 0 = no leaf. Compares the encoder's byte choices, not only the decoded colours.
static void color_bytes_rec(const DAG& dag, const HashDAGColors& colors, u32 level, u32 va, u32 colorNode, Path path,
                            std::vector<u32>& out) {
    const u32 ctl = dag.g().color_tree_levels;
    const u32 hdr = dag.node(va);
    const u8 m = child_mask(hdr);
    for (u8 c = 0; c < 8; ++c) {
        if (!(m & (1u << c))) continue;
        Path cp = path;
        cp.descend(c);
        const u32 child = dag.child_index(va, m, c);
        const u32 cn = colors.nodes[colorNode + c];
        if (level + 1 == ctl) {
            out.push_back(cp.p.x);
            out.push_back(cp.p.y);
            out.push_back(cp.p.z);
            const LeafView leaf = colors.get_leaf(cn);
            if (!leaf.d) {
                out.push_back(0);
            } else if (leaf.is_shared()) {
                out.push_back(2);
                out.push_back((u32)leaf.offset);
                out.push_back((u32)(leaf.offset >> 32));
            } else {
                const ColorLeaf& l = *leaf.d;
                out.push_back(1);
                out.push_back((u32)l.blocks.size());
                out.push_back((u32)l.weights.size());
                out.push_back((u32)l.macro.size());
                for (u64 b : l.blocks) {
                    out.push_back((u32)b);
                    out.push_back((u32)(b >> 32));
                }
                out.insert(out.end(), l.weights.begin(), l.weights.end());
                for (u64 b : l.macro) {
                    out.push_back((u32)b);
                    out.push_back((u32)(b >> 32));
                }
            }
        } else {
            if (cn == 0 && level > 0) continue;
            color_bytes_rec(dag, colors, level + 1, child, cn, cp, out);
        }
    }
}

uint64_t orc_color_leaf_bytes(void* d, void* c, uint32_t* out, uint64_t cap) {
    GUARD_BEGIN
    std::vector<u32> v;
    color_bytes_rec(*static_cast<DAG*>(d), *static_cast<HashDAGColors*>(c), 0, static_cast<DAG*>(d)->root, 0,
                    Path(0, 0, 0), v);
    if (out && cap >= v.size()) std::copy(v.begin(), v.end(), out);
    return v.size();
    GUARD_END(0)
}

uint64_t orc_color_canonical(void* d, void* c, uint32_t* out, uint64_t cap) {
    GUARD_BEGIN
    std::vector<u32> v;
    color_canon_rec(*static_cast<DAG*>(d), *static_cast<HashDAGColors*>(c), 0, static_cast<DAG*>(d)->root, 0,
                    Path(0, 0, 0), v);
    if (out && cap >= v.size()) std::copy(v.begin(), v.end(), out);
    return v.size();
    GUARD_END(0)
}

int orc_encode_colors(const uint32_t* colorBits, const uint8_t* weight, const uint8_t* bpw, uint64_t n, uint64_t* blocks,
                      uint32_t* weights, uint64_t* macro, uint64_t sizes[3], uint64_t capB, uint64_t capW,
                      uint64_t capM) {
    ColorLeafBuilder b;
    for (uint64_t i = 0; i < n; ++i) {
        CompressedColor c;
        c.colorBits = colorBits[i];
        c.weight = weight[i];
        c.bitsPerWeight = bpw[i];
        b.add(c);
    }
    ColorLeaf leaf;
    b.build(leaf);
    sizes[0] = leaf.blocks.size();
    sizes[1] = leaf.weights.size();
    sizes[2] = leaf.macro.size();
    if (sizes[0] > capB || sizes[1] > capW || sizes[2] > capM) return 1;
    std::copy(leaf.blocks.begin(), leaf.blocks.end(), blocks);
    std::copy(leaf.weights.begin(), leaf.weights.end(), weights);
    std::copy(leaf.macro.begin(), leaf.macro.end(), macro);
    return 0;
}

int orc_decode_color(const uint64_t* blocks, uint64_t nb, const uint32_t* weights, uint64_t nw, const uint64_t* macro,
                     uint64_t nm, uint64_t offset, uint64_t index, uint32_t out[2]) {
    ColorLeaf l;
    l.blocks.assign(blocks, blocks + nb);
    l.weights.assign(weights, weights + nw);
    l.macro.assign(macro, macro + nm);
    l.valid = nb > 0;
    LeafView v{&l, offset};
    if (!v.is_valid() || !v.is_valid_index(index)) return 1;
    const CompressedColor c = v.get_color(index);
    out[0] = c.colorBits;
    out[1] = (u32)c.weight | ((u32)c.bitsPerWeight << 8);
    return 0;
}

void orc_ray_params(const double position[3], const double rotation[9], const double bmin[3], const double bmax[3],
                    uint32_t levels, uint32_t W, uint32_t H, double out[12]) {
    const RayParams p = get_trace_params(position, rotation, bmin, bmax, levels, W, H);
    const double v[12] = {p.cameraPosition.x, p.cameraPosition.y, p.cameraPosition.z, p.rayMin.x, p.rayMin.y,
                          p.rayMin.z, p.rayDDx.x, p.rayDDx.y, p.rayDDx.z, p.rayDDy.x, p.rayDDy.y, p.rayDDy.z};
    std::copy(v, v + 12, out);
}

int orc_trace_paths(void* d, const double params[12], uint32_t W, uint32_t H, uint32_t y0, uint32_t y1, uint32_t* paths,
                    uint64_t* bytes, int threads) {
    const DAG& dag = *static_cast<DAG*>(d);
    const RayParams p = to_params(params);
    std::atomic<uint64_t> total{0};
    // row r of the image holds the ray of pixel y = H-1-r (tracer.cu:251)
    parallel_rows(y0, y1, threads, [&](uint32_t py) {
        uint64_t b = 0;
        const uint32_t row = H - 1 - py;
        for (uint32_t px = 0; px < W; ++px) {
            const U3 r = trace_path(dag, p, px, py, bytes ? &b : nullptr);
            uint32_t* o = paths + ((uint64_t)row * W + px) * 4;
            o[0] = r.x;
            o[1] = r.y;
            o[2] = r.z;
            o[3] = 0;
        }
        total += b;
    });
    if (bytes) *bytes = total.load();
    return 0;
}

int orc_trace_colors(void* d, void* c, uint32_t W, uint32_t H, uint32_t y0, uint32_t y1, const uint32_t* paths,
                     uint32_t* out, int threads) {
    (void)H;
    const DAG& dag = *static_cast<DAG*>(d);
    const HashDAGColors& colors = *static_cast<HashDAGColors*>(c);
    parallel_rows(y0, y1, threads, [&](uint32_t y) {
        for (uint32_t x = 0; x < W; ++x) {
            const uint32_t* p = paths + ((uint64_t)y * W + x) * 4;
            out[(uint64_t)y * W + x] = trace_color(dag, colors, U3{p[0], p[1], p[2]});
        }
    });
    return 0;
}

void orc_set_any_hit_order(int front_to_back) { set_any_hit_front_to_back(front_to_back != 0); }

int orc_trace_shadows(void* d, const double params[12], uint32_t W, uint32_t H, uint32_t y0, uint32_t y1,
                      const uint32_t* paths, uint32_t* colors, float bias, float fog, uint64_t* bytes, int threads) {
    const DAG& dag = *static_cast<DAG*>(d);
    const RayParams p = to_params(params);
    std::atomic<uint64_t> total{0};
    parallel_rows(y0, y1, threads, [&](uint32_t y) {
        uint64_t b = 0;
        for (uint32_t x = 0; x < W; ++x) {
            const uint32_t* pp = paths + ((uint64_t)y * W + x) * 4;
            uint32_t& c = colors[(uint64_t)y * W + x];
            c = shade_shadow(dag, p, W, H, x, y, U3{pp[0], pp[1], pp[2]}, c, bias, fog, bytes ? &b : nullptr);
            b += 24;  // path read (16 B) + colour read-modify-write (8 B)
        }
        total += b;
    });
    if (bytes) *bytes = total.load();
    return 0;
}

int orc_edit(void* d, void* c, uint32_t kind, const float center[3], float radius, uint64_t stats[3]) {
    return orc_edit_threads(d, c, kind, center, radius, 1, stats);
}

int orc_edit_threads(void* d, void* c, uint32_t kind, const float center[3], float radius, int threads,
                     uint64_t stats[3]) {
    GUARD_BEGIN
    Edit e{};
    e.kind = kind;
    e.center = {center[0], center[1], center[2]};
    e.radius = radius;
    const EditStats s = apply_edit(*static_cast<DAG*>(d), static_cast<HashDAGColors*>(c), e, nullptr, threads);
    if (stats) {
        stats[0] = s.numNodes;
        stats[1] = s.numVoxels;
        stats[2] = s.leavesRebuilt;
    }
    return 0;
    GUARD_END(-1)
}

int orc_edit_copy(void* d, void* c, const float center[3], float radius, const float source[3], const float dest[3],
                  uint64_t stats[3]) {
    return orc_edit_copy_kind(d, c, COPY, center, radius, source, dest, stats);
}

int orc_edit_copy_kind(void* d, void* c, uint32_t kind, const float center[3], float radius, const float source[3],
                       const float dest[3], uint64_t stats[3]) {
    GUARD_BEGIN
    if (kind != COPY && kind != COPY_DEC) throw std::runtime_error("oracle: not a copy editor kind");
    Edit e{};
    e.kind = kind;
    e.center = {center[0], center[1], center[2]};
    e.radius = radius;
    e.source = {source[0], source[1], source[2]};
    e.dest = {dest[0], dest[1], dest[2]};
    const EditStats s = apply_edit(*static_cast<DAG*>(d), static_cast<HashDAGColors*>(c), e, nullptr);
    if (stats) {
        stats[0] = s.numNodes;
        stats[1] = s.numVoxels;
        stats[2] = s.leavesRebuilt;
    }
    return 0;
    GUARD_END(-1)
}

int orc_edit_copy_tf(void* d, void* c, const float center[3], float radius, const float source[3],
                     const float dest[3], const double transform[9], int swirl, float swirl_period, uint64_t stats[3]) {
    GUARD_BEGIN
    Edit e{};
    e.kind = COPY;
    e.center = {center[0], center[1], center[2]};
    e.radius = radius;
    e.source = {source[0], source[1], source[2]};
    e.dest = {dest[0], dest[1], dest[2]};
    for (int i = 0; i < 9; ++i) e.transform[i] = transform[i];
    e.swirl = swirl != 0;
    e.swirlPeriod = swirl_period;
    const EditStats s = apply_edit(*static_cast<DAG*>(d), static_cast<HashDAGColors*>(c), e, nullptr);
    if (stats) {
        stats[0] = s.numNodes;
        stats[1] = s.numVoxels;
        stats[2] = s.leavesRebuilt;
    }
    return 0;
    GUARD_END(-1)
}

int orc_get_color(void* d, void* c, uint32_t x, uint32_t y, uint32_t z, uint32_t out[2]) {
    GUARD_BEGIN
    bool ok = false;
    const CompressedColor col = get_color_at(*static_cast<DAG*>(d), *static_cast<HashDAGColors*>(c), Path(x, y, z), &ok);
    out[0] = col.colorBits;
    out[1] = col.weight | (col.bitsPerWeight << 8);
    return ok ? 1 : 0;
    GUARD_END(-1)
}

int orc_gc(void* d, uint32_t max_level) {
    GUARD_BEGIN
    remove_stale_nodes(*static_cast<DAG*>(d), max_level);
    return 0;
    GUARD_END(-1)
}

// distinct node addresses reachable from the root per level 0..max_level (the valid-node
// sets of hash_dag.h:423-437)
int orc_dag_reachable(void* d, uint32_t max_level, uint64_t* out) {
    GUARD_BEGIN
    DAG& dag = *static_cast<DAG*>(d);
    std::vector<std::set<uint32_t>> valid(max_level + 1);
    valid[0].insert(dag.root);
    for (uint32_t level = 0; level < max_level; ++level)
        for (uint32_t idx : valid[level]) {
            const uint32_t* p = dag.t.ptr(idx);
            for (uint32_t o = 1; o < total_size(p[0]); ++o) valid[level + 1].insert(p[o]);
        }
    for (uint32_t l = 0; l <= max_level; ++l) out[l] = valid[l].size();
    return 0;
    GUARD_END(-1)
}

// HashTable::find_interior_node_in_bucket / find_leaf_node_in_bucket (hash_table.h:196-357) on
// one bucket of the table as it stands (leaf level: words = the two leaf words)
uint32_t orc_find_in_bucket(void* d, uint32_t level, uint32_t bucket, const uint32_t* words, uint32_t n) {
    DAG& dag = *static_cast<DAG*>(d);
    const uint32_t bs = dag.t.bucket_sizes[dag.g().bucket_global(level, bucket)];
    if (level == dag.g().leaf_level()) return dag.t.find_leaf_in_bucket(bs, bucket, (u64)words[0] | ((u64)words[1] << 32));
    return dag.t.find_interior_in_bucket(bs, level, bucket, n, words);
}

}  // extern "C"

int orc_perlin(int seed, float frequency, int octaves, uint32_t n, const float* xyz, float* out, uint8_t* perm) {
    PerlinFbm noise;
    noise.set_seed(seed);
    noise.frequency = frequency;
    noise.set_octaves(octaves);
    for (uint32_t i = 0; i < n; ++i) out[i] = noise.fractal(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    if (perm) {
        std::memcpy(perm, noise.perm, 512);
        std::memcpy(perm + 512, noise.perm12, 512);
    }
    return 0;
}
