// Files: the reference's BasicDAG and HashDAG dump formats, and a checkpoint of a context.
//
//   BasicDAG    BasicDAGFactory::save_dag_to_file / load_dag_from_file (basic_dag.cpp:89-117) with
//               FileWriter / FileReader (serializer.h:11-122): DAGInfo as six doubles, the level
//               count (uint32), the node words as a uint64 count and the words.
//   HashDAG     HashDAGFactory::save_dag_to_file / load_dag_from_file (hash_dag_factory.cpp:200-246):
//               DAGInfo, levels, firstNodeIndex, poolMaxSize, pageTableSize, poolTop (uint32 each),
//               then poolTop pages of the physical pool and the page table, without count prefixes.
//               The dump holds no bucket sizes and no colours, so it is loaded by walking the DAG
//               from its root into BasicDAG words and building the table from those.
//   colours     BasicDAGFactory::save/load_compressed_colors_to/from_file (basic_dag.cpp:34-89):
//               topLevels (uint32), enclosedLeaves, weights, blocks, macro blocks (uint64 count +
//               data each); the weight words are stored byte-swapped against their in-memory order
//               (ColorUtils::swap_byte_order, color_utils.h:85-105). Such a file pairs with a
//               BasicDAG whose headers count voxels at every level, through enclosedLeaves above
//               topLevels (BasicDAGColorsBase::get_leaves_count, basic_dag.h:60-72); the colour
//               tree over it is built as create_hash_dag_colors builds it
//               (hash_dag_factory.cpp:101-150).
//   checkpoint  this project's own format: the whole table image (with bucket sizes and full
//               nodes) and the colour state (tree, shared offsets, unique leaves, arena), so a
//               context can be saved and restored exactly; the undo history is not saved.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "hd_ctx.h"

namespace hd {
namespace {

struct File {
    FILE* f = nullptr;
    bool ok = true;
    File(const char* path, const char* mode) : f(std::fopen(path, mode)) { ok = f != nullptr; }
    ~File() {
        if (f) std::fclose(f);
    }
    void put(const void* p, uint64_t n) { ok = ok && (n == 0 || std::fwrite(p, 1, n, f) == n); }
    void get(void* p, uint64_t n) { ok = ok && (n == 0 || std::fread(p, 1, n, f) == n); }
    template <typename T>
    void put(T v) { put(&v, sizeof(T)); }
    template <typename T>
    T get() {
        T v{};
        get(&v, sizeof(T));
        return v;
    }
    template <typename T>
    void put_array(const T* p, uint64_t n) {  // FileWriter::write(StaticArray): uint64 size, data
        put<uint64_t>(n);
        put(p, n * sizeof(T));
    }
    template <typename T>
    bool get_array(std::vector<T>& v, uint64_t limit) {
        const uint64_t n = get<uint64_t>();
        if (!ok || n > limit) return ok = false;
        v.resize(n);
        get(v.data(), n * sizeof(T));
        return ok;
    }
    bool at_end() {
        const int c = std::fgetc(f);
        return c == EOF;
    }
};

void put_info(File& w, const double bmin[3], const double bmax[3]) {  // FileWriter::write(DAGInfo)
    for (int i = 0; i < 3; ++i) w.put<double>(bmin[i]);
    for (int i = 0; i < 3; ++i) w.put<double>(bmax[i]);
}

// The DAG reachable from `root` of a paged pool, as BasicDAG words (root at 0, every node once,
// children after their parent; interior nodes keep their header word, leaves are 2 words).
struct BasicFromPool {
    const std::vector<uint32_t>& pool;
    const std::vector<uint32_t>& pt;
    uint32_t pageSize, leafLevel;
    std::vector<uint32_t> out;
    std::unordered_map<uint64_t, uint32_t> memo;  // (level, address) -> BasicDAG offset

    uint32_t word(uint32_t va) const {
        const uint64_t page = va / pageSize;
        if (page >= pt.size() || pt[page] == 0) throw std::runtime_error("address on an unmapped page");
        const uint64_t at = (uint64_t)pt[page] * pageSize + va % pageSize;
        if (at >= pool.size()) throw std::runtime_error("page outside the pool");
        return pool[at];
    }
    uint32_t node(uint32_t level, uint32_t va) {
        const uint64_t key = ((uint64_t)level << 32) | va;
        auto it = memo.find(key);
        if (it != memo.end()) return it->second;
        const uint32_t at = (uint32_t)out.size();
        if (level == leafLevel) {
            out.push_back(word(va));
            out.push_back(word(va + 1));
        } else {
            const uint32_t hdr = word(va);
            const uint32_t n = (uint32_t)__builtin_popcount(hdr & 0xFFu);
            if (n == 0) throw std::runtime_error("interior node without children");
            out.push_back(hdr);
            out.resize(out.size() + n);
            for (uint32_t k = 0; k < n; ++k) {
                const uint32_t child = node(level + 1, word(va + 1 + k));
                out[at + 1 + k] = child;
            }
        }
        memo.emplace(key, at);
        if (out.size() >= (1ull << 32)) throw std::runtime_error("DAG too large for BasicDAG words");
        return at;
    }
};

// Voxel counts of a BasicDAG's nodes (children after their parent, leaves 2 words).
struct DagCounts {
    const uint32_t* words;
    uint64_t n;
    uint32_t leafLevel;
    std::unordered_map<uint32_t, uint64_t> memo;  // word index -> voxels
    uint32_t at(uint64_t i) const {
        if (i >= n) throw std::runtime_error("BasicDAG index out of range");
        return words[i];
    }
    uint64_t count(uint32_t level, uint32_t idx) {
        auto it = memo.find(idx);
        if (it != memo.end()) return it->second;
        uint64_t c = 0;
        if (level == leafLevel) {
            c = (uint64_t)__builtin_popcount(at(idx)) + (uint64_t)__builtin_popcount(at((uint64_t)idx + 1));
        } else {
            const uint32_t nk = (uint32_t)__builtin_popcount(at(idx) & 0xFFu);
            for (uint32_t k = 0; k < nk; ++k) c += count(level + 1, at((uint64_t)idx + 1 + k));
        }
        memo.emplace(idx, c);
        return c;
    }
};

constexpr char kCkptMagic[8] = {'H', 'D', 'A', 'G', 'C', 'K', 'P', 'T'};
constexpr uint32_t kCkptVersion = 1;

}  // namespace
}  // namespace hd

using namespace hd;

extern "C" {

hd_status hd_write_basic_dag_file(const char* path, const double bounds_min[3], const double bounds_max[3],
                                  uint32_t levels, const uint32_t* words, uint64_t n_words) {
    if (!path || !bounds_min || !bounds_max || (!words && n_words)) return fail(nullptr, HD_ERR_INVALID, "bad arguments");
    File w(path, "wb");
    if (!w.ok) return fail(nullptr, HD_ERR_INVALID, std::string("cannot open ") + path);
    put_info(w, bounds_min, bounds_max);
    w.put<uint32_t>(levels);
    w.put_array(words, n_words);
    if (!w.ok) return fail(nullptr, HD_ERR_INVALID, std::string("write failed: ") + path);
    return HD_OK;
}

hd_status hd_read_basic_dag_file(const char* path, double bounds[6], uint32_t* levels, uint32_t* words,
                                 uint64_t capacity, uint64_t* n_words) {
    if (!path || !n_words) return fail(nullptr, HD_ERR_INVALID, "bad arguments");
    File r(path, "rb");
    if (!r.ok) return fail(nullptr, HD_ERR_INVALID, std::string("cannot open ") + path);
    double b[6];
    for (double& v : b) v = r.get<double>();
    const uint32_t lv = r.get<uint32_t>();
    const uint64_t n = r.get<uint64_t>();
    if (!r.ok) return fail(nullptr, HD_ERR_INVALID, "truncated BasicDAG file");
    if (bounds) std::memcpy(bounds, b, sizeof(b));
    if (levels) *levels = lv;
    *n_words = n;
    if (!words) return HD_OK;
    if (capacity < n) return fail(nullptr, HD_ERR_INVALID, "word buffer too small");
    r.get(words, n * 4);
    if (!r.ok) return fail(nullptr, HD_ERR_INVALID, "truncated BasicDAG file");
    return HD_OK;
}

hd_status hd_save_hash_dag_file(hd_ctx* ctx, const char* path, const double bounds_min[3], const double bounds_max[3]) {
    if (!ctx || !path || !bounds_min || !bounds_max) return fail(ctx, HD_ERR_INVALID, "bad arguments");
    if (!ctx->loaded) return fail(ctx, HD_ERR_INVALID, "no DAG loaded");
    const Geom& g = ctx->g;
    hd_image_info info;
    hd_status st = hd_image_info_get(ctx, &info);
    if (st != HD_OK) return st;
    std::vector<uint32_t> pool((uint64_t)info.pool_top * g.page_size), pt(g.total_pages());
    st = download_pool(ctx, pool.data(), pool.size());
    if (st == HD_OK) st = hd_download_image(ctx, nullptr, pt.data(), nullptr, nullptr);
    if (st != HD_OK) return st;
    File w(path, "wb");
    if (!w.ok) return fail(ctx, HD_ERR_INVALID, std::string("cannot open ") + path);
    put_info(w, bounds_min, bounds_max);
    w.put<uint32_t>(g.levels);
    w.put<uint32_t>(ctx->root);
    w.put<uint32_t>(info.pool_pages);
    w.put<uint32_t>((uint32_t)pt.size());
    w.put<uint32_t>(info.pool_top);
    w.put(pool.data(), pool.size() * 4);
    w.put(pt.data(), pt.size() * 4);
    if (!w.ok) return fail(ctx, HD_ERR_INVALID, std::string("write failed: ") + path);
    return HD_OK;
}

hd_status hd_load_hash_dag_file(hd_ctx* ctx, const char* path, double bounds[6]) {
    if (!ctx || !path) return fail(ctx, HD_ERR_INVALID, "bad arguments");
    const Geom& g = ctx->g;
    std::vector<uint32_t> basic;
    {
        File r(path, "rb");
        if (!r.ok) return fail(ctx, HD_ERR_INVALID, std::string("cannot open ") + path);
        double b[6];
        for (double& v : b) v = r.get<double>();
        const uint32_t levels = r.get<uint32_t>(), root = r.get<uint32_t>();
        (void)r.get<uint32_t>();  // poolMaxSize
        const uint32_t ptSize = r.get<uint32_t>(), poolTop = r.get<uint32_t>();
        if (!r.ok) return fail(ctx, HD_ERR_INVALID, "truncated HashDAG file");
        if (levels != g.levels) return fail(ctx, HD_ERR_INVALID, "HashDAG file has another level count");
        if (ptSize == 0 || poolTop == 0 || ptSize > (1ull << 32) / g.page_size || poolTop > ptSize + 1)
            return fail(ctx, HD_ERR_INVALID, "HashDAG file sizes out of range");
        std::vector<uint32_t> pool((uint64_t)poolTop * g.page_size), pt(ptSize);
        r.get(pool.data(), pool.size() * 4);
        r.get(pt.data(), pt.size() * 4);
        if (!r.ok) return fail(ctx, HD_ERR_INVALID, "truncated HashDAG file");
        if (bounds) std::memcpy(bounds, b, sizeof(b));
        try {
            BasicFromPool w{pool, pt, g.page_size, g.leaf_level(), {}, {}};
            w.node(0, root);
            basic.swap(w.out);
        } catch (const std::exception& ex) {
            return fail(ctx, HD_ERR_INVALID, std::string("HashDAG file: ") + ex.what());
        }
    }
    return hd_load_basic_dag(ctx, basic.data(), basic.size());
}

hd_status hd_checkpoint_save(hd_ctx* ctx, const char* path) {
    if (!ctx || !path) return fail(ctx, HD_ERR_INVALID, "bad arguments");
    if (!ctx->loaded) return fail(ctx, HD_ERR_INVALID, "no DAG loaded");
    const Geom& g = ctx->g;
    hd_image_info info;
    hd_status st = hd_image_info_get(ctx, &info);
    if (st != HD_OK) return st;
    std::vector<uint32_t> pool((uint64_t)info.pool_top * g.page_size), pt(g.total_pages()), bs(g.total_buckets()),
        full(g.levels);
    st = download_pool(ctx, pool.data(), pool.size());
    if (st == HD_OK) st = hd_download_image(ctx, nullptr, pt.data(), bs.data(), full.data());
    if (st != HD_OK) return st;
    File w(path, "wb");
    if (!w.ok) return fail(ctx, HD_ERR_INVALID, std::string("cannot open ") + path);
    w.put(kCkptMagic, 8);
    w.put<uint32_t>(kCkptVersion);
    const uint32_t geom[8] = {g.levels,          g.page_size,       g.num_top_levels,  g.bucket_bits_top,
                              g.bucket_bits_low, g.bucket_size_top, g.bucket_size_low, g.color_tree_levels};
    w.put(geom, sizeof(geom));
    w.put<uint32_t>(info.pool_top);
    w.put<uint32_t>(ctx->root);
    w.put_array(pool.data(), pool.size());
    w.put_array(pt.data(), pt.size());
    w.put_array(bs.data(), bs.size());
    w.put_array(full.data(), full.size());
    w.put<uint32_t>(ctx->colors_loaded ? 1u : 0u);
    if (ctx->colors_loaded) {
        std::vector<uint32_t> nodes(ctx->n_cnodes);
        std::vector<uint64_t> offsets(ctx->n_coffsets), arena(ctx->arena_top);
        std::vector<LeafDesc> leaves(ctx->n_cleaves);
        HD_HIP(ctx, hipMemcpy(nodes.data(), ctx->cnodes, nodes.size() * 4, hipMemcpyDeviceToHost));
        if (!offsets.empty())
            HD_HIP(ctx, hipMemcpy(offsets.data(), ctx->coffsets, offsets.size() * 8, hipMemcpyDeviceToHost));
        if (!leaves.empty())
            HD_HIP(ctx, hipMemcpy(leaves.data(), ctx->cleaves, leaves.size() * sizeof(LeafDesc), hipMemcpyDeviceToHost));
        if (!arena.empty()) HD_HIP(ctx, hipMemcpy(arena.data(), ctx->arena, arena.size() * 8, hipMemcpyDeviceToHost));
        w.put_array(nodes.data(), nodes.size());
        w.put_array(offsets.data(), offsets.size());
        w.put_array(leaves.data(), leaves.size());
        w.put_array(arena.data(), arena.size());
        w.put(&ctx->main, sizeof(LeafDesc));
    }
    w.put(kCkptMagic, 8);
    if (!w.ok) return fail(ctx, HD_ERR_INVALID, std::string("write failed: ") + path);
    return HD_OK;
}

hd_status hd_checkpoint_load(hd_ctx* ctx, const char* path) {
    if (!ctx || !path) return fail(ctx, HD_ERR_INVALID, "bad arguments");
    const Geom& g = ctx->g;
    File r(path, "rb");
    if (!r.ok) return fail(ctx, HD_ERR_INVALID, std::string("cannot open ") + path);
    char magic[8];
    r.get(magic, 8);
    if (!r.ok || std::memcmp(magic, kCkptMagic, 8) != 0 || r.get<uint32_t>() != kCkptVersion)
        return fail(ctx, HD_ERR_INVALID, "not a checkpoint of this version");
    uint32_t geom[8];
    r.get(geom, sizeof(geom));
    const uint32_t want[8] = {g.levels,          g.page_size,       g.num_top_levels,  g.bucket_bits_top,
                              g.bucket_bits_low, g.bucket_size_top, g.bucket_size_low, g.color_tree_levels};
    if (!r.ok || std::memcmp(geom, want, sizeof(want)) != 0)
        return fail(ctx, HD_ERR_INVALID, "checkpoint geometry differs from the context's");
    const uint32_t poolTop = r.get<uint32_t>(), root = r.get<uint32_t>();
    if (!r.ok || poolTop == 0 || poolTop > ctx->cfg.pool_pages)
        return fail(ctx, HD_ERR_INVALID, "checkpoint pool larger than the context's page budget");
    std::vector<uint32_t> pool, pt, bs, full;
    r.get_array(pool, (uint64_t)poolTop * g.page_size);
    r.get_array(pt, g.total_pages());
    r.get_array(bs, g.total_buckets());
    r.get_array(full, g.levels);
    if (!r.ok || pool.size() != (uint64_t)poolTop * g.page_size || pt.size() != g.total_pages() ||
        bs.size() != g.total_buckets() || full.size() != g.levels)
        return fail(ctx, HD_ERR_INVALID, "truncated checkpoint (table)");
    const uint32_t hasColors = r.get<uint32_t>();
    std::vector<uint32_t> nodes;
    std::vector<uint64_t> offsets, arena;
    std::vector<LeafDesc> leaves;
    LeafDesc main{};
    if (hasColors) {
        r.get_array(nodes, 1ull << 32);
        r.get_array(offsets, 1ull << 32);
        r.get_array(leaves, 1ull << 31);
        r.get_array(arena, 1ull << 40);
        r.get(&main, sizeof(LeafDesc));
        if (!r.ok || nodes.size() < 8) return fail(ctx, HD_ERR_INVALID, "truncated checkpoint (colours)");
    }
    r.get(magic, 8);
    if (!r.ok || std::memcmp(magic, kCkptMagic, 8) != 0 || !r.at_end())
        return fail(ctx, HD_ERR_INVALID, "checkpoint does not end where its sizes say");
    if (hasColors) {
        const std::string why = validate_colors(g, nodes.data(), nodes.size(), offsets.size(), leaves.data(),
                                                leaves.size(), arena.size(), main);
        if (!why.empty()) return fail(ctx, HD_ERR_INVALID, "invalid checkpoint colours: " + why);
    }
    hd_status st = hd_upload_image(ctx, pool.data(), pt.data(), bs.data(), full.data(), poolTop, root);
    if (st != HD_OK) return st;
    ctx->undo.clear();
    ctx->redo.clear();
    if (!hasColors) return HD_OK;
    // colour state as it was: tree, shared offsets, unique leaves and the arena up to its top
    const uint64_t nodesCap = std::max<uint64_t>(ctx->cfg.color_nodes_cap, nodes.size() * 2 + (1u << 16));
    const uint64_t leavesCap = std::max<uint64_t>(ctx->cfg.color_leaves_cap, leaves.size() * 2 + 1024);
    const uint64_t arenaWords = std::max<uint64_t>(ctx->cfg.color_arena_words, arena.size() + arena.size() / 2 + (1u << 22));
    hipFree(ctx->cnodes);
    hipFree(ctx->coffsets);
    hipFree(ctx->cleaves);
    hipFree(ctx->arena);
    ctx->cnodes = nullptr;
    ctx->coffsets = nullptr;
    ctx->cleaves = nullptr;
    ctx->arena = nullptr;
    ctx->colors_loaded = false;
    HD_HIP(ctx, hipMalloc(&ctx->cnodes, nodesCap * 4));
    HD_HIP(ctx, hipMalloc(&ctx->coffsets, std::max<uint64_t>(offsets.size(), 1) * 8));
    HD_HIP(ctx, hipMalloc(&ctx->cleaves, leavesCap * sizeof(LeafDesc)));
    HD_HIP(ctx, hipMalloc(&ctx->arena, arenaWords * 8));
    ctx->cfg.color_nodes_cap = (uint32_t)nodesCap;
    ctx->cfg.color_leaves_cap = (uint32_t)leavesCap;
    ctx->cfg.color_arena_words = arenaWords;
    HD_HIP(ctx, hipMemsetAsync(ctx->cnodes, 0, nodesCap * 4, ctx->stream));
    HD_HIP(ctx, hipMemsetAsync(ctx->arena, 0, arenaWords * 8, ctx->stream));
    HD_HIP(ctx, hipMemcpyAsync(ctx->cnodes, nodes.data(), nodes.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    if (!offsets.empty())
        HD_HIP(ctx, hipMemcpyAsync(ctx->coffsets, offsets.data(), offsets.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    if (!leaves.empty())
        HD_HIP(ctx, hipMemcpyAsync(ctx->cleaves, leaves.data(), leaves.size() * sizeof(LeafDesc), hipMemcpyHostToDevice,
                                   ctx->stream));
    if (!arena.empty())
        HD_HIP(ctx, hipMemcpyAsync(ctx->arena, arena.data(), arena.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    const uint32_t tops[2] = {(uint32_t)nodes.size(), (uint32_t)leaves.size()};
    HD_HIP(ctx, hipMemcpyAsync(ctx->counters + 2, tops, sizeof(tops), hipMemcpyHostToDevice, ctx->stream));
    const uint64_t top64 = arena.size();
    HD_HIP(ctx, hipMemcpyAsync(ctx->counters + kArenaTopWord, &top64, 8, hipMemcpyHostToDevice, ctx->stream));
    HD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->main = main;
    ctx->arena_top = arena.size();
    ctx->n_cnodes = nodes.size();
    ctx->n_coffsets = offsets.size();
    ctx->n_cleaves = leaves.size();
    ctx->colors_loaded = true;
    ctx->ptr_gen++;
    return HD_OK;
}

hd_status hd_write_compressed_colors_file(const char* path, uint32_t top_levels, const uint64_t* enclosed,
                                          uint64_t n_enclosed, const uint32_t* weights, uint64_t n_weights,
                                          const uint64_t* blocks, uint64_t n_blocks, const uint64_t* macro,
                                          uint64_t n_macro) {
    if (!path || (!enclosed && n_enclosed) || (!weights && n_weights) || (!blocks && n_blocks) || (!macro && n_macro))
        return fail(nullptr, HD_ERR_INVALID, "bad arguments");
    File w(path, "wb");
    if (!w.ok) return fail(nullptr, HD_ERR_INVALID, std::string("cannot open ") + path);
    w.put<uint32_t>(top_levels);
    w.put_array(enclosed, n_enclosed);
    std::vector<uint32_t> swapped(weights, weights + n_weights);
    for (uint32_t& v : swapped) v = __builtin_bswap32(v);
    w.put_array(swapped.data(), n_weights);
    w.put_array(blocks, n_blocks);
    w.put_array(macro, n_macro);
    if (!w.ok) return fail(nullptr, HD_ERR_INVALID, std::string("write failed: ") + path);
    return HD_OK;
}

hd_status hd_read_compressed_colors_file(const char* path, uint32_t* top_levels, uint64_t sizes[4], uint64_t* enclosed,
                                         uint32_t* weights, uint64_t* blocks, uint64_t* macro) {
    if (!path || !sizes) return fail(nullptr, HD_ERR_INVALID, "bad arguments");
    File r(path, "rb");
    if (!r.ok) return fail(nullptr, HD_ERR_INVALID, std::string("cannot open ") + path);
    const bool fill = enclosed && weights && blocks && macro;
    const uint32_t top = r.get<uint32_t>();
    std::vector<uint64_t> e, b, m;
    std::vector<uint32_t> wv;
    const uint64_t limit = 1ull << 40;
    if (!(r.get_array(e, limit) && r.get_array(wv, limit) && r.get_array(b, limit) && r.get_array(m, limit)))
        return fail(nullptr, HD_ERR_INVALID, "truncated compressed colours file");
    const uint64_t n[4] = {e.size(), wv.size(), b.size(), m.size()};
    if (fill && (sizes[0] < n[0] || sizes[1] < n[1] || sizes[2] < n[2] || sizes[3] < n[3]))
        return fail(nullptr, HD_ERR_INVALID, "colour buffers too small");
    if (top_levels) *top_levels = top;
    for (int k = 0; k < 4; ++k) sizes[k] = n[k];
    if (!fill) return HD_OK;
    std::memcpy(enclosed, e.data(), n[0] * 8);
    for (uint64_t k = 0; k < n[1]; ++k) weights[k] = __builtin_bswap32(wv[k]);
    std::memcpy(blocks, b.data(), n[2] * 8);
    std::memcpy(macro, m.data(), n[3] * 8);
    return HD_OK;
}

hd_status hd_basic_dag_counts(const uint32_t* words, uint64_t n_words, uint32_t levels, uint32_t top_levels,
                              uint32_t* out_words, uint64_t* enclosed, uint64_t* n_enclosed) {
    if (!words || !n_words || !out_words || !n_enclosed || levels < 2 || levels > 32 || top_levels > levels - 2)
        return fail(nullptr, HD_ERR_INVALID, "bad arguments");
    try {
        DagCounts dc{words, n_words, levels - 2, {}};
        std::memcpy(out_words, words, n_words * 4);
        std::vector<uint64_t> encl;
        std::unordered_map<uint32_t, uint32_t> slot;  // node -> enclosedLeaves index
        std::vector<uint8_t> seen(n_words, 0);
        // every interior node once: its header counts its voxels (levels >= top_levels) or
        // holds its index into enclosedLeaves
        std::function<void(uint32_t, uint32_t)> walk = [&](uint32_t level, uint32_t idx) {
            if (level == dc.leafLevel || seen[idx]) return;
            seen[idx] = 1;
            const uint32_t hdr = dc.at(idx), nk = (uint32_t)__builtin_popcount(hdr & 0xFFu);
            const uint64_t c = dc.count(level, idx);
            uint64_t up = c;
            if (level < top_levels) {
                up = encl.size();
                encl.push_back(c);
            }
            if (up >= (1u << 24)) throw std::runtime_error("a voxel count does not fit a node header: raise top_levels");
            out_words[idx] = (hdr & 0xFFu) | (uint32_t)(up << 8);
            for (uint32_t k = 0; k < nk; ++k) walk(level + 1, dc.at((uint64_t)idx + 1 + k));
        };
        walk(0, 0);
        const uint64_t cap = *n_enclosed;
        *n_enclosed = encl.size();
        if (enclosed) {
            if (cap < encl.size()) return fail(nullptr, HD_ERR_INVALID, "enclosedLeaves buffer too small");
            std::memcpy(enclosed, encl.data(), encl.size() * 8);
        }
    } catch (const std::exception& ex) {
        return fail(nullptr, HD_ERR_INVALID, ex.what());
    }
    return HD_OK;
}

hd_status hd_colors_from_basic_dag(const uint32_t* words, uint64_t n_words, uint32_t levels, uint32_t color_tree_levels,
                                   uint32_t top_levels, const uint64_t* enclosed, uint64_t n_enclosed,
                                   uint32_t* scene_words, uint32_t* nodes, uint64_t* n_nodes, uint64_t* offsets,
                                   uint64_t* n_offsets) {
    const uint32_t ctl = color_tree_levels;
    if (!words || !n_words || !n_nodes || !n_offsets || levels < 3 || levels > 32 || ctl < 1 || ctl >= levels - 2 ||
        top_levels > ctl || (!enclosed && n_enclosed))
        return fail(nullptr, HD_ERR_INVALID, "bad arguments (the colour tree must end above the leaves' parents, "
                                             "and topLevels may not exceed it)");
    try {
        const uint32_t leafLevel = levels - 2;
        auto at = [&](uint64_t i) -> uint32_t {
            if (i >= n_words) throw std::runtime_error("BasicDAG index out of range");
            return words[i];
        };
        auto leaves_count = [&](uint32_t level, uint32_t node) -> uint64_t {
            const uint32_t up = node >> 8;
            if (level < top_levels) {
                if (up >= n_enclosed) throw std::runtime_error("enclosedLeaves index out of range");
                return enclosed[up];
            }
            return up;
        };
        // handle 0 stays "no leaf", so leaf k of the reference's builder is handle k + 1
        std::vector<uint32_t> tree;
        std::vector<uint64_t> leafOffsets{0};
        std::function<uint32_t(uint32_t, uint32_t, uint64_t)> build = [&](uint32_t level, uint32_t idx,
                                                                         uint64_t leavesCount) -> uint32_t {
            const uint32_t node = at(idx), mask = node & 0xFFu;
            const uint32_t colorIndex = (uint32_t)tree.size();
            uint32_t rank = 0;
            if (level == ctl - 1) {
                for (uint32_t i = 0; i < 8; ++i) {
                    tree.push_back((uint32_t)leafOffsets.size());
                    leafOffsets.push_back(leavesCount);
                    if (mask & (1u << i)) leavesCount += leaves_count(level + 1, at((uint64_t)at((uint64_t)idx + 1 + rank++)));
                }
            } else {
                tree.resize(tree.size() + 8, 0u);
                for (uint32_t i = 0; i < 8; ++i) {
                    if (!(mask & (1u << i))) continue;
                    const uint32_t child = at((uint64_t)idx + 1 + rank++);
                    const uint32_t ci = build(level + 1, child, leavesCount);
                    tree[colorIndex + i] = ci;
                    leavesCount += leaves_count(level + 1, at(child));
                }
            }
            return colorIndex;
        };
        build(0, 0, 0);
        const uint64_t capN = *n_nodes, capO = *n_offsets;
        *n_nodes = tree.size();
        *n_offsets = leafOffsets.size();
        if (!nodes && !offsets && !scene_words) return HD_OK;
        if (!nodes || !offsets || !scene_words || capN < tree.size() || capO < leafOffsets.size())
            return fail(nullptr, HD_ERR_INVALID, "colour tree buffers too small");
        std::memcpy(nodes, tree.data(), tree.size() * 4);
        std::memcpy(offsets, leafOffsets.data(), leafOffsets.size() * 8);
        // the scene's headers count voxels below the colour tree only (levels >= ctl)
        std::memcpy(scene_words, words, n_words * 4);
        std::vector<uint8_t> seen(n_words, 0);
        std::function<void(uint32_t, uint32_t)> norm = [&](uint32_t level, uint32_t idx) {
            if (level >= ctl || seen[idx]) return;
            seen[idx] = 1;
            const uint32_t hdr = at(idx), nk = (uint32_t)__builtin_popcount(hdr & 0xFFu);
            scene_words[idx] = hdr & 0xFFu;
            for (uint32_t k = 0; k < nk; ++k) norm(level + 1, at((uint64_t)idx + 1 + k));
        };
        norm(0, 0);
        (void)leafLevel;
    } catch (const std::exception& ex) {
        return fail(nullptr, HD_ERR_INVALID, ex.what());
    }
    return HD_OK;
}

}  // extern "C"
