// Host-side checks of images and colour state before they reach the device.
//
// A table image (hd_upload_image, hd_checkpoint_load) and a colour state (hd_checkpoint_load)
// come from files or callers. The kernels index device memory with the pointers they hold, so
// a corrupt image is turned into an error here, before any copy or launch: every page-table
// entry maps a distinct physical page below pool_top; every bucket size fits its bucket and
// its used pages are mapped; every node reachable from the root or from a full node lies
// inside its level's used bucket range with its children likewise; every colour-tree entry
// points at a node or leaf that exists, and every leaf's arrays lie inside the colour arena.
// The reference trusts its inputs (hash_dag_factory.cpp:225-246 reads a dump as is).
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hashdag_hip.h"
#include "hd_internal.h"

namespace hd {

namespace {

struct ImageView {
    const Geom& g;
    const uint32_t* pool;  // physical pages 0 .. pool_top - 1
    const uint32_t* page_table;
    const uint32_t* bucket_sizes;

    uint32_t word(uint32_t va) const {
        return pool[(uint64_t)page_table[va / g.page_size] * g.page_size + va % g.page_size];
    }
    // level of the virtual address range va falls in, or levels when outside the table
    uint32_t level_of(uint32_t va) const {
        uint64_t base = 0;
        for (uint32_t l = 0; l < g.levels; ++l) {
            const uint64_t size = (uint64_t)g.buckets_per_level(l) * g.bucket_size(l);
            if (va < base + size) return l;
            base += size;
        }
        return g.levels;
    }
    // a node of `len` words at va lies inside the used part of a bucket of `level`
    bool inside(uint32_t level, uint32_t va, uint32_t len) const {
        if (level_of(va) != level) return false;
        const uint32_t b0 = g.make_ptr(level, 0, 0);
        const uint32_t bucket = (va - b0) / g.bucket_size(level), pos = (va - b0) % g.bucket_size(level);
        return (uint64_t)pos + len <= bucket_sizes[g.bucket_global(level, bucket)];
    }
};

// Walks the DAG from `start` (a node of `level`) and checks every node it reaches.
std::string walk(const ImageView& v, uint32_t level, uint32_t start, std::vector<uint64_t>& seen) {
    const Geom& g = v.g;
    std::vector<std::pair<uint32_t, uint32_t>> stack{{level, start}};
    while (!stack.empty()) {
        const auto [l, va] = stack.back();
        stack.pop_back();
        // the level first: an address seen before must still belong to the level it is reached at
        if (v.level_of(va) != l) return "a pointer into another level's buckets";
        if ((seen[va >> 6] >> (va & 63)) & 1) continue;
        seen[va >> 6] |= 1ull << (va & 63);
        if (l == g.leaf_level()) {
            if (!v.inside(l, va, 2)) return "a leaf lies outside its level's used buckets";
            continue;
        }
        if (!v.inside(l, va, 1)) return "a node lies outside its level's used buckets";
        const uint32_t mask = v.word(va) & 0xFFu;
        const uint32_t n = (uint32_t)__builtin_popcount(mask);
        if (n == 0 && !(l == 0 && va == start)) return "an interior node without children";
        if (!v.inside(l, va, n + 1)) return "a node runs past its bucket's used words";
        for (uint32_t k = 1; k <= n; ++k) {
            const uint32_t child = v.word(va + k);
            if (child >= g.total_virtual()) return "a child pointer outside the virtual space";
            stack.push_back({l + 1, child});
        }
    }
    return "";
}

}  // namespace

std::string validate_image(const Geom& g, uint32_t pool_pages, const uint32_t* pool, const uint32_t* page_table,
                           const uint32_t* bucket_sizes, const uint32_t* full_nodes, uint32_t pool_top,
                           uint32_t root) {
    if (pool_top == 0 || pool_top > pool_pages) return "pool_top outside 1 .. the pool budget";
    std::vector<uint8_t> used(pool_top, 0);
    for (uint64_t p = 0; p < g.total_pages(); ++p) {
        const uint32_t phys = page_table[p];
        if (phys == 0) continue;
        if (phys >= pool_top) return "a page-table entry at or above pool_top";
        if (used[phys]) return "two virtual pages map one physical page";
        used[phys] = 1;
    }
    for (uint32_t l = 0; l < g.levels; ++l)
        for (uint32_t b = 0; b < g.buckets_per_level(l); ++b) {
            const uint32_t size = bucket_sizes[g.bucket_global(l, b)];
            if (size > g.bucket_size(l)) return "a bucket size above its bucket's capacity";
            const uint32_t page0 = g.make_ptr(l, b, 0) / g.page_size;
            for (uint32_t p = 0; p < (size + g.page_size - 1) / g.page_size; ++p)
                if (page_table[page0 + p] == 0) return "a bucket's used words lie on an unmapped page";
        }
    if (root >= g.total_virtual()) return "the root is not an address of the table";
    const ImageView v{g, pool, page_table, bucket_sizes};
    std::vector<uint64_t> seen((g.total_virtual() + 63) / 64, 0);
    std::string why = walk(v, 0, root, seen);
    if (!why.empty()) return "root: " + why;
    for (uint32_t l = 1; l <= g.leaf_level(); ++l) {  // level 0 has none until a GC (edits read levels >= 1)
        if (full_nodes[l] >= g.total_virtual()) return "a full node is not an address of the table";
        why = walk(v, l, full_nodes[l], seen);
        if (!why.empty()) return "full node of level " + std::to_string(l) + ": " + why;
    }
    return "";
}

std::string validate_colors(const Geom& g, const uint32_t* nodes, uint64_t n_nodes, uint64_t n_offsets,
                            const LeafDesc* leaves, uint64_t n_leaves, uint64_t arena_words, const LeafDesc& main) {
    auto leaf_ok = [&](const LeafDesc& d) {
        const uint64_t weightWords = 2 * arena_words;
        return d.blocks <= arena_words && d.n_blocks <= arena_words - d.blocks && d.macro <= arena_words &&
               d.n_macro <= arena_words - d.macro && d.weights32 <= weightWords &&
               d.n_weights <= weightWords - d.weights32;
    };
    if (!leaf_ok(main)) return "the main colour leaf lies outside the arena";
    for (uint64_t i = 0; i < n_leaves; ++i)
        if (!leaf_ok(leaves[i])) return "a colour leaf lies outside the arena";
    if (n_nodes < 8) return "the colour tree has no root node";
    if (n_offsets == 0) return "no shared colour offsets (entry 0 resolves to offset 0)";
    const uint32_t ctl = g.color_tree_levels;
    if (ctl == 0 || ctl > g.levels) return "colour tree levels outside 1 .. levels";
    // entries of the nodes at depth ctl - 1 are leaf handles, the others node indices; a node
    // is visited once (a tree), so the walk is bounded by the node count
    std::vector<uint8_t> seen(n_nodes, 0);
    std::vector<std::pair<uint64_t, uint32_t>> stack{{0, 0}};
    while (!stack.empty()) {
        const auto [node, depth] = stack.back();
        stack.pop_back();
        if (seen[node]) return "a colour-tree node is reached twice";
        seen[node] = 1;
        for (uint32_t c = 0; c < 8; ++c) {
            const uint32_t e = nodes[node + c];
            if (e == 0) continue;
            if (depth + 1 == ctl) {
                if (e & 0x80000000u) {
                    if ((e & 0x7FFFFFFFu) >= n_leaves) return "a colour leaf handle past the leaf table";
                } else if (e >= n_offsets) {
                    return "a shared colour leaf index past the offsets";
                }
            } else {
                if ((uint64_t)e + 8 > n_nodes) return "a colour-tree child index past the node array";
                stack.push_back({e, depth + 1});
            }
        }
    }
    return "";
}

}  // namespace hd

extern "C" hd_status hd_validate_image(const hd_config* cfg, const uint32_t* pool, const uint32_t* page_table,
                                       const uint32_t* bucket_sizes, const uint32_t* full_nodes, uint32_t pool_top,
                                       uint32_t root, char* why, uint32_t why_len) {
    if (!cfg || !pool || !page_table || !bucket_sizes || !full_nodes) return HD_ERR_INVALID;
    const hd::Geom g{cfg->levels,          cfg->page_size,       cfg->num_top_levels, cfg->bucket_bits_top,
                     cfg->bucket_bits_low, cfg->bucket_size_top, cfg->bucket_size_low, cfg->color_tree_levels};
    const std::string r = hd::validate_image(g, cfg->pool_pages, pool, page_table, bucket_sizes, full_nodes, pool_top,
                                             root);
    if (why && why_len) {
        std::strncpy(why, r.c_str(), why_len - 1);
        why[why_len - 1] = 0;
    }
    return r.empty() ? HD_OK : HD_ERR_INVALID;
}
