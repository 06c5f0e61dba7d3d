// Internal layout shared by the host code and the HIP kernels of the product library.
//
// Device memory layout (one context, one GPU):
//   vpool        uint32[total_virtual + pad]  node words addressed directly by their 32-bit
//                virtual address (hash_table.h:19-77 addressing). The reference translates
//                every access through a page table (hash_table.h:155-173); with 288 GB of
//                HBM the whole virtual space (< 2^32 words) is resident, so traversal does one
//                dependent load per word instead of two.
//   page_table   uint32[total_pages]  virtual page -> physical page number, kept for page
//                accounting (pool budget, allocation order) and the reference-layout image.
//   bucket_sizes uint32[total_buckets], full_nodes uint32[levels]
//   colour tree  nodes uint32[], offsets uint64[], leaf descriptors, one uint64 arena holding
//                every leaf's blocks / macro pairs / weights.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#ifndef HD_HOST_DEVICE
#if defined(__HIPCC__)
#define HD_HOST_DEVICE __host__ __device__ inline
#else
#define HD_HOST_DEVICE inline
#endif
#endif

namespace hd {

// HashTable geometry (hash_dag_globals.h:7-40), run-time.
struct Geom {
    uint32_t levels, page_size, num_top_levels, bucket_bits_top, bucket_bits_low, bucket_size_top, bucket_size_low,
        color_tree_levels;

    HD_HOST_DEVICE uint32_t leaf_level() const { return levels - 2; }
    HD_HOST_DEVICE uint32_t top() const { return num_top_levels < levels ? num_top_levels : levels; }
    HD_HOST_DEVICE uint32_t buckets_per_level(uint32_t l) const {
        return l < top() ? (1u << bucket_bits_top) : (1u << bucket_bits_low);
    }
    HD_HOST_DEVICE uint32_t bucket_size(uint32_t l) const { return l < top() ? bucket_size_top : bucket_size_low; }
    HD_HOST_DEVICE uint64_t total_buckets() const {
        return (uint64_t)top() * (1u << bucket_bits_top) + (uint64_t)(levels - top()) * (1u << bucket_bits_low);
    }
    HD_HOST_DEVICE uint64_t total_pages() const {
        return (uint64_t)top() * (1u << bucket_bits_top) * (bucket_size_top / page_size) +
               (uint64_t)(levels - top()) * (1u << bucket_bits_low) * (bucket_size_low / page_size);
    }
    HD_HOST_DEVICE uint64_t total_virtual() const { return total_pages() * page_size; }
    HD_HOST_DEVICE uint32_t bucket_global(uint32_t l, uint32_t b) const {
        return l < top() ? l * (1u << bucket_bits_top) + b
                         : top() * (1u << bucket_bits_top) + (l - top()) * (1u << bucket_bits_low) + b;
    }
    HD_HOST_DEVICE uint32_t make_ptr(uint32_t l, uint32_t b, uint32_t pos) const {
        const uint32_t base = l < top() ? (l * (1u << bucket_bits_top) + b) * bucket_size_top
                                        : top() * (1u << bucket_bits_top) * bucket_size_top +
                                              ((l - top()) * (1u << bucket_bits_low) + b) * bucket_size_low;
        return base + pos;
    }
    HD_HOST_DEVICE uint32_t bucket_from_hash(uint32_t l, uint32_t h) const { return h & (buckets_per_level(l) - 1); }
};

// hashes (utils.h:159-200): fmix64 for leaves, Murmur3-x86_32 body with a word-count finaliser
HD_HOST_DEVICE uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
HD_HOST_DEVICE uint64_t fmix64(uint64_t h) {
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return h;
}
HD_HOST_DEVICE uint32_t hash_leaf(uint64_t leaf) { return (uint32_t)fmix64(leaf); }
HD_HOST_DEVICE uint32_t hash_words(const uint32_t* w, uint32_t n) {
    uint32_t h = 0;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t k = w[i] * 0xcc9e2d51u;
        k = (k << 15) | (k >> 17);
        k *= 0x1b873593u;
        h ^= k;
        h = (h << 13) | (h >> 19);
        h = h * 5u + 0xe6546b64u;
    }
    return fmix32(h ^ n);
}

// A compressed colour leaf on the device: element offsets into the uint64 colour arena.
// blocks/macro are counted in 64-bit words; weights in 32-bit words from `weights32`
// (= 2 * its 64-bit arena offset). offset = shared offset, or ~0 for a unique leaf.
struct LeafDesc {
    uint64_t blocks, n_blocks;
    uint64_t macro, n_macro;
    uint64_t weights32, n_weights;
    uint64_t offset;
    uint64_t pad;
};

// Everything a kernel needs, passed by value.
struct DevDag {
    const uint32_t* vpool;
    uint32_t root;
    uint32_t levels;
    uint32_t leaf_level;
    uint32_t color_tree_levels;
    const uint32_t* root_ptr;  // when set, the root is read from device memory (graph replays)
};

struct DevColors {
    const uint32_t* nodes;
    const uint64_t* offsets;
    const LeafDesc* leaves;
    const uint64_t* arena;
    LeafDesc main;
    uint32_t valid;  // colour tree present
};

constexpr uint32_t kInvalidNode = 0xDFFFFFFu;  // C_InvalidNodeIndex (hash_dag_edits.h:17)

// Host-side reference-layout image (HashTable CPU data).
struct HostImage {
    Geom g;
    std::vector<uint32_t> pool;        // physical pages (grows as pages are allocated)
    std::vector<uint32_t> page_table;  // total_pages
    std::vector<uint32_t> bucket_sizes;
    std::vector<uint32_t> full_nodes;
    uint32_t pool_top = 1;
    uint32_t pool_pages = 0;
    uint32_t root = 0;
};

// hash_dag_factory.cpp:152-184 on the host (hd_factory.cpp). Throws std::runtime_error with
// "bucket"/"pool" in the message on overflow.
void load_from_basic_dag(HostImage& img, const uint32_t* data, uint64_t n);

// Checks of external inputs before they reach the device (hd_validate.cpp); "" when valid,
// else the reason. pool holds physical pages 0 .. pool_top - 1.
std::string validate_image(const Geom& g, uint32_t pool_pages, const uint32_t* pool, const uint32_t* page_table,
                           const uint32_t* bucket_sizes, const uint32_t* full_nodes, uint32_t pool_top,
                           uint32_t root);
std::string validate_colors(const Geom& g, const uint32_t* nodes, uint64_t n_nodes, uint64_t n_offsets,
                            const LeafDesc* leaves, uint64_t n_leaves, uint64_t arena_words, const LeafDesc& main);

}  // namespace hd
