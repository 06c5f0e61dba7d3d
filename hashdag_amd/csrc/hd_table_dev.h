// Device-side hash table append shared by edits and GC: HashTable::add_leaf_node /
// add_interior_node (hash_table.h:359-468) with concurrent appenders.
//
// A bucket's size is advanced with a CAS loop so that concurrent appenders reserve disjoint
// ranges; a node that does not fit the rest of its page starts on the next page (padding),
// and the appender whose range starts a page numbers it from poolTop (page 0 of the physical
// pool is never used, hash_table.h:821) unless the page is already mapped (after a GC).
#pragma once

#include <hip/hip_runtime.h>

#include "hd_internal.h"

namespace hd {

struct TableRef {
    Geom g;
    uint32_t* vpool;
    uint32_t* page_table;
    uint32_t* bucket_sizes;
    uint32_t* pool_top;
    uint32_t pool_pages;
};

constexpr uint32_t kAppendBucketFull = 0xFFFFFFFFu;
constexpr uint32_t kAppendPoolFull = 0xFFFFFFFEu;

__device__ inline bool table_map_page(const TableRef& T, uint32_t page) {
    if (atomicCAS(&T.page_table[page], 0u, 0xFFFFFFFFu) != 0u) return true;  // mapped already
    const uint32_t phys = atomicAdd(T.pool_top, 1u);
    if (phys >= T.pool_pages) {
        atomicExch(&T.page_table[page], 0u);
        return false;
    }
    atomicExch(&T.page_table[page], phys);
    return true;
}

// Appends `n` words to `bucket` of `level`; returns the virtual address, or one of the
// kAppend* codes on overflow (nothing written).
__device__ inline uint32_t table_append(const TableRef& T, uint32_t level, uint32_t bucket, const uint32_t* words,
                                        uint32_t n) {
    const Geom& g = T.g;
    const uint32_t ps = g.page_size;
    uint32_t* bsPtr = &T.bucket_sizes[g.bucket_global(level, bucket)];
    const uint32_t cap = g.bucket_size(level);
    uint32_t cur = __hip_atomic_load(bsPtr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t start;
    for (;;) {
        const uint32_t left = ps - cur % ps;
        start = (left != ps && left < n) ? cur + left : cur;
        if (start + n > cap) return kAppendBucketFull;
        const uint32_t prev = atomicCAS(bsPtr, cur, start + n);
        if (prev == cur) break;
        cur = prev;
    }
    const uint32_t va = g.make_ptr(level, bucket, start);
    if (start % ps == 0 && !table_map_page(T, va / ps)) return kAppendPoolFull;
    for (uint32_t k = 0; k < n; ++k) T.vpool[va + k] = words[k];
    return va;
}

}  // namespace hd
