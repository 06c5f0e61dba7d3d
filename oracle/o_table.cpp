// Oracle: hashes, paged bucketed hash table, BasicDAG loading, GC, canonical form.
// TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
#include "oracle_core.hpp"

#include <algorithm>
#include <map>
#include <set>
#include <stdexcept>

namespace orc {

// utils.h:159-167 (MurmurHash3 fmix32)
u32 murmurhash32(u32 h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

// utils.h:168-176 (MurmurHash3 fmix64)
u64 murmurhash64(u64 h) {
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return h;
}

// utils.h:179-200: MurmurHash3_x86_32 block loop over 32-bit words; the finaliser mixes in
// the WORD count n (textbook MurmurHash3 mixes in the byte count).
u32 murmurhash32xN(const u32* words, size_t n, u32 seed) {
    u32 h = seed;
    for (size_t i = 0; i < n; ++i) {
        u32 k = words[i];
        k *= 0xcc9e2d51u;
        k = (k << 15) | (k >> 17);
        k *= 0x1b873593u;
        h ^= k;
        h = (h << 13) | (h >> 19);
        h = h * 5u + 0xe6546b64u;
    }
    h ^= (u32)n;
    return murmurhash32(h);
}

u32 hash_leaf(u64 leaf) { return (u32)murmurhash64(leaf); }
u32 hash_interior(u32 n, const u32* node) { return murmurhash32xN(node, n, 0); }

// hash_table.cpp:5-71
void HashTable::create(const Geometry& geom, u32 pool_pages) {
    g = geom;
    pool_max_pages = pool_pages;
    pool.assign((u64)pool_pages * g.page_size, 0u);
    page_table.assign(g.total_pages(), 0u);
    bucket_sizes.assign(g.total_buckets(), 0u);
    full_nodes.assign(g.levels, 0u);
    pool_top = 1;
    overflow = false;
}

// hash_table.h:794-809
void HashTable::allocate_page(u32 page) {
    std::unique_lock<std::mutex> lk(pool_lock, std::defer_lock);
    if (locks) lk.lock();
    if (pool_top >= pool_max_pages) {
        overflow = true;
        throw std::runtime_error("oracle: pool exhausted");
    }
    page_table[page] = pool_top++;
}

// hash_table.h:196-274 (leaf scan, 2-word stride, page by page)
u32 HashTable::find_leaf_in_bucket(u32 bucketSize, u32 bucket, u64 leaf) const {
    const u32 level = g.leaf_level();
    const u32 base = g.make_ptr(level, bucket, 0);
    for (u32 pindex = 0; pindex < bucketSize; pindex += g.page_size) {
        const u32 end = std::min(bucketSize, pindex + g.page_size);
        for (u32 index = pindex; index < end; index += 2) {
            const u32* p = ptr(base + index);
            const u64 bits = (u64)p[0] | ((u64)p[1] << 32);
            if (bits == leaf) return base + index;
        }
    }
    return NOT_FOUND;
}

// hash_table.h:276-357. Only nodes whose start satisfies start + n < pageEnd are compared,
// and a page whose start + n >= pageEnd ends the whole search.
u32 HashTable::find_interior_in_bucket(u32 bucketSize, u32 level, u32 bucket, u32 n, const u32* node) const {
    const u32 base = g.make_ptr(level, bucket, 0);
    for (u32 pindex = 0; pindex < bucketSize; pindex += g.page_size) {
        u32 pageEnd = std::min(bucketSize, pindex + g.page_size);
        if (pindex + n >= pageEnd) return NOT_FOUND;
        pageEnd -= n;
        for (u32 index = pindex; index < pageEnd;) {
            const u32* p = ptr(base + index);
            if (std::memcmp(node, p, n * sizeof(u32)) == 0) return base + index;
            index += total_size(p[0]);
        }
    }
    return NOT_FOUND;
}

// hash_table.h:359-400
u32 HashTable::add_leaf(u64 leaf, u32 hash) {
    const u32 level = g.leaf_level();
    const u32 bucket = g.bucket_from_hash(level, hash);
    u32& bs = bucket_size_ref(level, bucket);
    if (bs + 2 > g.bucket_size(level)) {
        overflow = true;
        throw std::runtime_error("oracle: leaf bucket full");
    }
    const u32 va = g.make_ptr(level, bucket, bs);
    const u32 page = va / g.page_size;
    if (bs % g.page_size == 0 && !page_allocated(page)) allocate_page(page);
    u32* p = ptr(va);
    p[0] = (u32)leaf;
    p[1] = (u32)(leaf >> 32);
    bs += 2;
    return va;
}

// hash_table.h:401-468 (a node never straddles a page: pad to the next page)
u32 HashTable::add_interior(u32 level, u32 n, const u32* node, u32 hash) {
    const u32 bucket = g.bucket_from_hash(level, hash);
    u32& bsRef = bucket_size_ref(level, bucket);
    u32 bs = bsRef;
    const u32 spaceLeft = g.page_size - (bs % g.page_size);
    u32 va;
    if (spaceLeft == g.page_size || spaceLeft < n) {
        if (spaceLeft != g.page_size) bs += spaceLeft;
        if (bs + n > g.bucket_size(level)) {
            overflow = true;
            throw std::runtime_error("oracle: interior bucket full");
        }
        va = g.make_ptr(level, bucket, bs);
        if (!page_allocated(va / g.page_size)) allocate_page(va / g.page_size);
    } else {
        if (bs + n > g.bucket_size(level)) {
            overflow = true;
            throw std::runtime_error("oracle: interior bucket full");
        }
        va = g.make_ptr(level, bucket, bs);
    }
    u32* p = ptr(va);
    for (u32 i = 0; i < n; ++i) p[i] = node[i];
    bsRef = bs + n;
    return va;
}

// hash_table.h:470-514
u32 HashTable::find_or_add_leaf(u64 leaf) {
    const u32 hash = hash_leaf(leaf);
    const u32 bucket = g.bucket_from_hash(g.leaf_level(), hash);
    std::unique_lock<std::mutex> lk;
    if (locks) lk = std::unique_lock<std::mutex>((*locks)[g.bucket_global(g.leaf_level(), bucket) % locks->size()]);
    const u32 bs = bucket_sizes[g.bucket_global(g.leaf_level(), bucket)];
    u32 r = find_leaf_in_bucket(bs, bucket, leaf);
    if (r == NOT_FOUND) r = add_leaf(leaf, hash);
    return r;
}

// hash_table.h:515-560
u32 HashTable::find_or_add_interior(u32 level, u32 n, const u32* node) {
    const u32 hash = hash_interior(n, node);
    const u32 bucket = g.bucket_from_hash(level, hash);
    std::unique_lock<std::mutex> lk;
    if (locks) lk = std::unique_lock<std::mutex>((*locks)[g.bucket_global(level, bucket) % locks->size()]);
    const u32 bs = bucket_sizes[g.bucket_global(level, bucket)];
    u32 r = find_interior_in_bucket(bs, level, bucket, n, node);
    if (r == NOT_FOUND) r = add_interior(level, n, node, hash);
    return r;
}

// hash_table.h:749-789
void HashTable::add_full_node(u32 level) {
    if (level == g.leaf_level()) {
        const u64 leaf = ~0ull;
        full_nodes[level] = add_leaf(leaf, hash_leaf(leaf));
    } else {
        const u32 count = level < g.color_tree_levels ? 0u : (1u << (3 * (g.levels - level)));
        u32 node[9];
        node[0] = (count << 8) | 0xFFu;
        for (int c = 0; c < 8; ++c) node[1 + c] = full_nodes[level + 1];
        full_nodes[level] = add_interior(level, 9, node, hash_interior(9, node));
    }
}

u64 HashTable::used_words(u32 level) const {
    u64 s = 0;
    for (u32 b = 0; b < g.buckets_per_level(level); ++b) s += bucket_sizes[g.bucket_global(level, b)];
    return s;
}

// hash_dag_factory.cpp:5-99 (post-order recursion with a memo map) and :152-184
static u32 create_hash_dag(DAG& dag, const u32* data, std::vector<u32>& memo, u32 level, u32 index) {
    const Geometry& g = dag.t.g;
    if (level == g.leaf_level()) {
        const u64 leaf = (u64)data[index] | ((u64)data[index + 1] << 32);
        return dag.t.add_leaf(leaf, hash_leaf(leaf));
    }
    const u32 header = data[index];
    const u8 mask = child_mask(header);
    u32 buf[9];
    u32 n = 0;
    buf[n++] = header;
    for (u8 i = 0; i < 8; ++i) {
        if (!(mask & (1u << i))) continue;
        const u32 childIndex = data[index + child_offset(mask, i)];
        u32 v = memo[childIndex];
        if (v == 0) {
            v = create_hash_dag(dag, data, memo, level + 1, childIndex);
            memo[childIndex] = v + 1;
        } else {
            v--;
        }
        buf[n++] = v;
    }
    return dag.t.add_interior(level, n, buf, hash_interior(n, buf));
}

void load_from_basic_dag(DAG& dag, const Geometry& g, u32 pool_pages, const u32* data, u64 n) {
    dag.t.create(g, pool_pages);
    for (u32 level = g.leaf_level(); level > 0; --level) dag.t.add_full_node(level);
    std::vector<u32> memo(n, 0u);
    dag.root = create_hash_dag(dag, data, memo, 0, 0);
}

// hash_dag.h:407-547. Live nodes are iterated in ascending address order (the reference
// iterates unordered_sets, whose order is implementation-defined).
void remove_stale_nodes(DAG& dag, u32 maxLevel) {
    HashTable& t = dag.t;
    const Geometry& g = t.g;
    std::vector<std::set<u32>> valid(maxLevel + 1);
    valid[0].insert(dag.root);
    for (u32 level = 0; level < maxLevel; ++level) {
        for (u32 idx : valid[level]) {
            const u32* p = t.ptr(idx);
            for (u32 o = 1; o < total_size(p[0]); ++o) valid[level + 1].insert(p[o]);
        }
    }
    std::map<u32, u32> map, newMap;
    for (u32 level = maxLevel; level != (u32)-1; --level) {
        const bool isLeaf = level == g.leaf_level();
        std::vector<u32> temp;
        std::vector<u32> keys;
        newMap.clear();
        for (u32 idx : valid[level]) {
            u32* p = t.ptr(idx);
            const u32 n = isLeaf ? 2 : total_size(p[0]);
            if (level != maxLevel)
                for (u32 o = 1; o < n; ++o) p[o] = map.at(p[o]);
            newMap[idx] = (u32)temp.size();
            for (u32 o = 0; o < n; ++o) temp.push_back(p[o]);
            keys.push_back(idx);
        }
        for (u32 b = 0; b < g.buckets_per_level(level); ++b) t.bucket_size_ref(level, b) = 0;
        t.add_full_node(level);
        for (u32 oldIndex : keys) {
            u32& v = newMap.at(oldIndex);
            if (isLeaf) {
                const u64 leaf = (u64)temp[v] | ((u64)temp[v + 1] << 32);
                v = t.add_leaf(leaf, hash_leaf(leaf));
            } else {
                const u32* np = &temp[v];
                const u32 n = total_size(np[0]);
                v = t.add_interior(level, n, np, hash_interior(n, np));
            }
        }
        map.swap(newMap);
    }
    dag.root = map.at(dag.root);
}

// Canonical form: nodes interned per level in first-appearance DFS order from the root
// (children visited 0..7). Leaves: 64-bit mask. Interior at level >= color_tree_levels:
// full header (mask | count << 8) and canonical child ids. Interior below: mask only (the
// upper header bits there are not content: edits write 0, imports carry BasicDAG bits).
// Output: for each level, [count, then per node: words...] concatenated.
std::vector<u32> canonicalize(const DAG& dag) {
    const Geometry& g = dag.g();
    const u32 L = g.leaf_level();
    std::vector<std::unordered_map<u32, u32>> addrToCanon(L + 1);
    std::vector<std::map<std::vector<u32>, u32>> intern(L + 1);
    std::vector<std::vector<u32>> serial(L + 1);
    std::vector<u32> counts(L + 1, 0);
    // Iterative post-order DFS.
    struct Frame { u32 level, va, next; std::vector<u32> kids; };
    std::vector<Frame> stack;
    auto canon_of = [&](u32 level, u32 va, u32& out) -> bool {
        auto it = addrToCanon[level].find(va);
        if (it == addrToCanon[level].end()) return false;
        out = it->second;
        return true;
    };
    auto intern_node = [&](u32 level, std::vector<u32> key) -> u32 {
        auto it = intern[level].find(key);
        if (it != intern[level].end()) return it->second;
        const u32 id = counts[level]++;
        intern[level].emplace(key, id);
        serial[level].insert(serial[level].end(), key.begin(), key.end());
        return id;
    };
    u32 rootCanon = 0;
    stack.push_back({0, dag.root, 0, {}});
    while (!stack.empty()) {
        Frame& f = stack.back();
        if (f.level == L) {
            const u64 leaf = dag.leaf(f.va);
            const u32 id = intern_node(L, {(u32)leaf, (u32)(leaf >> 32)});
            addrToCanon[L][f.va] = id;
            stack.pop_back();
            continue;
        }
        const u32 hdr = dag.node(f.va);
        const u32 nk = popc(child_mask(hdr));
        if (f.next < nk) {
            const u32 childVa = *dag.t.ptr(f.va + 1 + f.next);
            u32 c;
            if (canon_of(f.level + 1, childVa, c)) {
                f.kids.push_back(c);
                f.next++;
            } else {
                stack.push_back({f.level + 1, childVa, 0, {}});
            }
            continue;
        }
        std::vector<u32> key;
        key.push_back(f.level >= g.color_tree_levels ? hdr : (hdr & 0xFFu));
        key.insert(key.end(), f.kids.begin(), f.kids.end());
        const u32 lvl = f.level, va = f.va;
        const u32 id = intern_node(lvl, key);
        addrToCanon[lvl][va] = id;
        stack.pop_back();
        if (!stack.empty()) {
            stack.back().kids.push_back(id);
            stack.back().next++;
        } else {
            rootCanon = id;
        }
    }
    std::vector<u32> out;
    out.push_back(L + 1);
    out.push_back(rootCanon);
    for (u32 l = 0; l <= L; ++l) {
        out.push_back(counts[l]);
        out.push_back((u32)serial[l].size());
        out.insert(out.end(), serial[l].begin(), serial[l].end());
    }
    return out;
}

// dag_utils.h:138-172
bool get_value(const DAG& dag, Path path) {
    const Geometry& g = dag.g();
    u32 va = dag.root;
    for (u32 level = 0; level < g.levels; ++level) {
        if (level < g.leaf_level()) {
            const u32 hdr = dag.node(va);
            const u8 mask = child_mask(hdr);
            const u8 c = path.child_index(level + 1, g.levels);
            if (!(mask & (1u << c))) return false;
            va = dag.child_index(va, mask, c);
        } else {
            const u64 leaf = dag.leaf(va);
            const u32 bit = ((path.p.x & 1) ? 4 : 0) | ((path.p.y & 1) ? 2 : 0) | ((path.p.z & 1) ? 1 : 0) |
                            ((path.p.x & 2) ? 32 : 0) | ((path.p.y & 2) ? 16 : 0) | ((path.p.z & 2) ? 8 : 0);
            return (leaf >> bit) & 1;
        }
    }
    return true;
}

// dag_utils.h:174-252 (inclusive-max box overlap quirk kept)
static bool is_empty_impl(const DAG& dag, u32 va, Path path, u32 level, u32 maxLevel, U3 start, U3 size) {
    const Geometry& g = dag.g();
    if (level == maxLevel) return false;
    if (level < g.leaf_level()) {
        const u32 shift = g.levels - level;
        const U3 bmin{path.p.x << shift, path.p.y << shift, path.p.z << shift};
        const u32 ext = (u32)((1 << shift) - 1);
        const U3 bmax{bmin.x + ext, bmin.y + ext, bmin.z + ext};
        const U3 imin = start;
        const U3 imax{start.x + size.x, start.y + size.y, start.z + size.z};
        const bool overlap = !(bmin.x >= imax.x || bmin.y >= imax.y || bmin.z >= imax.z || bmax.x <= imin.x ||
                               bmax.y <= imin.y || bmax.z <= imin.z || imin.x >= bmax.x || imin.y >= bmax.y ||
                               imin.z >= bmax.z || imax.x <= bmin.x || imax.y <= bmin.y || imax.z <= bmin.z);
        if (!overlap) return true;
    }
    const u32 hdr = dag.node(va);
    const u8 mask = child_mask(hdr);
    u32 off = 1;
    for (u8 c = 0; c < 8; ++c) {
        Path np = path;
        np.descend(c);
        if (mask & (1u << c)) {
            const u32 cva = *dag.t.ptr(va + off++);
            if (!is_empty_impl(dag, cva, np, level + 1, maxLevel, start, size)) return false;
        }
    }
    return true;
}

bool is_empty_region(const DAG& dag, u32 maxLevel, U3 start, U3 size) {
    return is_empty_impl(dag, dag.root, Path(0, 0, 0), 0, maxLevel, start, size);
}

}  // namespace orc
