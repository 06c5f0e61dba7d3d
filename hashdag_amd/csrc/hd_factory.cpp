// Host-side construction of the reference-layout hash table from a BasicDAG
// (HashDAGFactory::load_from_DAG, hash_dag_factory.cpp:5-99,152-184): full nodes first
// (hash_table.h:749-789), then every BasicDAG node appended post-order with a memo map.
// Appends follow HashTable::add_leaf_node / add_interior_node (hash_table.h:359-468):
// bucket = hash & (buckets - 1), new pages taken from poolTop, interior nodes never
// straddle a page. The image is uploaded to the device by hd_load_basic_dag.
#include <stdexcept>

#include "hd_internal.h"

namespace hd {

namespace {

struct Appender {
    HostImage& img;
    const Geom& g;

    uint32_t* word(uint32_t va) {
        return &img.pool[(uint64_t)img.page_table[va / g.page_size] * g.page_size + va % g.page_size];
    }
    void new_page(uint32_t page) {
        if (img.pool_top >= img.pool_pages) throw std::runtime_error("pool exhausted");
        img.page_table[page] = img.pool_top++;
        const uint64_t need = (uint64_t)img.pool_top * g.page_size;
        if (img.pool.size() < need) img.pool.resize(std::max<uint64_t>(need, img.pool.size() * 3 / 2), 0u);
    }
    uint32_t add_leaf(uint64_t leaf) {
        const uint32_t level = g.leaf_level();
        const uint32_t bucket = g.bucket_from_hash(level, hash_leaf(leaf));
        uint32_t& bs = img.bucket_sizes[g.bucket_global(level, bucket)];
        if (bs + 2 > g.bucket_size(level)) throw std::runtime_error("leaf bucket full");
        const uint32_t va = g.make_ptr(level, bucket, bs);
        if (bs % g.page_size == 0 && img.page_table[va / g.page_size] == 0) new_page(va / g.page_size);
        uint32_t* p = word(va);
        p[0] = (uint32_t)leaf;
        p[1] = (uint32_t)(leaf >> 32);
        bs += 2;
        return va;
    }
    uint32_t add_interior(uint32_t level, const uint32_t* node, uint32_t n) {
        const uint32_t bucket = g.bucket_from_hash(level, hash_words(node, n));
        uint32_t& bsRef = img.bucket_sizes[g.bucket_global(level, bucket)];
        uint32_t bs = bsRef;
        const uint32_t left = g.page_size - bs % g.page_size;
        if (left == g.page_size || left < n) {
            if (left != g.page_size) bs += left;
            if (bs + n > g.bucket_size(level)) throw std::runtime_error("interior bucket full");
            const uint32_t va = g.make_ptr(level, bucket, bs);
            if (img.page_table[va / g.page_size] == 0) new_page(va / g.page_size);
        } else if (bs + n > g.bucket_size(level)) {
            throw std::runtime_error("interior bucket full");
        }
        const uint32_t va = g.make_ptr(level, bucket, bs);
        uint32_t* p = word(va);
        for (uint32_t i = 0; i < n; ++i) p[i] = node[i];
        bsRef = bs + n;
        return va;
    }
    void add_full(uint32_t level) {
        if (level == g.leaf_level()) {
            img.full_nodes[level] = add_leaf(~0ull);
            return;
        }
        const uint32_t count = level < g.color_tree_levels ? 0u : (1u << (3 * (g.levels - level)));
        uint32_t node[9];
        node[0] = (count << 8) | 0xFFu;
        for (int c = 0; c < 8; ++c) node[1 + c] = img.full_nodes[level + 1];
        img.full_nodes[level] = add_interior(level, node, 9);
    }
};

}  // namespace

void load_from_basic_dag(HostImage& img, const uint32_t* data, uint64_t n) {
    const Geom& g = img.g;
    img.page_table.assign(g.total_pages(), 0u);
    img.bucket_sizes.assign(g.total_buckets(), 0u);
    img.full_nodes.assign(g.levels, 0u);
    img.pool.assign((uint64_t)g.page_size * 1024, 0u);
    img.pool_top = 1;
    Appender a{img, g};
    for (uint32_t level = g.leaf_level(); level > 0; --level) a.add_full(level);

    // Iterative post-order DFS over the BasicDAG with a memo (BasicDAG index -> va + 1).
    std::vector<uint32_t> memo(n, 0u);
    struct Frame { uint32_t level, index, next, n; uint32_t buf[9]; };
    std::vector<Frame> stack;
    stack.reserve(64);
    auto push = [&](uint32_t level, uint32_t index) {
        Frame f;
        f.level = level;
        f.index = index;
        f.next = 0;
        f.n = 1;
        f.buf[0] = data[index];
        stack.push_back(f);
    };
    uint32_t result = 0;
    push(0, 0);
    while (!stack.empty()) {
        Frame& f = stack.back();
        if (f.level == g.leaf_level()) {
            const uint64_t leaf = (uint64_t)data[f.index] | ((uint64_t)data[f.index + 1] << 32);
            result = a.add_leaf(leaf);
            stack.pop_back();
        } else {
            const uint32_t mask = f.buf[0] & 0xFFu;
            bool descended = false;
            while (f.next < 8) {
                const uint32_t c = f.next;
                if (!(mask & (1u << c))) {
                    f.next++;
                    continue;
                }
                const uint32_t childIndex = data[f.index + __builtin_popcount(mask & ((1u << c) - 1u)) + 1];
                if (memo[childIndex]) {
                    f.buf[f.n++] = memo[childIndex] - 1;
                    f.next++;
                    continue;
                }
                push(f.level + 1, childIndex);
                descended = true;
                break;
            }
            if (descended) continue;
            result = a.add_interior(f.level, f.buf, f.n);
            stack.pop_back();
        }
        if (!stack.empty()) {
            // a child just finished: record it in the parent and advance
            Frame& p = stack.back();
            const uint32_t mask = p.buf[0] & 0xFFu;
            const uint32_t c = p.next;
            const uint32_t childIndex = data[p.index + __builtin_popcount(mask & ((1u << c) - 1u)) + 1];
            memo[childIndex] = result + 1;
            p.buf[p.n++] = result;
            p.next++;
        }
    }
    img.root = result;
}

}  // namespace hd
