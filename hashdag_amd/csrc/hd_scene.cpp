// Synthetic terrain scene generator: task loop and assembly of the BasicDAG and colours.
// (Noise, colour encoder, sharded node sets and the per-cube builder: hd_scene_detail.h.)
//
// Work unit: one column stack of S x S columns. For every cube of that stack at the cube
// level (the task level, or the colour-tree level when colours need finer subtrees) the
// builder deduplicates nodes into the per-level sets and, with colours, encodes the cube's
// voxels in DFS (Morton) order into a task-local colour buffer. The cubes are then sorted by
// Morton key; siblings are contiguous, so the levels above the cube level are built by one
// linear pass per level, the shared main leaf is the concatenation of the cube chunks in
// that order (each starting on a macro-block boundary), and the colour tree is built by a
// recursion over key ranges in the reference factory's DFS allocation order
// (hash_dag_factory.cpp:101-150).
#include "hd_scene_detail.h"

namespace hd {

namespace {

struct CubeRec {
    uint64_t key;     // Morton key at the cube level
    uint32_t id;      // packed node id at the cube level
    uint32_t count;   // voxels
    uint32_t task;    // task index (colour buffer owner)
    uint32_t chunk;   // chunk index in the task's colour buffer
};

struct ChunkRef {
    uint64_t blockStart, nBlocks, weightStart, nWeights, macroStart, nMacro;
};

struct TaskColors {
    std::vector<uint64_t> blocks;
    std::vector<uint32_t> weights;
    std::vector<uint32_t> macroBlock, macroWeight;
    std::vector<ChunkRef> chunks;
};

inline uint64_t morton(uint32_t x, uint32_t y, uint32_t z, uint32_t bits) {
    uint64_t k = 0;
    for (int b = (int)bits - 1; b >= 0; --b)
        k = (k << 3) | (((x >> b) & 1u) << 2) | (((y >> b) & 1u) << 1) | ((z >> b) & 1u);
    return k;
}

void append_chunk(TaskColors& tc, const Encoder& enc) {
    ChunkRef r;
    r.blockStart = tc.blocks.size();
    r.nBlocks = enc.blocks.size();
    r.weightStart = tc.weights.size();
    r.nWeights = enc.weights.size();
    r.macroStart = tc.macroBlock.size();
    r.nMacro = enc.macroStartBlock.size();
    for (const auto& b : enc.blocks)
        tc.blocks.push_back(((uint64_t)b.colorBits << 32) | block_header(b.weightOffset, b.bpw, b.index));
    for (uint32_t w : enc.weights) tc.weights.push_back(__builtin_bswap32(w));
    tc.macroBlock.insert(tc.macroBlock.end(), enc.macroStartBlock.begin(), enc.macroStartBlock.end());
    tc.macroWeight.insert(tc.macroWeight.end(), enc.macroStartWeight.begin(), enc.macroStartWeight.end());
    tc.chunks.push_back(r);
}

template <typename Fn>
void parallel_for(uint64_t n, uint32_t threads, Fn fn) {
    std::atomic<uint64_t> next{0};
    std::vector<std::thread> th;
    for (uint32_t t = 0; t < threads; ++t)
        th.emplace_back([&]() {
            for (;;) {
                const uint64_t i = next.fetch_add(1);
                if (i >= n) return;
                fn(i);
            }
        });
    for (auto& t : th) t.join();
}

}  // namespace

int32_t terrain_height(const TerrainParams& p, uint32_t x, uint32_t z) {
    Terrain t(p);
    if (x >= t.F || z >= t.F) return -1;
    return t.height(x, z);
}

Scene generate_terrain(const TerrainParams& prm) {
    const uint32_t D = prm.levels;
    if (D < 4 || D > 17) throw std::runtime_error("scene: depth must be in [4, 17]");
    Terrain T(prm);
    const uint32_t E = 1u << D;
    if (T.F == 0 || T.F > E || (T.F & (T.F - 1))) throw std::runtime_error("scene: footprint must be a power of two <= 2^D");
    const uint32_t L = D - 2;
    const uint32_t ctl = prm.color_tree_levels;
    const bool colors = prm.colors && L > ctl;
    uint32_t logS = std::min<uint32_t>(5, D - 1);
    if (colors) logS = std::max(logS, D - ctl);
    logS = std::min(logS, 31u - __builtin_clz(T.F));
    const uint32_t S = 1u << logS;
    const uint32_t Tl = D - logS;
    const uint32_t tiles = T.F / S;
    if (colors && Tl > ctl) throw std::runtime_error("scene: footprint too small for colour subtrees");
    const uint32_t cubeLevel = (colors && Tl < ctl) ? ctl : Tl;
    const uint32_t sub = cubeLevel - Tl;  // cube level below the task level
    const uint32_t nthreads = prm.threads ? prm.threads : std::max(1u, std::thread::hardware_concurrency());

    Builder B(T, ctl, colors);
    const uint64_t ntasks = (uint64_t)tiles * tiles;
    std::vector<std::vector<CubeRec>> taskCubes(ntasks);
    std::vector<TaskColors> taskColors(colors ? ntasks : 0);
    std::atomic<uint64_t> voxels{0};
    std::atomic<bool> failed{false};
    std::string err;
    std::mutex errMu;

    parallel_for(ntasks, nthreads, [&](uint64_t t) {
        if (failed) return;
        try {
            thread_local Builder::Cols cols;
            const uint32_t tx = (uint32_t)(t % tiles), tz = (uint32_t)(t / tiles);
            B.columns(cols, tx * S, tz * S, S);
            const int32_t mn = cols.mipLo[logS][0], mx = cols.mipHi[logS][0];
            if (mx < 0) return;
            const uint32_t n = 1u << sub;
            for (int64_t cy = mn >> logS; cy <= (mx >> logS); ++cy) {
                for (uint64_t m = 0; m < (uint64_t)n * n * n; ++m) {
                    uint32_t sx = 0, sy = 0, sz = 0;
                    for (uint32_t b = 0; b < sub; ++b) {
                        const uint32_t c = (uint32_t)(m >> (3 * (sub - 1 - b))) & 7u;
                        sx = (sx << 1) | ((c >> 2) & 1u);
                        sy = (sy << 1) | ((c >> 1) & 1u);
                        sz = (sz << 1) | (c & 1u);
                    }
                    const uint32_t gx = (tx << sub) + sx, gy = ((uint32_t)cy << sub) + sy, gz = (tz << sub) + sz;
                    Encoder enc;
                    uint32_t cnt = 0;
                    const uint32_t id = B.build(cols, cubeLevel, gx, gy, gz, cnt, colors ? &enc : nullptr);
                    if (id == ~0u) continue;
                    voxels += cnt;
                    CubeRec r{morton(gx, gy, gz, cubeLevel), id, cnt, (uint32_t)t, 0};
                    if (colors) {
                        r.chunk = (uint32_t)taskColors[t].chunks.size();
                        append_chunk(taskColors[t], enc);
                    }
                    taskCubes[t].push_back(r);
                }
            }
        } catch (const std::exception& e) {
            std::lock_guard<std::mutex> lk(errMu);
            err = e.what();
            failed = true;
        }
    });
    if (failed) throw std::runtime_error(err);

    std::vector<CubeRec> cubes;
    {
        uint64_t n = 0;
        for (auto& v : taskCubes) n += v.size();
        cubes.reserve(n);
        for (auto& v : taskCubes) {
            cubes.insert(cubes.end(), v.begin(), v.end());
            std::vector<CubeRec>().swap(v);
        }
    }
    if (cubes.empty()) throw std::runtime_error("scene: empty terrain");
    std::sort(cubes.begin(), cubes.end(), [](const CubeRec& a, const CubeRec& b) { return a.key < b.key; });

    // ---- levels above the cube level: siblings are contiguous in Morton order ----
    {
        struct Item { uint64_t key; uint32_t id; uint64_t count; };
        std::vector<Item> cur(cubes.size());
        for (size_t i = 0; i < cubes.size(); ++i) cur[i] = {cubes[i].key, cubes[i].id, cubes[i].count};
        for (uint32_t level = cubeLevel; level-- > 0;) {
            std::vector<Item> next;
            size_t i = 0;
            while (i < cur.size()) {
                const uint64_t pk = cur[i].key >> 3;
                uint32_t node[9];
                uint32_t nn = 1, mask = 0;
                uint64_t total = 0;
                while (i < cur.size() && (cur[i].key >> 3) == pk) {
                    mask |= 1u << (cur[i].key & 7u);
                    node[nn++] = cur[i].id;
                    total += cur[i].count;
                    ++i;
                }
                if (level >= ctl && total >= (1u << 24)) throw std::runtime_error("scene: voxel count overflows header");
                node[0] = mask | (level >= ctl ? (uint32_t)(total << 8) : 0u);
                next.push_back({pk, B.sets[level]->insert(node, nn), total});
            }
            cur.swap(next);
        }
        if (cur.size() != 1) throw std::runtime_error("scene: root level must hold one node");
    }

    // ---- lay out the BasicDAG: level by level, shard by shard ----
    Scene sc;
    sc.levels = D;
    sc.color_tree_levels = ctl;
    sc.num_voxels = voxels.load();
    sc.level_nodes.assign(L + 1, 0);
    sc.level_words.assign(L + 1, 0);
    std::vector<std::array<std::vector<uint32_t>, kShards>> pos(L + 1);
    uint64_t total = 0;
    for (uint32_t l = 0; l <= L; ++l) {
        const uint64_t levelStart = total;
        for (uint32_t s = 0; s < kShards; ++s) {
            Shard& sh = B.sets[l]->shards[s];
            pos[l][s].resize(sh.count);
            for (uint32_t i = 0; i < sh.count; ++i) {
                pos[l][s][i] = (uint32_t)total;
                const uint32_t len = (i + 1 < sh.count ? sh.start[i + 1] : (uint32_t)sh.words.size()) - sh.start[i];
                total += len;
            }
            sc.level_nodes[l] += sh.count;
        }
        sc.level_words[l] = total - levelStart;
        if (total >= 0xFFFFFFF0ull) throw std::runtime_error("scene: BasicDAG exceeds 32-bit indices");
    }
    if (sc.level_nodes[0] != 1) throw std::runtime_error("scene: root level must hold one node");
    sc.dag.resize(total);
    parallel_for((uint64_t)(L + 1) * kShards, nthreads, [&](uint64_t it) {
        const uint32_t l = (uint32_t)(it / kShards), s = (uint32_t)(it % kShards);
        Shard& sh = B.sets[l]->shards[s];
        for (uint32_t i = 0; i < sh.count; ++i) {
            const uint32_t st = sh.start[i];
            const uint32_t len = (i + 1 < sh.count ? sh.start[i + 1] : (uint32_t)sh.words.size()) - st;
            uint32_t* out = &sc.dag[pos[l][s][i]];
            out[0] = sh.words[st];
            if (l == L) {
                out[1] = sh.words[st + 1];
            } else {
                for (uint32_t j = 1; j < len; ++j) {
                    const uint32_t id = sh.words[st + j];
                    out[j] = pos[l + 1][id % kShards][id / kShards];
                }
            }
        }
    });

    if (!colors) {
        sc.color_nodes.assign(8, 0u);
        return sc;
    }

    // ---- shared main leaf: chunks in Morton order, each starting a new macro block ----
    const size_t nc = cubes.size();
    std::vector<uint64_t> bStart(nc + 1, 0), wStart(nc + 1, 0), mStart(nc + 1, 0);
    for (size_t i = 0; i < nc; ++i) {
        const ChunkRef& r = taskColors[cubes[i].task].chunks[cubes[i].chunk];
        bStart[i + 1] = bStart[i] + r.nBlocks;
        wStart[i + 1] = wStart[i] + r.nWeights;
        mStart[i + 1] = mStart[i] + r.nMacro;
    }
    sc.blocks.resize(bStart[nc]);
    sc.weights.resize(wStart[nc]);
    sc.macro.resize(2 * mStart[nc]);
    parallel_for(nc, nthreads, [&](uint64_t i) {
        const TaskColors& tc = taskColors[cubes[i].task];
        const ChunkRef& r = tc.chunks[cubes[i].chunk];
        std::copy(tc.blocks.begin() + r.blockStart, tc.blocks.begin() + r.blockStart + r.nBlocks,
                  sc.blocks.begin() + bStart[i]);
        std::copy(tc.weights.begin() + r.weightStart, tc.weights.begin() + r.weightStart + r.nWeights,
                  sc.weights.begin() + wStart[i]);
        for (uint64_t m = 0; m < r.nMacro; ++m) {
            sc.macro[2 * (mStart[i] + m)] = bStart[i] + tc.macroBlock[r.macroStart + m];
            sc.macro[2 * (mStart[i] + m) + 1] = wStart[i] * 32 + tc.macroWeight[r.macroStart + m];
        }
    });
    std::vector<TaskColors>().swap(taskColors);

    // ---- colour tree: DFS pre-order allocation, 8 entries per node; level ctl-1 entries
    // are handles into the offsets table (handle 0 reserved: an edit reads it as "no leaf") ----
    sc.color_offsets.push_back(0);
    std::function<uint32_t(uint32_t, size_t, size_t)> rec = [&](uint32_t level, size_t lo, size_t hi) -> uint32_t {
        const uint32_t idx = (uint32_t)sc.color_nodes.size();
        sc.color_nodes.resize(idx + 8, 0u);
        const uint32_t shift = 3 * (ctl - level - 1);  // child digit position in the cube key
        size_t i = lo;
        while (i < hi) {
            const uint32_t c = (uint32_t)(cubes[i].key >> shift) & 7u;
            size_t j = i;
            while (j < hi && ((cubes[j].key >> shift) & 7u) == c) ++j;
            if (level + 1 == ctl) {
                sc.color_nodes[idx + c] = (uint32_t)sc.color_offsets.size();
                sc.color_offsets.push_back(mStart[i] * 16384);
            } else {
                const uint32_t ci = rec(level + 1, i, j);
                sc.color_nodes[idx + c] = ci;
            }
            i = j;
        }
        return idx;
    };
    rec(0, 0, nc);
    return sc;
}

}  // namespace hd
