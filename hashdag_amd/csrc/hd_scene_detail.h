// Synthetic terrain scene generator (host, multi-threaded).
//
// Terrain: h(x,z) = floor(F * (0.5 + A * fbm(x*f, 0, z*f))), f = frequency / F, fbm = 3D simplex
// noise (Gustavson's formulation) summed over `octaves` with lacunarity 2 and gain 0.5,
// normalised to [-1, 1]; permutation from a seeded splitmix64 shuffle. Voxels (x, y, z) with
// lo(x,z) <= y <= h(x,z) are set, lo = min(h, 1 + min of the in-range 4-neighbours' h): a
// watertight shell, at least one voxel thick.
//
// The DAG is built per column stack of S x S columns in parallel, deduplicated per level
// through sharded hash sets, then laid out as a BasicDAG array (root first, level by level).
// Colours: one compressed leaf per level-`color_tree_levels` subtree, encoded with the
// reference's block/weight/macro-block rules (variable_weight_size_colors.h:557-721), and
// concatenated into one shared main leaf with every subtree starting on a macro-block
// boundary; the colour tree holds, per subtree, a handle to its offset.
#pragma once
#include "hd_scene.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>

namespace hd {

namespace {

inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

inline uint32_t mix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

struct Simplex {
    uint8_t perm[512];
    double off[8][3];
    explicit Simplex(uint64_t seed) {
        uint8_t p[256];
        for (int i = 0; i < 256; ++i) p[i] = (uint8_t)i;
        uint64_t s = seed;
        for (int i = 255; i > 0; --i) {
            const int j = (int)(splitmix64(s) % (uint64_t)(i + 1));
            std::swap(p[i], p[j]);
        }
        for (int i = 0; i < 512; ++i) perm[i] = p[i & 255];
        for (auto& o : off)
            for (double& c : o) c = (double)(splitmix64(s) % 100000) / 1000.0;
    }
    static double grad(int h, double x, double y, double z) {
        // 12 cube-edge gradients
        switch (h % 12) {
            case 0: return x + y;
            case 1: return -x + y;
            case 2: return x - y;
            case 3: return -x - y;
            case 4: return x + z;
            case 5: return -x + z;
            case 6: return x - z;
            case 7: return -x - z;
            case 8: return y + z;
            case 9: return -y + z;
            case 10: return y - z;
            default: return -y - z;
        }
    }
    double noise(double x, double y, double z) const {
        const double F3 = 1.0 / 3.0, G3 = 1.0 / 6.0;
        const double s = (x + y + z) * F3;
        const int i = (int)std::floor(x + s), j = (int)std::floor(y + s), k = (int)std::floor(z + s);
        const double t = (i + j + k) * G3;
        const double x0 = x - (i - t), y0 = y - (j - t), z0 = z - (k - t);
        int i1, j1, k1, i2, j2, k2;
        if (x0 >= y0) {
            if (y0 >= z0) { i1 = 1; j1 = 0; k1 = 0; i2 = 1; j2 = 1; k2 = 0; }
            else if (x0 >= z0) { i1 = 1; j1 = 0; k1 = 0; i2 = 1; j2 = 0; k2 = 1; }
            else { i1 = 0; j1 = 0; k1 = 1; i2 = 1; j2 = 0; k2 = 1; }
        } else {
            if (y0 < z0) { i1 = 0; j1 = 0; k1 = 1; i2 = 0; j2 = 1; k2 = 1; }
            else if (x0 < z0) { i1 = 0; j1 = 1; k1 = 0; i2 = 0; j2 = 1; k2 = 1; }
            else { i1 = 0; j1 = 1; k1 = 0; i2 = 1; j2 = 1; k2 = 0; }
        }
        const double x1 = x0 - i1 + G3, y1 = y0 - j1 + G3, z1 = z0 - k1 + G3;
        const double x2 = x0 - i2 + 2 * G3, y2 = y0 - j2 + 2 * G3, z2 = z0 - k2 + 2 * G3;
        const double x3 = x0 - 1 + 3 * G3, y3 = y0 - 1 + 3 * G3, z3 = z0 - 1 + 3 * G3;
        const int ii = i & 255, jj = j & 255, kk = k & 255;
        double n = 0;
        double t0 = 0.6 - x0 * x0 - y0 * y0 - z0 * z0;
        if (t0 > 0) { t0 *= t0; n += t0 * t0 * grad(perm[ii + perm[jj + perm[kk]]], x0, y0, z0); }
        double t1 = 0.6 - x1 * x1 - y1 * y1 - z1 * z1;
        if (t1 > 0) { t1 *= t1; n += t1 * t1 * grad(perm[ii + i1 + perm[jj + j1 + perm[kk + k1]]], x1, y1, z1); }
        double t2 = 0.6 - x2 * x2 - y2 * y2 - z2 * z2;
        if (t2 > 0) { t2 *= t2; n += t2 * t2 * grad(perm[ii + i2 + perm[jj + j2 + perm[kk + k2]]], x2, y2, z2); }
        double t3 = 0.6 - x3 * x3 - y3 * y3 - z3 * z3;
        if (t3 > 0) { t3 *= t3; n += t3 * t3 * grad(perm[ii + 1 + perm[jj + 1 + perm[kk + 1]]], x3, y3, z3); }
        return 32.0 * n;
    }
    double fbm(double x, double y, double z, uint32_t octaves) const {
        double sum = 0, amp = 1, norm = 0, freq = 1;
        for (uint32_t o = 0; o < octaves; ++o) {
            sum += amp * noise(x * freq + off[o & 7][0], y * freq + off[o & 7][1], z * freq + off[o & 7][2]);
            norm += amp;
            amp *= 0.5;
            freq *= 2.0;
        }
        return std::max(-1.0, std::min(1.0, sum / norm));
    }
};

struct Terrain {
    TerrainParams p;
    uint32_t F;
    Simplex noise;
    explicit Terrain(const TerrainParams& prm) : p(prm), F(prm.footprint ? prm.footprint : (1u << prm.levels)), noise(prm.seed) {}
    int32_t height(int64_t x, int64_t z) const {
        const double f = p.frequency / (double)F;
        const double v = noise.fbm((double)x * f, 0.0, (double)z * f, p.octaves);
        double h = std::floor((double)F * (0.5 + p.amplitude * v));
        if (h < 0) h = 0;
        if (h > F - 1) h = F - 1;
        return (int32_t)h;
    }
};

// ---- compressed colour encoder (reference semantics, variable_weight_size_colors.h:557-721) ----
struct Encoder {
    struct Block { uint32_t bpw, colorBits, weightOffset, index; };
    std::vector<Block> blocks;
    std::vector<uint32_t> weights;
    std::vector<uint32_t> macroStartWeight, macroStartBlock;
    uint32_t numWeights = 0, targetBitPos = 0, colorIndex = 0;
    uint32_t lastBpw = 0, lastColorBits = 0, lastMacroStartColor = 0, lastMacroStartWeight = 0;

    void add_weight(uint32_t w, uint32_t bpw) {
        const uint32_t target = (targetBitPos + 32 - bpw) % 64;
        const uint64_t ww = (uint64_t)w << target;
        if (target > 32 - bpw) {
            weights.back() |= (uint32_t)(ww >> 32);
            targetBitPos -= bpw;
        }
        if (target < 32) {
            weights.push_back((uint32_t)ww);
            targetBitPos = target;
        }
        numWeights += bpw;
    }
    void add(uint32_t colorBits, uint32_t weight, uint32_t bpw) {
        if (colorIndex % 16384 == 0) {
            lastMacroStartColor = colorIndex;
            lastMacroStartWeight = numWeights;
            macroStartWeight.push_back(numWeights);
            macroStartBlock.push_back((uint32_t)blocks.size());
        }
        if (colorIndex % 16384 == 0 || lastColorBits != colorBits || lastBpw != bpw) {
            lastColorBits = colorBits;
            lastBpw = bpw;
            blocks.push_back({bpw, colorBits, numWeights - lastMacroStartWeight, colorIndex - lastMacroStartColor});
        }
        if (bpw > 0) add_weight(weight, bpw);
        colorIndex++;
    }
};

inline uint32_t block_header(uint32_t weightOffset, uint32_t bpw, uint32_t index) {
    if (bpw == 0) weightOffset = 0xFFFF;
    else bpw--;
    return (weightOffset << 16) | (bpw << 14) | index;
}

inline uint16_t to565(float r, float g, float b) {
    auto cl = [](float v) { return v < 0.f ? 0.f : (v > 1.f ? 1.f : v); };
    return (uint16_t)(((uint16_t)(cl(r) * 31.0f)) | ((uint16_t)(cl(g) * 63.0f) << 5) | ((uint16_t)(cl(b) * 31.0f) << 11));
}
inline uint32_t to101210(float r, float g, float b) {
    auto cl = [](float v) { return v < 0.f ? 0.f : (v > 1.f ? 1.f : v); };
    return ((uint32_t)(cl(r) * 1023.0f)) | ((uint32_t)(cl(g) * 4095.0f) << 10) | ((uint32_t)(cl(b) * 1023.0f) << 22);
}

// Colour of a voxel: one block style per 8^3 cell (min/max rgb565 from the altitude band plus
// seeded noise, 0-4 bits per weight), per-voxel weight from a position hash.
struct ColorFn {
    uint32_t seed, F;
    void operator()(uint32_t x, uint32_t y, uint32_t z, uint32_t& colorBits, uint32_t& weight, uint32_t& bpw) const {
        const uint32_t cell = mix32(seed ^ mix32((x >> 3) * 73856093u ^ (y >> 3) * 19349663u ^ (z >> 3) * 83492791u));
        const float t = (float)y / (float)F;
        float r, g, b;
        if (t < 0.35f) { r = 0.20f; g = 0.50f; b = 0.15f; }
        else if (t < 0.55f) { r = 0.45f; g = 0.35f; b = 0.20f; }
        else if (t < 0.75f) { r = 0.50f; g = 0.50f; b = 0.52f; }
        else { r = 0.95f; g = 0.95f; b = 0.97f; }
        const float j = (float)(cell & 0xFF) / 255.0f * 0.2f - 0.1f;
        bpw = (cell >> 8) % 5;
        if (bpw == 0) {
            colorBits = to101210(r + j, g + j, b + j);
            weight = 0;
        } else {
            const uint16_t lo = to565(r + j - 0.08f, g + j - 0.08f, b + j - 0.08f);
            const uint16_t hi = to565(r + j + 0.08f, g + j + 0.08f, b + j + 0.08f);
            colorBits = (uint32_t)lo | ((uint32_t)hi << 16);
            weight = mix32(seed * 31u + x * 2654435761u + y * 40503u + z * 97u) & ((1u << bpw) - 1);
        }
    }
};

// ---- sharded concurrent node sets ----
constexpr uint32_t kShards = 256;

struct Shard {
    std::mutex mu;
    std::vector<uint32_t> words;
    std::vector<uint32_t> start;   // per node: offset into words
    std::vector<uint32_t> hashes;  // per node
    std::vector<uint32_t> slots;   // 0 = empty, else local index + 1
    uint32_t count = 0;

    void grow() {
        std::vector<uint32_t> ns(slots.empty() ? 1024 : slots.size() * 2, 0u);
        const uint32_t m = (uint32_t)ns.size() - 1;
        for (uint32_t i = 0; i < count; ++i) {
            uint32_t s = (hashes[i] >> 8) & m;
            while (ns[s]) s = (s + 1) & m;
            ns[s] = i + 1;
        }
        slots.swap(ns);
    }
    uint32_t find_or_insert(const uint32_t* key, uint32_t n, uint32_t h) {
        std::lock_guard<std::mutex> lock(mu);
        if ((uint64_t)(count + 1) * 2 > slots.size()) grow();
        const uint32_t m = (uint32_t)slots.size() - 1;
        uint32_t s = (h >> 8) & m;
        while (slots[s]) {
            const uint32_t li = slots[s] - 1;
            if (hashes[li] == h) {
                const uint32_t st = start[li];
                const uint32_t len = (li + 1 < count ? start[li + 1] : (uint32_t)words.size()) - st;
                if (len == n && std::memcmp(&words[st], key, n * 4) == 0) return li;
            }
            s = (s + 1) & m;
        }
        const uint32_t li = count++;
        start.push_back((uint32_t)words.size());
        hashes.push_back(h);
        words.insert(words.end(), key, key + n);
        slots[s] = li + 1;
        return li;
    }
};

struct LevelSet {
    Shard shards[kShards];
    // returns a packed id: local * kShards + shard
    uint32_t insert(const uint32_t* key, uint32_t n) {
        uint32_t h = 0x9747b28cu ^ n;
        for (uint32_t i = 0; i < n; ++i) h = mix32(h ^ key[i]) + 0x9e3779b9u;
        h = mix32(h);
        const uint32_t s = h & (kShards - 1);
        const uint32_t li = shards[s].find_or_insert(key, n, h);
        if (li >= (1u << 24)) throw std::runtime_error("scene: too many nodes in one shard");
        return li * kShards + s;
    }
};

struct ColorChunk {
    std::vector<uint64_t> blocks;
    std::vector<uint32_t> weights;
    std::vector<uint32_t> macroWeight, macroBlock;
    uint32_t numColors = 0;
};

struct CubeResult {
    uint32_t level = 0;
    uint32_t cx = 0, cy = 0, cz = 0;
    uint32_t id = 0;     // packed id at `level`
    uint32_t count = 0;
};

struct Builder {
    const Terrain& T;
    uint32_t D, L, ctl;
    bool colors;
    std::vector<LevelSet*> sets;  // per level
    ColorFn colorFn;

    // per-task column data
    struct Cols {
        uint32_t S = 0, x0 = 0, z0 = 0;
        std::vector<int32_t> lo, hi;                   // S*S
        std::vector<std::vector<int32_t>> mipLo, mipHi;  // k = 0..log2 S
    };

    Builder(const Terrain& t, uint32_t ctl_, bool col)
        : T(t), D(t.p.levels), L(t.p.levels - 2), ctl(ctl_), colors(col), colorFn{t.p.color_seed, t.F} {
        sets.resize(L + 1);
        for (auto& s : sets) s = new LevelSet();
    }
    ~Builder() {
        for (auto* s : sets) delete s;
    }

    void columns(Cols& c, uint32_t x0, uint32_t z0, uint32_t S) const {
        c.S = S;
        c.x0 = x0;
        c.z0 = z0;
        const uint32_t F = T.F;
        const uint32_t W = S + 2;
        std::vector<int32_t> h((size_t)W * W, -1);
        for (uint32_t dz = 0; dz < W; ++dz)
            for (uint32_t dx = 0; dx < W; ++dx) {
                const int64_t x = (int64_t)x0 + dx - 1, z = (int64_t)z0 + dz - 1;
                if (x < 0 || z < 0 || x >= F || z >= F) continue;
                h[(size_t)dz * W + dx] = T.height(x, z);
            }
        c.lo.assign((size_t)S * S, INT32_MAX);
        c.hi.assign((size_t)S * S, -1);
        for (uint32_t dz = 0; dz < S; ++dz)
            for (uint32_t dx = 0; dx < S; ++dx) {
                const int32_t hh = h[(size_t)(dz + 1) * W + dx + 1];
                if (hh < 0) continue;
                int32_t mn = INT32_MAX;
                const int32_t nb[4] = {h[(size_t)(dz + 1) * W + dx], h[(size_t)(dz + 1) * W + dx + 2],
                                       h[(size_t)dz * W + dx + 1], h[(size_t)(dz + 2) * W + dx + 1]};
                for (int32_t v : nb)
                    if (v >= 0) mn = std::min(mn, v);
                const int32_t lo = mn == INT32_MAX ? hh : std::min(hh, 1 + mn);
                c.lo[(size_t)dz * S + dx] = lo;
                c.hi[(size_t)dz * S + dx] = hh;
            }
        // min/max pyramids
        uint32_t K = 0;
        while ((1u << K) < S) K++;
        c.mipLo.assign(K + 1, {});
        c.mipHi.assign(K + 1, {});
        c.mipLo[0] = c.lo;
        c.mipHi[0] = c.hi;
        for (uint32_t k = 1; k <= K; ++k) {
            const uint32_t s = S >> k, ps = S >> (k - 1);
            c.mipLo[k].assign((size_t)s * s, 0);
            c.mipHi[k].assign((size_t)s * s, 0);
            for (uint32_t z = 0; z < s; ++z)
                for (uint32_t x = 0; x < s; ++x) {
                    int32_t mn = INT32_MAX, mx = -1;
                    for (uint32_t q = 0; q < 4; ++q) {
                        const size_t i = (size_t)(2 * z + (q >> 1)) * ps + 2 * x + (q & 1);
                        mn = std::min(mn, c.mipLo[k - 1][i]);
                        mx = std::max(mx, c.mipHi[k - 1][i]);
                    }
                    c.mipLo[k][(size_t)z * s + x] = mn;
                    c.mipHi[k][(size_t)z * s + x] = mx;
                }
        }
    }

    // Build the cube at `level` with cube coords (cx, cy, cz); returns packed id or ~0u (empty).
    uint32_t build(const Cols& c, uint32_t level, uint32_t cx, uint32_t cy, uint32_t cz, uint32_t& count,
                   Encoder* enc) {
        const uint32_t k = D - level;  // cube side = 2^k
        const uint32_t lx = (cx << k) - c.x0, lz = (cz << k) - c.z0;
        const uint32_t s = c.S >> k;
        const int32_t mn = c.mipLo[k][(size_t)(lz >> k) * s + (lx >> k)];
        const int32_t mx = c.mipHi[k][(size_t)(lz >> k) * s + (lx >> k)];
        const int64_t y0 = (int64_t)cy << k, y1 = y0 + (1ll << k) - 1;
        count = 0;
        if (mx < 0 || mx < y0 || mn > y1) return ~0u;
        if (level == L) {
            uint64_t leaf = 0;
            for (uint32_t b = 0; b < 64; ++b) {
                const uint32_t c1 = b >> 3, c2 = b & 7;
                const uint32_t dx = ((c1 >> 2) & 1) * 2 + ((c2 >> 2) & 1);
                const uint32_t dy = ((c1 >> 1) & 1) * 2 + ((c2 >> 1) & 1);
                const uint32_t dz = (c1 & 1) * 2 + (c2 & 1);
                const size_t ci = (size_t)(lz + dz) * c.S + lx + dx;
                const int32_t y = (int32_t)(y0 + dy);
                if (c.lo[ci] <= y && y <= c.hi[ci]) {
                    leaf |= 1ull << b;
                    if (enc) {
                        uint32_t cb, w, bpw;
                        colorFn(c.x0 + lx + dx, (uint32_t)y, c.z0 + lz + dz, cb, w, bpw);
                        enc->add(cb, w, bpw);
                    }
                }
            }
            if (!leaf) return ~0u;
            count = (uint32_t)__builtin_popcountll(leaf);
            const uint32_t key[2] = {(uint32_t)leaf, (uint32_t)(leaf >> 32)};
            return sets[L]->insert(key, 2);
        }
        uint32_t node[9];
        uint32_t n = 1;
        uint32_t mask = 0;
        uint64_t total = 0;
        for (uint32_t ch = 0; ch < 8; ++ch) {
            uint32_t cc = 0;
            const uint32_t id = build(c, level + 1, 2 * cx + ((ch >> 2) & 1), 2 * cy + ((ch >> 1) & 1),
                                      2 * cz + (ch & 1), cc, enc);
            if (id == ~0u) continue;
            mask |= 1u << ch;
            node[n++] = id;
            total += cc;
        }
        if (!mask) return ~0u;
        if (level >= ctl && total >= (1u << 24)) throw std::runtime_error("scene: voxel count overflows header");
        node[0] = mask | (level >= ctl ? (uint32_t)(total << 8) : 0u);
        count = (uint32_t)std::min<uint64_t>(total, 0xFFFFFFFFu);
        return sets[level]->insert(node, n);
    }
};

}  // namespace


}  // namespace hd
