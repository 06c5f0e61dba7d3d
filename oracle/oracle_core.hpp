// Oracle: CPU restatement of the HashDAG reference algorithms on the hot path.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product (hashdag_amd/) links, imports or
// calls this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg use it, and only as the checker / CPU baseline.
//
// Every function restates the semantics of the cited reference file:line (paths are
// relative to the reference's src/ directory). It is written from the reference's
// behaviour, not copied from it. Floating point follows strict IEEE-754 evaluation in
// the reference's source operation order with contraction disabled (-ffp-contract=off).
//
// Parity pinning (what checks each part) is listed in oracle_capi.h.
#pragma once

#include <cstdint>
#include <cstring>
#include <cmath>
#include <mutex>
#include <vector>
#include <string>
#include <unordered_map>
#include <memory>

namespace orc {

using u8 = uint8_t;
using u16 = uint16_t;
using u32 = uint32_t;
using u64 = uint64_t;

// ---------------------------------------------------------------------------------------------
// Small vector types with the reference's component-wise operator semantics (cuda_math.h:26-80).
// ---------------------------------------------------------------------------------------------
struct F3 { float x, y, z; };
struct D3 { double x, y, z; };
struct U3 { u32 x, y, z; };

inline F3 f3(float a) { return {a, a, a}; }
inline F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 operator*(F3 a, F3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline F3 operator/(F3 a, F3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline F3 operator*(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline F3 operator+(F3 a, float s) { return {a.x + s, a.y + s, a.z + s}; }
inline F3 operator-(F3 a, float s) { return {a.x - s, a.y - s, a.z - s}; }
inline F3 operator/(F3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(D3 a, D3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline D3 operator/(D3 a, D3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline D3 operator*(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline D3 operator+(D3 a, double s) { return {a.x + s, a.y + s, a.z + s}; }
inline D3 operator/(D3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
// dot/length/normalize: cuda_math.h:40,51-52 (left-to-right sum, normalize = (1/len)*v)
inline float dotf(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double dotd(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float lengthf(F3 a) { return (float)std::sqrt((double)dotf(a, a)); }
inline double lengthd(D3 a) { return std::sqrt(dotd(a, a)); }
inline F3 normalizef(F3 a) { return (1.0f / lengthf(a)) * a; }
inline D3 normalized(D3 a) { return (1.0 / lengthd(a)) * a; }
inline F3 to_f3(D3 d) { return {(float)d.x, (float)d.y, (float)d.z}; }
inline D3 to_d3(F3 f) { return {(double)f.x, (double)f.y, (double)f.z}; }
inline float tmax3(F3 v) { float m = v.y > v.z ? v.y : v.z; return v.x > m ? v.x : m; }  // constexpr_max chain
inline float tmin3(F3 v) { float m = v.y < v.z ? v.y : v.z; return v.x < m ? v.x : m; }
template <typename T> inline T tclamp(T f, T a, T b) { return f < a ? a : (f > b ? b : f); }

// ---------------------------------------------------------------------------------------------
// Bit utilities and hashes (utils.h:10-174)
// ---------------------------------------------------------------------------------------------
inline u32 popc(u32 a) { return (u32)__builtin_popcount(a); }
inline u32 popcll(u64 a) { return (u32)__builtin_popcountll(a); }
inline u8 child_mask(u32 node) { return (u8)(node & 0xff); }
inline u32 total_size(u32 node) { return popc(child_mask(node)) + 1; }
inline u32 child_offset(u8 mask, u8 child) { return popc(mask & ((1u << child) - 1u)) + 1; }
u32 murmurhash32(u32 h);
u64 murmurhash64(u64 h);
u32 murmurhash32xN(const u32* words, size_t n, u32 seed = 0);

// Leaf = 64-bit two-level mask (base_dag.h:6-59)
inline u8 leaf_first_mask(u64 leaf) {
    u8 m = 0;
    for (int i = 0; i < 8; ++i)
        if ((leaf >> (8 * i)) & 0xffu) m |= (u8)(1u << i);
    return m;
}
inline u8 leaf_second_mask(u64 leaf, u8 firstChild) {
    u32 word = (firstChild & 4) ? (u32)(leaf >> 32) : (u32)leaf;
    return (u8)(word >> ((firstChild & 3) * 8));
}

// Path (path.h:6-81)
struct Path {
    U3 p{0, 0, 0};
    Path() = default;
    Path(u32 x, u32 y, u32 z) : p{x, y, z} {}
    void ascend(u32 n) { p.x >>= n; p.y >>= n; p.z >>= n; }
    void descend(u8 c) {
        p.x = (p.x << 1) | ((c & 4u) >> 2);
        p.y = (p.y << 1) | ((c & 2u) >> 1);
        p.z = (p.z << 1) | (c & 1u);
    }
    F3 as_position(u32 shift = 0) const { return {(float)(p.x << shift), (float)(p.y << shift), (float)(p.z << shift)}; }
    u8 child_index(u32 level, u32 total) const {
        u32 s = total - level;
        return (u8)((((p.x >> s) & 1) ? 4 : 0) | (((p.y >> s) & 1) ? 2 : 0) | (((p.z >> s) & 1) ? 1 : 0));
    }
    bool is_null() const { return p.x == 0 && p.y == 0 && p.z == 0; }
};

// ---------------------------------------------------------------------------------------------
// HashTable geometry and addressing (hash_dag_globals.h:7-40, hash_table.h:19-77).
// The geometry is a run-time struct here; the reference fixes it at compile time.
// ---------------------------------------------------------------------------------------------
struct Geometry {
    u32 levels = 16;             // SCENE_DEPTH
    u32 page_size = 512;         // PAGE_SIZE
    u32 num_top_levels = 9;      // BUCKETS_NUM_TOP_LEVELS
    u32 bucket_bits_top = 10;    // BUCKETS_BITS_FOR_TOP_LEVELS
    u32 bucket_bits_low = 16;    // BUCKETS_BITS_FOR_LOW_LEVELS
    u32 bucket_size_top = 1024;  // BUCKETS_SIZE_FOR_TOP_LEVELS
    u32 bucket_size_low = 4096;  // BUCKETS_SIZE_FOR_LOW_LEVELS
    u32 color_tree_levels = 10;  // C_colorTreeLevels

    u32 leaf_level() const { return levels - 2; }
    u32 top() const { return num_top_levels < levels ? num_top_levels : levels; }
    u32 buckets_per_level(u32 l) const { return l < top() ? (1u << bucket_bits_top) : (1u << bucket_bits_low); }
    u32 bucket_size(u32 l) const { return l < top() ? bucket_size_top : bucket_size_low; }
    u64 total_buckets() const { return (u64)top() * (1u << bucket_bits_top) + (u64)(levels - top()) * (1u << bucket_bits_low); }
    u64 total_pages() const {
        return (u64)top() * (1u << bucket_bits_top) * (bucket_size_top / page_size) +
               (u64)(levels - top()) * (1u << bucket_bits_low) * (bucket_size_low / page_size);
    }
    u64 total_virtual() const { return total_pages() * page_size; }
    u32 bucket_global(u32 l, u32 b) const {
        return l < top() ? l * (1u << bucket_bits_top) + b
                         : top() * (1u << bucket_bits_top) + (l - top()) * (1u << bucket_bits_low) + b;
    }
    u32 make_ptr(u32 l, u32 b, u32 pos) const {
        u32 idx;
        if (l < top()) idx = (l * (1u << bucket_bits_top) + b) * bucket_size_top;
        else idx = top() * (1u << bucket_bits_top) * bucket_size_top + ((l - top()) * (1u << bucket_bits_low) + b) * bucket_size_low;
        return idx + pos;
    }
    u32 bucket_from_hash(u32 l, u32 h) const { return h & (buckets_per_level(l) - 1); }
};

u32 hash_leaf(u64 leaf);                            // hash_table.h:80-84
u32 hash_interior(u32 n, const u32* node);          // hash_table.h:85-89

constexpr u32 NOT_FOUND = 0xFFFFFFFFu;
constexpr u32 C_InvalidNodeIndex = 0xDFFFFFFu;      // hash_dag_edits.h:15-17
constexpr u32 C_IgnoreColorIndex = 0xFFFFFFFu;
constexpr u32 C_RemoveColorIndex = 0xEFFFFFFu;

// Paged bucketed hash table, CPU layout: pool of physical pages + page table
// (hash_table.h:137-898, hash_table.cpp:5-71). No Bloom filter (result-neutral,
// hash_table.h:563-659). find_or_add takes a bucket's lock when `locks` is set (threaded edit
// tasks, as the reference's per-bucket mutexes); pages are then numbered under `pool_lock`.
struct HashTable {
    Geometry g;
    std::vector<u32> pool;         // pool_max_pages * page_size
    std::vector<u32> page_table;   // total_pages
    std::vector<u32> bucket_sizes; // total_buckets
    std::vector<u32> full_nodes;   // per level (ADD_FULL_NODES_FIRST)
    u32 pool_top = 1;              // page 0 is never used (hash_table.h:821)
    u32 pool_max_pages = 0;
    bool overflow = false;         // bucket or pool overflow observed (the reference does not check)
    std::vector<std::mutex>* locks = nullptr;  // striped bucket locks while edit tasks run threaded
    std::mutex pool_lock;

    void create(const Geometry& geom, u32 pool_pages);
    u32* ptr(u32 va) { return &pool[(u64)page_table[va / g.page_size] * g.page_size + va % g.page_size]; }
    const u32* ptr(u32 va) const { return &pool[(u64)page_table[va / g.page_size] * g.page_size + va % g.page_size]; }
    bool page_allocated(u32 page) const { return page_table[page] != 0; }
    void allocate_page(u32 page);
    u32& bucket_size_ref(u32 l, u32 b) { return bucket_sizes[g.bucket_global(l, b)]; }

    u32 find_leaf_in_bucket(u32 bucketSize, u32 bucket, u64 leaf) const;
    u32 find_interior_in_bucket(u32 bucketSize, u32 level, u32 bucket, u32 n, const u32* node) const;
    u32 add_leaf(u64 leaf, u32 hash);
    u32 add_interior(u32 level, u32 n, const u32* node, u32 hash);
    u32 find_or_add_leaf(u64 leaf);
    u32 find_or_add_interior(u32 level, u32 n, const u32* node);
    void add_full_node(u32 level);
    u64 used_words(u32 level) const;
};

// ---------------------------------------------------------------------------------------------
// Colour codec (variable_weight_size_colors.h, color_utils.h)
// ---------------------------------------------------------------------------------------------
constexpr u32 kColorsPerMacro = 16 * 1024;

u16 float3_to_rgb565(F3 c);
F3 rgb565_to_float3(u16 c);
u32 float3_to_rgb888(F3 c);
F3 rgb888_to_float3(u32 c);
F3 rgb101210_to_float3(u32 c);
u32 float3_to_rgb101210(F3 c);

struct CompressedColor {
    u32 colorBits = 0;
    u8 weight = 0;
    u8 bitsPerWeight = 0;
    F3 min_color() const { return rgb565_to_float3((u16)(colorBits & 0xFFFF)); }
    F3 max_color() const { return rgb565_to_float3((u16)((colorBits >> 16) & 0xFFFF)); }
    float get_weight() const;
    void set_single_color(F3 c) { bitsPerWeight = 0; colorBits = float3_to_rgb101210(c); }
    F3 get_color() const;
};

struct ColorLeaf {
    static constexpr u64 kUnique = ~0ull;
    u64 offset = 0;                 // offset into the shared main leaf; kUnique for unique leaves
    std::vector<u64> blocks;
    std::vector<u32> weights;       // words stored byte-swapped (CFG_COLOR_SWAP_BYTE_ORDER)
    std::vector<u64> macro;         // [first block, weight bit offset] per macro block
    bool valid = false;             // StaticArray::is_valid of blocks
};

// A colour leaf as the tracer/editor sees it: data plus the shared offset (or kUnique).
struct LeafView {
    const ColorLeaf* d = nullptr;
    u64 offset = ColorLeaf::kUnique;
    bool is_shared() const { return offset != ColorLeaf::kUnique; }
    bool is_valid() const { return d && d->valid; }
    bool is_valid_index(u64 idx) const;
    u64 macro_last_block(u64 m) const;
    u32 binary_search_blocks(u64 colorIndex) const;
    CompressedColor color_for_block(u32 block, u64 colorIndex) const;
    CompressedColor get_color(u64 colorIndex) const;
    template <typename B> void copy_colors(B& builder, u64 start, u64 count) const;
};

u8 extract_bits(u32 bits, const std::vector<u32>& words, u64 bitPtr);

// Encoder (variable_weight_size_colors.h:557-721)
struct ColorLeafBuilder {
    struct Block { u32 bpw, colorBits, weightOffset, index; };
    struct Macro { u32 startWeight, startBlock; };
    std::vector<Block> blocks;
    std::vector<u32> weights;
    std::vector<Macro> macros;
    u32 numWeights = 0, targetBitPos = 0, colorIndex = 0;
    u32 lastBpw = 0, lastColorBits = 0, lastMacroStartColor = 0, lastMacroStartWeight = 0;

    void add_weight(u8 w, u8 bpw);
    void add(const CompressedColor& c);
    void add_large_single_color(F3 c, u32 n);
    void build(ColorLeaf& leaf) const;
};

// Colour tree over the top levels (hash_dag_colors.h:9-198)
struct HashDAGColors {
    std::vector<u32> nodes;
    std::vector<ColorLeaf> leaves;  // unique leaves (handles with bit 31)
    std::vector<u64> offsets;       // shared handles -> offset into main leaf
    ColorLeaf mainLeaf;
    LeafView get_leaf(u32 handle) const;
    u32 allocate_interior_node() { nodes.push_back(0); return (u32)nodes.size() - 1; }
    u32 allocate_leaf() { ColorLeaf l; l.offset = ColorLeaf::kUnique; leaves.push_back(l); return (u32)(leaves.size() - 1) | 0x80000000u; }
};

// ---------------------------------------------------------------------------------------------
// DAG = HashTable + root (hash_dag.h:214-248)
// ---------------------------------------------------------------------------------------------
struct DAG {
    HashTable t;
    u32 root = 0;
    const Geometry& g() const { return t.g; }
    u32 node(u32 va) const { return *t.ptr(va); }
    u32 child_index(u32 va, u8 mask, u8 child) const { return *t.ptr(va + child_offset(mask, child)); }
    u64 leaf(u32 va) const { const u32* p = t.ptr(va); return (u64)p[0] | ((u64)p[1] << 32); }
};

// BasicDAG -> HashDAG (hash_dag_factory.cpp:5-99,152-184)
void load_from_basic_dag(DAG& dag, const Geometry& g, u32 pool_pages, const u32* data, u64 n);

// ---------------------------------------------------------------------------------------------
// Tracer (tracer.cu, dag_tracer.cu:71-113)
// ---------------------------------------------------------------------------------------------
struct RayParams { D3 cameraPosition, rayMin, rayDDx, rayDDy; };
RayParams get_trace_params(const double position[3], const double rotation[9], const double boundsMin[3],
                           const double boundsMax[3], u32 levels, u32 W, u32 H);
U3 trace_path(const DAG& dag, const RayParams& p, u32 px, u32 py, u64* bytes);
u32 trace_color(const DAG& dag, const HashDAGColors& colors, U3 path);
u32 shade_shadow(const DAG& dag, const RayParams& p, u32 W, u32 H, u32 px, u32 py, U3 path, u32 colorIn,
                 float bias, float fog, u64* bytes);
bool intersect_any(const DAG& dag, F3 origin, F3 dir, u64* bytes);
void set_any_hit_front_to_back(bool on);
F3 sun_direction();

// ---------------------------------------------------------------------------------------------
// Editors (hash_dag_editors.h) and edits (hash_dag_edits.h, hash_dag.h:255-405)
// ---------------------------------------------------------------------------------------------
enum EditKind : u32 {
    SPHERE_ADD = 0, SPHERE_CARVE = 1, BOX_ADD = 2, BOX_CARVE = 3, SPHERE_PAINT = 4, COPY = 5, FILL = 6,
    COPY_DEC = 7,  // CopyEditorWithDecompression (hash_dag_editors.h:437-616)
    NOISE_ADD = 8, NOISE_CARVE = 9  // SphereNoiseEditor(isAdding) (hash_dag_editors.h:762-809)
};

// FastNoise's 3D Perlin fractal (FN_DECIMAL = float, quintic interpolation, FBM) as
// SphereNoiseEditor configures it: SetSeed, SetFrequency, SetFractalOctaves
// (FastNoise.cpp:195-237, 328-333, 760-797, 841-888; defaults FastNoise.h:222-230).
struct PerlinFbm {
    u8 perm[512], perm12[512];
    float frequency = 0.01f, lacunarity = 2.f, gain = 0.5f, bounding = 0.f;
    int octaves = 3;
    void set_seed(int seed);
    void set_octaves(int n);
    float single(u8 offset, float x, float y, float z) const;
    float fractal(float x, float y, float z) const;  // GetPerlinFractal
};
struct Edit {
    u32 kind;
    F3 center;
    float radius;
    F3 source;   // copy only
    F3 dest;     // copy only
    // copy only: Matrix3x3 transform (row-major D00 .. D22, FP64) and the swirl
    // (hash_dag_editors.h:332-379); identity / off by default
    double transform[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    bool swirl = false;
    float swirlPeriod = 1.f;
};
struct EditStats { u64 numNodes = 0, numVoxels = 0, leavesRebuilt = 0; };

struct UndoFrame { u32 root; std::vector<std::pair<u32, std::pair<ColorLeaf, ColorLeaf>>> leaves; };

// One edit on (dag, colors); colors may be null (EDITS_ENABLE_COLORS=0).
EditStats apply_edit(DAG& dag, HashDAGColors* colors, const Edit& e, UndoFrame* undo, int threads = 1);

// GC (hash_dag.h:407-547); deterministic sorted iteration instead of unordered_set order.
void remove_stale_nodes(DAG& dag, u32 maxLevel);

// Canonical DAG serialisation (parity definition for DAG identity)
std::vector<u32> canonicalize(const DAG& dag);

// DAGUtils (dag_utils.h:138-523) used by copy and tests
bool get_value(const DAG& dag, Path path);
CompressedColor get_color_at(const DAG& dag, const HashDAGColors& colors, Path path, bool* ok);
bool is_empty_region(const DAG& dag, u32 maxLevel, U3 start, U3 size);

}  // namespace orc
