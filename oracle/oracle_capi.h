/* Oracle C API (ctypes) over the CPU restatement in oracle_core.hpp.
 *
 * TEST INFRASTRUCTURE ONLY: called from tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg, never from the product.
 *
 * Parity pinning of the oracle (which independent check fixes each part):
 *   hashes            MurmurHash3_x86_32 / x64_128 published vectors (fmix32, fmix64, the
 *                     word-count finaliser relation)            tests/test_oracle_hashes.py
 *   leaf masks, Path  decoded grid and get_value == the grid the terrain heights define
 *                                                                tests/test_oracle_table.py
 *   colour codec      decode(encode(seq)) == seq, bit-stream layout vs a direct big-endian
 *                     bit-stream model                          tests/test_oracle_colors.py
 *   hash table        every node in the bucket its (published-vector) hash selects, no page
 *                     straddle, children in the child level's range, full nodes
 *                                                                tests/test_oracle_table.py
 *   load_from_DAG     decoded voxel set == dense generator grid tests/test_oracle_table.py
 *   trace_paths       brute-force DDA nearest hit agrees except where the reference's DFS
 *                     order/epsilon differ (counted), mask properties, root rejection
 *                                                                tests/test_oracle_tracer.py
 *                     exact hit order: a reference-order DFS over the dense occupancy
 *                     pyramid with exact FP64 slab tests gives the same voxel on every ray
 *                     away from a decision boundary              tests/test_oracle_pins.py
 *   trace_colors      colour index == Morton rank of the voxel in its level-10 subtree
 *                     (dense-grid brute force)                   tests/test_oracle_tracer.py
 *                     sky / magenta / checker fallbacks in float32 tests/test_oracle_pins.py
 *   encoder bytes     blocks, weight words and macro words == a Python encoder written from
 *                     the format                                 tests/test_oracle_pins.py
 *   paint colours     the quantised blend of each touched voxel's old colour, written out
 *                     in float32                                 tests/test_oracle_pins.py
 *   trace_shadows     any-hit result == brute-force segment/voxel test on a dense grid;
 *                     the same result with either any-hit child order  tests/test_oracle_tracer.py
 *   edits             dense-grid brute force of sphere/box/paint voxel sets and colours, and of
 *                     copies (translations, and transforms that permute / mirror / scale
 *                     axes exactly)                              tests/test_oracle_edits.py
 *   fill              dense-grid flood fill with the buffer's loaded set, bounds and colour
 *                     scan written out independently                tests/test_oracle_edits.py
 *   copy swirl        dense-grid brute force turning each offset about y in FP64, on every
 *                     voxel whose source coordinate is 1e-3 or more from an integer
 *                                                                tests/test_oracle_pins.py
 *   GC                canonical form unchanged, idempotent, used words == reachable words +
 *                     full node + padding        tests/test_oracle_table.py, test_oracle_edits.py
 *   interior find     the walk rule (start + n < min(bucket size, page end)) on synthetic
 *                     buckets whose nodes end at page and bucket ends, against a walk written
 *                     from hash_table.h:276-357                   tests/test_oracle_pins.py
 *   decompressing copy dense-grid brute force with the buffer mapping and the empty-source
 *                     checks                                     tests/test_oracle_copy_dec.py
 *   noise editor      FastNoise's Perlin fractal: the seeded permutation against std::mt19937_64
 *                     (pinned by the standard's 10000th value), zero at lattice points, the
 *                     selected gradient next to them, continuity across cell faces, the bounded
 *                     octave sum; the editor against a dense-grid brute force
 *                                                                tests/test_oracle_noise.py
 *   fog (density > 0) a density that saturates the fog: every pixel its fog colour, the sky
 *                     colour exactly away from the sun; a density with exp(..) = 1/2 at a
 *                     pixel: the midpoint of its fog-free and fog colours  tests/test_oracle_pins.py
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_geometry {
    uint32_t levels, page_size, num_top_levels, bucket_bits_top, bucket_bits_low, bucket_size_top, bucket_size_low,
        color_tree_levels;
} orc_geometry;

const char* orc_last_error(void);

/* hashes */
uint32_t orc_murmurhash32(uint32_t h);
uint64_t orc_murmurhash64(uint64_t h);
uint32_t orc_murmurhash32xN(const uint32_t* words, uint64_t n, uint32_t seed);

/* DAG */
void* orc_dag_from_basic(const orc_geometry* g, uint32_t pool_pages, const uint32_t* data, uint64_t n);
void* orc_dag_from_image(const orc_geometry* g, uint32_t pool_pages, const uint32_t* pool, const uint32_t* page_table,
                         const uint32_t* bucket_sizes, const uint32_t* full_nodes, uint32_t pool_top, uint32_t root);
void orc_dag_free(void* dag);
uint32_t orc_dag_root(void* dag);
uint32_t orc_dag_pool_top(void* dag);
/* copies the reference-layout image out; arrays sized per geometry / pool_pages */
int orc_dag_export(void* dag, uint32_t* pool, uint32_t* page_table, uint32_t* bucket_sizes, uint32_t* full_nodes);
uint64_t orc_canonical(void* dag, uint32_t* out, uint64_t cap); /* returns required length */
int orc_get_value(void* dag, uint32_t x, uint32_t y, uint32_t z);
/* dense occupancy of the whole 2^levels cube (small depths only), x-major [x][y][z] */
int orc_dense(void* dag, uint8_t* out);

/* colours: tree nodes, shared offsets and the main leaf */
void* orc_colors_new(const uint32_t* nodes, uint64_t n_nodes, const uint64_t* offsets, uint64_t n_offsets,
                     const uint64_t* blocks, uint64_t n_blocks, const uint32_t* weights, uint64_t n_weights,
                     const uint64_t* macro, uint64_t n_macro);
/* appends a unique leaf (weights in stored byte-swapped order); returns its handle (bit 31 set) */
uint32_t orc_colors_add_unique(void* colors, const uint64_t* blocks, uint64_t n_blocks, const uint32_t* weights,
                               uint64_t n_weights, const uint64_t* macro, uint64_t n_macro);
/* appends a leaf slot holding the shared main leaf at `offset` (what undo restores into a
 * slot that was shared before the edit, hash_dag.h:81-88); returns its handle */
uint32_t orc_colors_add_shared(void* colors, uint64_t offset);
void orc_colors_free(void* colors);
uint64_t orc_colors_num_nodes(void* colors);
uint64_t orc_colors_num_leaves(void* colors);
int orc_colors_nodes(void* colors, uint32_t* out);
/* sizes / copy of unique leaf i: sizes[0..2] = n_blocks, n_weights, n_macro */
int orc_colors_leaf_sizes(void* colors, uint32_t i, uint64_t sizes[3]);
int orc_colors_leaf_copy(void* colors, uint32_t i, uint64_t* blocks, uint32_t* weights, uint64_t* macro);
/* decoded colour sequence per level-(color_tree_levels) subtree in DFS order: for each subtree
 * [x, y, z, n, then n x (colorBits, weight | bpw << 8)]; returns required length */
uint64_t orc_color_canonical(void* dag, void* colors, uint32_t* out, uint64_t cap);
/* raw leaf words per level-(color_tree_levels) subtree in DFS order: [x, y, z, kind], then for
 * kind 1 (unique) n_blocks, n_weights, n_macro and the stored words, for kind 2 (shared) the
 * offset as two words; returns required length */
uint64_t orc_color_leaf_bytes(void* dag, void* colors, uint32_t* out, uint64_t cap);

/* colour codec round trip: encode n colours (colorBits[i], weight[i], bpw[i]) with the
 * reference encoder; sizes out = n_blocks, n_weights, n_macro */
int orc_encode_colors(const uint32_t* colorBits, const uint8_t* weight, const uint8_t* bpw, uint64_t n,
                      uint64_t* blocks, uint32_t* weights, uint64_t* macro, uint64_t sizes[3], uint64_t cap_blocks,
                      uint64_t cap_weights, uint64_t cap_macro);
/* decode colour index i of a leaf given as arrays; out = colorBits, weight | bpw << 8 */
int orc_decode_color(const uint64_t* blocks, uint64_t n_blocks, const uint32_t* weights, uint64_t n_weights,
                     const uint64_t* macro, uint64_t n_macro, uint64_t offset, uint64_t index, uint32_t out[2]);

/* tracer; params = 12 doubles: cameraPosition, rayMin, rayDDx, rayDDy */
void orc_ray_params(const double position[3], const double rotation[9], const double bmin[3], const double bmax[3],
                    uint32_t levels, uint32_t W, uint32_t H, double out[12]);
/* rows [y0, y1) of the W x H paths image (uint4 per pixel, row r = pixel y H-1-r) */
int orc_trace_paths(void* dag, const double params[12], uint32_t W, uint32_t H, uint32_t y0, uint32_t y1,
                    uint32_t* paths, uint64_t* bytes, int threads);
int orc_trace_colors(void* dag, void* colors, uint32_t W, uint32_t H, uint32_t y0, uint32_t y1, const uint32_t* paths,
                     uint32_t* colors_out, int threads);
/* child order of the any-hit (shadow) walk: 0 = the reference's highest-bit-first order
 * (default), 1 = front to back along the ray. The shadow result is the same either way; the
 * bytes read differ. Process-wide. */
void orc_set_any_hit_order(int front_to_back);
int orc_trace_shadows(void* dag, const double params[12], uint32_t W, uint32_t H, uint32_t y0, uint32_t y1,
                      const uint32_t* paths, uint32_t* colors_inout, float bias, float fog, uint64_t* bytes,
                      int threads);

/* edits: kind 0 sphere add, 1 sphere carve, 2 box add, 3 box carve, 4 sphere paint, 6 fill.
 * stats out: num nodes, num voxels, colour leaves rebuilt. colors may be NULL. */
int orc_edit(void* dag, void* colors, uint32_t kind, const float center[3], float radius, uint64_t stats[3]);
/* the same with the edit tasks (one per touched colour subtree, hash_dag.h:330-356) run on
 * `threads` threads under striped bucket locks; canonically the same result */
int orc_edit_threads(void* dag, void* colors, uint32_t kind, const float center[3], float radius, int threads,
                     uint64_t stats[3]);
/* copy editor (CopyEditorWithoutDecompression, identity transform): box of `radius` around
 * dest + (center - source), filled from source + (p - dest) of the DAG before the edit */
int orc_edit_copy(void* dag, void* colors, const float center[3], float radius, const float source[3],
                  const float dest[3], uint64_t stats[3]);
/* kind 5 = the copy above, 7 = CopyEditorWithDecompression (hash_dag_editors.h:437-616): the
 * source box read through a voxel buffer of 2 ceil(radius) per axis from truncate(center -
 * radius), nodes above levels - 4 whose source region is empty left unedited */
int orc_edit_copy_kind(void* dag, void* colors, uint32_t kind, const float center[3], float radius,
                       const float source[3], const float dest[3], uint64_t stats[3]);
/* the same with the copy's Matrix3x3 transform (row-major, FP64) and swirl (swirl != 0: the
 * relative position turns about y by angle y / swirl_period, hash_dag_editors.h:364-379) */
int orc_edit_copy_tf(void* dag, void* colors, const float center[3], float radius, const float source[3],
                     const float dest[3], const double transform[9], int swirl, float swirl_period, uint64_t stats[3]);
/* DAGUtils::get_color at a voxel (out: colorBits, weight | bpw << 8); returns 1 if the voxel is set */
int orc_get_color(void* dag, void* colors, uint32_t x, uint32_t y, uint32_t z, uint32_t out[2]);
int orc_gc(void* dag, uint32_t max_level);
/* distinct node addresses reachable from the root, per level 0..max_level */
int orc_dag_reachable(void* dag, uint32_t max_level, uint64_t* out);
/* the reference's in-bucket find on one bucket (hash_table.h:196-357): address or 0xFFFFFFFF */
uint32_t orc_find_in_bucket(void* dag, uint32_t level, uint32_t bucket, const uint32_t* words, uint32_t n);
/* FastNoise::GetPerlinFractal (FastNoise.cpp:760-797) with SetSeed(seed), SetFrequency(frequency),
 * SetFractalOctaves(octaves) at n points xyz[3 * i]; octaves = 1 with frequency 1 gives
 * SinglePerlin(perm[0], x, y, z). perm (optional, 1024 bytes): m_perm then m_perm12. */
int orc_perlin(int seed, float frequency, int octaves, uint32_t n, const float* xyz, float* out, uint8_t* perm);

#ifdef __cplusplus
}
#endif
