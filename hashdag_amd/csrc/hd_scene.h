// Synthetic scene generation: seeded simplex-noise heightfield terrain voxelised as a
// watertight shell, built directly into the reference's input formats (a BasicDAG node
// array, basic_dag.h:11-45, plus a colour tree and a shared compressed colour leaf,
// hash_dag_colors.h:9-198 / variable_weight_size_colors.h:157-555).
#pragma once

#include <cstdint>
#include <vector>

namespace hd {

struct TerrainParams {
    uint32_t levels = 12;           // scene depth D (extent 2^D)
    uint32_t footprint = 0;         // terrain footprint in voxels along x and z (0 = 2^D)
    uint64_t seed = 1337;           // permutation seed
    uint32_t octaves = 4;
    double amplitude = 0.3;         // relative to the footprint
    double frequency = 4.0;         // noise periods across the footprint
    uint32_t color_tree_levels = 10;
    bool colors = true;             // build colours (needs leaf level > color_tree_levels)
    uint32_t color_seed = 7;
    uint32_t threads = 0;           // 0 = hardware concurrency
};

struct Scene {
    uint32_t levels = 0;
    uint32_t color_tree_levels = 10;
    std::vector<uint32_t> dag;       // BasicDAG words; root at 0; header = mask | count << 8 at
                                     // levels >= color_tree_levels, mask only above
    std::vector<uint64_t> level_nodes;  // unique nodes per level
    std::vector<uint64_t> level_words;  // BasicDAG words per level
    uint64_t num_voxels = 0;
    // colour tree (empty when colours are off: a single all-zero root node is used)
    std::vector<uint32_t> color_nodes;
    std::vector<uint64_t> color_offsets;
    std::vector<uint64_t> blocks;
    std::vector<uint32_t> weights;   // stored byte-swapped, as in device memory
    std::vector<uint64_t> macro;
};

// Heights of the terrain (used by tests and by surface-point sampling for edits).
int32_t terrain_height(const TerrainParams& p, uint32_t x, uint32_t z);

Scene generate_terrain(const TerrainParams& p);

}  // namespace hd
