"""Synthetic terrain scenes (seeded simplex heightfield shell + compressed colours).

The scene comes back in the reference's input formats: a BasicDAG word array
(basic_dag.h:11-45) and a colour tree with a shared compressed main leaf
(hash_dag_colors.h:9-198), ready for HashDAG.load (HashDAGFactory::load_from_DAG +
load_colors_from_DAG).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from ._lib import check, hd_terrain_params, lib, ptr


@dataclass
class TerrainParams:
    levels: int = 12
    footprint: int = 0            # 0 = full 2^levels extent
    seed: int = 1337
    octaves: int = 4
    amplitude: float = 0.3
    frequency: float = 4.0
    color_tree_levels: int = 10
    colors: bool = True
    color_seed: int = 7
    threads: int = 0

    def to_c(self) -> hd_terrain_params:
        return hd_terrain_params(self.levels, self.footprint, self.seed, self.octaves, self.amplitude, self.frequency,
                                 self.color_tree_levels, 1 if self.colors else 0, self.color_seed, self.threads)


@dataclass
class Scene:
    levels: int
    color_tree_levels: int
    dag: np.ndarray                    # uint32 BasicDAG words, root at 0
    color_nodes: np.ndarray            # uint32, 8 entries per colour-tree node
    color_offsets: np.ndarray          # uint64 shared-leaf offsets per handle
    blocks: np.ndarray                 # uint64 main-leaf blocks
    weights: np.ndarray                # uint32 main-leaf weight words (stored byte-swapped)
    macro: np.ndarray                  # uint64 main-leaf macro-block pairs
    num_voxels: int
    level_nodes: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint64))
    params: TerrainParams | None = None
    level_words: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint64))


def _cache_path(params: TerrainParams, cache_dir: str) -> str:
    key = (f"t{params.levels}_f{params.footprint}_s{params.seed}_o{params.octaves}_a{params.amplitude!r}"
           f"_q{params.frequency!r}_c{params.color_tree_levels}_{int(params.colors)}_k{params.color_seed}")
    return os.path.join(cache_dir, key + ".npz")


def generate_terrain(params: TerrainParams, cache_dir: str | None = None) -> Scene:
    """Builds the scene of `params`. cache_dir (default: $HD_SCENE_CACHE, unset = no cache)
    keeps generated scenes as .npz files keyed by every parameter that shapes the scene, so
    later processes (profiler passes of one benchmark) load instead of regenerating."""
    cache_dir = cache_dir if cache_dir is not None else os.environ.get("HD_SCENE_CACHE")
    if cache_dir:
        path = _cache_path(params, cache_dir)
        if os.path.exists(path):
            with np.load(path) as z:
                return Scene(params.levels, params.color_tree_levels, z["dag"], z["nodes"], z["offsets"],
                             z["blocks"], z["weights"], z["macro"], int(z["num_voxels"]), z["level_nodes"], params,
                             z["level_words"])
        sc = _generate(params)
        os.makedirs(cache_dir, exist_ok=True)
        tmp = f"{path}.{os.getpid()}.tmp.npz"
        np.savez(tmp, dag=sc.dag, nodes=sc.color_nodes, offsets=sc.color_offsets, blocks=sc.blocks,
                 weights=sc.weights, macro=sc.macro, num_voxels=np.uint64(sc.num_voxels),
                 level_nodes=sc.level_nodes, level_words=sc.level_words)
        os.replace(tmp, path)
        return sc
    return _generate(params)


def _generate(params: TerrainParams) -> Scene:
    p = params.to_c()
    handle = C.c_void_p()
    check(lib.hd_scene_generate(C.byref(p), C.byref(handle)))
    try:
        sizes = np.zeros(7, np.uint64)
        check(lib.hd_scene_sizes(handle, ptr(sizes, C.c_uint64)))
        dag = np.zeros(int(sizes[0]), np.uint32)
        nodes = np.zeros(int(sizes[1]), np.uint32)
        offsets = np.zeros(int(sizes[2]), np.uint64)
        blocks = np.zeros(int(sizes[3]), np.uint64)
        weights = np.zeros(int(sizes[4]), np.uint32)
        macro = np.zeros(int(sizes[5]), np.uint64)
        check(lib.hd_scene_copy(handle, ptr(dag), ptr(nodes), ptr(offsets, C.c_uint64), ptr(blocks, C.c_uint64),
                                ptr(weights), ptr(macro, C.c_uint64)))
        per_level = np.zeros(2 * (params.levels - 1), np.uint64)
        check(lib.hd_scene_level_nodes(handle, ptr(per_level, C.c_uint64)))
    finally:
        lib.hd_scene_free(handle)
    n = params.levels - 1
    return Scene(params.levels, params.color_tree_levels, dag, nodes, offsets, blocks, weights, macro,
                 int(sizes[6]), per_level[:n].copy(), params, per_level[n:].copy())


def terrain_height(params: TerrainParams, x: int, z: int) -> int:
    p = params.to_c()
    return int(lib.hd_terrain_height(C.byref(p), x, z))


def dense_from_heights(params: TerrainParams) -> np.ndarray:
    """Dense [x][y][z] occupancy of the terrain shell (small depths only), from the heights."""
    E = 1 << params.levels
    F = params.footprint or E
    h = np.full((E + 2, E + 2), -1, np.int64)
    for x in range(F):
        for z in range(F):
            h[x + 1, z + 1] = terrain_height(params, x, z)
    grid = np.zeros((E, E, E), np.uint8)
    for x in range(F):
        for z in range(F):
            hh = h[x + 1, z + 1]
            nb = [h[x, z + 1], h[x + 2, z + 1], h[x + 1, z], h[x + 1, z + 2]]
            nb = [v for v in nb if v >= 0]
            lo = min(hh, 1 + min(nb)) if nb else hh
            grid[x, lo:hh + 1, z] = 1
    return grid
