"""Small seeded test scenes shared by the CPU and GPU tests (cached per process)."""
from __future__ import annotations

import functools

from hashdag_amd import Geometry, TerrainParams, generate_terrain

# (levels, footprint, color_tree_levels): small geometries keep the virtual space small
SMALL = {
    "d7": (7, 128, 3),
    "d8": (8, 256, 4),
    "d8s": (8, 64, 4),
    "d10": (10, 512, 6),
    "d10w": (10, 512, 3),   # colour subtrees of 128^3: leaves of several 16384-colour macro blocks
    "d13": (13, 256, 10),   # reference colour-tree depth (10) with a small footprint
}


def geometry_for(levels: int, ctl: int, small: bool = True) -> Geometry:
    if not small:
        return Geometry(levels=levels, color_tree_levels=ctl)
    top = min(9, max(1, levels - 3))
    return Geometry(levels=levels, num_top_levels=top, bucket_bits_top=6, bucket_bits_low=10,
                    bucket_size_top=1024, bucket_size_low=2048, color_tree_levels=ctl)


@functools.lru_cache(maxsize=None)
def scene(name: str, seed: int = 1337):
    levels, footprint, ctl = SMALL[name]
    p = TerrainParams(levels=levels, footprint=footprint, color_tree_levels=ctl, colors=True, seed=seed, threads=4)
    return p, generate_terrain(p), geometry_for(levels, ctl)


POOL_PAGES = 1 << 14
