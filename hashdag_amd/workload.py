"""Seeded workloads over a terrain scene: cameras above the terrain and surface points.

Cameras: xz uniform in the middle 80 % of the terrain footprint, height 5-20 % of the
footprint above the terrain surface, yaw uniform in [0, 2 pi), pitch uniform in
[-45, -5] degrees. Surface points: uniform (x, z), y = terrain height.
"""
from __future__ import annotations

import math

import numpy as np

from .scene import TerrainParams, terrain_height
from .tracer import CameraView


def cameras(params: TerrainParams, n: int, seed: int = 1) -> list[CameraView]:
    rng = np.random.default_rng(seed)
    F = params.footprint or (1 << params.levels)
    out = []
    for _ in range(n):
        x = rng.uniform(0.1 * F, 0.9 * F)
        z = rng.uniform(0.1 * F, 0.9 * F)
        h = terrain_height(params, int(x), int(z))
        y = h + rng.uniform(0.05, 0.20) * F
        yaw = rng.uniform(0.0, 2.0 * math.pi)
        pitch = math.radians(rng.uniform(-45.0, -5.0))
        out.append(CameraView.look((x, y, z), yaw, pitch))
    return out


def surface_points(params: TerrainParams, n: int, seed: int = 2) -> np.ndarray:
    rng = np.random.default_rng(seed)
    F = params.footprint or (1 << params.levels)
    pts = np.zeros((n, 3), np.float32)
    for i in range(n):
        x = int(rng.integers(0, F))
        z = int(rng.integers(0, F))
        pts[i] = (x, terrain_height(params, x, z), z)
    return pts


# editor kinds of the C ABI (hd_edit.kind)
SPHERE_ADD, SPHERE_CARVE, BOX_ADD, BOX_CARVE, SPHERE_PAINT, COPY = range(6)
FILL, COPY_DEC = 6, 7


def _centre(rng, params: TerrainParams):
    F = params.footprint or (1 << params.levels)
    x = int(rng.integers(0, F))
    z = int(rng.integers(0, F))
    return (float(x), float(terrain_height(params, x, z)), float(z))


def sphere_edits(params: TerrainParams, n: int, seed: int = 3, rmin: float = 8.0, rmax: float = 64.0):
    """n sphere edits, alternating add / carve, radius uniform in [rmin, rmax], centred at
    seeded integer surface points: [(kind, (x, y, z), radius), ...]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        c = _centre(rng, params)
        r = float(np.float32(rng.uniform(rmin, rmax)))
        out.append((SPHERE_ADD if i % 2 == 0 else SPHERE_CARVE, c, r))
    return out


def mixed_edits(params: TerrainParams, n: int, seed: int = 4, rmin: float = 2.0, rmax: float = 256.0,
                mix=(0.7, 0.2, 0.1)):
    """n mixed edits: fractions mix = (sphere add/carve, box add/carve, sphere paint), add and
    carve equally likely, radius log-uniform in [rmin, rmax], seeded integer surface centres."""
    rng = np.random.default_rng(seed)
    p = np.array(mix, np.float64) / sum(mix)
    out = []
    for _ in range(n):
        c = _centre(rng, params)
        r = float(np.float32(np.exp(rng.uniform(np.log(rmin), np.log(rmax)))))
        cls = int(rng.choice(3, p=p))
        add = bool(rng.integers(0, 2))
        if cls == 0:
            kind = SPHERE_ADD if add else SPHERE_CARVE
        elif cls == 1:
            kind = BOX_ADD if add else BOX_CARVE
        else:
            kind = SPHERE_PAINT
        out.append((kind, c, r))
    return out


def instances(params: TerrainParams, n: int, seed: int = 6, hmin: float = 16.0, hmax: float = 512.0):
    """n instanced objects standing on the terrain: boxes and spheres (equally likely) of
    height log-uniform in [hmin, hmax], centred one radius above a seeded surface point."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        x, y, z = _centre(rng, params)
        h = float(np.exp(rng.uniform(np.log(hmin), np.log(hmax))))
        r = float(np.float32(h / 2))
        kind = BOX_ADD if rng.integers(0, 2) else SPHERE_ADD
        out.append((kind, (x, float(np.float32(y + r)), z), r))
    return out


def copy_edits(params: TerrainParams, n: int, seed: int = 7, radius: float = 256.0, decompress: bool = False):
    """n copies of the (2 radius)^3 region around a seeded surface point to another seeded
    surface point: [(COPY, source, radius, source, dest), ...] (translation only); with
    decompress the kind is COPY_DEC (CopyEditorWithDecompression)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        src = _centre(rng, params)
        dst = _centre(rng, params)
        out.append((COPY_DEC if decompress else COPY, src, float(radius), src, dst))
    return out


def carves(params: TerrainParams, n: int, seed: int = 8, rmin: float = 512.0, rmax: float = 1024.0):
    """n sphere carves of radius uniform in [rmin, rmax] at seeded surface points."""
    rng = np.random.default_rng(seed)
    return [(SPHERE_CARVE, _centre(rng, params), float(np.float32(rng.uniform(rmin, rmax)))) for _ in range(n)]
