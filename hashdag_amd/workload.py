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
