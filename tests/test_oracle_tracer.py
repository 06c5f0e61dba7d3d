"""Pins the oracle tracer (tracer.cu:7-697) with independent brute force on a dense grid:

* primary hits: an Amanatides-Woo DDA through the dense occupancy grid gives the nearest
  voxel; the reference's DFS over slab masks (1e-4 epsilon, first intersected child in ray
  order) must agree on almost every pixel (ties on voxel edges may differ) and on every
  miss/hit decision away from grazing rays;
* colour indexing: the colour index trace_colors reaches must be the Morton rank of the hit
  voxel among the set voxels of its colour subtree, looked up through the colour tree;
* shadows: the any-hit shadow ray, marched by DDA from the same FP32 start point, decides
  lit / shadowed identically on almost every pixel.
"""
import ctypes as C
import math

import numpy as np
import pytest

from hashdag_amd import DAGInfo
from hashdag_amd.workload import cameras
from oracle import Oracle, OracleColors, OracleDAG, orc
from tests.scenes import POOL_PAGES, scene


def ray_dir(params, px, py):
    cam, rmin, ddx, ddy = params[0:3], params[3:6], params[6:9], params[9:12]
    v = [rmin[i] + px * ddx[i] + py * ddy[i] - cam[i] for i in range(3)]
    length = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    s = 1.0 / length
    return np.array([s * c for c in v], np.float64)


def dda_first(grid, o, d, any_hit=False, t_start=0.0):
    """First set voxel along o + t d (t >= t_start) in a dense grid; None if none."""
    E = grid.shape[0]
    inv = np.array([1.0 / c if c != 0 else math.inf for c in d])
    t0, t1 = t_start, math.inf
    for a in range(3):
        if d[a] == 0:
            if not (0 <= o[a] <= E):
                return None
            continue
        ta, tb = (0 - o[a]) * inv[a], (E - o[a]) * inv[a]
        t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
    if t0 > t1:
        return None
    p = o + d * (t0 + 1e-9)
    v = [min(E - 1, max(0, int(math.floor(p[a])))) for a in range(3)]
    step = [1 if d[a] > 0 else -1 for a in range(3)]
    tmax = []
    tdelta = []
    for a in range(3):
        if d[a] == 0:
            tmax.append(math.inf)
            tdelta.append(math.inf)
        else:
            nb = v[a] + (1 if d[a] > 0 else 0)
            tmax.append((nb - o[a]) * inv[a])
            tdelta.append(abs(inv[a]))
    while True:
        if grid[v[0], v[1], v[2]]:
            return tuple(v)
        a = int(np.argmin(tmax))
        v[a] += step[a]
        if not (0 <= v[a] < E):
            return None
        tmax[a] += tdelta[a]


@pytest.fixture(scope="module")
def setup():
    p, sc, g = scene("d7")
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    return p, sc, g, od, oc, od.dense()


def test_primary_hits_match_dda(setup):
    p, sc, g, od, oc, grid = setup
    W, H = 48, 32
    agree = total = 0
    for cam in cameras(p, 3, seed=21):
        prm = Oracle.ray_params(cam, DAGInfo.unit(g.levels), g.levels, W, H)
        paths = Oracle.trace_paths(od, prm, W, H)
        o = np.array([float(np.float32(c)) for c in prm[0:3]])
        for py in range(H):
            for px in range(W):
                d = ray_dir(prm, px, py).astype(np.float32).astype(np.float64)
                hit = dda_first(grid, o, d)
                got = tuple(int(c) for c in paths[H - 1 - py, px, :3])
                want = hit if hit is not None and hit != (0, 0, 0) else (0, 0, 0)
                total += 1
                agree += got == want
    assert agree / total > 0.98, f"only {agree}/{total} primary hits agree with the DDA"


def colour_handle(nodes, ctl, levels, v):
    idx = 0
    for level in range(1, ctl + 1):
        s = levels - level
        c = (((v[0] >> s) & 1) << 2) | (((v[1] >> s) & 1) << 1) | ((v[2] >> s) & 1)
        idx = int(nodes[idx + c])
        if level < ctl and idx == 0:
            return None
    return idx


def morton_rank(grid, v, levels, ctl):
    s = levels - ctl
    base = [c >> s << s for c in v]
    n = 1 << s
    sub = grid[base[0]:base[0] + n, base[1]:base[1] + n, base[2]:base[2] + n]
    xs, ys, zs = np.nonzero(sub)

    def key(x, y, z):
        k = 0
        for b in range(s - 1, -1, -1):
            k = (k << 3) | (((x >> b) & 1) << 2) | (((y >> b) & 1) << 1) | ((z >> b) & 1)
        return k

    mine = key(v[0] - base[0], v[1] - base[1], v[2] - base[2])
    return sum(1 for x, y, z in zip(xs, ys, zs) if key(x, y, z) < mine)


def to_rgb888(r, g, b):
    f = np.float32
    vals = [min(max(f(c), f(0)), f(1)) for c in (r, g, b)]
    return int(vals[0] * f(255)) | (int(vals[1] * f(255)) << 8) | (int(vals[2] * f(255)) << 16) | 0xFF000000


def colour_rgb(colorBits, weight, bpw):
    f = np.float32
    if bpw == 0:
        return to_rgb888(f(colorBits & 0x3FF) / f(1023), f((colorBits >> 10) & 0xFFF) / f(4095),
                         f((colorBits >> 22) & 0x3FF) / f(1023))
    wf = f(weight) / f((1 << bpw) - 1)
    lo = [f(colorBits & 0x1F) / f(31), f((colorBits >> 5) & 0x3F) / f(63), f((colorBits >> 11) & 0x1F) / f(31)]
    hi = [f((colorBits >> 16) & 0x1F) / f(31), f((colorBits >> 21) & 0x3F) / f(63), f((colorBits >> 27) & 0x1F) / f(31)]
    return to_rgb888(*[a * (f(1) - wf) + b * wf for a, b in zip(lo, hi)])


def test_colour_index_is_morton_rank(setup):
    p, sc, g, od, oc, grid = setup
    W, H = 40, 30
    cam = cameras(p, 1, seed=23)[0]
    prm = Oracle.ray_params(cam, DAGInfo.unit(g.levels), g.levels, W, H)
    paths = Oracle.trace_paths(od, prm, W, H)
    cols = Oracle.trace_colors(od, oc, paths, W, H)
    out = np.zeros(2, np.uint32)
    checked = 0
    for y in range(H):
        for x in range(W):
            v = tuple(int(c) for c in paths[y, x, :3])
            if v == (0, 0, 0):
                continue
            h = colour_handle(sc.color_nodes, g.color_tree_levels, g.levels, v)
            rank = morton_rank(grid, v, g.levels, g.color_tree_levels)
            off = int(sc.color_offsets[h])
            orc.orc_decode_color(sc.blocks.ctypes.data_as(C.POINTER(C.c_uint64)), sc.blocks.size,
                                 sc.weights.ctypes.data_as(C.POINTER(C.c_uint32)), sc.weights.size,
                                 sc.macro.ctypes.data_as(C.POINTER(C.c_uint64)), sc.macro.size, off, rank,
                                 out.ctypes.data_as(C.POINTER(C.c_uint32)))
            want = colour_rgb(int(out[0]), int(out[1]) & 0xFF, int(out[1]) >> 8)
            assert int(cols[y, x]) == want, (x, y, v)
            checked += 1
    assert checked > 100


def test_shadow_decisions_match_dda(setup):
    p, sc, g, od, oc, grid = setup
    W, H = 40, 30
    f = np.float32
    sun = np.array([f(0.3), f(1.0), f(0.5)], np.float32)
    n = f(1.0) / np.sqrt(np.float32(sun[0] * sun[0] + sun[1] * sun[1] + sun[2] * sun[2]))
    sun = (n * sun).astype(np.float32)
    agree = total = 0
    for cam in cameras(p, 2, seed=29):
        prm = Oracle.ray_params(cam, DAGInfo.unit(g.levels), g.levels, W, H)
        paths = Oracle.trace_paths(od, prm, W, H)
        white = np.full((H, W), 0xFFFFFFFF, np.uint32)
        shaded = Oracle.trace_shadows(od, prm, paths, white, W, H, 1.0, 0.0)
        for y in range(H):
            for x in range(W):
                v = np.array(paths[y, x, :3], np.float64)
                if not v.any():
                    continue
                cd = ray_dir(prm, x, H - 1 - y)
                cam_p = prm[0:3]
                t0 = (v - cam_p) / cd
                t1 = (v + 1 - cam_p) / cd
                tm = np.max(np.minimum(t0, t1))
                hit = (cam_p + cd * tm).astype(np.float32)
                start = (hit + f(1.0) * sun).astype(np.float64)
                blocked = dda_first(grid, start, sun.astype(np.float64)) is not None
                lit = (int(shaded[y, x]) & 0xFF) == 255
                total += 1
                agree += blocked != lit
    assert agree / total > 0.97, f"only {agree}/{total} shadow decisions agree with the DDA"


def test_shadow_result_independent_of_child_order(setup):
    """The any-hit walk's result does not depend on the child order: the reference's
    highest-bit-first order and the device's front-to-back order shade every pixel the same,
    while front to back reads fewer bytes on this terrain (shadowed rays stop at the nearest
    occluder)."""
    p, sc, g, od, oc, grid = setup
    W, H = 64, 48
    bytes_by_order = []
    for cam in cameras(p, 3, seed=31):
        prm = Oracle.ray_params(cam, DAGInfo.unit(g.levels), g.levels, W, H)
        paths = Oracle.trace_paths(od, prm, W, H)
        cols = Oracle.trace_colors(od, oc, paths, W, H)
        outs = []
        for ftb in (False, True):
            Oracle.set_any_hit_front_to_back(ftb)
            try:
                out, b = Oracle.trace_shadows(od, prm, paths, cols, W, H, 1.0, 0.0, count=True)
            finally:
                Oracle.set_any_hit_front_to_back(False)
            outs.append(out)
            bytes_by_order.append(b)
        np.testing.assert_array_equal(outs[0], outs[1])
    assert sum(bytes_by_order[1::2]) < sum(bytes_by_order[0::2])
