"""ctypes wrapper of liboracle.so (the oracle C API in oracle_capi.h). Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
u8p = C.POINTER(C.c_uint8)
dblp = C.POINTER(C.c_double)
fltp = C.POINTER(C.c_float)


class orc_geometry(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("levels", "page_size", "num_top_levels", "bucket_bits_top",
                                           "bucket_bits_low", "bucket_size_top", "bucket_size_low",
                                           "color_tree_levels")]


_SIGS = {
    "orc_last_error": (C.c_char_p, []),
    "orc_murmurhash32": (C.c_uint32, [C.c_uint32]),
    "orc_murmurhash64": (C.c_uint64, [C.c_uint64]),
    "orc_murmurhash32xN": (C.c_uint32, [u32p, C.c_uint64, C.c_uint32]),
    "orc_dag_from_basic": (C.c_void_p, [C.POINTER(orc_geometry), C.c_uint32, u32p, C.c_uint64]),
    "orc_dag_from_image": (C.c_void_p, [C.POINTER(orc_geometry), C.c_uint32, u32p, u32p, u32p, u32p, C.c_uint32,
                                        C.c_uint32]),
    "orc_dag_free": (None, [C.c_void_p]),
    "orc_dag_root": (C.c_uint32, [C.c_void_p]),
    "orc_dag_pool_top": (C.c_uint32, [C.c_void_p]),
    "orc_dag_export": (C.c_int, [C.c_void_p, u32p, u32p, u32p, u32p]),
    "orc_canonical": (C.c_uint64, [C.c_void_p, u32p, C.c_uint64]),
    "orc_get_value": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]),
    "orc_dense": (C.c_int, [C.c_void_p, u8p]),
    "orc_colors_new": (C.c_void_p, [u32p, C.c_uint64, u64p, C.c_uint64, u64p, C.c_uint64, u32p, C.c_uint64, u64p,
                                    C.c_uint64]),
    "orc_colors_add_unique": (C.c_uint32, [C.c_void_p, u64p, C.c_uint64, u32p, C.c_uint64, u64p, C.c_uint64]),
    "orc_colors_add_shared": (C.c_uint32, [C.c_void_p, C.c_uint64]),
    "orc_colors_free": (None, [C.c_void_p]),
    "orc_colors_num_nodes": (C.c_uint64, [C.c_void_p]),
    "orc_colors_num_leaves": (C.c_uint64, [C.c_void_p]),
    "orc_colors_nodes": (C.c_int, [C.c_void_p, u32p]),
    "orc_colors_leaf_sizes": (C.c_int, [C.c_void_p, C.c_uint32, u64p]),
    "orc_colors_leaf_copy": (C.c_int, [C.c_void_p, C.c_uint32, u64p, u32p, u64p]),
    "orc_color_canonical": (C.c_uint64, [C.c_void_p, C.c_void_p, u32p, C.c_uint64]),
    "orc_color_leaf_bytes": (C.c_uint64, [C.c_void_p, C.c_void_p, u32p, C.c_uint64]),
    "orc_encode_colors": (C.c_int, [u32p, u8p, u8p, C.c_uint64, u64p, u32p, u64p, u64p, C.c_uint64, C.c_uint64,
                                    C.c_uint64]),
    "orc_decode_color": (C.c_int, [u64p, C.c_uint64, u32p, C.c_uint64, u64p, C.c_uint64, C.c_uint64, C.c_uint64,
                                   u32p]),
    "orc_ray_params": (None, [dblp, dblp, dblp, dblp, C.c_uint32, C.c_uint32, C.c_uint32, dblp]),
    "orc_trace_paths": (C.c_int, [C.c_void_p, dblp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, u32p, u64p,
                                  C.c_int]),
    "orc_trace_colors": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, u32p,
                                   u32p, C.c_int]),
    "orc_set_any_hit_order": (None, [C.c_int]),
    "orc_trace_shadows": (C.c_int, [C.c_void_p, dblp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, u32p, u32p,
                                    C.c_float, C.c_float, u64p, C.c_int]),
    "orc_edit": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, fltp, C.c_float, u64p]),
    "orc_edit_threads": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, fltp, C.c_float, C.c_int, u64p]),
    "orc_gc": (C.c_int, [C.c_void_p, C.c_uint32]),
    "orc_edit_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p]),
    "orc_edit_copy_kind": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_float, C.c_void_p,
                                     C.c_void_p, C.c_void_p]),
    "orc_edit_copy_tf": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_int, C.c_float, C.c_void_p]),
    "orc_get_color": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]),
    "orc_dag_reachable": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "orc_find_in_bucket": (C.c_uint32, [C.c_void_p, C.c_uint32, C.c_uint32, u32p, C.c_uint32]),
    "orc_perlin": (C.c_int, [C.c_int, C.c_float, C.c_int, C.c_uint32, fltp, fltp, u8p]),
}


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE, "-j8"], check=True)
    return LIB_PATH


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


orc = _load()


def _p(a, t=C.c_uint32):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def _err():
    return orc.orc_last_error().decode()


def geom(g) -> orc_geometry:
    return orc_geometry(g.levels, g.page_size, g.num_top_levels, g.bucket_bits_top, g.bucket_bits_low,
                        g.bucket_size_top, g.bucket_size_low, g.color_tree_levels)


class OracleDAG:
    def __init__(self, handle, g, pool_pages):
        self.h = handle
        self.geometry = g
        self.pool_pages = pool_pages

    @staticmethod
    def from_basic(g, words: np.ndarray, pool_pages: int) -> "OracleDAG":
        words = np.ascontiguousarray(words, np.uint32)
        gg = geom(g)
        h = orc.orc_dag_from_basic(C.byref(gg), pool_pages, _p(words), words.size)
        if not h:
            raise RuntimeError(_err())
        return OracleDAG(h, g, pool_pages)

    @staticmethod
    def from_image(img, pool_pages: int) -> "OracleDAG":
        gg = geom(img.geometry)
        pool = np.zeros(pool_pages * img.geometry.page_size, np.uint32)
        n = min(pool.size, img.pool.size)
        pool[:n] = img.pool[:n]
        h = orc.orc_dag_from_image(C.byref(gg), pool_pages, _p(pool), _p(np.ascontiguousarray(img.page_table)),
                                   _p(np.ascontiguousarray(img.bucket_sizes)),
                                   _p(np.ascontiguousarray(img.full_nodes)), img.pool_top, img.root)
        if not h:
            raise RuntimeError(_err())
        return OracleDAG(h, img.geometry, pool_pages)

    def __del__(self):
        if getattr(self, "h", None):
            orc.orc_dag_free(self.h)
            self.h = None

    @property
    def root(self) -> int:
        return orc.orc_dag_root(self.h)

    def export(self):
        g = self.geometry
        pool = np.zeros(self.pool_pages * g.page_size, np.uint32)
        pt = np.zeros(g.total_pages(), np.uint32)
        bs = np.zeros(g.total_buckets(), np.uint32)
        full = np.zeros(g.levels, np.uint32)
        orc.orc_dag_export(self.h, _p(pool), _p(pt), _p(bs), _p(full))
        return pool, pt, bs, full, orc.orc_dag_pool_top(self.h)

    def canonical(self) -> np.ndarray:
        n = orc.orc_canonical(self.h, None, 0)
        if n == 0:
            raise RuntimeError(_err())
        out = np.zeros(n, np.uint32)
        orc.orc_canonical(self.h, _p(out), n)
        return out

    def dense(self) -> np.ndarray:
        E = 1 << self.geometry.levels
        out = np.zeros((E, E, E), np.uint8)
        if orc.orc_dense(self.h, _p(out, C.c_uint8)) != 0:
            raise RuntimeError(_err())
        return out

    def get_value(self, x, y, z) -> bool:
        return bool(orc.orc_get_value(self.h, x, y, z))

    def edit(self, kind: int, center, radius: float, colors: "OracleColors | None" = None, threads: int = 1):
        c = (C.c_float * 3)(*center)
        stats = np.zeros(3, np.uint64)
        r = orc.orc_edit_threads(self.h, colors.h if colors else None, kind, c, radius, threads,
                                 _p(stats, C.c_uint64))
        if r != 0:
            raise RuntimeError(_err())
        return [int(v) for v in stats]

    def edit_copy(self, center, radius: float, source, dest, colors: "OracleColors", transform=None,
                  swirl_period=None, decompress=False):
        """CopyEditor; transform: 3x3 row-major matrix (FP64), swirl_period: None = no swirl;
        decompress: CopyEditorWithDecompression (no transform or swirl)."""
        c, s, d = (C.c_float * 3)(*center), (C.c_float * 3)(*source), (C.c_float * 3)(*dest)
        stats = np.zeros(3, np.uint64)
        if decompress:
            rc = orc.orc_edit_copy_kind(self.h, colors.h, 7, c, radius, s, d, stats.ctypes.data)
        elif transform is None and swirl_period is None:
            rc = orc.orc_edit_copy(self.h, colors.h, c, radius, s, d, stats.ctypes.data)
        else:
            t = np.ascontiguousarray(np.eye(3) if transform is None else transform, np.float64).reshape(9)
            rc = orc.orc_edit_copy_tf(self.h, colors.h, c, radius, s, d, t.ctypes.data, 0 if swirl_period is None else 1,
                                      1.0 if swirl_period is None else swirl_period, stats.ctypes.data)
        if rc != 0:
            raise RuntimeError(_err())
        return [int(v) for v in stats]

    def color(self, colors: "OracleColors", x: int, y: int, z: int):
        out = np.zeros(2, np.uint32)
        r = orc.orc_get_color(self.h, colors.h, x, y, z, out.ctypes.data)
        if r < 0:
            raise RuntimeError(_err())
        return bool(r), int(out[0]), int(out[1])

    def gc(self, max_level: int) -> None:
        if orc.orc_gc(self.h, max_level) != 0:
            raise RuntimeError(_err())

    def reachable(self, max_level: int) -> np.ndarray:
        out = np.zeros(max_level + 1, np.uint64)
        if orc.orc_dag_reachable(self.h, max_level, out.ctypes.data) != 0:
            raise RuntimeError(_err())
        return out


class OracleColors:
    def __init__(self, nodes, offsets, blocks, weights, macro):
        nodes = np.ascontiguousarray(nodes, np.uint32)
        offsets = np.ascontiguousarray(offsets, np.uint64)
        blocks = np.ascontiguousarray(blocks, np.uint64)
        weights = np.ascontiguousarray(weights, np.uint32)
        macro = np.ascontiguousarray(macro, np.uint64)
        self.h = orc.orc_colors_new(_p(nodes), nodes.size, _p(offsets, C.c_uint64), offsets.size,
                                    _p(blocks, C.c_uint64), blocks.size, _p(weights), weights.size,
                                    _p(macro, C.c_uint64), macro.size)

    @staticmethod
    def from_scene(scene) -> "OracleColors":
        return OracleColors(scene.color_nodes, scene.color_offsets, scene.blocks, scene.weights, scene.macro)

    def __del__(self):
        if getattr(self, "h", None):
            orc.orc_colors_free(self.h)
            self.h = None

    def nodes(self) -> np.ndarray:
        out = np.zeros(orc.orc_colors_num_nodes(self.h), np.uint32)
        orc.orc_colors_nodes(self.h, _p(out))
        return out

    def num_leaves(self) -> int:
        return orc.orc_colors_num_leaves(self.h)

    def leaf(self, i: int):
        sizes = np.zeros(3, np.uint64)
        orc.orc_colors_leaf_sizes(self.h, i, _p(sizes, C.c_uint64))
        b = np.zeros(int(sizes[0]), np.uint64)
        w = np.zeros(int(sizes[1]), np.uint32)
        m = np.zeros(int(sizes[2]), np.uint64)
        orc.orc_colors_leaf_copy(self.h, i, _p(b, C.c_uint64), _p(w), _p(m, C.c_uint64))
        return b, w, m

    def canonical(self, dag: OracleDAG) -> np.ndarray:
        n = orc.orc_color_canonical(dag.h, self.h, None, 0)
        out = np.zeros(n, np.uint32)
        if n:
            orc.orc_color_canonical(dag.h, self.h, _p(out), n)
        return out

    def leaf_bytes(self, dag: OracleDAG) -> np.ndarray:
        """The stored words of every colour subtree's leaf in DFS order (encoder byte choices
        included), see orc_color_leaf_bytes."""
        n = orc.orc_color_leaf_bytes(dag.h, self.h, None, 0)
        out = np.zeros(n, np.uint32)
        if n:
            orc.orc_color_leaf_bytes(dag.h, self.h, _p(out), n)
        return out


class Oracle:
    """Stateless oracle entry points."""

    @staticmethod
    def ray_params(camera, info, levels, W, H) -> np.ndarray:
        out = np.zeros(12, np.float64)
        orc.orc_ray_params((C.c_double * 3)(*camera.position), (C.c_double * 9)(*camera.rotation),
                           (C.c_double * 3)(*info.bounds_min), (C.c_double * 3)(*info.bounds_max), levels, W, H,
                           _p(out, C.c_double))
        return out

    @staticmethod
    def trace_paths(dag: OracleDAG, params, W, H, rows=None, threads=8, count=False):
        params = np.ascontiguousarray(params, np.float64)
        out = np.zeros((H, W, 4), np.uint32)
        y0, y1 = rows if rows else (0, H)
        b = np.zeros(1, np.uint64)
        orc.orc_trace_paths(dag.h, _p(params, C.c_double), W, H, y0, y1, _p(out), _p(b, C.c_uint64) if count else None,
                            threads)
        return (out, int(b[0])) if count else out

    @staticmethod
    def trace_colors(dag: OracleDAG, colors: OracleColors, paths, W, H, rows=None, threads=8):
        out = np.zeros((H, W), np.uint32)
        y0, y1 = rows if rows else (0, H)
        paths = np.ascontiguousarray(paths, np.uint32)
        orc.orc_trace_colors(dag.h, colors.h, W, H, y0, y1, _p(paths), _p(out), threads)
        return out

    @staticmethod
    def set_any_hit_front_to_back(on: bool) -> None:
        """Child order of the shadow walk: the reference's highest-bit-first order (default) or
        front to back along the ray (the device's). The result is the same; the bytes differ."""
        orc.orc_set_any_hit_order(1 if on else 0)

    @staticmethod
    def trace_shadows(dag: OracleDAG, params, paths, colors_in, W, H, bias=1.0, fog=0.0, rows=None, threads=8,
                      count=False):
        params = np.ascontiguousarray(params, np.float64)
        paths = np.ascontiguousarray(paths, np.uint32)
        out = np.ascontiguousarray(colors_in, np.uint32).copy()
        y0, y1 = rows if rows else (0, H)
        b = np.zeros(1, np.uint64)
        orc.orc_trace_shadows(dag.h, _p(params, C.c_double), W, H, y0, y1, _p(paths), _p(out), bias, fog,
                              _p(b, C.c_uint64) if count else None, threads)
        return (out, int(b[0])) if count else out


def perlin_fractal(seed: int, frequency: float, octaves: int, xyz, with_perm: bool = False):
    """The oracle's FastNoise Perlin fractal at points xyz (N x 3 float32); optionally the
    permutation tables (m_perm, m_perm12)."""
    pts = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    out = np.zeros(len(pts), np.float32)
    perm = np.zeros(1024, np.uint8) if with_perm else None
    orc.orc_perlin(seed, frequency, octaves, len(pts), _p(pts, C.c_float), _p(out, C.c_float),
                   _p(perm, C.c_uint8) if with_perm else None)
    return (out, perm[:512], perm[512:]) if with_perm else out
