"""Device-resident HashDAG: the hash table (pool, page table, buckets), the root and the
colour tree, behind the C ABI. Mirrors the reference's HashDAG / HashTable / HashDAGColors
surface (hash_dag.h:214-636, hash_table.h:137-898, hash_dag_colors.h:9-198).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import check, hd_config, hd_edit, hd_edit_stats, hd_image_info, lib, ptr
from .scene import Scene


@dataclass
class Geometry:
    """HashTable geometry (hash_dag_globals.h:7-40; defaults of typedefs.h:202-238)."""
    levels: int = 16
    page_size: int = 512
    num_top_levels: int = 9
    bucket_bits_top: int = 10
    bucket_bits_low: int = 16
    bucket_size_top: int = 1024
    bucket_size_low: int = 4096
    color_tree_levels: int = 10

    def __post_init__(self):
        # full-node headers count 8^(levels - level) voxels in 24 bits from the colour-tree
        # level down (hash_table.h:749-789)
        if self.color_tree_levels < 1 or self.color_tree_levels + 7 < self.levels:
            raise ValueError(f"color_tree_levels {self.color_tree_levels} must be at least levels - 7 "
                             f"({self.levels - 7}) and positive")

    @property
    def leaf_level(self) -> int:
        return self.levels - 2

    def top(self) -> int:
        return min(self.num_top_levels, self.levels)

    def total_buckets(self) -> int:
        return self.top() * (1 << self.bucket_bits_top) + (self.levels - self.top()) * (1 << self.bucket_bits_low)

    def total_pages(self) -> int:
        return (self.top() * (1 << self.bucket_bits_top) * (self.bucket_size_top // self.page_size) +
                (self.levels - self.top()) * (1 << self.bucket_bits_low) * (self.bucket_size_low // self.page_size))

    def total_virtual(self) -> int:
        return self.total_pages() * self.page_size

    def capacity(self, level: int) -> int:
        if level < self.top():
            return (1 << self.bucket_bits_top) * self.bucket_size_top
        return (1 << self.bucket_bits_low) * self.bucket_size_low

    @staticmethod
    def fit(levels: int, level_words, color_tree_levels: int = 10, top_fill: float = 0.4,
            low_fill: float = 0.6, headroom: float = 1.0, bucket_scale: int = 1) -> "Geometry":
        """Bucket geometry for measured per-level node words: the reference's bucket counts,
        with the reference's bucket sizes when they hold the scene and larger buckets (in that
        order of preference) when they do not. A level is 'top' while its words fit `top_fill`
        of the top capacity (the reference's 9 when all do); every level must fit its fill
        limit, and the virtual space must stay below 2^32 words. More than the reference's 9
        top levels are used only when no geometry with 9 fits (deep scenes whose uniform low
        levels would not fit 32-bit addresses). `headroom` scales the words for later growth.
        `bucket_scale` (a power of two) multiplies the bucket counts and divides the bucket
        sizes by it (down to one page): the same capacity in more, shorter buckets, so a find
        scans fewer words."""
        last = None
        if bucket_scale < 1 or bucket_scale & (bucket_scale - 1):
            raise ValueError("bucket_scale must be a power of two")
        shift = bucket_scale.bit_length() - 1
        sizes = ((1024, 4096), (2048, 4096), (4096, 4096), (1024, 8192), (2048, 8192), (4096, 8192))
        for max_top, (size_top, size_low) in [(9, sz) for sz in sizes] + [(levels - 1, sz) for sz in sizes]:
            st = max(size_top >> shift, 512)
            sl = max(size_low >> shift, 512)
            g = Geometry(levels=levels, color_tree_levels=color_tree_levels, bucket_size_top=st, bucket_size_low=sl,
                         bucket_bits_top=10 + (size_top // st).bit_length() - 1,
                         bucket_bits_low=16 + (size_low // sl).bit_length() - 1)
            top_cap = (1 << g.bucket_bits_top) * g.bucket_size_top
            n_top = 0
            while n_top < min(max_top, levels) and n_top < len(level_words) and \
                    level_words[n_top] * headroom <= top_fill * top_cap:
                n_top += 1
            g.num_top_levels = n_top
            bad = [lvl for lvl, w in enumerate(level_words)
                   if w * headroom > (low_fill if lvl >= n_top else top_fill) * g.capacity(lvl)]
            if bad:
                last = f"level {bad[0]}: {int(level_words[bad[0]])} words do not fit the bucket geometry"
                continue
            if g.total_virtual() >= 1 << 32:
                last = "virtual address space exceeds 32 bits"
                continue
            return g
        raise ValueError(last)


@dataclass
class HostImage:
    """A HashTable in the reference's CPU layout (hash_table.h:829-843)."""
    geometry: Geometry
    pool: np.ndarray          # uint32 physical pages
    page_table: np.ndarray    # uint32
    bucket_sizes: np.ndarray  # uint32
    full_nodes: np.ndarray    # uint32
    pool_top: int
    root: int


def fill_edit(e: hd_edit, ed) -> None:
    """hd_edit from (kind, center, radius[, source, dest[, transform[, swirl_period]]])."""
    kind, center, radius = ed[:3]
    e.kind = int(kind)
    e.center[:] = [float(c) for c in center]
    e.radius = float(radius)
    if len(ed) > 3:
        e.source[:] = [float(c) for c in ed[3]]
        e.dest[:] = [float(c) for c in ed[4]]
    if len(ed) > 5 and ed[5] is not None:
        e.transform[:] = [float(v) for v in np.asarray(ed[5], np.float64).reshape(9)]
    if len(ed) > 6 and ed[6] is not None:
        e.swirl = 1
        e.swirl_period = float(ed[6])


class HashDAG:
    """One device-resident HashDAG (+ colours) on one GPU."""

    def __init__(self, geometry: Geometry, pool_pages: int | None = None, device: int = 0,
                 color_leaves_cap: int = 1 << 20):
        self.geometry = geometry
        cfg = hd_config()
        lib.hd_default_config(geometry.levels, C.byref(cfg))
        cfg.page_size = geometry.page_size
        cfg.num_top_levels = geometry.num_top_levels
        cfg.bucket_bits_top = geometry.bucket_bits_top
        cfg.bucket_bits_low = geometry.bucket_bits_low
        cfg.bucket_size_top = geometry.bucket_size_top
        cfg.bucket_size_low = geometry.bucket_size_low
        cfg.color_tree_levels = geometry.color_tree_levels
        # the device pool maps the whole virtual space, so by default every page may be handed
        # out (the reference's budget, main.cpp:46, sized its CPU pool)
        cfg.pool_pages = pool_pages if pool_pages is not None else min(geometry.total_pages(), 0xFFFFFFFF)
        cfg.device = device
        cfg.color_leaves_cap = color_leaves_cap
        self.pool_pages = cfg.pool_pages
        h = C.c_void_p()
        check(lib.hd_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.has_colors = False

    # -- lifetime --
    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.hd_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    # -- construction: HashDAGFactory::load_from_DAG / load_colors_from_DAG --
    def load_basic_dag(self, words: np.ndarray) -> None:
        words = np.ascontiguousarray(words, np.uint32)
        check(lib.hd_load_basic_dag(self._h, ptr(words), words.size), self._h)

    def load_colors(self, nodes, offsets, blocks, weights, macro) -> None:
        nodes = np.ascontiguousarray(nodes, np.uint32)
        offsets = np.ascontiguousarray(offsets, np.uint64)
        blocks = np.ascontiguousarray(blocks, np.uint64)
        weights = np.ascontiguousarray(weights, np.uint32)
        macro = np.ascontiguousarray(macro, np.uint64)
        check(lib.hd_load_colors(self._h, ptr(nodes), nodes.size, ptr(offsets, C.c_uint64), offsets.size,
                                 ptr(blocks, C.c_uint64), blocks.size, ptr(weights), weights.size,
                                 ptr(macro, C.c_uint64), macro.size), self._h)
        self.has_colors = True

    def load_scene(self, scene: Scene) -> None:
        self.load_basic_dag(scene.dag)
        self.load_colors(scene.color_nodes, scene.color_offsets, scene.blocks, scene.weights, scene.macro)

    # -- root (firstNodeIndex) --
    @property
    def root(self) -> int:
        r = C.c_uint32()
        check(lib.hd_get_root(self._h, C.byref(r)), self._h)
        return r.value

    @root.setter
    def root(self, value: int) -> None:
        check(lib.hd_set_root(self._h, value), self._h)

    # -- image exchange --
    def info(self) -> hd_image_info:
        i = hd_image_info()
        check(lib.hd_image_info_get(self._h, C.byref(i)), self._h)
        return i

    def download(self) -> HostImage:
        i = self.info()
        g = self.geometry
        pool = np.zeros(self.pool_pages * g.page_size, np.uint32)
        pt = np.zeros(i.total_pages, np.uint32)
        bs = np.zeros(i.total_buckets, np.uint32)
        full = np.zeros(g.levels, np.uint32)
        check(lib.hd_download_image(self._h, ptr(pool), ptr(pt), ptr(bs), ptr(full)), self._h)
        return HostImage(g, pool, pt, bs, full, i.pool_top, i.root)

    def upload(self, img: HostImage) -> None:
        pool = np.ascontiguousarray(img.pool, np.uint32)
        check(lib.hd_upload_image(self._h, ptr(pool), ptr(np.ascontiguousarray(img.page_table, np.uint32)),
                                  ptr(np.ascontiguousarray(img.bucket_sizes, np.uint32)),
                                  ptr(np.ascontiguousarray(img.full_nodes, np.uint32)), img.pool_top, img.root),
              self._h)

    # -- edits: HashDAG::edit_threads<Editor> (hash_dag.h:255-405) --
    def edit_batch(self, edits, fused: bool = False, stream=None) -> hd_edit_stats:
        """Applies [(kind, (x, y, z), radius), ...] in order. A copy adds source and dest, and
        optionally a 3x3 transform and a swirl period: (5, center, radius, source, dest[,
        transform[, swirl_period]]). fused=True runs each run of editors between copies in one
        pass."""
        arr = (hd_edit * max(len(edits), 1))()
        for i, ed in enumerate(edits):
            fill_edit(arr[i], ed)
        st = hd_edit_stats()
        check(lib.hd_edit_batch(self._h, arr, len(edits), 1 if fused else 0, C.byref(st), stream), self._h)
        return st

    def edit_batch_static(self, edits, stream=None) -> hd_edit_stats:
        """One fused pass of at most HD_STATIC_MAX_EDITS sphere / box / paint editors in the
        library's static device scratch: allocates nothing, records no undo frame and clears
        the undo history."""
        arr = (hd_edit * max(len(edits), 1))()
        for i, ed in enumerate(edits):
            fill_edit(arr[i], ed)
        st = hd_edit_stats()
        check(lib.hd_edit_batch_static(self._h, arr, len(edits), C.byref(st), stream), self._h)
        return st

    def set_edit_launch(self, workgroup_max_items: int = 0, persist_max_edits: int = 0,
                        blocks_per_cu: int = 2) -> None:
        """Passes of at most persist_max_edits editors run as one cooperative launch; otherwise
        a pass with at most workgroup_max_items work items on every level runs in one
        single-workgroup launch, and any other pass launches each step on its own."""
        check(lib.hd_set_edit_launch(self._h, workgroup_max_items, persist_max_edits, blocks_per_cu), self._h)

    def autotune_edit_launch(self, sample, reps: int = 1) -> dict:
        """Times one pass per sample editor with each small-pass threshold (0, 32, 128, 512, 2048
        work items per level) and keeps the fastest (hd_autotune_edit_launch)."""
        arr = (hd_edit * max(len(sample), 1))()
        for i, ed in enumerate(sample):
            fill_edit(arr[i], ed)
        pick = C.c_uint32()
        ms = (C.c_float * 5)()
        check(lib.hd_autotune_edit_launch(self._h, arr, len(sample), reps, C.byref(pick), ms), self._h)
        return {"workgroup_max_items": pick.value,
                "ms_per_edit": {str(c): round(float(v), 4) for c, v in zip((0, 32, 128, 512, 2048), ms)}}

    def edit_launch(self) -> tuple[int, int, int]:
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib.hd_get_edit_launch(self._h, C.byref(a), C.byref(b), C.byref(c)), self._h)
        return a.value, b.value, c.value

    def count_edit_bytes(self, on: bool):
        """Starts (on) or stops counting the edit passes' algorithmic bytes; stopping returns
        {descent, find, nodes_written, colours_staged, leaves_written} in bytes."""
        out = np.zeros(5, np.uint64)
        check(lib.hd_edit_count_bytes(self._h, 1 if on else 0, ptr(out, C.c_uint64)), self._h)
        if on:
            return None
        keys = ("descent", "find", "nodes_written", "colours_staged", "leaves_written")
        return {k: int(v) for k, v in zip(keys, out)}

    def edit(self, kind: int, center, radius: float) -> hd_edit_stats:
        return self.edit_batch([(kind, center, radius)])

    def copy(self, center, radius: float, source, dest, transform=None, swirl_period=None,
             decompress: bool = False) -> hd_edit_stats:
        """CopyEditor (hash_dag_editors.h:332-435): the box of radius * max|T (1,1,1)| around
        dest + T (center - source) gets the voxels (and colours) at source + T^-1 (p - dest),
        the relative position first swirled about y when swirl_period is given.

        decompress=True: CopyEditorWithDecompression (hash_dag_editors.h:437-616), no transform
        or swirl; the source is read through its voxel buffer at truncate(center - radius)."""
        if decompress:
            if transform is not None or swirl_period is not None:
                raise ValueError("the decompressing copy takes no transform or swirl")
            return self.edit_batch([(7, center, radius, source, dest)])
        return self.edit_batch([(5, center, radius, source, dest, transform, swirl_period)])

    def remove_stale_nodes(self, max_level: int | None = None, stream=None, keep_undo: bool = False,
                           extra_roots=None) -> dict:
        """GC (hash_dag.h:407-547); the engine calls it with levels - 2 (engine.cpp:182).

        Without options the undo history is cleared, as the engine does. keep_undo keeps the
        nodes of every undo / redo frame alive and the history usable; extra_roots (level-0
        node addresses) are kept alive too and returned as "roots", rewritten to their
        addresses after GC."""
        st = np.zeros(4, np.uint64)
        lvl = self.geometry.leaf_level if max_level is None else max_level
        roots = np.ascontiguousarray(extra_roots if extra_roots is not None else [], np.uint32)
        if keep_undo or roots.size:
            check(lib.hd_gc_roots(self._h, lvl, 1 if keep_undo else 0, ptr(roots, C.c_uint32), roots.size,
                                  ptr(st, C.c_uint64), stream), self._h)
        else:
            check(lib.hd_gc(self._h, lvl, ptr(st, C.c_uint64), stream), self._h)
        return {"live_nodes": int(st[0]), "words_before": int(st[1]), "words_after": int(st[2]),
                "roots": [int(r) for r in roots]}

    # -- files --
    def save_checkpoint(self, path: str) -> None:
        """The table image and colour state, restorable exactly by load_checkpoint."""
        check(lib.hd_checkpoint_save(self._h, path.encode()), self._h)

    def load_checkpoint(self, path: str) -> None:
        check(lib.hd_checkpoint_load(self._h, path.encode()), self._h)

    def save_hash_dag_file(self, path: str, bounds_min, bounds_max) -> None:
        """HashDAGFactory::save_dag_to_file (hash_dag_factory.cpp:200-223)."""
        bmin, bmax = (C.c_double * 3)(*bounds_min), (C.c_double * 3)(*bounds_max)
        check(lib.hd_save_hash_dag_file(self._h, path.encode(), bmin, bmax), self._h)

    def load_hash_dag_file(self, path: str):
        """HashDAGFactory::load_dag_from_file (hash_dag_factory.cpp:225-246): the dumped DAG,
        re-interned into this table; returns the bounds (min, max)."""
        b = (C.c_double * 6)()
        check(lib.hd_load_hash_dag_file(self._h, path.encode(), b), self._h)
        return tuple(b[:3]), tuple(b[3:])

    def undo(self) -> None:
        check(lib.hd_undo(self._h), self._h)

    def redo(self) -> None:
        check(lib.hd_redo(self._h), self._h)

    def clear_undo(self) -> None:
        check(lib.hd_undo_clear(self._h), self._h)

    def undo_depth(self):
        u, r = C.c_uint32(), C.c_uint32()
        check(lib.hd_undo_depth(self._h, C.byref(u), C.byref(r)), self._h)
        return u.value, r.value

    def reserve_edit_scratch(self, items_per_level: int, edit_list_words: int, colors: int) -> None:
        check(lib.hd_reserve_edit_scratch(self._h, items_per_level, edit_list_words, colors), self._h)

    # -- colours --
    def color_nodes(self) -> np.ndarray:
        n = C.c_uint64()
        check(lib.hd_colors_info(self._h, C.byref(n), None, None), self._h)
        out = np.zeros(n.value, np.uint32)
        check(lib.hd_download_color_nodes(self._h, ptr(out)), self._h)
        return out

    def color_leaf(self, handle: int):
        sizes = np.zeros(4, np.uint64)
        check(lib.hd_color_leaf_sizes(self._h, handle, ptr(sizes, C.c_uint64)), self._h)
        blocks = np.zeros(int(sizes[0]), np.uint64)
        weights = np.zeros(int(sizes[1]), np.uint32)
        macro = np.zeros(int(sizes[2]), np.uint64)
        check(lib.hd_download_color_leaf(self._h, handle, ptr(blocks, C.c_uint64), ptr(weights),
                                         ptr(macro, C.c_uint64)), self._h)
        return blocks, weights, macro, int(sizes[3])
