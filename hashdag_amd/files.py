"""The reference's BasicDAG and compressed-colours file formats (BasicDAGFactory,
basic_dag.cpp:34-117; serializer.h:11-122), through the library, and the colour tree the
reference's factory builds over them (create_hash_dag_colors, hash_dag_factory.cpp:101-150)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib, ptr


def write_basic_dag_file(path: str, bounds_min, bounds_max, levels: int, words) -> None:
    words = np.ascontiguousarray(words, np.uint32)
    bmin, bmax = (C.c_double * 3)(*bounds_min), (C.c_double * 3)(*bounds_max)
    check(lib.hd_write_basic_dag_file(path.encode(), bmin, bmax, levels, ptr(words), words.size))


def read_basic_dag_file(path: str):
    """(bounds_min, bounds_max, levels, words)."""
    b = (C.c_double * 6)()
    levels, n = C.c_uint32(), C.c_uint64()
    check(lib.hd_read_basic_dag_file(path.encode(), b, C.byref(levels), None, 0, C.byref(n)))
    words = np.zeros(n.value, np.uint32)
    check(lib.hd_read_basic_dag_file(path.encode(), b, C.byref(levels), ptr(words), words.size, C.byref(n)))
    return tuple(b[:3]), tuple(b[3:]), int(levels.value), words


def write_compressed_colors_file(path: str, top_levels: int, enclosed, weights, blocks, macro) -> None:
    """BasicDAGFactory::save_compressed_colors_to_file (basic_dag.cpp:34-52). `weights` in memory
    order (the order colour leaves hold them); the file holds them byte-swapped."""
    e = np.ascontiguousarray(enclosed, np.uint64)
    w = np.ascontiguousarray(weights, np.uint32)
    b = np.ascontiguousarray(blocks, np.uint64)
    m = np.ascontiguousarray(macro, np.uint64)
    check(lib.hd_write_compressed_colors_file(path.encode(), top_levels, ptr(e, C.c_uint64), e.size, ptr(w), w.size,
                                              ptr(b, C.c_uint64), b.size, ptr(m, C.c_uint64), m.size))


def read_compressed_colors_file(path: str):
    """(top_levels, enclosed, weights, blocks, macro) with the weights in memory order
    (load_compressed_colors_from_file, basic_dag.cpp:55-89)."""
    top = C.c_uint32()
    sizes = np.zeros(4, np.uint64)
    check(lib.hd_read_compressed_colors_file(path.encode(), C.byref(top), ptr(sizes, C.c_uint64), None, None, None,
                                             None))
    e, w = np.zeros(int(sizes[0]), np.uint64), np.zeros(int(sizes[1]), np.uint32)
    b, m = np.zeros(int(sizes[2]), np.uint64), np.zeros(int(sizes[3]), np.uint64)
    check(lib.hd_read_compressed_colors_file(path.encode(), C.byref(top), ptr(sizes, C.c_uint64), ptr(e, C.c_uint64),
                                             ptr(w), ptr(b, C.c_uint64), ptr(m, C.c_uint64)))
    return int(top.value), e, w, b, m


def basic_dag_counts(words, levels: int, top_levels: int):
    """(words, enclosed): the BasicDAG with voxel-count headers, through `enclosed` above
    top_levels (BasicDAGColorsBase::get_leaves_count, basic_dag.h:60-72)."""
    words = np.ascontiguousarray(words, np.uint32)
    out = np.zeros_like(words)
    n = C.c_uint64(0)
    check(lib.hd_basic_dag_counts(ptr(words), words.size, levels, top_levels, ptr(out), None, C.byref(n)))
    enclosed = np.zeros(n.value, np.uint64)
    check(lib.hd_basic_dag_counts(ptr(words), words.size, levels, top_levels, ptr(out), ptr(enclosed, C.c_uint64),
                                  C.byref(n)))
    return out, enclosed


def colors_from_basic_dag(words, levels: int, color_tree_levels: int, top_levels: int, enclosed):
    """(scene_words, color_nodes, color_offsets): the colour tree create_hash_dag_colors builds
    (hash_dag_factory.cpp:101-150), in this library's handle convention, and the words with the
    headers above the colour tree reduced to the child mask."""
    words = np.ascontiguousarray(words, np.uint32)
    enclosed = np.ascontiguousarray(enclosed, np.uint64)
    nn, no = C.c_uint64(0), C.c_uint64(0)
    check(lib.hd_colors_from_basic_dag(ptr(words), words.size, levels, color_tree_levels, top_levels,
                                       ptr(enclosed, C.c_uint64), enclosed.size, None, None, C.byref(nn), None,
                                       C.byref(no)))
    sw, nodes, offs = np.zeros_like(words), np.zeros(nn.value, np.uint32), np.zeros(no.value, np.uint64)
    check(lib.hd_colors_from_basic_dag(ptr(words), words.size, levels, color_tree_levels, top_levels,
                                       ptr(enclosed, C.c_uint64), enclosed.size, ptr(sw), ptr(nodes), C.byref(nn),
                                       ptr(offs, C.c_uint64), C.byref(no)))
    return sw, nodes, offs


def load_reference_scene(dag_path: str, colors_path: str, color_tree_levels: int):
    """A Scene from the reference's BasicDAG and compressed-colours files (what
    HashDAGFactory::load_from_DAG reads, hash_dag_factory.cpp:152-198)."""
    from .scene import Scene

    _, _, levels, words = read_basic_dag_file(dag_path)
    top, enclosed, weights, blocks, macro = read_compressed_colors_file(colors_path)
    sw, nodes, offs = colors_from_basic_dag(words, levels, color_tree_levels, top, enclosed)
    voxels = int(enclosed[words[0] >> 8]) if top > 0 else int(words[0] >> 8)
    return Scene(levels=levels, color_tree_levels=color_tree_levels, dag=sw, color_nodes=nodes, color_offsets=offs,
                 blocks=blocks, weights=weights, macro=macro, num_voxels=voxels)
