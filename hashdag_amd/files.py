"""The reference's BasicDAG file format (BasicDAGFactory::save_dag_to_file / load_dag_from_file,
basic_dag.cpp:89-117; serializer.h:11-122), through the library."""
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
