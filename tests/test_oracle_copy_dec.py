"""Pins the oracle's CopyEditorWithDecompression (hash_dag_editors.h:437-616, the copy the
reference's checked-in script_definitions.h selects) with a dense-grid brute force written from
the reference text:

* the destination box is BoxEditorBase(center - source + dest, radius);
* a voxel the box touches reads the buffer voxel p = truncate(position - dest + source -
  copyMin), copyMin = center - radius, when p lies inside the 2 ceil(radius) buffer; the buffer
  holds the DAG before the edit from truncate(copyMin); new value = old | buffer[p], newly set
  voxels take the buffer voxel's colour;
* COPY_EMPTY_CHECKS: a node above levels - 4 is left unedited when DAGUtils::is_empty finds its
  source region [truncate(start - dest + source), + side) empty: no existing node one level up
  overlaps it in the reference's inclusive-max test (dag_utils.h:174-252);
* "num voxels" is the box volume; the colour leaves rebuilt are the level-color_tree_levels
  subtrees that the box touches, that no empty check prunes and that hold a voxel afterwards.
"""
import numpy as np
import pytest

from oracle import OracleColors, OracleDAG
from tests.scenes import POOL_PAGES, scene
from tests.test_oracle_edits import BOX_ADD, per_voxel_colours, touched

f32 = np.float32


def to_u32(v):
    """float -> uint32 as x86-64 code converts it (truncate to int64, keep the low 32 bits)."""
    return int(np.trunc(np.float64(v))) & 0xFFFFFFFF


def pyramid(grid):
    occ = [grid.astype(bool)]
    while occ[0].shape[0] > 1:
        a = occ[0]
        n = a.shape[0] // 2
        occ.insert(0, a.reshape(n, 2, n, 2, n, 2).any(axis=(1, 3, 5)))
    return occ


def source_empty(occ, levels, level, start, size):
    """is_empty(level, start, size): the walk tests nodes above `level` against [start, start +
    size) (inclusive max corner) and reports "not empty" as soon as it reaches a node of
    `level`, untested; overlap only narrows going down, so the region is empty exactly when no
    existing node of level - 1 (each has a child) overlaps it. The root level is never empty."""
    if level == 0:
        return False
    s = 1 << (levels - level + 1)
    nodes = np.argwhere(occ[level - 1]).astype(np.int64) * s
    if nodes.size == 0:
        return True
    ov = np.ones(len(nodes), bool)
    for a in range(3):
        mn = nodes[:, a]
        mx = mn + s - 1
        ov &= (mn < start[a] + size) & (mx > start[a])
    return not ov.any()


def node_passes(occ, levels, level, pos, source, dest):
    """COPY_EMPTY_CHECKS part of should_edit for the node at `pos` (node coordinates) of level."""
    if level >= levels - 4:
        return True
    depth = levels - level
    start = [f32(c << depth) for c in pos]
    off = [f32(f32(start[a] - f32(dest[a])) + f32(source[a])) for a in range(3)]
    if not all(o >= 0 for o in off):
        return True
    return not source_empty(occ, levels, level, [to_u32(o) for o in off], 1 << depth)


COPIES = [
    # center, radius, source, dest, scene
    ((40.0, 60.0, 40.0), 10.0, (40.0, 60.0, 40.0), (90.0, 80.0, 90.0), "d7"),
    ((30.0, 55.0, 90.0), 7.0, (20.0, 10.0, 30.0), (75.0, 20.0, 20.0), "d7"),  # a mostly empty source
    ((64.5, 60.0, 63.5), 6.5, (64.0, 60.25, 64.0), (30.0, 70.0, 100.75), "d7"),  # fractional: the buffer window decides
    ((100.0, 130.0, 120.0), 20.0, (100.0, 130.0, 120.0), (160.0, 150.0, 60.0), "d8"),
    # sources above the terrain: the empty checks prune some (d7) or all (d8) touched subtrees
    ((64.0, 115.0, 64.0), 16.0, (64.0, 115.0, 64.0), (38.0, 51.0, 77.0), "d7"),
    ((128.0, 230.0, 128.0), 16.0, (128.0, 230.0, 128.0), (77.0, 102.0, 154.0), "d8"),
]


@pytest.mark.parametrize("center,radius,source,dest,name", COPIES)
def test_copy_with_decompression_brute_force(center, radius, source, dest, name):
    p, sc, g = scene(name)
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    before, cols_before = per_voxel_colours(od, oc, g)
    occ = pyramid(before)
    stats = od.edit_copy(center, radius, source, dest, oc, decompress=True)
    after, cols_after = per_voxel_colours(od, oc, g)
    levels, E = g.levels, 1 << g.levels
    box_c = tuple(float(f32(f32(f32(c) - f32(s)) + f32(d))) for c, s, d in zip(center, source, dest))
    t = touched(BOX_ADD, box_c, radius, E)
    copy_min = [f32(f32(c) - f32(radius)) for c in center]
    int_min = [to_u32(c) for c in copy_min]
    size = 2 * int(np.ceil(radius))

    def buffer_source(v):
        q = [f32(f32(f32(f32(v[a]) - f32(dest[a])) + f32(source[a])) - copy_min[a]) for a in range(3)]
        b = [to_u32(x) for x in q]
        if not all(x < size for x in b):
            return None
        return tuple((int_min[a] + b[a]) & 0xFFFFFFFF & (E - 1) for a in range(3))

    def allowed(v):
        return all(node_passes(occ, levels, lv, tuple(c >> (levels - lv) for c in v), source, dest)
                   for lv in range(levels - 4))

    expect = before.copy()
    src_of = {}
    for v in zip(*np.nonzero(t)):
        v = tuple(int(c) for c in v)
        if before[v] or not allowed(v):
            continue
        s = buffer_source(v)
        if s is not None and before[s]:
            expect[v] = 1
            src_of[v] = s
    assert np.array_equal(after, expect)
    assert stats[1] == int(round((2 * radius) ** 3))
    for v, col in cols_after.items():
        if v in src_of:
            assert col == cols_before[src_of[v]], v
        else:
            assert col == cols_before[v], v
    # colour leaves rebuilt
    ctl = g.color_tree_levels
    side = 1 << (levels - ctl)
    rebuilt = 0
    for S in np.argwhere(np.ones((1 << ctl,) * 3, bool)):
        lo = S * side
        ok = all(float(f32(lo[a])) < box_c[a] + radius and float(f32(lo[a] + side)) > box_c[a] - radius
                 for a in range(3))
        if not ok:
            continue
        if not all(node_passes(occ, levels, lv, tuple(int(c) >> (ctl - lv) for c in S), source, dest)
                   for lv in range(min(ctl + 1, levels - 4))):
            continue
        rebuilt += bool(after[lo[0]:lo[0] + side, lo[1]:lo[1] + side, lo[2]:lo[2] + side].any())
    assert stats[2] == rebuilt
