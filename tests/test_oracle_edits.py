"""Pins the oracle's edit path (hash_dag_edits.h:106-643 with the editors of
hash_dag_editors.h) with dense-grid brute force:

* voxel sets: sphere/box add/carve change exactly the voxels whose unit box the editor's
  shape touches (distance from the centre to the voxel box < r for spheres, open-interval
  overlap for boxes), whatever early exits (is_full / is_empty) the recursion takes;
* edited-voxel count (the "num voxels" stat) = number of touched voxels in the cube;
* colours: every touched voxel that is set afterwards gets the editor colour (sphere red,
  box magenta, both as rgb101210 single colours; paint = old colour lerped to blue), every
  other set voxel keeps its colour, in the Morton order of its colour subtree;
* GC keeps the canonical DAG and colours.
"""
import numpy as np
import pytest

from oracle import OracleColors, OracleDAG
from tests.scenes import POOL_PAGES, scene

SPHERE_ADD, SPHERE_CARVE, BOX_ADD, BOX_CARVE, SPHERE_PAINT = range(5)


def touched(kind, center, radius, E):
    c = np.array(center, np.float64)
    idx = np.arange(E, dtype=np.float64)
    x, y, z = np.meshgrid(idx, idx, idx, indexing="ij")
    if kind in (BOX_ADD, BOX_CARVE):
        lo, hi = c - radius, c + radius
        return ((lo[0] < x + 1) & (x < hi[0]) & (lo[1] < y + 1) & (y < hi[1]) & (lo[2] < z + 1) & (z < hi[2]))
    d2 = np.zeros_like(x)
    for a, g in enumerate((x, y, z)):
        d = np.maximum(0.0, np.maximum(g - c[a], c[a] - (g + 1)))
        d2 += d * d
    return d2 < radius * radius


def rgb101210(r, g, b):
    f = np.float32
    cl = lambda v: min(max(f(v), f(0)), f(1))  # noqa: E731
    return int(cl(r) * f(1023)) | (int(cl(g) * f(4095)) << 10) | (int(cl(b) * f(1023)) << 22)


def colours_by_voxel(canon, levels, ctl):
    """Parse orc_color_canonical output into {(x,y,z) voxel: (colorBits, weight|bpw<<8)} using
    the Morton order of each colour subtree and the dense grid of that subtree."""
    out = {}
    i = 0
    while i < canon.size:
        x, y, z, n = (int(v) for v in canon[i:i + 4])
        i += 4
        out[(x, y, z)] = [(int(canon[i + 2 * k]), int(canon[i + 2 * k + 1])) for k in range(n)]
        i += 2 * n
    return out


def morton_voxels(grid, sub, levels, ctl):
    s = levels - ctl
    n = 1 << s
    base = [c << s for c in sub]
    cube = grid[base[0]:base[0] + n, base[1]:base[1] + n, base[2]:base[2] + n]
    xs, ys, zs = np.nonzero(cube)

    def key(t):
        x, y, z = t
        k = 0
        for b in range(s - 1, -1, -1):
            k = (k << 3) | (((x >> b) & 1) << 2) | (((y >> b) & 1) << 1) | ((z >> b) & 1)
        return k

    vox = sorted(zip(xs.tolist(), ys.tolist(), zs.tolist()), key=key)
    return [(base[0] + a, base[1] + b, base[2] + c) for a, b, c in vox]


def per_voxel_colours(od, oc, g):
    canon = colours_by_voxel(oc.canonical(od), g.levels, g.color_tree_levels)
    grid = od.dense()
    res = {}
    for sub, cols in canon.items():
        vox = morton_voxels(grid, sub, g.levels, g.color_tree_levels)
        assert len(vox) == len(cols)
        res.update(zip(vox, cols))
    return grid, res


EDITS = [
    (SPHERE_ADD, (40.0, 60.0, 50.0), 9.0),
    (SPHERE_CARVE, (64.0, 62.0, 64.0), 14.0),
    (BOX_ADD, (20.0, 70.0, 90.0), 6.0),
    (BOX_CARVE, (100.0, 60.0, 30.0), 11.0),
    (SPHERE_ADD, (70.0, 40.0, 70.0), 30.0),   # large: is_full early exits
    (SPHERE_CARVE, (30.0, 50.0, 30.0), 25.0),  # large: is_empty early exits
    (SPHERE_PAINT, (64.0, 60.0, 64.0), 20.0),
]


@pytest.mark.parametrize("kind,center,radius", EDITS)
def test_edit_voxels_count_and_colours(kind, center, radius):
    p, sc, g = scene("d7")
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    before, cols_before = per_voxel_colours(od, oc, g)
    stats = od.edit(kind, center, radius, oc)
    after, cols_after = per_voxel_colours(od, oc, g)
    E = 1 << g.levels
    t = touched(kind, center, radius, E)
    if kind in (SPHERE_ADD, BOX_ADD):
        expect = before | t
    elif kind in (SPHERE_CARVE, BOX_CARVE):
        expect = before & ~t
    else:
        expect = before
    assert np.array_equal(after, expect.astype(np.uint8))
    assert stats[1] == int(t.sum())
    single = {SPHERE_ADD: rgb101210(1, 0, 0), BOX_ADD: rgb101210(1, 0, 1)}
    for v, col in cols_after.items():
        if t[v] and kind in single:
            assert col == (single[kind], 0), v
        elif t[v] and kind == SPHERE_PAINT:
            assert col[1] >> 8 == 0  # repainted voxels are single colours
        else:
            assert col == cols_before[v], v


def test_paint_colour_value():
    """SpherePaintEditor: lerp(old, blue, clamp(1 - |p - c| / r, 0, 0.1)) (hash_dag_editors.h:323-329),
    computed here in float32 from the old decoded colour."""
    p, sc, g = scene("d7")
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    _, before = per_voxel_colours(od, oc, g)
    center, radius = (64.0, 60.0, 64.0), 20.0
    od.edit(SPHERE_PAINT, center, radius, oc)
    _, after = per_voxel_colours(od, oc, g)
    f = np.float32
    t = touched(SPHERE_PAINT, center, radius, 1 << g.levels)
    checked = 0
    for v, (bits, wb) in before.items():
        if not t[v]:
            continue
        w, bpw = wb & 0xFF, wb >> 8
        if bpw == 0:
            old = [f(bits & 0x3FF) / f(1023), f((bits >> 10) & 0xFFF) / f(4095), f((bits >> 22) & 0x3FF) / f(1023)]
        else:
            wf = f(w) / f((1 << bpw) - 1)
            lo = [f(bits & 0x1F) / f(31), f((bits >> 5) & 0x3F) / f(63), f((bits >> 11) & 0x1F) / f(31)]
            hi = [f((bits >> 16) & 0x1F) / f(31), f((bits >> 21) & 0x3F) / f(63), f((bits >> 27) & 0x1F) / f(31)]
            old = [a * (f(1) - wf) + b * wf for a, b in zip(lo, hi)]
        d = np.array(v, np.float32) - np.array(center, np.float32)
        length = f(np.sqrt(np.float32(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])))
        k = min(max(f(1) - length / f(radius), f(0)), f(0.1))
        new = [a * (f(1) - k) + b * k for a, b in zip(old, (f(0), f(0), f(1)))]
        assert after[v] == (rgb101210(*new), 0), v
        checked += 1
    assert checked > 50


def test_edit_sequence_then_gc_keeps_canonical_form():
    p, sc, g = scene("d8")
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    rng = np.random.default_rng(3)
    for i in range(12):
        kind = int(rng.integers(0, 5))
        center = tuple(float(v) for v in rng.integers(20, 236, size=3))
        od.edit(kind, center, float(rng.integers(3, 30)), oc)
    canon = od.canonical()
    ccanon = oc.canonical(od)
    dense = od.dense()
    od.gc(g.leaf_level)
    assert np.array_equal(od.canonical(), canon)
    assert np.array_equal(oc.canonical(od), ccanon)
    assert np.array_equal(od.dense(), dense)
    # after GC every node left in the table is reachable: bucket words == reachable words
    pool, pt, bs, full, top = od.export()
    assert int(bs.sum()) <= int(sum(canon.size for _ in [0])) * 4


COPIES = [
    # center, radius, source, dest (integer valued; the source box lies inside the cube)
    ((40.0, 60.0, 40.0), 10.0, (40.0, 60.0, 40.0), (90.0, 80.0, 90.0)),
    ((30.0, 55.0, 90.0), 7.0, (20.0, 10.0, 30.0), (75.0, 20.0, 20.0)),
    ((64.0, 64.0, 64.0), 16.0, (60.0, 40.0, 20.0), (20.0, 50.0, 100.0)),
]


@pytest.mark.parametrize("center,radius,source,dest", COPIES)
def test_copy_brute_force(center, radius, source, dest):
    """CopyEditorWithoutDecompression (hash_dag_editors.h:332-435), identity transform: the
    box of `radius` around dest + (center - source) gets old | before[v - dest + source];
    newly set voxels take the colour of their source voxel, all others keep theirs; the
    edited-voxel count is the box volume."""
    p, sc, g = scene("d7")
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    before, cols_before = per_voxel_colours(od, oc, g)
    stats = od.edit_copy(center, radius, source, dest, oc)
    after, cols_after = per_voxel_colours(od, oc, g)
    E = 1 << g.levels
    box_c = tuple(d + (c - s) for c, s, d in zip(center, source, dest))
    t = touched(BOX_ADD, box_c, radius, E)
    off = np.array(source, np.int64) - np.array(dest, np.int64)
    expect = before.copy()
    for v in zip(*np.nonzero(t)):
        src = tuple(int(a) for a in np.array(v) + off)
        assert all(0 <= s < E for s in src)
        if before[src]:
            expect[v] = 1
    assert np.array_equal(after, expect)
    assert stats[1] == int((2 * radius) ** 3)
    for v, col in cols_after.items():
        if t[v] and not before[v]:
            src = tuple(int(a) for a in np.array(v) + off)
            assert col == cols_before[src], v
        else:
            assert col == cols_before[v], v


TRANSFORMED_COPIES = [
    # center, radius, source, dest, transform: axis swaps, mirrors and scales by powers of two
    # keep every step of destination_to_source exact, so the brute force below is exact too
    ((40.0, 60.0, 40.0), 10.0, (40.0, 60.0, 40.0), (90.0, 80.0, 90.0), [[0, 0, 1], [0, 1, 0], [1, 0, 0]]),
    ((30.0, 55.0, 90.0), 7.0, (30.0, 50.0, 60.0), (75.0, 30.0, 20.0), [[-1, 0, 0], [0, 1, 0], [0, 0, 1]]),
    ((64.0, 60.0, 64.0), 6.0, (64.0, 60.0, 64.0), (30.0, 70.0, 100.0), [[2, 0, 0], [0, 2, 0], [0, 0, 2]]),
    ((50.0, 58.0, 70.0), 12.0, (50.0, 58.0, 70.0), (90.0, 40.0, 30.0), [[0.5, 0, 0], [0, 0, 0.5], [0, 0.5, 0]]),
]


@pytest.mark.parametrize("center,radius,source,dest,transform", TRANSFORMED_COPIES)
def test_transformed_copy_brute_force(center, radius, source, dest, transform):
    """CopyEditorWithoutDecompression with a Matrix3x3 transform (hash_dag_editors.h:332-379):
    the box of radius * max|T (1,1,1)| around dest + T (center - source) gets
    old | before[floor(source + T^-1 (v - dest))], newly set voxels take their source's colour."""
    p, sc, g = scene("d7")
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    before, cols_before = per_voxel_colours(od, oc, g)
    T = np.array(transform, np.float64)
    stats = od.edit_copy(center, radius, source, dest, oc, transform=T)
    after, cols_after = per_voxel_colours(od, oc, g)
    E = 1 << g.levels
    box_c = np.array(dest) + T @ (np.array(center) - np.array(source))
    box_r = radius * np.abs(T @ np.ones(3)).max()
    t = touched(BOX_ADD, tuple(box_c), box_r, E)
    Tinv = np.linalg.inv(T)

    def src_of(v):
        s_ = np.array(source) + Tinv @ (np.array(v, np.float64) - np.array(dest))
        assert (s_ >= 0).all() and (s_ < E).all()
        return tuple(int(a) for a in np.floor(s_))

    expect = before.copy()
    for v in zip(*np.nonzero(t)):
        if before[src_of(v)]:
            expect[v] = 1
    assert np.array_equal(after, expect)
    assert stats[1] == int((2 * box_r) ** 3)
    for v, col in cols_after.items():
        if t[v] and not before[v]:
            assert col == cols_before[src_of(v)], v
        else:
            assert col == cols_before[v], v


FILLS = [
    # center (integer valued), radius: above the surface, below the shell, on a solid voxel,
    # a fractional radius
    ("above", 8.0),
    ("below", 7.0),
    ("solid", 6.0),
    ("above", 5.5),
]


def brute_force_fill(before, colours_before, center, radius):
    """FillEditorColors (hash_dag_editors.h:129-222, 624-768) on a dense grid: 6-connected
    flood fill of empty voxels from the box centre (set even if solid) over buffer voxels
    loaded from world voxel floor(min + p); the voxels it checked are 'loaded'. The edit region
    is every voxel overlapping (min + minEdit, min + maxEdit) (bounds from the min corners);
    there voxel w becomes set if buffer voxel p = floor(w - min) is filled or solid. A voxel set
    from a filled buffer voxel takes the colour of the last loaded solid voxel of an x, y, z
    scan over the filled voxels' bounding box. Returns the grid and {voxel: colour} for the
    voxels set from filled buffer voxels."""
    from collections import deque
    f32 = np.float32
    fmin = np.array(center, np.float32) - f32(radius)
    size = 2 * int(np.ceil(f32(radius)))
    world = lambda q: tuple(int(f32(fmin[a]) + f32(q[a])) for a in range(3))
    solid = lambda q: bool(before[world(q)])
    c0 = (size // 2,) * 3
    filled, loaded, todo = {c0}, set(), deque([c0])
    while todo:
        q = todo.popleft()
        for a in range(3):
            for d in (-1, 1):
                n = list(q)
                n[a] += d
                n = tuple(n)
                if not all(0 <= v < size for v in n) or n in filled:
                    continue
                loaded.add(n)
                if not solid(n):
                    filled.add(n)
                    todo.append(n)
    lo = [min(q[a] for q in filled) for a in range(3)]
    hi = [max(q[a] for q in filled) for a in range(3)]
    cols, color = {}, (0, 0)
    for x in range(lo[0], hi[0] + 1):
        for y in range(lo[1], hi[1] + 1):
            for z in range(lo[2], hi[2] + 1):
                q = (x, y, z)
                if q in filled:
                    cols[q] = color
                elif q in loaded and solid(q):
                    color = colours_before[world(q)]
    bmin = [f32(fmin[a] + f32(lo[a])) for a in range(3)]
    bmax = [f32(fmin[a] + f32(hi[a])) for a in range(3)]
    after, new_cols = before.copy(), {}
    ranges = [range(max(0, int(np.floor(bmin[a])) - 1), min(before.shape[a], int(np.ceil(bmax[a])) + 1))
              for a in range(3)]
    for w in ((x, y, z) for x in ranges[0] for y in ranges[1] for z in ranges[2]):
        if not all(f32(w[a] + 1) > bmin[a] and f32(w[a]) < bmax[a] for a in range(3)):
            continue
        q = tuple(int(f32(f32(w[a]) - fmin[a])) for a in range(3))
        if not all(0 <= v < size for v in q):
            continue
        if q in filled or solid(q):
            if not before[w] and q in filled:
                new_cols[w] = cols[q]
            after[w] = 1
    return after, new_cols


@pytest.mark.parametrize("where,radius", FILLS)
def test_fill_brute_force(where, radius):
    from hashdag_amd.scene import terrain_height
    p, sc, g = scene("d7")
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    before, cols_before = per_voxel_colours(od, oc, g)
    x, z = 60, 70
    h = terrain_height(p, x, z)
    y = {"above": h + 4, "below": h - 4, "solid": h}[where]
    center = (float(x), float(y), float(z))
    want, want_cols = brute_force_fill(before, cols_before, center, radius)
    od.edit(6, center, radius, oc)
    after, cols_after = per_voxel_colours(od, oc, g)
    assert np.array_equal(after, want)
    assert int(after.sum()) > int(before.sum()) or where == "solid"
    for v, col in cols_after.items():
        if v in want_cols:
            assert col == want_cols[v], v
        elif v in cols_before:
            assert col == cols_before[v], v


def test_swirl_copy_moves_voxels_by_rotation():
    """Swirl (hash_dag_editors.h:366-372): with a period so long that every angle in the box is
    below one voxel's worth of turn, the swirled copy equals the plain copy; with a shorter period
    the copied set differs but keeps its size within the box's rotation error."""
    p, sc, g = scene("d7")
    center, radius, source, dest = (40.0, 60.0, 40.0), 10.0, (40.0, 60.0, 40.0), (90.0, 80.0, 90.0)
    grids = []
    for period in (None, 1e9, 30.0):
        od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
        oc = OracleColors.from_scene(sc)
        od.edit_copy(center, radius, source, dest, oc, swirl_period=period)
        grids.append(od.dense())
    assert np.array_equal(grids[0], grids[1])
    assert not np.array_equal(grids[0], grids[2])


def test_color_lookup_matches_canonical_order():
    """DAGUtils::get_color (dag_utils.h:423-523) == the colour at the voxel's Morton rank."""
    p, sc, g = scene("d7")
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    grid, cols = per_voxel_colours(od, oc, g)
    rng = np.random.default_rng(5)
    keys = list(cols)
    for k in rng.choice(len(keys), 300):
        v = keys[int(k)]
        ok, bits, wb = od.color(oc, *v)
        assert ok and (bits, wb) == cols[v]
    ok, bits, wb = od.color(oc, 0, 127, 0)  # empty (above the terrain)
    assert not ok and (bits, wb) == (0, 0)


def test_threaded_edit_tasks_same_result():
    """The edit tasks run on several threads under striped bucket locks (the reference's thread
    pool with per-bucket mutexes, hash_dag.h:330-356) give the canonical DAG, colours and counts
    of the single-threaded run."""
    p, sc, g = scene("d10")
    runs = []
    for threads in (1, 6):
        od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
        oc = OracleColors.from_scene(sc)
        rng = np.random.default_rng(5)
        stats = []
        for _ in range(8):
            kind = int(rng.integers(0, 5))
            center = tuple(float(v) for v in rng.integers(60, 450, size=3))
            stats.append(od.edit(kind, center, float(rng.integers(6, 60)), oc, threads=threads)[1:])
        runs.append((od.canonical(), oc.canonical(od), oc.leaf_bytes(od), stats))
    a, b = runs
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert a[3] == b[3]
