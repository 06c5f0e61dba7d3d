"""GPU edits vs the oracle: after every edit (plain form) and after a fused batch, the device
DAG must be canonically identical to the oracle's (same unique node set, same root) and every
colour subtree must decode to the same colour sequence; the edited-voxel counts must agree.
"""
import numpy as np
import pytest

from tests.scenes import POOL_PAGES, scene

pytestmark = pytest.mark.gpu

SPHERE_ADD, SPHERE_CARVE, BOX_ADD, BOX_CARVE, SPHERE_PAINT = range(5)


def random_edits(p, n, seed, rmax=None):
    rng = np.random.default_rng(seed)
    F = p.footprint or (1 << p.levels)
    rmax = rmax or max(4, F // 8)
    from hashdag_amd import terrain_height
    out = []
    for _ in range(n):
        x, z = (int(v) for v in rng.integers(0, F, size=2))
        y = terrain_height(p, x, z)
        kind = int(rng.choice([SPHERE_ADD, SPHERE_CARVE, BOX_ADD, BOX_CARVE, SPHERE_PAINT],
                              p=[0.3, 0.3, 0.1, 0.1, 0.2]))
        r = float(np.exp(rng.uniform(np.log(2.0), np.log(rmax))))
        out.append((kind, (float(x), float(y), float(z)), float(round(r))))
    return out


def gpu_state_in_oracle(dag, sc):
    """Downloads the device table and colour state into oracle objects (checker only)."""
    from oracle import OracleColors, OracleDAG

    img = dag.download()
    od = OracleDAG.from_image(img, dag.pool_pages)
    oc = OracleColors(dag.color_nodes(), sc.color_offsets, sc.blocks, sc.weights, sc.macro)
    import ctypes as C

    from hashdag_amd._lib import lib
    n = C.c_uint64()
    lib.hd_colors_info(dag.handle, None, C.byref(n), None)
    from oracle.binding import _p, orc
    for i in range(n.value):
        b, w, m, off = dag.color_leaf(i | 0x80000000)
        if off == (1 << 64) - 1:
            h = orc.orc_colors_add_unique(oc.h, _p(b, C.c_uint64), b.size, _p(w), w.size, _p(m, C.c_uint64), m.size)
        else:  # a slot restored to the shared main leaf by undo
            h = orc.orc_colors_add_shared(oc.h, off)
        assert h == (i | 0x80000000)
    return od, oc


def make(name):
    from hashdag_amd import HashDAG
    from oracle import OracleColors, OracleDAG

    p, sc, g = scene(name)
    dag = HashDAG(g, pool_pages=POOL_PAGES)
    dag.load_scene(sc)
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    return p, sc, g, dag, od, oc


def assert_same(dag, sc, od, oc, what):
    """Canonical DAG, decoded colours per subtree, and the stored words of every subtree's leaf
    (blocks, weights, macro blocks of rebuilt leaves; the shared offset of the others)."""
    god, goc = gpu_state_in_oracle(dag, sc)
    assert np.array_equal(god.canonical(), od.canonical()), f"{what}: DAG differs"
    assert np.array_equal(goc.canonical(god), oc.canonical(od)), f"{what}: colours differ"
    gb, ob = goc.leaf_bytes(god), oc.leaf_bytes(od)
    assert gb.size == ob.size and np.array_equal(gb, ob), f"{what}: colour leaf words differ"


@pytest.mark.parametrize("name", ["d8", "d13", "d10w"])
def test_plain_edits_match_oracle(name):
    p, sc, g, dag, od, oc = make(name)
    for k, (kind, center, r) in enumerate(random_edits(p, 10, seed=7)):
        st = dag.edit(kind, center, r)
        ost = od.edit(kind, center, r, oc)
        assert st.num_voxels == ost[1], f"edit {k}: voxel count {st.num_voxels} != {ost[1]}"
        assert st.leaves_rebuilt == ost[2], f"edit {k}: leaves rebuilt {st.leaves_rebuilt} != {ost[2]}"
        assert_same(dag, sc, od, oc, f"edit {k} {(kind, center, r)}")


@pytest.mark.parametrize("name", ["d8", "d13", "d10w"])
def test_fused_batch_matches_sequential_oracle(name):
    p, sc, g, dag, od, oc = make(name)
    edits = random_edits(p, 40, seed=11)
    st = dag.edit_batch(edits, fused=True)
    total = 0
    for kind, center, r in edits:
        total += od.edit(kind, center, r, oc)[1]
    assert st.passes == 1
    assert st.num_voxels == total
    assert_same(dag, sc, od, oc, "fused batch")


@pytest.mark.parametrize("name", ["d8", "d13"])
def test_fused_batch_of_many_editors_matches_sequential_oracle(name):
    """A pass over 200 editors: the top levels' items hold edit lists longer than one
    wavefront (their expand runs a wavefront per child slot), with large boxes that cover whole
    nodes at list positions past the first 64 (the last cover decides which edits the child
    keeps)."""
    p, sc, g, dag, od, oc = make(name)
    F = p.footprint or (1 << p.levels)
    edits = random_edits(p, 200, seed=29, rmax=max(4, F // 32))
    (_, c1, _), (_, c2, _) = edits[100], edits[170]
    edits[100] = (BOX_ADD, c1, float(F // 4))
    edits[170] = (BOX_CARVE, c2, float(F // 8))
    st = dag.edit_batch(edits, fused=True)
    total = 0
    for kind, center, r in edits:
        total += od.edit(kind, center, r, oc)[1]
    assert st.passes == 1
    assert st.num_voxels == total
    assert_same(dag, sc, od, oc, "fused batch of 200 editors")


def test_undo_redo_restore_roots_and_colours():
    p, sc, g, dag, od, oc = make("d13")
    from oracle import OracleColors, OracleDAG
    base_dag = od.canonical()
    base_col = oc.canonical(od)
    edits = random_edits(p, 5, seed=3)
    for kind, center, r in edits:
        dag.edit(kind, center, r)
        od.edit(kind, center, r, oc)
    final_dag, final_col = od.canonical(), oc.canonical(od)
    for _ in edits:
        dag.undo()
    god, goc = gpu_state_in_oracle(dag, sc)
    assert np.array_equal(god.canonical(), base_dag)
    assert np.array_equal(goc.canonical(god), base_col)
    for _ in edits:
        dag.redo()
    god, goc = gpu_state_in_oracle(dag, sc)
    assert np.array_equal(god.canonical(), final_dag)
    assert np.array_equal(goc.canonical(god), final_col)


def test_trace_after_edits_bit_exact():
    import torch
    from hashdag_amd import DAGInfo, DAGTracer, ray_params
    from hashdag_amd.workload import cameras
    from oracle import Oracle

    p, sc, g, dag, od, oc = make("d13")
    for kind, center, r in random_edits(p, 8, seed=5):
        dag.edit(kind, center, r)
        od.edit(kind, center, r, oc)
    W, H = 128, 72
    info = DAGInfo.unit(g.levels)
    tr = DAGTracer(W, H)
    for cam in cameras(p, 2, seed=13):
        prm = ray_params(cam, info, g.levels, W, H)
        tr.resolve_paths(cam, info, dag, params=prm)
        tr.resolve_colors(dag)
        tr.resolve_shadows(cam, info, dag, 1.0, 0.0, params=prm)
        torch.cuda.synchronize()
        gpu = tr.colors.cpu().numpy().view(np.uint32)
        op = Oracle.ray_params(cam, info, g.levels, W, H)
        paths = Oracle.trace_paths(od, op, W, H)
        cols = Oracle.trace_colors(od, oc, paths, W, H)
        ref = Oracle.trace_shadows(od, op, paths, cols, W, H, 1.0, 0.0)
        assert np.array_equal(gpu, ref), f"{int((gpu != ref).sum())} pixels differ"


@pytest.mark.parametrize("fused", [True, False])
def test_scratch_growth_keeps_results(fused):
    """Passes that overflow a deliberately tiny scratch (and colour leaf table) are re-run on
    a larger one; the uploaded edits must survive the re-reserve and the result must match."""
    from hashdag_amd import HashDAG
    p, sc, g, _, od, oc = make("d13")
    dag = HashDAG(g, pool_pages=POOL_PAGES, color_leaves_cap=16)  # colour slots grow too
    dag.load_scene(sc)
    dag.reserve_edit_scratch(64, 256, 1024)
    edits = random_edits(p, 24, seed=17)
    st = dag.edit_batch(edits, fused=fused)
    total = 0
    for kind, center, r in edits:
        total += od.edit(kind, center, r, oc)[1]
    assert st.num_voxels == total
    assert_same(dag, sc, od, oc, f"after scratch growth (fused={fused})")


COPY = 5


def copy_edits(p, n, seed):
    """Copies of terrain boxes (radius 4-24) to seeded destinations nearby, integer valued."""
    from hashdag_amd import terrain_height
    rng = np.random.default_rng(seed)
    F = p.footprint or (1 << p.levels)
    out = []
    for _ in range(n):
        r = float(rng.integers(4, 25))
        x, z = (int(v) for v in rng.integers(32, F - 32, size=2))
        y = terrain_height(p, x, z)
        src = (float(x), float(y), float(z))
        dx, dz = (int(v) for v in rng.integers(-24, 25, size=2))
        dst = (float(x + dx), float(y + int(rng.integers(8, 40))), float(z + dz))
        out.append((COPY, src, r, src, dst))
    return out


@pytest.mark.parametrize("fused", [False, True])
def test_copy_matches_oracle(fused):
    """CopyEditor (identity transform) alone and between other editors: same DAG, colours and
    "num voxels" (box volume for copies) as the oracle applying the edits in order."""
    p, sc, g, dag, od, oc = make("d13")
    edits = []
    others = random_edits(p, 9, seed=23)
    copies = copy_edits(p, 3, seed=29)
    for k in range(3):
        edits += others[3 * k:3 * k + 3] + [copies[k]]
    st = dag.edit_batch(edits, fused=fused)
    total = 0
    for ed in edits:
        if ed[0] == COPY:
            total += od.edit_copy(ed[1], ed[2], ed[3], ed[4], oc)[1]
        else:
            total += od.edit(ed[0], ed[1], ed[2], oc)[1]
    assert st.num_voxels == total
    if fused:
        assert st.passes == 4  # each copy starts a new pass: [o o o] [C o o o] [C o o o] [C]
    assert_same(dag, sc, od, oc, f"copies (fused={fused})")


@pytest.mark.parametrize("fused", [False, True])
def test_decompressing_copy_matches_oracle(fused):
    """CopyEditorWithDecompression (hash_dag_editors.h:437-616) between other editors: the
    buffer read from truncate(center - radius) with fractional radii and centres off the source,
    and the empty-source checks above levels - 4 (copies reaching into the air above the
    terrain and past the scene's low corner)."""
    p, sc, g, dag, od, oc = make("d13")
    from hashdag_amd.workload import COPY_DEC
    others = random_edits(p, 6, seed=41)
    copies = [(COPY_DEC, c, 36.0, s, d) for _, c, _, s, d in copy_edits(p, 3, seed=43)]
    # a fractional radius, a centre off the source point, and a source box past the low corner
    _, c0, _, s0, d0 = copies[0]
    copies[0] = (COPY_DEC, (c0[0] + 3.0, c0[1] + 17.5, c0[2] - 2.0), 20.5, s0, d0)
    _, c2, _, _, d2 = copies[2]
    copies[2] = (COPY_DEC, (10.0, c2[1], 12.0), 33.25, (10.0, c2[1], 12.0), d2)
    edits = []
    for k in range(3):
        edits += others[2 * k:2 * k + 2] + [copies[k]]
    st = dag.edit_batch(edits, fused=fused)
    total = 0
    for ed in edits:
        if ed[0] == COPY_DEC:
            total += od.edit_copy(ed[1], ed[2], ed[3], ed[4], oc, decompress=True)[1]
        else:
            total += od.edit(ed[0], ed[1], ed[2], oc)[1]
    assert st.num_voxels == total
    if fused:
        assert st.passes == 4
    assert_same(dag, sc, od, oc, f"decompressing copies (fused={fused})")


def _rotation(deg, axis):
    a = np.radians(deg)
    c, s = np.cos(a), np.sin(a)
    m = {"y": [[c, 0, s], [0, 1, 0], [-s, 0, c]], "z": [[c, -s, 0], [s, c, 0], [0, 0, 1]]}[axis]
    return np.array(m, np.float64)


TRANSFORMS = [
    np.array([[0, 0, 1], [0, 1, 0], [1, 0, 0]], np.float64),        # axis swap
    np.array([[-1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64),       # mirror
    np.diag([2.0, 2.0, 2.0]),                                        # scale up
    np.diag([0.5, 0.5, 0.5]) @ _rotation(30, "y"),                   # rotate and scale down
    _rotation(45, "z") * 1.1,                                        # the replays' kind of transform
]


def test_transformed_copy_matches_oracle():
    """CopyEditor with a Matrix3x3 transform (hash_dag_editors.h:332-379): the FP64 inverse and
    matrix-vector products run the same operations on both sides, so the result is identical."""
    p, sc, g, dag, od, oc = make("d13")
    copies = copy_edits(p, len(TRANSFORMS), seed=31)
    edits = [c + (T,) for c, T in zip(copies, TRANSFORMS)]
    st = dag.edit_batch(edits, fused=True)
    total = 0
    for ed in edits:
        total += od.edit_copy(ed[1], ed[2], ed[3], ed[4], oc, transform=ed[5])[1]
    assert st.passes == len(edits)
    assert st.num_voxels == total
    assert_same(dag, sc, od, oc, "transformed copies")


def test_swirl_copy_close_to_oracle():
    """Swirled copies: the reference turns by host cosf / sinf, the device by FP64 cos / sin
    rounded to FP32; the two may differ in the last bit of a source coordinate, which moves a
    voxel only when that coordinate sits within an ulp of an integer. The edited DAGs must
    agree on all but a tiny fraction of the voxels of the copied boxes (observed: none)."""
    from tests.test_edit_gpu import gpu_state_in_oracle

    p, sc, g, dag, od, oc = make("d8")
    copies = copy_edits(p, 3, seed=37)
    edits = [c + (None, period) for c, period in zip(copies, (12.0, 40.0, 7.5))]
    dag.edit_batch(edits, fused=True)
    for ed in edits:
        od.edit_copy(ed[1], ed[2], ed[3], ed[4], oc, swirl_period=ed[6])
    god, _ = gpu_state_in_oracle(dag, sc)
    a, b = god.dense(), od.dense()
    diff = int((a != b).sum())
    boxes = sum(int((2 * ed[2]) ** 3) for ed in edits)
    assert diff <= max(1, boxes // 100000), f"{diff} voxels differ"


def fill_edits(p, seed):
    """Fills above the surface, below the terrain shell and on a solid voxel, integer and
    fractional radii."""
    from hashdag_amd import terrain_height
    rng = np.random.default_rng(seed)
    F = p.footprint or (1 << p.levels)
    out = []
    for dy, r in ((5, 9.0), (-5, 7.0), (0, 6.0), (6, 5.5), (-3, 12.0)):
        x, z = (int(v) for v in rng.integers(40, F - 40, size=2))
        y = terrain_height(p, x, z) + dy
        out.append((6, (float(x), float(y), float(z)), r))
    return out


@pytest.mark.parametrize("fused", [False, True])
def test_fill_matches_oracle(fused):
    """FillEditorColors (hash_dag_editors.h:624-768): flood fill of the empty voxels around the
    centre within its box, colours carried from the last loaded solid voxel; alone and between
    other editors (each fill starts a new pass, built from the DAG the edits before it left)."""
    p, sc, g, dag, od, oc = make("d13")
    fills = fill_edits(p, seed=61)
    others = random_edits(p, 6, seed=62)
    edits = [fills[0], others[0], others[1], fills[1], fills[2], others[2], fills[3], others[3], fills[4]]
    st = dag.edit_batch(edits, fused=fused)
    total = 0
    for ed in edits:
        total += od.edit(ed[0], ed[1], ed[2], oc)[1]
    assert st.num_voxels == total
    if fused:
        assert st.passes == 5  # [F o o] [F] [F o] [F o] [F]
    assert_same(dag, sc, od, oc, f"fills (fused={fused})")


def _leaf_words(dag, g):
    """Words used in the leaf level's buckets."""
    img = dag.download()
    L = g.leaf_level
    b0 = g.top() * (1 << g.bucket_bits_top) + (L - g.top()) * (1 << g.bucket_bits_low) if L >= g.top() \
        else L * (1 << g.bucket_bits_top)
    nb = (1 << g.bucket_bits_low) if L >= g.top() else (1 << g.bucket_bits_top)
    return int(img.bucket_sizes[b0:b0 + nb].astype(np.uint64).sum())


def test_leaf_find_reuses_existing_leaves():
    """The leaf-level find (tag bytes per slot, then the leaf) finds every leaf the table
    holds: re-running an undone batch, and re-applying a batch after GC (which rebuilds the
    tags), appends no leaf words; the DAG matches the oracle throughout."""
    p, sc, g, dag, od, oc = make("d13")
    edits = [e for e in random_edits(p, 24, seed=31) if e[0] != SPHERE_PAINT]
    dag.edit_batch(edits, fused=True)
    for kind, center, r in edits:
        od.edit(kind, center, r, oc)
    after = _leaf_words(dag, g)
    dag.undo()
    dag.edit_batch(edits, fused=True)
    assert _leaf_words(dag, g) == after, "re-running the batch appended leaves the table held"
    assert_same(dag, sc, od, oc, "re-run batch")
    dag.remove_stale_nodes()
    after_gc = _leaf_words(dag, g)
    dag.edit_batch(edits, fused=True)  # idempotent: every leaf it builds exists
    for kind, center, r in edits:
        od.edit(kind, center, r, oc)
    assert _leaf_words(dag, g) == after_gc, "after GC the batch appended leaves the table held"
    assert_same(dag, sc, od, oc, "batch after GC")


NOISE_ADD, NOISE_CARVE = 8, 9


@pytest.mark.parametrize("fused", [False, True])
def test_noise_matches_oracle(fused):
    """SphereNoiseEditor adding and carving (each starts its own pass, reading the DAG the
    editors before it left), mixed with sphere / paint editors: voxel counts, DAG, colours and
    leaf words against the oracle."""
    from hashdag_amd import terrain_height
    p, sc, g, dag, od, oc = make("d13")
    F = p.footprint
    edits = []
    for k, (x, z, r) in enumerate(((60, 70, 6.0), (140, 90, 9.5), (200, 180, 4.0), (90, 200, 12.0))):
        y = float(terrain_height(p, x, z) + (3 if k % 2 == 0 else 0))
        kind = NOISE_CARVE if k == 2 else NOISE_ADD
        edits.append((kind, (float(x), y, float(z)), r))
        edits.append((SPHERE_PAINT if k % 2 else SPHERE_ADD, (float(x + 5), y, float(z + 3)), 5.0))
    assert all(0 < c[0] < F and 0 < c[2] < F for _, c, _ in edits)
    st = dag.edit_batch(edits, fused=fused)
    total = 0
    for kind, center, r in edits:
        total += od.edit(kind, center, r, oc)[1]
    assert st.passes == (4 if fused else len(edits))
    assert st.num_voxels == total
    assert_same(dag, sc, od, oc, f"noise fused={fused}")


def test_edit_byte_counters():
    """Counting a pass's algorithmic bytes leaves its result unchanged (DAG, colours and leaf
    words against the oracle); written node words and staged colours agree with the pass's own
    statistics; an idle count reads zero."""
    p, sc, g, dag, od, oc = make("d13")
    edits = random_edits(p, 30, seed=41)
    dag.count_edit_bytes(True)
    st = dag.edit_batch(edits, fused=True)
    b = dag.count_edit_bytes(False)
    for kind, center, r in edits:
        od.edit(kind, center, r, oc)
    assert_same(dag, sc, od, oc, "counted batch")
    assert b["nodes_written"] == 4 * st.new_words
    # markers and explicit colours: at least a marker per 64 colours, at most every colour and
    # its chunk's marker, 24 B each
    assert 24 * (st.colors_encoded // 64) <= b["colours_staged"] <= 24 * (st.colors_encoded + st.colors_encoded // 64 + 64 * 4096)
    assert b["descent"] > b["nodes_written"] > 0
    assert b["find"] > 0 and b["leaves_written"] > 0
    dag.count_edit_bytes(True)
    assert all(v == 0 for v in dag.count_edit_bytes(False).values())
    st2 = dag.edit_batch(random_edits(p, 5, seed=43), fused=True)  # not counted
    dag.count_edit_bytes(True)
    assert all(v == 0 for v in dag.count_edit_bytes(False).values())
    assert st2.num_voxels > 0


def test_fused_batch_with_more_shorter_buckets():
    """Twice the buckets at half the size (the bench's bucket-count axis): the same DAG,
    colours and leaf words as the oracle on the same geometry."""
    from hashdag_amd import Geometry, HashDAG
    from oracle import OracleColors, OracleDAG
    p, sc, g0 = scene("d13")
    g = Geometry(levels=g0.levels, num_top_levels=g0.num_top_levels, bucket_bits_top=g0.bucket_bits_top + 1,
                 bucket_bits_low=g0.bucket_bits_low + 1, bucket_size_top=g0.bucket_size_top // 2,
                 bucket_size_low=g0.bucket_size_low // 2, color_tree_levels=g0.color_tree_levels)
    dag = HashDAG(g, pool_pages=POOL_PAGES)
    dag.load_scene(sc)
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    edits = random_edits(p, 40, seed=13)
    st = dag.edit_batch(edits, fused=True)
    assert st.num_voxels == sum(od.edit(kind, center, r, oc)[1] for kind, center, r in edits)
    assert_same(dag, sc, od, oc, "fused batch, 2x buckets")
    dag.remove_stale_nodes()
    od.gc(g.levels - 2)
    assert_same(dag, sc, od, oc, "after GC, 2x buckets")
