"""Edit passes run as one launch (every step of the pass in sequence: in one workgroup with a
block barrier between steps, or cooperatively over every CU with a grid-wide barrier), or small
passes in one workgroup and the others a launch per step, against the oracle, as the level-synchronous form is tested: plain edits one by one and fused batches on
scenes with one and several macro blocks per colour subtree, passes with copy / fill / noise
editors (the per-voxel colour kernel), the binary depth-12 configuration, overflow errors, and
undo / redo."""
import numpy as np
import pytest

from tests.scenes import scene
from tests.test_edit_gpu import COPY, assert_same, copy_edits, fill_edits, make, random_edits

pytestmark = pytest.mark.gpu

ALL = 1 << 30  # every pass
WORKGROUP = (ALL, 0, 2)     # every step in one workgroup
MIXED = (256, 0, 2)         # passes of at most 256 items per level in one workgroup, the rest per step
COOPERATIVE = (0, ALL, 2)


@pytest.mark.parametrize("name,launch", [("d13", WORKGROUP), ("d10w", WORKGROUP), ("d8", WORKGROUP),
                                         ("d13", MIXED), ("d10w", (64, 0, 2)), ("d8", MIXED),
                                         ("d13", COOPERATIVE), ("d10w", (0, ALL, 1)), ("d8", (0, ALL, 4))])
def test_persistent_plain_and_fused_match_oracle(name, launch):
    p, sc, g, dag, od, oc = make(name)
    dag.set_edit_launch(*launch)
    assert dag.edit_launch() == launch
    for k, (kind, center, r) in enumerate(random_edits(p, 8, seed=71)):
        st = dag.edit(kind, center, r)
        ost = od.edit(kind, center, r, oc)
        assert st.num_voxels == ost[1], f"edit {k}: voxel count {st.num_voxels} != {ost[1]}"
        assert st.leaves_rebuilt == ost[2], f"edit {k}: leaves rebuilt {st.leaves_rebuilt} != {ost[2]}"
    assert_same(dag, sc, od, oc, "persistent plain edits")
    edits = random_edits(p, 40, seed=73)
    st = dag.edit_batch(edits, fused=True)
    assert st.passes == 1
    assert st.num_voxels == sum(od.edit(kind, center, r, oc)[1] for kind, center, r in edits)
    assert_same(dag, sc, od, oc, "persistent fused batch")


@pytest.mark.parametrize("launch", [WORKGROUP, MIXED, COOPERATIVE])
def test_persistent_copy_fill_noise_passes(launch):
    from hashdag_amd import terrain_height
    p, sc, g, dag, od, oc = make("d13")
    dag.set_edit_launch(*launch)
    others = random_edits(p, 6, seed=75)
    copies = copy_edits(p, 2, seed=76)
    fills = fill_edits(p, seed=77)[:2]
    x, z = 120, 140
    noise = (8, (float(x), float(terrain_height(p, x, z) + 2), float(z)), 7.0)
    edits = [others[0], copies[0], others[1], fills[0], others[2], noise, others[3], copies[1], fills[1], others[4]]
    st = dag.edit_batch(edits, fused=True)
    total = 0
    for ed in edits:
        if ed[0] == COPY:
            total += od.edit_copy(ed[1], ed[2], ed[3], ed[4], oc)[1]
        else:
            total += od.edit(ed[0], ed[1], ed[2], oc)[1]
    assert st.num_voxels == total
    assert st.passes == 6  # [o] [C o] [F o] [N o] [C] [F o]
    assert_same(dag, sc, od, oc, "persistent copy / fill / noise passes")


@pytest.mark.parametrize("launch", [WORKGROUP, MIXED, COOPERATIVE])
def test_persistent_binary_depth12(launch):
    from hashdag_amd import Geometry, HashDAG, TerrainParams, generate_terrain
    from hashdag_amd.workload import sphere_edits
    from oracle import OracleDAG

    params = TerrainParams(levels=12, seed=1337, colors=False, color_tree_levels=10)
    sc = generate_terrain(params)
    g = Geometry.fit(12, sc.level_words, color_tree_levels=10, headroom=1.5)
    dag = HashDAG(g, pool_pages=1 << 19)
    dag.load_basic_dag(sc.dag)
    dag.set_edit_launch(*launch)
    od = OracleDAG.from_basic(g, sc.dag, 1 << 19)
    for k, (kind, center, r) in enumerate(sphere_edits(params, 20, seed=1337 + 9)):
        st = dag.edit(kind, center, r)
        assert st.num_voxels == od.edit(kind, center, r, None, threads=8)[1], f"edit {k}"
    god = OracleDAG.from_image(dag.download(), dag.pool_pages)
    assert np.array_equal(god.canonical(), od.canonical())
    dag.close()


@pytest.mark.parametrize("launch", [WORKGROUP, COOPERATIVE])
def test_persistent_overflow_and_undo(launch):
    """A persistent pass that overflows a bucket reports it and keeps the previous state; undo /
    redo of persistent passes restore roots and colours."""
    from hashdag_amd import HashDAG
    from hashdag_amd._lib import HD_ERR_BUCKET_FULL, HashDAGError
    from oracle import OracleColors, OracleDAG
    from tests.test_overflow_gpu import BIG, SMALL, _frame, _geometry

    p, sc, _ = scene("d8")
    g = _geometry(6)
    dag = HashDAG(g, pool_pages=1 << 14)
    dag.load_scene(sc)
    dag.set_edit_launch(*launch)
    od = OracleDAG.from_basic(g, sc.dag, 1 << 14)
    oc = OracleColors.from_scene(sc)
    before = _frame(dag, p)
    with pytest.raises(HashDAGError) as e:
        dag.edit(*BIG)
    assert e.value.status == HD_ERR_BUCKET_FULL
    assert np.array_equal(_frame(dag, p), before)
    assert_same(dag, sc, od, oc, "after the failed persistent pass")
    dag.remove_stale_nodes()  # what the failed pass appended
    dag.edit(*SMALL)
    od.edit(*SMALL, oc)
    assert_same(dag, sc, od, oc, "small edit after the failure")
    after = _frame(dag, p)
    dag.undo()
    assert np.array_equal(_frame(dag, p), before)
    dag.redo()
    assert np.array_equal(_frame(dag, p), after)
    dag.close()


def test_edit_launch_autotune_keeps_results():
    """The edit launch autotuner times its candidates on applied-and-undone sample edits, keeps
    one of them, and leaves the DAG as it was; edits after it match the oracle."""
    from tests.test_graph_gpu import canonical
    p, sc, g, dag, od, oc = make("d13")
    before = canonical(dag)
    got = dag.autotune_edit_launch(random_edits(p, 6, seed=79), reps=1)
    assert got["workgroup_max_items"] in (0, 32, 128, 512, 2048)
    assert dag.edit_launch()[0] == got["workgroup_max_items"]
    assert all(v > 0 for v in got["ms_per_edit"].values())
    assert np.array_equal(canonical(dag), before)
    for kind, center, r in random_edits(p, 6, seed=81):
        assert dag.edit(kind, center, r).num_voxels == od.edit(kind, center, r, oc)[1]
    assert_same(dag, sc, od, oc, "edits after the launch autotuner")
