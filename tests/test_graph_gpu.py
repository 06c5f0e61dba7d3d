"""HIP-graph replays of edit batch + trace: every replay (with eager work and synchronisation
between replays) must give the same frame, voxel count and final DAG as the eager calls, and
the oracle applying the same edits one by one must trace the same frames and reach the same DAG
and colour leaves."""
import numpy as np
import pytest

from tests.scenes import POOL_PAGES, scene
from tests.test_edit_gpu import random_edits

pytestmark = pytest.mark.gpu


def canonical(dag):
    from oracle import OracleDAG
    return OracleDAG.from_image(dag.download(), dag.pool_pages).canonical()


@pytest.mark.parametrize("tiny_scratch,schedule,workgroup", [(False, (0, 0), False), (True, (0, 0), False),
                                                        (False, (2, 4), False), (False, (1, 16), False),
                                                        (True, (0, 0), True), (False, (3, 4 | (20 << 16)), False)])
def test_graph_replays_match_eager(tiny_scratch, schedule, workgroup):
    import torch

    from hashdag_amd import DAGInfo, DAGTracer, HashDAG, ray_params
    from hashdag_amd.graph import FrameGraph
    from hashdag_amd.workload import cameras
    from oracle import Oracle, OracleColors, OracleDAG
    from tests.test_edit_gpu import assert_same
    p, sc, g = scene("d13")
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    W, H = 120, 80
    info = DAGInfo.unit(g.levels)
    ga = HashDAG(g, pool_pages=POOL_PAGES)
    ga.load_scene(sc)
    if tiny_scratch:  # the first replays overflow: eager fallback, growth, re-capture
        ga.reserve_edit_scratch(64, 256, 1024)
    if workgroup:  # the captured pass (and the eager fallback) run in one workgroup
        ga.set_edit_launch(1 << 30)
    eb = HashDAG(g, pool_pages=POOL_PAGES)
    eb.load_scene(sc)
    ta, tb = DAGTracer(W, H), DAGTracer(W, H)
    ta.set_schedule(ga, *schedule)  # the graph traces with this schedule, the eager side per tile
    graph = FrameGraph(ga, ta, max_edits=32)
    cams = cameras(p, 6, seed=31)
    replays = 0
    for k, cam in enumerate(cams):
        edits = [] if k == 2 else random_edits(p, 12, seed=100 + k)
        prm = ray_params(cam, info, g.levels, W, H)
        st, ran = graph.launch(edits, prm)
        replays += ran
        sb = eb.edit_batch(edits, fused=True) if edits else None
        tb.resolve_paths(cam, info, eb, params=prm)
        tb.resolve_colors(eb)
        tb.resolve_shadows(cam, info, eb, 1.0, 0.0, params=prm)
        torch.cuda.synchronize()
        assert np.array_equal(ta.colors.cpu().numpy(), tb.colors.cpu().numpy()), f"frame {k} differs"
        ovox = sum(od.edit(kind, c, r, oc)[1] for kind, c, r in edits)
        if sb is not None:
            assert st.num_voxels == sb.num_voxels == ovox
        op = Oracle.ray_params(cam, info, g.levels, W, H)
        opaths = Oracle.trace_paths(od, op, W, H)
        oframe = Oracle.trace_shadows(od, op, opaths, Oracle.trace_colors(od, oc, opaths, W, H), W, H, 1.0, 0.0)
        assert np.array_equal(ta.colors.cpu().numpy().view(np.uint32), oframe), f"frame {k} differs from the oracle"
    assert replays >= (3 if tiny_scratch else len(cams))
    assert np.array_equal(canonical(ga), canonical(eb))
    assert_same(ga, sc, od, oc, "after the graph replays")
    # undo through graph-committed frames restores the scene
    for _ in range(ga.undo_depth()[0]):
        ga.undo()
    ref = HashDAG(g, pool_pages=POOL_PAGES)
    ref.load_scene(sc)
    assert np.array_equal(canonical(ga), canonical(ref))
    graph.close()


def test_edit_only_graph_matches_oracle():
    """An edit-only graph (the pass without a trace) replayed once per edit, with an overflow
    on its first replays (eager fallback, growth, re-capture): voxel counts, DAG, colours and
    leaf words against the oracle; undo restores the loaded scene."""
    from hashdag_amd import HashDAG
    from hashdag_amd.graph import FrameGraph
    from oracle import OracleColors, OracleDAG
    from tests.test_edit_gpu import assert_same
    p, sc, g = scene("d13")
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    dag = HashDAG(g, pool_pages=POOL_PAGES)
    dag.load_scene(sc)
    dag.reserve_edit_scratch(64, 256, 1024)
    before = canonical(dag)
    graph = FrameGraph(dag, None, 1)
    replays = 0
    for k, e in enumerate(random_edits(p, 10, seed=57)):
        st, ran = graph.launch([e])
        replays += ran
        assert st.num_voxels == od.edit(e[0], e[1], e[2], oc)[1], f"edit {k}"
    assert replays >= 5
    assert_same(dag, sc, od, oc, "edit-only graph")
    for _ in range(dag.undo_depth()[0]):
        dag.undo()
    assert np.array_equal(canonical(dag), before)
    graph.close()
    dag.close()
