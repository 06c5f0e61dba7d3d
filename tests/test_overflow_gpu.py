"""Capacity errors on the device edit path. The reference writes past a full bucket or page
budget without a check (hash_table.h:393, 461, 794-809); the device pass reports
HD_ERR_BUCKET_FULL / HD_ERR_POOL_FULL and keeps the previous root, colour leaves and traced
frame. The geometries are sized so the scene fits and one large sphere add does not: the
oracle throws on the same edit (checked here too), a smaller one fits on both sides.
"""
import numpy as np
import pytest

from tests.scenes import scene

pytestmark = pytest.mark.gpu

BIG = (0, (128.0, 140.0, 128.0), 90.0)    # overflows both geometries below
SMALL = (0, (60.0, 139.0, 190.0), 12.0)   # a surface point; fits both geometries


def _geometry(bits_low):
    from hashdag_amd import Geometry
    return Geometry(levels=8, num_top_levels=5, bucket_bits_top=6, bucket_bits_low=bits_low, bucket_size_top=1024,
                    bucket_size_low=512, color_tree_levels=4)


def _frame(dag, p, W=96, H=64):
    import torch

    from hashdag_amd import DAGInfo, DAGTracer, ray_params
    from hashdag_amd.workload import cameras
    info = DAGInfo.unit(dag.geometry.levels)
    cam = cameras(p, 1, seed=21)[0]
    prm = ray_params(cam, info, dag.geometry.levels, W, H)
    tr = DAGTracer(W, H)
    tr.resolve_paths(cam, info, dag, params=prm)
    tr.resolve_colors(dag)
    tr.resolve_shadows(cam, info, dag, 1.0, 0.0, params=prm)
    torch.cuda.synchronize()
    return tr.colors.cpu().numpy().copy()


@pytest.mark.parametrize("case", ["bucket", "pool"])
def test_overflow_keeps_previous_state(case):
    from hashdag_amd import HashDAG
    from hashdag_amd._lib import HD_ERR_BUCKET_FULL, HD_ERR_POOL_FULL, HashDAGError
    from oracle import OracleColors, OracleDAG
    from tests.test_edit_gpu import assert_same

    p, sc, _ = scene("d8")
    if case == "bucket":
        g, pool_pages, want = _geometry(6), 1 << 14, HD_ERR_BUCKET_FULL
    else:
        g = _geometry(7)
        probe = OracleDAG.from_basic(g, sc.dag, 1 << 14)
        pool_pages, want = probe.export()[4] + 12, HD_ERR_POOL_FULL  # SMALL takes 9 pages, BIG 26
    od = OracleDAG.from_basic(g, sc.dag, pool_pages)
    oc = OracleColors.from_scene(sc)
    with pytest.raises(RuntimeError):  # the oracle overflows on the same edit
        OracleDAG.from_basic(g, sc.dag, pool_pages).edit(*BIG, OracleColors.from_scene(sc))

    dag = HashDAG(g, pool_pages=pool_pages)
    dag.load_scene(sc)
    root = dag.root
    before = _frame(dag, p)
    print(f"[overflow {case}] large edit", flush=True)
    for fused in (False, True):
        with pytest.raises(HashDAGError) as ei:
            dag.edit_batch([SMALL, BIG] if fused else [BIG], fused=fused)
        assert ei.value.status == want, f"status {ei.value.status}, expected {want}"
        assert dag.root == root
        assert dag.undo_depth() == (0, 0)
        assert_same(dag, sc, od, oc, f"{case} overflow (fused={fused})")
        assert np.array_equal(_frame(dag, p), before), "traced frame changed by a failed pass"
    if case == "pool":  # the failed passes mapped the spare pages into buckets of their own
        dag.close()
        return
    # the table stays usable: collect what the failed passes appended, then a smaller edit
    # matches the oracle
    dag.remove_stale_nodes()
    st = dag.edit(*SMALL)
    ost = od.edit(*SMALL, oc)
    assert st.num_voxels == ost[1]
    assert_same(dag, sc, od, oc, f"{case}: edit after the failed passes")
    dag.close()
