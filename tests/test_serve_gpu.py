"""Serving (hd_serve_run): every trace request's frame is the same served alone (requests
one at a time, in arrival order, edits one pass per editor) and in a mix (continuous
batching: edit runs fused, traces concurrent on several slots), and the final DAGs are
canonically identical."""
import numpy as np
import pytest

from tests.scenes import POOL_PAGES, scene

pytestmark = pytest.mark.gpu

SMALL_RES = ((128, 72), (96, 64), (160, 90), (64, 64))


def alone(sc, g, reqs):
    """Reference run: each request by itself, in arrival order."""
    import torch

    from hashdag_amd import DAGInfo, DAGTracer, HashDAG
    from hashdag_amd._lib import HD_REQ_TRACE
    from hashdag_amd.serve import frame_checksum
    dag = HashDAG(g, pool_pages=POOL_PAGES)
    dag.load_scene(sc)
    info = DAGInfo.unit(g.levels)
    sums = {}
    for i, r in enumerate(reqs):
        if r.kind == HD_REQ_TRACE:
            tr = DAGTracer(r.width, r.height)
            tr.resolve_paths(None, info, dag, params=r.params)
            tr.resolve_colors(dag)
            tr.resolve_shadows(None, info, dag, 1.0, 0.0, params=r.params)
            torch.cuda.synchronize()
            sums[i] = frame_checksum(dag, tr.colors, r.width, r.height)
        else:
            dag.edit_batch(r.edits, fused=False)
    return dag, sums


def canonical(dag):
    from oracle import OracleDAG
    return OracleDAG.from_image(dag.download(), dag.pool_pages).canonical()


@pytest.mark.parametrize("slots,reserve", [(4, 1024), (1, 1024), (3, 1 << 13)])
def test_served_frames_match_alone(slots, reserve):
    from hashdag_amd import HashDAG
    from hashdag_amd.serve import request_stream, serve
    p, sc, g = scene("d13")
    reqs = request_stream(p, 48, rate=2000.0, seed=9, resolutions=SMALL_RES, edits_per_request=(1, 8), rmax=24.0)
    dag = HashDAG(g, pool_pages=POOL_PAGES)
    dag.load_scene(sc)
    # all requests at once: the largest batches; reserve 1 << 13 pages forces GC before edits
    res = serve(dag, reqs, trace_slots=slots, realtime=False, reserve_pages=reserve)
    assert (res["status"] == 0).all()
    assert res["batch"].max() + 1 < len(reqs)  # requests were batched
    ref, sums = alone(sc, g, reqs)
    for i, s in sums.items():
        assert res["checksum"][i] == s, f"request {i} ({reqs[i].width}x{reqs[i].height})"
    assert np.array_equal(canonical(dag), canonical(ref))


def test_realtime_stream_latencies():
    from hashdag_amd import HashDAG
    from hashdag_amd.serve import latency_summary, request_stream, serve
    p, sc, g = scene("d13")
    reqs = request_stream(p, 60, rate=600.0, seed=4, resolutions=SMALL_RES, edits_per_request=(1, 8), rmax=24.0)
    dag = HashDAG(g, pool_pages=POOL_PAGES)
    dag.load_scene(sc)
    res = serve(dag, reqs, trace_slots=4, realtime=True)
    arrival = np.array([r.arrival_s for r in reqs])
    assert (res["status"] == 0).all()
    assert (res["start_s"] >= arrival - 1e-6).all()  # nothing starts before it arrives
    assert (res["done_s"] >= res["start_s"]).all()
    summ = latency_summary(reqs, res)
    assert summ["trace_count"] + summ["edit_count"] == len(reqs)
    assert summ["trace_p99_ms"] >= summ["trace_p50_ms"] > 0
