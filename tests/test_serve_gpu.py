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
    from hashdag_amd._lib import HD_REQ_REDO, HD_REQ_TRACE, HD_REQ_UNDO
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
        elif r.kind == HD_REQ_UNDO:
            dag.undo()
        elif r.kind == HD_REQ_REDO:
            dag.redo()
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


def synthetic_replay(p, frames, seed):
    """A replay in the reference's CSV format on the d13 terrain (recorded at depth 13): sphere /
    cube / paint edits, copies and fills at surface points, undo and redo, a camera per frame."""
    from hashdag_amd import terrain_height
    from hashdag_amd.workload import cameras
    rng = np.random.default_rng(seed)
    F = p.footprint or (1 << p.levels)
    cams = cameras(p, frames, seed=seed)
    lines = []
    for k in range(frames):
        cam = cams[k]
        lines.append("SetLocation," + ",".join(f"{v:.6f}" for v in cam.position))
        lines.append("SetRotation," + ",".join(f"{v:.6f}" for v in cam.rotation))
        x, z = (int(v) for v in rng.integers(40, F - 40, size=2))
        y = terrain_height(p, x, z)
        r = float(rng.integers(3, 20))
        u = rng.random()
        if u < 0.3:
            lines.append(f"EditSphere,{x},{y},{z},{r},{'true' if rng.random() < 0.5 else 'false'}")
        elif u < 0.5:
            lines.append(f"EditCube,{x},{y},{z},{r},{'TRUE' if rng.random() < 0.5 else 'FALSE'}")
        elif u < 0.6:
            lines.append(f"EditPaint,{x},{y},{z},{r}")
        elif u < 0.7:
            dx, dz = (int(v) for v in rng.integers(-30, 31, size=2))
            lines.append(f"EditCopy,{x},{y},{z},{x},{y},{z},{x + dx},{y + 12},{z + dz},{r}")
        elif u < 0.75:
            lines.append(f"EditFill,{x},{y + 4},{z},{r / 2}")
        elif u < 0.85:
            lines.append("Undo")
        else:
            lines.append("Redo")
        lines.append("EndFrame")
    return "\n".join(lines) + "\n"


def test_served_replay_matches_alone_and_oracle():
    """A replay served as requests (edits, undo / redo, a trace per frame at its camera): the
    frames equal those of the requests run one at a time, and the final DAG equals the oracle
    applying the edits that the undo / redo history leaves in effect."""
    from hashdag_amd import HashDAG
    from hashdag_amd.replay import COPY, Replay, replay_requests
    from hashdag_amd.serve import serve
    from oracle import OracleColors, OracleDAG
    from tests.test_edit_gpu import assert_same

    p, sc, g = scene("d13")
    rep = Replay.parse(synthetic_replay(p, 40, seed=17), 13, 13)
    assert rep.counts()["undo"] > 0 and rep.counts()["redo"] > 0
    reqs = replay_requests(rep, 96, 64, g.levels, fps=500.0)
    dag = HashDAG(g, pool_pages=POOL_PAGES)
    dag.load_scene(sc)
    # one pass per edit request, so every undo step is one replay edit, as in the reference
    res = serve(dag, reqs, trace_slots=3, realtime=False, keep_undo=True, max_batch_edits=1)
    assert (res["status"] == 0).all()
    ref, sums = alone(sc, g, reqs)
    for i, s in sums.items():
        assert res["checksum"][i] == s, f"request {i}"
    # the edits in effect: undo moves the last one to the redo stack, redo moves it back, a
    # new edit clears the redo stack (HashDAGUndoRedo, hash_dag.h:47-120)
    done, undone = [], []
    for a in rep.actions:
        if a.edit is not None:
            done.append(a.edit)
            undone.clear()
        elif a.type == 4 and done:
            undone.append(done.pop())
        elif a.type == 5 and undone:
            done.append(undone.pop())
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    for ed in done:
        if ed[0] == COPY:
            od.edit_copy(ed[1], ed[2], ed[3], ed[4], oc, transform=ed[5], swirl_period=ed[6])
        else:
            od.edit(ed[0], ed[1], ed[2], oc)
    assert_same(dag, sc, od, oc, "served replay")
