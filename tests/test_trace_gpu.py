"""GPU parity of the tracer kernels against the oracle: per-pixel paths, colours and
shadowed colours must be bit-exact; algorithmic byte counters must agree."""
import numpy as np
import pytest

from tests.scenes import POOL_PAGES, scene

pytestmark = pytest.mark.gpu


def _setup(name):
    import torch  # noqa: F401
    from hashdag_amd import HashDAG
    from oracle import OracleColors, OracleDAG

    p, sc, g = scene(name)
    dag = HashDAG(g, pool_pages=POOL_PAGES)
    dag.load_scene(sc)
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    return p, sc, g, dag, od, oc


@pytest.mark.parametrize("name", ["d7", "d8", "d10", "d13"])
def test_image_matches_oracle_layout(name):
    """load_from_DAG on the host + upload/download round trip == the oracle's table, word for word."""
    p, sc, g, dag, od, oc = _setup(name)
    img = dag.download()
    pool, pt, bs, full, top = od.export()
    assert img.root == od.root
    assert img.pool_top == top
    assert np.array_equal(img.page_table, pt)
    assert np.array_equal(img.bucket_sizes, bs)
    assert np.array_equal(img.full_nodes, full)
    assert np.array_equal(img.pool[: top * g.page_size], pool[: top * g.page_size])


@pytest.mark.parametrize("name,W,H,ncam", [("d7", 64, 48, 4), ("d8", 96, 64, 4), ("d10", 128, 72, 3),
                                           ("d13", 160, 90, 3)])
def test_trace_parity(name, W, H, ncam):
    import torch
    from hashdag_amd import DAGInfo, DAGTracer, ray_params
    from hashdag_amd.workload import cameras
    from oracle import Oracle

    p, sc, g, dag, od, oc = _setup(name)
    info = DAGInfo.unit(g.levels)
    tracer = DAGTracer(W, H)
    for cam in cameras(p, ncam, seed=11):
        prm = ray_params(cam, info, g.levels, W, H)
        oparams = Oracle.ray_params(cam, info, g.levels, W, H)
        assert np.array_equal(np.array(prm.as_list()), oparams), "ray setup differs"
        tracer.resolve_paths(cam, info, dag, params=prm)
        tracer.resolve_colors(dag)
        torch.cuda.synchronize()
        gpaths = tracer.paths.cpu().numpy().view(np.uint32)
        gcols = tracer.colors.cpu().numpy().view(np.uint32)
        tracer.resolve_shadows(cam, info, dag, 1.0, 0.0, params=prm)
        torch.cuda.synchronize()
        gshad = tracer.colors.cpu().numpy().view(np.uint32)

        opaths = Oracle.trace_paths(od, oparams, W, H)
        assert np.array_equal(gpaths, opaths), f"paths differ at {np.argwhere((gpaths != opaths).any(-1))[:5]}"
        hits = (opaths[..., :3] != 0).any(-1).mean()
        assert hits > 0.05, "camera sees almost nothing"
        ocols = Oracle.trace_colors(od, oc, opaths, W, H)
        assert np.array_equal(gcols, ocols), f"colours differ at {np.argwhere(gcols != ocols)[:5]}"
        oshad = Oracle.trace_shadows(od, oparams, opaths, ocols, W, H, 1.0, 0.0)
        assert np.array_equal(gshad, oshad), f"shadows differ at {np.argwhere(gshad != oshad)[:5]}"


def test_byte_counters_match_oracle():
    import torch
    from hashdag_amd import DAGInfo, DAGTracer, ray_params
    from hashdag_amd.workload import cameras
    from oracle import Oracle

    p, sc, g, dag, od, oc = _setup("d10")
    W, H = 128, 72
    info = DAGInfo.unit(g.levels)
    tracer = DAGTracer(W, H)
    cam = cameras(p, 1, seed=5)[0]
    prm = ray_params(cam, info, g.levels, W, H)
    tracer.resolve_paths(cam, info, dag, params=prm)
    tracer.resolve_colors(dag)
    b = tracer.count_bytes(dag, prm)
    torch.cuda.synchronize()
    oparams = Oracle.ray_params(cam, info, g.levels, W, H)
    opaths, ob = Oracle.trace_paths(od, oparams, W, H, count=True)
    ocols = Oracle.trace_colors(od, oc, opaths, W, H)
    Oracle.set_any_hit_front_to_back(True)  # the device walks shadow rays front to back
    try:
        _, osb = Oracle.trace_shadows(od, oparams, opaths, ocols, W, H, 1.0, 0.0, count=True)
    finally:
        Oracle.set_any_hit_front_to_back(False)
    assert b[0] == ob
    assert b[2] == osb
    assert b[1] > 0


@pytest.mark.parametrize("schedule", [(0, 0), (1, 16), (2, 3), (3, 2 | (9 << 16))])
def test_fog_close_to_oracle(schedule):
    """With fog the shading uses exp/pow; device and host libm may differ in the last ulp,
    so RGB8 bytes are allowed to differ by at most 1 (both tracer schedules)."""
    import torch
    from hashdag_amd import DAGInfo, DAGTracer, ray_params
    from hashdag_amd.workload import cameras
    from oracle import Oracle

    p, sc, g, dag, od, oc = _setup("d8")
    W, H = 96, 64
    info = DAGInfo.unit(g.levels)
    tracer = DAGTracer(W, H)
    tracer.set_schedule(dag, *schedule)
    cam = cameras(p, 1, seed=3)[0]
    prm = ray_params(cam, info, g.levels, W, H)
    tracer.resolve_paths(cam, info, dag, params=prm)
    tracer.resolve_colors(dag)
    tracer.resolve_shadows(cam, info, dag, 1.0, 2000.0, params=prm)
    torch.cuda.synchronize()
    g8 = tracer.colors.cpu().numpy().view(np.uint8).reshape(H, W, 4).astype(int)
    oparams = Oracle.ray_params(cam, info, g.levels, W, H)
    opaths = Oracle.trace_paths(od, oparams, W, H)
    ocols = Oracle.trace_colors(od, oc, opaths, W, H)
    o8 = Oracle.trace_shadows(od, oparams, opaths, ocols, W, H, 1.0, 2000.0).view(np.uint8).reshape(H, W, 4).astype(int)
    assert np.abs(g8 - o8).max() <= 1


@pytest.mark.parametrize("tile", [(16, 4, 0, 0, 0), (4, 16, 1, 0, 0), (64, 1, 0, 0, 0), (8, 8, 1, 0, 0),
                                  (2, 32, 1, 0, 0), (8, 8, 0, 1, 16), (16, 4, 1, 1, 8), (64, 1, 0, 1, 1),
                                  (32, 2, 1, 1, 64), (2, 32, 0, 1, 24), (8, 8, 0, 2, 1), (8, 8, 1, 2, 6),
                                  (32, 2, 0, 2, 20), (16, 4, 1, 2, 128), (8, 8, 0, 3, (24, 128)),
                                  (16, 4, 0, 3, (1, 3)), (64, 1, 0, 3, (5, 400)), (8, 8, 1, 3, (8, 40)), "auto"])
def test_tile_shapes_bit_exact(tile):
    """Every wavefront tile shape, node fetch strategy and schedule (per tile; persistent
    wavefronts refilled at the given idle-lane count; per tile with walks beyond the given step
    count searched a wavefront per ray, down to every ray at 1 step; per tile to A steps with the
    saved walks resumed to B, down to 1 and 3), and the autotuner's pick, traces the same
    frame."""
    import torch
    from hashdag_amd import DAGInfo, DAGTracer, ray_params
    from hashdag_amd.tracer import autotune
    from hashdag_amd.workload import cameras
    from oracle import Oracle

    p, sc, g, dag, od, oc = _setup("d13")
    W, H = 100, 60  # not a multiple of the tile sizes
    info = DAGInfo.unit(g.levels)
    tracer = DAGTracer(W, H)
    cams = cameras(p, 2, seed=21)
    prms = [ray_params(c, info, g.levels, W, H) for c in cams]
    if tile == "auto":
        got = autotune(tracer, dag, prms, reps=1)
        assert got["tile"] in {"8x8", "16x4", "4x16", "32x2", "2x32", "64x1"}
    else:
        tracer.set_tile(dag, tile[0], tile[1])
        tracer.set_fetch(dag, tile[2])
        if tile[3] == 3:
            tracer.set_schedule(dag, 3, tile[4][0], resume_to=tile[4][1])
        else:
            tracer.set_schedule(dag, tile[3], tile[4] if tile[3] else 16)
    for cam, prm in zip(cams, prms):
        tracer.resolve_paths(cam, info, dag, params=prm)
        tracer.resolve_colors(dag)
        tracer.resolve_shadows(cam, info, dag, 1.0, 0.0, params=prm)
        torch.cuda.synchronize()
        gpu = tracer.colors.cpu().numpy().view(np.uint32)
        op = Oracle.ray_params(cam, info, g.levels, W, H)
        paths = Oracle.trace_paths(od, op, W, H)
        assert np.array_equal(tracer.paths.cpu().numpy().view(np.uint32).reshape(paths.shape), paths)
        ref = Oracle.trace_shadows(od, op, paths, Oracle.trace_colors(od, oc, paths, W, H), W, H, 1.0, 0.0)
        assert np.array_equal(gpu, ref)
