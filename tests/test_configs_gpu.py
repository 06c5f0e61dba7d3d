"""GPU parity at the benchmark configurations' own shapes.

- The traced workload at full size: the 2^15 coloured terrain at 1920x1080, four of the
  bench's 64 seeded cameras; paths, colours and shadowed colours bit-exact against the oracle
  (tracer.cu:145-697).
- The CPU-plain workload: the 2^12 binary terrain with the colour tree down to the leaf level
  (no colour leaves; hash_dag_edits.h:181-185), 100 sphere edits (50 add / 50 carve, radius
  uniform 8-64) one pass each; canonical DAG and edited-voxel counts against the oracle.
- The depth-17 workload on a small footprint: instanced boxes and spheres in one fused pass,
  translation copies, large carves replayed through a captured frame graph, GC keeping the undo
  history, and a 3840x2160 frame; DAG, colour leaves and frame against the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _frames_match(dag, od, oc, cams, levels, W, H, what):
    """Each camera's paths, colours and shaded colours against the oracle, traced with one
    wavefront per tile, with persistent wavefronts, and per tile with walks beyond 128 (and 24)
    steps searched a wavefront per ray."""
    import torch

    from hashdag_amd import DAGInfo, DAGTracer, ray_params
    from oracle import Oracle

    info = DAGInfo.unit(levels)
    tracer = DAGTracer(W, H)
    for k, cam in enumerate(cams):
        prm = ray_params(cam, info, levels, W, H)
        oparams = Oracle.ray_params(cam, info, levels, W, H)
        assert np.array_equal(np.array(prm.as_list()), oparams), f"{what}: ray setup differs"
        opaths = Oracle.trace_paths(od, oparams, W, H, threads=16)
        assert (opaths[..., :3] != 0).any(-1).mean() > 0.05, f"{what} camera {k} sees almost nothing"
        ocols = Oracle.trace_colors(od, oc, opaths, W, H, threads=16)
        oshad = Oracle.trace_shadows(od, oparams, opaths, ocols, W, H, 1.0, 0.0, threads=16)
        for schedule in ((0, 0), (1, 16), (2, 128), (2, 24), (3, 32 | (128 << 16))):
            print(f"[{what}] camera {k} {W}x{H} schedule={schedule}", flush=True)
            tracer.set_schedule(dag, *schedule)
            tracer.resolve_paths(cam, info, dag, params=prm)
            tracer.resolve_colors(dag)
            torch.cuda.synchronize()
            gpaths = tracer.paths.cpu().numpy().view(np.uint32)
            gcols = tracer.colors.cpu().numpy().view(np.uint32)
            tracer.resolve_shadows(cam, info, dag, 1.0, 0.0, params=prm)
            torch.cuda.synchronize()
            gshad = tracer.colors.cpu().numpy().view(np.uint32)
            sched = f"schedule {schedule}"
            bad = (gpaths != opaths).any(-1)
            assert not bad.any(), f"{what} camera {k} {sched}: {int(bad.sum())} paths differ, first {np.argwhere(bad)[:3]}"
            assert np.array_equal(gcols, ocols), f"{what} camera {k} {sched}: {int((gcols != ocols).sum())} colours differ"
            assert np.array_equal(gshad, oshad), \
                f"{what} camera {k} {sched}: {int((gshad != oshad).sum())} shaded pixels differ"
            del tracer
            tracer = DAGTracer(W, H)
    tracer.set_schedule(dag, 0)


def test_full_size_frames_bit_exact():
    """2^15 coloured terrain, 1920x1080, bench cameras 0, 21, 42 and 63 (seed 1337)."""
    from hashdag_amd import Geometry, HashDAG, TerrainParams, generate_terrain
    from hashdag_amd.workload import cameras
    from oracle import OracleColors, OracleDAG

    D, W, H = 15, 1920, 1080
    params = TerrainParams(levels=D, seed=1337, colors=True)
    sc = generate_terrain(params)
    try:
        g = Geometry.fit(D, sc.level_words, headroom=1.5)
    except ValueError:
        g = Geometry.fit(D, sc.level_words)
    dag = HashDAG(g)
    dag.load_scene(sc)
    od = OracleDAG.from_basic(g, sc.dag, 0x8FFFFFFF // 512 // 4)
    oc = OracleColors.from_scene(sc)
    cams = cameras(params, 64, seed=1337)
    _frames_match(dag, od, oc, [cams[i] for i in (0, 21, 42, 63)], D, W, H, "2^15 1080p")
    dag.close()


def test_binary_depth12_plain_sphere_edits():
    """2^12 binary terrain (colours off, colour tree down to the leaf level), 100 sphere edits
    (50 add / 50 carve, radius uniform 8-64) at seeded surface points, one pass each."""
    from hashdag_amd import Geometry, HashDAG, TerrainParams, generate_terrain
    from hashdag_amd.workload import sphere_edits
    from oracle import OracleDAG

    params = TerrainParams(levels=12, seed=1337, colors=False, color_tree_levels=10)
    sc = generate_terrain(params)
    g = Geometry.fit(12, sc.level_words, color_tree_levels=10, headroom=1.5)
    assert g.color_tree_levels == g.leaf_level
    dag = HashDAG(g, pool_pages=1 << 19)
    dag.load_basic_dag(sc.dag)
    od = OracleDAG.from_basic(g, sc.dag, 1 << 19)
    edits = sphere_edits(params, 100, seed=1337 + 9)
    assert sum(k == 0 for k, _, _ in edits) == 50
    for k, (kind, center, r) in enumerate(edits):
        st = dag.edit(kind, center, r)
        ost = od.edit(kind, center, r, None, threads=8)
        assert st.num_voxels == ost[1], f"edit {k}: voxel count {st.num_voxels} != {ost[1]}"
        assert st.leaves_rebuilt == 0
        if k in (9, 99):
            from oracle import OracleDAG as OD
            god = OD.from_image(dag.download(), dag.pool_pages)
            assert np.array_equal(god.canonical(), od.canonical()), f"after edit {k}: DAG differs"
    dag.close()


def test_depth17_instances_copies_carves_graph_gc_4k():
    import torch

    from hashdag_amd import DAGInfo, DAGTracer, HashDAG, TerrainParams, generate_terrain, ray_params
    from hashdag_amd.graph import FrameGraph
    from hashdag_amd.workload import cameras, carves, copy_edits, instances
    from oracle import OracleColors, OracleDAG
    from tests.scenes import geometry_for
    from tests.test_edit_gpu import assert_same

    D = 17
    params = TerrainParams(levels=D, footprint=512, seed=1337, colors=True, color_tree_levels=10, threads=8)
    sc = generate_terrain(params)
    g = geometry_for(D, 10)
    pool_pages = 1 << 15
    dag = HashDAG(g, pool_pages=pool_pages)
    dag.load_scene(sc)
    od = OracleDAG.from_basic(g, sc.dag, pool_pages)
    oc = OracleColors.from_scene(sc)

    print("[depth 17] instances", flush=True)
    inst = instances(params, 40, seed=1342, hmin=16.0, hmax=128.0)
    st = dag.edit_batch(inst, fused=True)
    assert st.passes == 1
    assert st.num_voxels == sum(od.edit(k, c, r, oc)[1] for k, c, r in inst)
    assert_same(dag, sc, od, oc, "depth 17 instances")

    print("[depth 17] copies", flush=True)
    cps = copy_edits(params, 3, seed=1343, radius=32.0)
    st = dag.edit_batch(cps, fused=True)
    assert st.passes == 3
    voxels = 0
    for _, center, r, src, dst in cps:
        voxels += od.edit_copy(center, r, src, dst, oc)[1]
    assert st.num_voxels == voxels
    assert_same(dag, sc, od, oc, "depth 17 copies")

    print("[depth 17] carves through the frame graph", flush=True)
    W, H = 3840, 2160
    cam = cameras(params, 1, seed=1344)[0]
    info = DAGInfo.unit(D)
    prm = ray_params(cam, info, D, W, H)
    tr = DAGTracer(W, H)
    cv = carves(params, 4, seed=1345, rmin=64.0, rmax=128.0)
    before = OracleDAG.from_image(dag.download(), dag.pool_pages).canonical()
    dag.edit_batch(cv, fused=True)  # sizes the scratch the graph bakes in
    dag.undo()
    fg = FrameGraph(dag, tr, 8)
    fg.launch([], prm)
    st, replayed = fg.launch(cv, prm)
    torch.cuda.synchronize()
    fg.close()
    assert replayed
    assert st.num_voxels == sum(od.edit(k, c, r, oc)[1] for k, c, r in cv)
    assert_same(dag, sc, od, oc, "depth 17 carves (graph)")

    print("[depth 17] GC keeping undo", flush=True)
    depth = dag.undo_depth()
    dag.remove_stale_nodes(keep_undo=True)
    assert dag.undo_depth() == depth
    assert_same(dag, sc, od, oc, "depth 17 after GC")

    print("[depth 17] 3840x2160 frame", flush=True)
    _frames_match(dag, od, oc, [cam], D, W, H, "depth 17 4K")
    dag.undo()
    after_undo = OracleDAG.from_image(dag.download(), dag.pool_pages).canonical()
    assert np.array_equal(after_undo, before), "undo after GC does not restore the DAG before the carves"
    dag.redo()
    assert_same(dag, sc, od, oc, "depth 17 redo after GC")
    dag.close()
