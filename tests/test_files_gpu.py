"""Checkpoint of a device context (saved and restored exactly, edits continue identically) and
the reference's HashDAG dump (hash_dag_factory.cpp:200-246): its layout, and loading it back as
the same DAG."""
import struct

import numpy as np
import pytest

from tests.scenes import POOL_PAGES, scene
from tests.test_edit_gpu import assert_same, make, random_edits

pytestmark = pytest.mark.gpu


def frame(dag, g, p):
    import torch
    from hashdag_amd import DAGInfo, DAGTracer, ray_params
    from hashdag_amd.serve import frame_checksum
    from hashdag_amd.workload import cameras
    info = DAGInfo.unit(g.levels)
    cam = cameras(p, 1, seed=3)[0]
    prm = ray_params(cam, info, g.levels, 128, 72)
    tr = DAGTracer(128, 72)
    tr.resolve_paths(None, info, dag, params=prm)
    tr.resolve_colors(dag)
    tr.resolve_shadows(None, info, dag, 1.0, 0.0, params=prm)
    torch.cuda.synchronize()
    return frame_checksum(dag, tr.colors, 128, 72)


def test_checkpoint_restores_table_and_colours(tmp_path):
    from hashdag_amd import HashDAG
    p, sc, g, dag, od, oc = make("d13")
    for kind, center, r in random_edits(p, 8, seed=51):
        dag.edit(kind, center, r)
        od.edit(kind, center, r, oc)
    path = str(tmp_path / "ctx.ckpt")
    dag.save_checkpoint(path)
    back = HashDAG(g, pool_pages=POOL_PAGES)
    back.load_checkpoint(path)
    a, b = dag.download(), back.download()
    assert a.root == b.root and a.pool_top == b.pool_top
    assert np.array_equal(a.page_table, b.page_table) and np.array_equal(a.bucket_sizes, b.bucket_sizes)
    assert np.array_equal(a.full_nodes, b.full_nodes)
    assert np.array_equal(a.pool[: a.pool_top * g.page_size], b.pool[: b.pool_top * g.page_size])
    assert_same(back, sc, od, oc, "restored checkpoint")
    assert frame(back, g, p) == frame(dag, g, p)
    # both keep editing the same way
    for kind, center, r in random_edits(p, 4, seed=52):
        back.edit(kind, center, r)
        od.edit(kind, center, r, oc)
    assert_same(back, sc, od, oc, "edits after restore")


def test_checkpoint_rejects_other_geometry(tmp_path):
    from hashdag_amd import HashDAG
    from hashdag_amd._lib import HashDAGError
    p, sc, g, dag, od, oc = make("d8")
    path = str(tmp_path / "d8.ckpt")
    dag.save_checkpoint(path)
    _, _, g13 = scene("d13")
    other = HashDAG(g13, pool_pages=POOL_PAGES)
    with pytest.raises(HashDAGError, match="geometry"):
        other.load_checkpoint(path)


def test_hash_dag_dump_layout_and_reload(tmp_path):
    from hashdag_amd import HashDAG
    p, sc, g, dag, od, oc = make("d13")
    for kind, center, r in random_edits(p, 6, seed=53):
        dag.edit(kind, center, r)
        od.edit(kind, center, r, oc)
    path = str(tmp_path / "scene.hash_dag.dag.bin")
    dag.save_hash_dag_file(path, (0.0, 0.0, 0.0), (8192.0, 8192.0, 8192.0))
    raw = open(path, "rb").read()
    img = dag.download()
    info = struct.unpack_from("<6d", raw, 0)
    levels, root, pool_max, pt_size, pool_top = struct.unpack_from("<5I", raw, 48)
    assert info == (0.0, 0.0, 0.0, 8192.0, 8192.0, 8192.0)
    assert (levels, root, pool_top, pt_size) == (g.levels, img.root, img.pool_top, g.total_pages())
    assert pool_max == dag.info().pool_pages
    pool_words = pool_top * g.page_size
    assert len(raw) == 68 + 4 * pool_words + 4 * pt_size
    assert np.array_equal(np.frombuffer(raw, np.uint32, pool_words, 68), img.pool[:pool_words])
    assert np.array_equal(np.frombuffer(raw, np.uint32, pt_size, 68 + 4 * pool_words), img.page_table)
    # loaded back: the same DAG (re-interned into a fresh table)
    back = HashDAG(g, pool_pages=POOL_PAGES)
    bounds = back.load_hash_dag_file(path)
    assert bounds == ((0.0, 0.0, 0.0), (8192.0, 8192.0, 8192.0))
    from oracle import OracleDAG
    assert np.array_equal(OracleDAG.from_image(back.download(), back.pool_pages).canonical(), od.canonical())


@pytest.mark.parametrize("name,top", [("d10w", 3), ("d13", 5)])
def test_reference_scene_files_trace_like_the_scene(tmp_path, name, top):
    """The reference's BasicDAG + compressed-colours files of a scene, loaded through the colour
    tree create_hash_dag_colors builds, trace to the same frame as the scene itself."""
    from hashdag_amd import HashDAG
    from hashdag_amd.files import load_reference_scene
    from tests.test_files import reference_format_files
    p, sc, g, od, canon, words, enclosed, dag_path, col_path = reference_format_files(tmp_path, name, top)
    back = load_reference_scene(dag_path, col_path, g.color_tree_levels)
    a = HashDAG(g, pool_pages=POOL_PAGES)
    a.load_scene(sc)
    b = HashDAG(g, pool_pages=POOL_PAGES)
    b.load_scene(back)
    assert frame(b, g, p) == frame(a, g, p)
    # and edits on it give the oracle's DAG and colours
    from oracle import OracleColors
    oc = OracleColors(back.color_nodes, back.color_offsets, back.blocks, back.weights, back.macro)
    for kind, center, r in random_edits(p, 6, seed=61):
        b.edit(kind, center, r)
        od.edit(kind, center, r, oc)
    assert_same(b, back, od, oc, "edits on the loaded reference files")
