"""The edit entry point that allocates nothing (hd_edit_batch_static) against hd_edit_batch
(fused) on the same scene: two calls in a row, the second with fewer editors. After each call
both contexts hold canonically the same DAG, the same colours per subtree and the same stored
colour-leaf words, with the same edited-voxel and rebuilt-leaf counts; both also match the
oracle applying the same editors one by one. A pass larger than the static scratch fails with
HD_ERR_OOM and leaves the DAG and colours as they were.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_static_batch_matches_fused_batch_twice():
    from hashdag_amd import HashDAG, terrain_height
    from tests.scenes import POOL_PAGES
    from tests.test_edit_gpu import assert_same, gpu_state_in_oracle, make, random_edits

    p, sc, g, dag, od, oc = make("d13")
    other = HashDAG(g, pool_pages=POOL_PAGES)
    other.load_scene(sc)
    for call, (n, seed, rmax) in enumerate(((12, 101, 8), (4, 102, 6))):
        print(f"[no-alloc] call {call}: {n} editors (seed {seed}, radius <= {rmax})", flush=True)
        edits = random_edits(p, n, seed=seed, rmax=rmax)
        st = other.edit_batch_static(edits)
        ref = dag.edit_batch(edits, fused=True)
        total = sum(od.edit(kind, center, r, oc)[1] for kind, center, r in edits)
        assert st.passes == 1 and ref.passes == 1
        assert st.num_voxels == ref.num_voxels == total, f"call {call}: voxel counts differ"
        assert st.leaves_rebuilt == ref.leaves_rebuilt, f"call {call}: leaves rebuilt differ"
        assert st.colors_encoded == ref.colors_encoded, f"call {call}: colours encoded differ"
        assert other.undo_depth() == (0, 0)
        assert_same(other, sc, od, oc, f"static call {call}")
        assert_same(dag, sc, od, oc, f"fused call {call}")
    # a pass beyond the static scratch (an add sphere of radius 60 touches far more than 4096
    # leaves) fails and keeps the DAG and its colours as they were
    print("[no-alloc] oversized pass", flush=True)
    before, bcol = gpu_state_in_oracle(other, sc)
    x, z = 128, 128
    big = [(0, (float(x), float(terrain_height(p, x, z)), float(z)), 60.0)]
    from hashdag_amd._lib import HashDAGError
    try:
        other.edit_batch_static(big)
        raise AssertionError("oversized static pass did not fail")
    except HashDAGError as e:
        assert e.status == -5, e  # HD_ERR_OOM
    after, acol = gpu_state_in_oracle(other, sc)
    assert np.array_equal(after.canonical(), before.canonical())
    assert np.array_equal(acol.canonical(after), bcol.canonical(before))
    other.close()
    dag.close()


def test_static_batch_rejects_oversized_input():
    import ctypes as C

    from hashdag_amd._lib import HD_ERR_INVALID, hd_edit, hd_edit_stats, lib
    from tests.test_edit_gpu import make

    p, sc, g, dag, od, oc = make("d8")
    n = 257  # HD_STATIC_MAX_EDITS + 1
    arr = (hd_edit * n)()
    for i in range(n):
        arr[i].kind = 0
        arr[i].radius = 2.0
    st = hd_edit_stats()
    assert lib.hd_edit_batch_static(dag.handle, arr, n, C.byref(st), None) == HD_ERR_INVALID
    assert b"HD_STATIC_MAX_EDITS" in lib.hd_last_error(dag.handle)
    dag.close()
