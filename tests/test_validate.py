"""Host checks of external images (hd_validate_image, run by hd_upload_image and
hd_checkpoint_load before any copy): a valid image passes; each kind of corruption that would
send a kernel outside its buffers is rejected with a reason. CPU only: the check needs no
device. The image comes from the oracle's table built from a test scene."""
import ctypes as C

import numpy as np
import pytest

from tests.scenes import POOL_PAGES, scene


def _config(g, pool_pages):
    from hashdag_amd._lib import hd_config, lib
    cfg = hd_config()
    lib.hd_default_config(g.levels, C.byref(cfg))
    for k in ("page_size", "num_top_levels", "bucket_bits_top", "bucket_bits_low", "bucket_size_top",
              "bucket_size_low", "color_tree_levels"):
        setattr(cfg, k, getattr(g, k))
    cfg.pool_pages = pool_pages
    return cfg


def validate(g, pool, pt, bs, full, top, root):
    from hashdag_amd._lib import lib, ptr
    cfg = _config(g, POOL_PAGES)
    why = C.create_string_buffer(256)
    arrs = [np.ascontiguousarray(a, np.uint32) for a in (pool, pt, bs, full)]
    st = lib.hd_validate_image(C.byref(cfg), *(ptr(a) for a in arrs), top, root, why, 256)
    return st, why.value.decode()


@pytest.fixture(scope="module")
def image():
    from oracle import OracleDAG
    p, sc, g = scene("d8")
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    pool, pt, bs, full, top = od.export()
    return g, pool[: top * g.page_size].copy(), pt, bs, full, top, od.root


def test_valid_image_passes(image):
    g, pool, pt, bs, full, top, root = image
    assert validate(g, pool, pt, bs, full, top, root) == (0, "")


def _first_child_word(g, pool, pt, root):
    """Physical index of the root's first child pointer."""
    va = root + 1
    return int(pt[va // g.page_size]) * g.page_size + va % g.page_size


@pytest.mark.parametrize("what", ["page_above_top", "page_twice", "bucket_over", "unmapped_used_page", "root_out",
                                  "child_out", "child_wrong_level", "full_out", "pool_top_zero"])
def test_corrupt_image_rejected(image, what):
    g, pool, pt, bs, full, top, root = image
    pool, pt, bs, full = pool.copy(), pt.copy(), bs.copy(), full.copy()
    mapped = np.flatnonzero(pt)
    if what == "page_above_top":
        pt[mapped[0]] = top
    elif what == "page_twice":
        pt[mapped[1]] = pt[mapped[0]]
    elif what == "bucket_over":
        bs[int(np.flatnonzero(bs)[0])] = g.bucket_size_top + 1
    elif what == "unmapped_used_page":
        pt[mapped[-1]] = 0
    elif what == "root_out":
        root = g.total_virtual()
    elif what == "child_out":
        pool[_first_child_word(g, pool, pt, root)] = g.total_virtual() + 5
    elif what == "child_wrong_level":
        pool[_first_child_word(g, pool, pt, root)] = root  # a level-0 address as a level-1 child
    elif what == "full_out":
        full[g.leaf_level] = 0xFFFFFFF0
    elif what == "pool_top_zero":
        top = 0
    st, why = validate(g, pool, pt, bs, full, top, root)
    assert st == -1 and why, f"{what} passed"


def test_color_tree_levels_range():
    """A full node's 24-bit voxel count limits the colour tree to at least levels - 7
    (hash_table.h:749-789)."""
    from hashdag_amd import Geometry
    Geometry(levels=17, color_tree_levels=10)
    with pytest.raises(ValueError):
        Geometry(levels=17, color_tree_levels=9)
    with pytest.raises(ValueError):
        Geometry(levels=12, color_tree_levels=0)
