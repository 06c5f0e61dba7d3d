"""The reference's BasicDAG file layout (basic_dag.cpp:89-117 with serializer.h:11-122) written
and read by the library, checked against the layout decoded independently with struct."""
import struct

import numpy as np
import pytest

from hashdag_amd._lib import HashDAGError
from hashdag_amd.files import read_basic_dag_file, write_basic_dag_file
from tests.scenes import scene


def test_basic_dag_file_layout_and_round_trip(tmp_path):
    p, sc, g = scene("d8")
    path = str(tmp_path / "scene.basic_dag.dag.bin")
    bmin, bmax = (-1.5, 0.0, 2.0), (255.0, 256.0, 257.5)
    write_basic_dag_file(path, bmin, bmax, g.levels, sc.dag)
    raw = open(path, "rb").read()
    info = struct.unpack_from("<6d", raw, 0)
    levels, n = struct.unpack_from("<IQ", raw, 48)
    assert info == bmin + bmax and levels == g.levels and n == sc.dag.size
    assert len(raw) == 48 + 4 + 8 + 4 * n
    assert np.array_equal(np.frombuffer(raw, np.uint32, n, 60), sc.dag)
    got = read_basic_dag_file(path)
    assert got[0] == bmin and got[1] == bmax and got[2] == g.levels
    assert np.array_equal(got[3], sc.dag)


def test_basic_dag_file_truncated(tmp_path):
    path = str(tmp_path / "t.bin")
    write_basic_dag_file(path, (0, 0, 0), (1, 1, 1), 5, np.arange(10, dtype=np.uint32))
    raw = open(path, "rb").read()
    open(path, "wb").write(raw[:-4])
    with pytest.raises(HashDAGError, match="truncated"):
        read_basic_dag_file(path)
    with pytest.raises(HashDAGError, match="cannot open"):
        read_basic_dag_file(str(tmp_path / "missing.bin"))


# ---- compressed colours (basic_dag.cpp:34-89) and the colour tree over them ----

def test_compressed_colors_file_layout_and_round_trip(tmp_path):
    from hashdag_amd.files import read_compressed_colors_file, write_compressed_colors_file
    rng = np.random.default_rng(5)
    e = rng.integers(0, 1 << 40, 7, dtype=np.uint64)
    w = rng.integers(0, 1 << 32, 33, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 1 << 63, 11, dtype=np.uint64)
    m = rng.integers(0, 1 << 20, 4, dtype=np.uint64)
    path = str(tmp_path / "scene.compressed_colors.variable.bin")
    write_compressed_colors_file(path, 3, e, w, b, m)
    raw = open(path, "rb").read()
    # FileWriter: uint32 topLevels, then per array a uint64 count and the data
    top, ne = struct.unpack_from("<IQ", raw, 0)
    assert top == 3 and ne == e.size
    at = 12
    assert np.array_equal(np.frombuffer(raw, np.uint64, ne, at), e)
    at += 8 * ne
    nw = struct.unpack_from("<Q", raw, at)[0]
    at += 8
    assert nw == w.size
    on_disk = np.frombuffer(raw, np.uint32, nw, at)
    assert np.array_equal(on_disk, w.byteswap())  # stored in the default (swapped) byte order
    at += 4 * nw
    nb = struct.unpack_from("<Q", raw, at)[0]
    assert nb == b.size and np.array_equal(np.frombuffer(raw, np.uint64, nb, at + 8), b)
    at += 8 + 8 * nb
    nm = struct.unpack_from("<Q", raw, at)[0]
    assert nm == m.size and np.array_equal(np.frombuffer(raw, np.uint64, nm, at + 8), m)
    assert len(raw) == at + 8 + 8 * nm
    got = read_compressed_colors_file(path)
    assert got[0] == 3
    for x, y in zip(got[1:], (e, w, b, m)):
        assert np.array_equal(x, y)


def test_compressed_colors_file_truncated(tmp_path):
    from hashdag_amd.files import read_compressed_colors_file, write_compressed_colors_file
    path = str(tmp_path / "c.bin")
    write_compressed_colors_file(path, 1, np.ones(3, np.uint64), np.ones(5, np.uint32), np.ones(2, np.uint64),
                                 np.ones(2, np.uint64))
    raw = open(path, "rb").read()
    open(path, "wb").write(raw[:-5])
    with pytest.raises(HashDAGError, match="truncated"):
        read_compressed_colors_file(path)


def reference_format_files(tmp_path, name, top_levels):
    """The scene written as the reference's two files: the BasicDAG with voxel-count headers
    and the compressed colours of every voxel in DFS order as one leaf (encoded by the oracle's
    restatement of the reference encoder, from the oracle's decode of the scene's colours)."""
    import ctypes as C

    from hashdag_amd.files import basic_dag_counts, write_compressed_colors_file
    from oracle import OracleColors, OracleDAG
    from oracle.binding import _p, orc
    from tests.scenes import POOL_PAGES
    p, sc, g = scene(name)
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    canon = oc.canonical(od)
    bits, wb = [], []
    i = 0
    while i < canon.size:  # per colour subtree in DFS order: x, y, z, n, then n (colour, weight | bpw << 8)
        n = int(canon[i + 3])
        seq = canon[i + 4: i + 4 + 2 * n].reshape(n, 2)
        bits.append(seq[:, 0])
        wb.append(seq[:, 1])
        i += 4 + 2 * n
    bits = np.concatenate(bits).astype(np.uint32)
    wb = np.concatenate(wb).astype(np.uint32)
    w8, bpw = (wb & 0xFF).astype(np.uint8), (wb >> 8).astype(np.uint8)
    n = bits.size
    capB, capW, capM = n + 16, n + 16, 2 * (n // 16384 + 2)
    blocks, weights, macro = np.zeros(capB, np.uint64), np.zeros(capW, np.uint32), np.zeros(capM, np.uint64)
    sizes = np.zeros(3, np.uint64)
    assert orc.orc_encode_colors(_p(bits), _p(w8, C.c_uint8), _p(bpw, C.c_uint8), n, _p(blocks, C.c_uint64),
                                 _p(weights), _p(macro, C.c_uint64), _p(sizes, C.c_uint64), capB, capW, capM) == 0
    blocks, weights, macro = blocks[: sizes[0]], weights[: sizes[1]], macro[: sizes[2]]
    words, enclosed = basic_dag_counts(sc.dag, g.levels, top_levels)
    dag_path, col_path = str(tmp_path / "s.basic_dag.dag.bin"), str(tmp_path / "s.basic_dag.compressed_colors.bin")
    write_basic_dag_file(dag_path, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), g.levels, words)
    write_compressed_colors_file(col_path, top_levels, enclosed, weights, blocks, macro)
    return p, sc, g, od, canon, words, enclosed, dag_path, col_path


@pytest.mark.parametrize("name,top", [("d8", 0), ("d8", 2), ("d10w", 3), ("d13", 5)])
def test_reference_scene_files_give_the_same_scene(tmp_path, name, top):
    from hashdag_amd.files import load_reference_scene
    from oracle import OracleColors
    p, sc, g, od, canon, words, enclosed, dag_path, col_path = reference_format_files(tmp_path, name, top)
    # headers: counts below top_levels' indirection, the subtree's voxels
    assert enclosed.size > 0 if top else enclosed.size == 0
    root_count = int(enclosed[words[0] >> 8]) if top else int(words[0] >> 8)
    assert root_count == sc.num_voxels
    back = load_reference_scene(dag_path, col_path, g.color_tree_levels)
    assert back.num_voxels == sc.num_voxels
    # the headers above the colour tree reduce to the child mask: the scene's own words
    assert np.array_equal(back.dag, sc.dag)
    # the colour tree built as create_hash_dag_colors does, over the one contiguous leaf,
    # decodes to the scene's colours subtree by subtree
    oc2 = OracleColors(back.color_nodes, back.color_offsets, back.blocks, back.weights, back.macro)
    assert np.array_equal(oc2.canonical(od), canon)
    # handle 0 stays "no leaf"; the builder's leaves follow it
    assert back.color_offsets[0] == 0 and back.color_offsets.size > 1


def full_chain(levels):
    """A BasicDAG of full nodes: every interior node has 8 children, all the same node."""
    words = []
    for level in range(levels - 2):
        child = len(words) + 9
        words += [0xFF] + [child] * 8
    words += [0xFFFFFFFF, 0xFFFFFFFF]
    return np.array(words, np.uint32)


def test_basic_dag_counts_of_full_nodes():
    from hashdag_amd.files import basic_dag_counts
    words = full_chain(6)  # levels 0..3 interior, leaf level 4: 64 * 8^k voxels at depth 4 - k
    out, enc = basic_dag_counts(words, 6, 0)
    assert [int(out[9 * k]) >> 8 for k in range(4)] == [64 * 8 ** (4 - k) for k in range(4)]
    assert [int(out[9 * k]) & 0xFF for k in range(4)] == [0xFF] * 4 and enc.size == 0
    out, enc = basic_dag_counts(words, 6, 2)  # levels 0 and 1 through enclosedLeaves
    assert list(enc) == [64 * 8 ** 4, 64 * 8 ** 3]
    assert int(out[0]) >> 8 == 0 and int(out[9]) >> 8 == 1 and int(out[18]) >> 8 == 64 * 8 ** 2


def test_basic_dag_counts_needs_indirection_above_24_bits():
    from hashdag_amd.files import basic_dag_counts
    words = full_chain(9)  # root: 64 * 8^7 = 2^27 voxels
    with pytest.raises(HashDAGError, match="does not fit"):
        basic_dag_counts(words, 9, 0)
    out, enc = basic_dag_counts(words, 9, 3)
    assert list(enc) == [2 ** 27, 2 ** 24, 2 ** 21] and int(out[27]) >> 8 == 2 ** 18
