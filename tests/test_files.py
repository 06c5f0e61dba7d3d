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
