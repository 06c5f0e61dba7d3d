"""Replay parsing (ReplayManager::load_csv, replay.h:565-607, and the scaling of replay.cpp:5)
against an independent reading of the same CSV files, and the request streams built from it.
The fixtures under tests/golden/replays are the reference's own replay files."""
import os

import numpy as np
import pytest

from hashdag_amd._lib import HashDAGError
from hashdag_amd.replay import COPY, Replay, replay_requests
from hashdag_amd.serve import HD_REQ_EDIT, HD_REQ_REDO, HD_REQ_TRACE, HD_REQ_UNDO

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(f for f in os.listdir(os.path.join(HERE, "golden", "replays")) if f.endswith(".csv"))

KINDS = {("EditSphere", True): 0, ("EditSphere", False): 1, ("EditCube", True): 2, ("EditCube", False): 3}


def expected_edits(path, scale):
    """Edits read with the csv rules written out independently: float32 cells times a float32
    scale, booleans 'true' / 'TRUE', trailing empty cells dropped."""
    f32 = np.float32
    out = []
    for line in open(path):
        cells = line.rstrip("\n").split(",")
        while cells and cells[-1] == "":
            cells.pop()
        name, d = cells[0], cells[1:]
        v = [f32(c) for c in d if c not in ("true", "false", "TRUE", "FALSE")]
        if name in ("EditSphere", "EditCube"):
            add = d[4] in ("true", "TRUE")
            out.append((KINDS[(name, add)], tuple(float(x * scale) for x in v[:3]), float(v[3] * scale)))
        elif name in ("EditPaint", "EditFill"):
            out.append((4 if name == "EditPaint" else 6, tuple(float(x * scale) for x in v[:3]), float(v[3] * scale)))
        elif name == "EditCopy":
            T = np.eye(3) if len(d) == 10 else np.array([float(x) for x in v[10:19]]).reshape(3, 3)
            swirl = None if len(d) == 10 or d[19] not in ("true", "TRUE") else float(v[19])
            out.append((COPY, tuple(float(x * scale) for x in v[0:3]), float(v[9] * scale),
                        tuple(float(x * scale) for x in v[3:6]), tuple(float(x * scale) for x in v[6:9]), T, swirl))
    return out


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("scene_depth", [17, 16, 13])
def test_replay_parse_matches_independent_reading(name, scene_depth):
    path = os.path.join(HERE, "golden", "replays", name)
    rep = Replay.load(path, scene_depth)
    lines = [ln.split(",")[0].strip() for ln in open(path) if ln.strip()]
    assert len(rep.actions) == len(lines)
    assert rep.frames == lines.count("EndFrame")
    c = rep.counts()
    assert c["undo"] == lines.count("Undo") and c["redo"] == lines.count("Redo")
    assert c["edits"] == sum(lines.count(k) for k in ("EditSphere", "EditCube", "EditPaint", "EditCopy", "EditFill"))
    scale = np.float32(2.0 ** scene_depth) / np.float32(2.0 ** 17)
    got = [a.edit for a in rep.actions if a.edit is not None and a.type == 3]
    want = expected_edits(path, scale)
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g[:3] == w[:3]
        if g[0] == COPY:
            assert g[3:5] == w[3:5] and np.array_equal(g[5], w[5]) and g[6] == w[6]
    # every action sits in the frame its EndFrame closes
    frames = [a.frame for a in rep.actions]
    assert frames == sorted(frames) and (not frames or frames[-1] == rep.frames - 1)


@pytest.mark.parametrize("text,msg", [
    ("EditSphere,1,2,3\n", "expected 5"),
    ("EditSphere,1,2,3,4,yes\n", "not a boolean"),
    ("Teleport,1,2,3\n", "unknown action"),
    ("SetLocation,1,x,3\n", "not a number"),
    ("EndFrame", "newline"),
])
def test_replay_parse_rejects_malformed_lines(text, msg):
    with pytest.raises(HashDAGError, match=msg):
        Replay.parse(text, 13)


def test_replay_trailing_empty_cells_and_missing_transform():
    rep = Replay.parse("EditSphere,1.5,2,3,4,TRUE,,,,\nEditCopy,1,2,3,4,5,6,7,8,9,2\nUndo\nRedo\nEndFrame\n", 17)
    assert rep.frames == 1 and len(rep.actions) == 5
    assert rep.actions[0].edit == (0, (1.5, 2.0, 3.0), 4.0)
    cp = rep.actions[1].edit
    assert cp[:5] == (COPY, (1.0, 2.0, 3.0), 2.0, (4.0, 5.0, 6.0), (7.0, 8.0, 9.0))
    assert np.array_equal(cp[5], np.eye(3)) and cp[6] is None


def test_replay_requests_follow_frames():
    text = ("SetLocation,1,2,3\nSetRotation,1,0,0,0,1,0,0,0,1\nEditSphere,10,10,10,4,true\nEndFrame\n"
            "EditCube,20,20,20,3,false\nUndo\nRedo\nEndFrame\nEditFill,1,1,1,1\nEndFrame\n")
    rep = Replay.parse(text, 13, 13)
    reqs = replay_requests(rep, 64, 32, 13, fps=10.0)
    kinds = [r.kind for r in reqs]
    assert kinds == [HD_REQ_EDIT, HD_REQ_TRACE, HD_REQ_EDIT, HD_REQ_UNDO, HD_REQ_REDO, HD_REQ_TRACE, HD_REQ_EDIT,
                     HD_REQ_TRACE]
    assert [r.arrival_s for r in reqs] == [0.0, 0.0, 0.1, 0.1, 0.1, 0.1, 0.2, 0.2]
    assert reqs[0].edits == [(0, (10.0, 10.0, 10.0), 4.0)]
    assert reqs[6].edits == [(6, (1.0, 1.0, 1.0), 1.0)]
