"""Independent checks that pin oracle functions whose exact behaviour a brute-force nearest-hit
or a round trip leaves open:

* trace_paths' hit order: a DFS over the dense occupancy pyramid that visits children in the
  reference's order (child ^ order, child = 0..7, tracer.cu:7-17) and decides every child by an
  exact FP64 ray / box slab test in place of the FP32 mask with its epsilon (tracer.cu:19-136).
  On every ray that never comes within 2e-3 voxels of a decision boundary it must give the
  oracle's hit voxel, the reference's (0,0,0)-means-miss quirk included (path.h).
* ColorLeafBuilder's bytes (variable_weight_size_colors.h:557-721): a Python encoder written
  from the format (a macro block per 16384 colours holding its first block and first weight
  bit; a block per run of equal colour bits and bits per weight inside a macro block, header =
  weight bit offset in the macro block << 16 | (bpw - 1) << 14 | colour index in the macro
  block, 0xFFFF for bpw 0; weights as one MSB-first bit stream in 32-bit words stored
  byte-swapped) must give the oracle encoder's blocks, weight words and macro words exactly.
* trace_colors' fallbacks (tracer.cu:270-330): sky = float3_to_rgb888((187, 242, 250) / 255)
  for a null path, 0xFF00FF (no alpha) for a path through a missing child, the checker
  float3_to_rgb888((1, b, 1 - b)), b = (x ^ y ^ z) & 1, for a missing colour node; evaluated
  here in IEEE float32 with numpy.
* the interior find's walk (hash_table.h:276-357) on synthetic buckets whose nodes end
  exactly at page ends and at the bucket end, against a walk written here from the
  reference text: per page, compare nodes while start < min(bucket size, page end) - n,
  stepping by each node's size; a page where pageStart + n >= that end ends the search.
* the copy swirl (hash_dag_editors.h:364-379): a dense-grid brute force turning each offset
  about y in FP64, away from the voxels whose source coordinate is within 1e-3 of an integer.
* applyFog (tracer.cu:551-572): a saturating density (every pixel its fog colour) and, per
  pixel, the density at which exp gives 1/2 (the midpoint of the fog-free and fog colours).
"""
import ctypes as C
import math

import numpy as np
import pytest

from hashdag_amd import CameraView, DAGInfo, Geometry
from hashdag_amd.dag import HostImage
from hashdag_amd.workload import cameras
from oracle import Oracle, OracleColors, OracleDAG, orc
from oracle.binding import _p
from tests.scenes import POOL_PAGES, scene

# ---------------------------------------------------------------------------------------------
# trace_paths: reference-order DFS with exact slab tests
# ---------------------------------------------------------------------------------------------


def pyramid(grid):
    occ = [grid.astype(bool)]
    while occ[0].shape[0] > 1:
        a = occ[0]
        n = a.shape[0] // 2
        occ.insert(0, a.reshape(n, 2, n, 2, n, 2).any(axis=(1, 3, 5)))
    return occ


def slab(o, d, lo, hi):
    te, tx = -math.inf, math.inf
    for a in range(3):
        t0 = (lo[a] - o[a]) / d[a]
        t1 = (hi[a] - o[a]) / d[a]
        te = max(te, min(t0, t1))
        tx = min(tx, max(t0, t1))
    return max(te, 0.0), tx


class Grazing(Exception):
    pass


def dfs_first_hit(occ, o, d, delta):
    """First voxel reached by the reference-order DFS, or None; raises Grazing when a decision
    on an existing node lies within delta of its boundary."""
    levels = len(occ) - 1
    order = (4 if d[0] < 0 else 0) | (2 if d[1] < 0 else 0) | (1 if d[2] < 0 else 0)
    E = float(1 << levels)
    te, tx = slab(o, d, (0.0, 0.0, 0.0), (E, E, E))
    if abs(tx - te) < delta:
        raise Grazing()
    if not te < tx:
        return None

    def visit(level, pos):
        s = float(1 << (levels - level - 1))
        for c in range(8):
            k = c ^ order
            cp = (2 * pos[0] + ((k >> 2) & 1), 2 * pos[1] + ((k >> 1) & 1), 2 * pos[2] + (k & 1))
            if not occ[level + 1][cp]:
                continue
            lo = (cp[0] * s, cp[1] * s, cp[2] * s)
            a, b = slab(o, d, lo, (lo[0] + s, lo[1] + s, lo[2] + s))
            if abs(b - a) < delta:
                raise Grazing()
            if a < b:
                if level + 1 == levels:
                    return cp
                r = visit(level + 1, cp)
                if r is not None:
                    return r
        return None

    return visit(0, (0, 0, 0))


def fp32_ray(params, px, py):
    """The reference's FP32 origin and direction of pixel (px, py) (tracer.cu:160-170)."""
    cam, rmin, ddx, ddy = params[0:3], params[3:6], params[6:9], params[9:12]
    v = [rmin[i] + px * ddx[i] + py * ddy[i] - cam[i] for i in range(3)]
    s = 1.0 / math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    d = [float(np.float32(s * c)) for c in v]
    o = [float(np.float32(c)) for c in cam]
    return o, d


@pytest.mark.parametrize("name,ncam", [("d7", 3), ("d8s", 2)])
def test_dfs_order_matches_exact_reference_order_walk(name, ncam):
    p, sc, g = scene(name)
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    occ = pyramid(od.dense())
    W, H = 40, 30
    agree = grazing = total = hits = 0
    for cam in cameras(p, ncam, seed=41):
        prm = Oracle.ray_params(cam, DAGInfo.unit(g.levels), g.levels, W, H)
        paths = Oracle.trace_paths(od, prm, W, H)
        for py in range(H):
            for px in range(W):
                o, d = fp32_ray(prm, px, py)
                total += 1
                try:
                    hit = dfs_first_hit(occ, o, d, 2e-3)
                except Grazing:
                    grazing += 1
                    continue
                want = hit if hit is not None and hit != (0, 0, 0) else (0, 0, 0)
                got = tuple(int(c) for c in paths[H - 1 - py, px, :3])
                hits += want != (0, 0, 0)
                assert got == want, f"pixel ({px}, {py}): oracle {got}, reference-order walk {want}"
                agree += 1
    assert grazing <= 0.02 * total, f"{grazing} of {total} rays graze a boundary"
    assert hits > 0.3 * agree


# ---------------------------------------------------------------------------------------------
# ColorLeafBuilder bytes
# ---------------------------------------------------------------------------------------------

MACRO = 16384


def model_encode(colors):
    """(blocks, weight words as stored, macro words) of a colour sequence [(colorBits, weight, bpw)]."""
    blocks, macro, stream = [], [], []
    macro_color = macro_bit = 0
    last = None
    for i, (cb, w, bpw) in enumerate(colors):
        if i % MACRO == 0:
            macro_color, macro_bit = i, len(stream)
            macro += [len(blocks), len(stream)]
        if i % MACRO == 0 or last != (cb, bpw):
            off = 0xFFFF if bpw == 0 else len(stream) - macro_bit
            assert off <= 0xFFFF
            header = (off << 16) | ((bpw - 1 if bpw else 0) << 14) | (i - macro_color)
            blocks.append((cb << 32) | header)
            last = (cb, bpw)
        stream += [(w >> k) & 1 for k in range(bpw - 1, -1, -1)]
    words = []
    for k in range(0, len(stream), 32):
        bits = stream[k:k + 32] + [0] * (32 - len(stream[k:k + 32]))
        v = 0
        for b in bits:
            v = (v << 1) | b
        words.append(int.from_bytes(v.to_bytes(4, "big"), "little"))  # stored byte-swapped
    return (np.array(blocks, np.uint64), np.array(words, np.uint32), np.array(macro, np.uint64))


def oracle_encode(colors):
    n = len(colors)
    cb = np.array([c[0] for c in colors], np.uint32)
    w = np.array([c[1] for c in colors], np.uint8)
    b = np.array([c[2] for c in colors], np.uint8)
    cap = n + 16
    blocks, weights, macro = np.zeros(cap, np.uint64), np.zeros(cap, np.uint32), np.zeros(2 * cap, np.uint64)
    sizes = np.zeros(3, np.uint64)
    assert orc.orc_encode_colors(_p(cb), _p(w, C.c_uint8), _p(b, C.c_uint8), n, _p(blocks, C.c_uint64), _p(weights),
                                 _p(macro, C.c_uint64), _p(sizes, C.c_uint64), cap, cap, 2 * cap) == 0
    return blocks[:sizes[0]], weights[:sizes[1]], macro[:sizes[2]]


def random_colors(rng, n, run, bpws):
    out = []
    palette = [int(v) for v in rng.integers(0, 1 << 32, size=6, dtype=np.uint64)]
    cb, bpw = palette[0], int(rng.choice(bpws))
    for _ in range(n):
        if rng.random() < 1.0 / run:
            cb, bpw = palette[int(rng.integers(0, len(palette)))], int(rng.choice(bpws))
        out.append((cb, int(rng.integers(0, 1 << bpw)) if bpw else 0, bpw))
    return out


@pytest.mark.parametrize("n,run,bpws,seed", [
    (1, 4, (0,), 1), (1, 4, (3,), 2), (37, 3, (0, 1, 2, 3, 4), 3), (500, 40, (0, 4), 4),
    (MACRO, 200, (0, 1, 2, 3, 4), 5),          # exactly one macro block
    (MACRO + 1, 1e9, (4,), 6),                 # one colour into the second macro block, one run
    (3 * MACRO + 77, 500, (0, 2, 3, 4), 7),    # runs across macro boundaries
    (40000, 2, (1, 2, 3, 4), 8),               # short runs, weight offsets near 16 bits
])
def test_encoder_bytes_match_format_model(n, run, bpws, seed):
    colors = random_colors(np.random.default_rng(seed), n, run, bpws)
    got = oracle_encode(colors)
    want = model_encode(colors)
    for name, a, b in zip(("blocks", "weights", "macro"), got, want):
        assert a.size == b.size and np.array_equal(a, b), f"{name} differ"


def test_encoder_full_macro_block_of_4_bit_weights():
    """16384 colours of 4-bit weights in one macro block: the last block's weight offset is
    65532 bits, the largest the 16-bit header field holds besides the bpw-0 marker."""
    colors = [(0x12345678 + (i % 2), i & 15, 4) for i in range(MACRO)]
    for a, b in zip(oracle_encode(colors), model_encode(colors)):
        assert np.array_equal(a, b)
    assert int(model_encode(colors)[0][-1]) & 0xFFFFFFFF == (65532 << 16) | (3 << 14) | (MACRO - 1)


# ---------------------------------------------------------------------------------------------
# trace_colors fallbacks
# ---------------------------------------------------------------------------------------------


def rgb888(r, g, b):
    f = np.float32
    cl = [min(max(f(c), f(0)), f(1)) for c in (r, g, b)]
    return int(np.uint32(cl[0] * f(255))) | (int(np.uint32(cl[1] * f(255))) << 8) | \
        (int(np.uint32(cl[2] * f(255))) << 16) | 0xFF000000


def test_fallback_colours():
    f = np.float32
    p, sc, g = scene("d7")
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    grid = od.dense()
    set_vox = [tuple(int(c) for c in v) for v in np.argwhere(grid)[::97][:20]]
    # voxels in empty 4^3 leaf blocks: a path to one meets an absent child on the way down (the
    # walk checks child masks down to the leaf level only, tracer.cu:305-312)
    leaf_occ = pyramid(grid)[g.levels - 2]
    empty = [tuple(int(c) for c in v) for v in np.argwhere(grid == 0)[::4099]
             if tuple(v) != (0, 0, 0) and not leaf_occ[tuple(v >> 2)]][:20]
    W = 64
    paths = np.zeros((1, W, 4), np.uint32)
    cases = set_vox + empty
    for k, v in enumerate(cases):
        paths[0, k, :3] = v
    # paths[0, len(cases):] stay (0, 0, 0): the sky
    sky = rgb888(f(187) / f(255), f(242) / f(255), f(250) / f(255))
    assert sky == 0xFFFAF2BB
    # with the scene's colours: absent children give magenta without alpha
    cols = Oracle.trace_colors(od, OracleColors.from_scene(sc), paths, W, 1)
    for k, v in enumerate(empty):
        assert int(cols[0, len(set_vox) + k]) == 0xFF00FF, v
    assert all(int(c) == sky for c in cols[0, len(cases):])
    # no colour tree below the root: every set voxel takes the checker at level 1
    bare = OracleColors(np.zeros(8, np.uint32), np.zeros(1, np.uint64), np.zeros(0, np.uint64),
                        np.zeros(0, np.uint32), np.zeros(0, np.uint64))
    cols = Oracle.trace_colors(od, bare, paths, W, 1)
    for k, v in enumerate(set_vox):
        b = f((v[0] ^ v[1] ^ v[2]) & 1)
        assert int(cols[0, k]) == rgb888(f(1), b, f(1) - b), v
    assert all(int(c) == sky for c in cols[0, len(cases):])
    assert len(set_vox) == 20 and len(empty) >= 10


# ---------------------------------------------------------------------------------------------
# interior find walk on synthetic buckets
# ---------------------------------------------------------------------------------------------


def walk_find(bucket_words, bucket_size, page, node):
    """hash_table.h:276-357 on a bucket given as its words (index = position in the bucket)."""
    n = len(node)
    for pstart in range(0, bucket_size, page):
        end = min(bucket_size, pstart + page)
        if pstart + n >= end:
            return None
        end -= n
        i = pstart
        while i < end:
            if list(bucket_words[i:i + n]) == list(node):
                return i
            i += bin(int(bucket_words[i]) & 0xFF).count("1") + 1
    return None


def build_bucket(rng, page, nodes_per_page_fill, bucket_pages):
    """Interior nodes appended with the page padding rule so that some end exactly at a page
    end and the last one exactly at the bucket end (when the sizes allow)."""
    words, starts, nodes = [], [], []
    size = 0
    for pg in range(bucket_pages):
        left = page
        while left > 0:
            # pick a node size: fill the page exactly when 2..9 words are left
            n = left if 2 <= left <= 9 and rng.random() < nodes_per_page_fill else int(rng.integers(2, 10))
            if n > left:
                break
            mask = 0
            while bin(mask).count("1") != n - 1:
                mask = int(rng.integers(1, 256))
            node = [mask | (int(rng.integers(0, 1 << 20)) << 8)] + [int(v) for v in rng.integers(1, 1 << 30, n - 1)]
            starts.append(size)
            nodes.append(node)
            words += node
            size += n
            left -= n
        if left and pg + 1 < bucket_pages:  # padding to the next page (what the append rule leaves)
            words += [0] * left
            size += left
    return words, starts, nodes, size  # the bucket ends where its last node ends


@pytest.mark.parametrize("seed,fill,pages", [(1, 1.0, 1), (2, 1.0, 3), (3, 0.5, 4), (4, 0.0, 2), (5, 1.0, 2)])
def test_interior_find_walk(seed, fill, pages):
    rng = np.random.default_rng(seed)
    page = 64
    g = Geometry(levels=6, page_size=page, num_top_levels=4, bucket_bits_top=2, bucket_bits_low=2,
                 bucket_size_top=page * 4, bucket_size_low=page * 4, color_tree_levels=3)
    level, bucket = 1, 2
    words, starts, nodes, size = build_bucket(rng, page, fill, pages)
    assert size <= g.bucket_size_top
    # the table: every page of the bucket mapped in order, the bucket size set
    pt = np.zeros(g.total_pages(), np.uint32)
    first_page = (level * (1 << g.bucket_bits_top) + bucket) * g.bucket_size_top // page
    pool = np.zeros((pages + 2) * page, np.uint32)
    for k in range((size + page - 1) // page):
        pt[first_page + k] = 1 + k
        chunk = words[k * page:(k + 1) * page]
        pool[(1 + k) * page:(1 + k) * page + len(chunk)] = chunk
    bs = np.zeros(g.total_buckets(), np.uint32)
    bs[level * (1 << g.bucket_bits_top) + bucket] = size
    od = OracleDAG.from_image(HostImage(g, pool, pt, bs, np.zeros(g.levels, np.uint32), pages + 1, 0), pages + 2)
    base_va = first_page * page
    queries = nodes + [[nodes[0][0]] + [w ^ 1 for w in nodes[0][1:]]]  # plus one absent node
    found = 0
    for node in queries:
        arr = np.array(node, np.uint32)
        got = orc.orc_find_in_bucket(od.h, level, bucket, _p(arr), len(node))
        want = walk_find(words, size, page, node)
        assert got == (0xFFFFFFFF if want is None else base_va + want), (node, got, want)
        found += want is not None
    # the rule leaves some nodes unfound (the last node of a page, the bucket's tail)
    assert 0 < found < len(nodes)


# ---------------------------------------------------------------------------------------------
# SpherePaintEditor's quantised colour (hash_dag_editors.h:315-330, color_utils.h:9-60)
# ---------------------------------------------------------------------------------------------


def decode_compressed(bits, wb):
    """CompressedColor::get_color (variable_weight_size_colors.h:144-154) in float32:
    rgb101210 for 0 bits per weight, else lerp(rgb565 min, rgb565 max, w / (2^bpw - 1))."""
    f = np.float32
    w, bpw = wb & 0xFF, wb >> 8
    if bpw == 0:
        return [f(bits & 0x3FF) / f(1023), f((bits >> 10) & 0xFFF) / f(4095), f((bits >> 22) & 0x3FF) / f(1023)]
    lo = [f(bits & 0x1F) / f(31), f((bits >> 5) & 0x3F) / f(63), f((bits >> 11) & 0x1F) / f(31)]
    hi = [f((bits >> 16) & 0x1F) / f(31), f((bits >> 21) & 0x3F) / f(63), f((bits >> 27) & 0x1F) / f(31)]
    t = f(w) / f((1 << bpw) - 1)
    return [a * (f(1) - t) + b * t for a, b in zip(lo, hi)]


def paint_expected(old_bits, old_wb, pos, center, radius):
    """set_single_color(lerp(old, (0, 0, 1), clamp(1 - length(pos - center) / radius, 0, 0.1)))
    with lerp(a, b, f) = a (1 - f) + b f and float3_to_rgb101210's clamp and truncation."""
    f = np.float32
    d = [f(p) - f(c) for p, c in zip(pos, center)]
    length = np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2], dtype=np.float32)
    k = f(1) - length / f(radius)
    k = f(0) if k < f(0) else (f(0.1) if k > f(0.1) else k)
    old = decode_compressed(old_bits, old_wb)
    new = [a * (f(1) - k) + b * k for a, b in zip(old, (f(0), f(0), f(1)))]
    cl = [min(max(c, f(0)), f(1)) for c in new]
    return int(np.uint32(cl[0] * f(1023))) | (int(np.uint32(cl[1] * f(4095))) << 10) | \
        (int(np.uint32(cl[2] * f(1023))) << 22)


def test_paint_quantisation():
    """Every voxel a paint sphere touches gets the quantised blend of its old colour (single or
    rgb565 min / max with weights); the others keep theirs. Paint edits twice so the second
    blends colours the first one wrote."""
    from hashdag_amd import terrain_height
    p, sc, g = scene("d8")
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    grid = od.dense()
    x, z = 120, 130
    y = terrain_height(p, x, z)
    checked = blended = 0
    for center, radius in (((float(x), float(y), float(z)), 9.0), ((float(x + 3), float(y + 1), float(z - 2)), 6.5)):
        lo = [int(math.floor(c - radius)) - 1 for c in center]
        hi = [int(math.ceil(c + radius)) + 1 for c in center]
        box = [tuple(v) for v in np.argwhere(grid[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]) + lo]
        before = {v: od.color(oc, *v) for v in box}
        od.edit(4, center, radius, oc)
        for v in box:
            s0, bits0, wb0 = before[v]
            s1, bits1, wb1 = od.color(oc, *v)
            assert s0 and s1
            # should_edit on the voxel's unit box (SphereEditorBase, hash_dag_editors.h:112-126)
            f = np.float32
            dist = f(radius) * f(radius)
            for a in range(3):
                c, v0, v1 = f(center[a]), f(v[a]), f(v[a] + 1)
                if c < v0:
                    dist -= (c - v0) * (c - v0)
                elif c > v1:
                    dist -= (c - v1) * (c - v1)
            if dist > 0:
                assert (bits1, wb1) == (paint_expected(bits0, wb0, v, center, radius), 0), v
                blended += int(bits1 != bits0 or wb0 != 0)
            else:
                assert (bits1, wb1) == (bits0, wb0), v
            checked += 1
    assert checked > 500 and blended > 200


# ---------------------------------------------------------------------------------------------
# copy swirl: dense-grid brute force with the rotation in FP64
# ---------------------------------------------------------------------------------------------

SWIRLS = [
    # center, radius, source, dest, swirl period, transform (None = identity)
    ((40.0, 60.0, 40.0), 10.0, (40.0, 60.0, 40.0), (90.0, 80.0, 90.0), 30.0, None),
    ((64.0, 64.0, 64.0), 12.0, (60.0, 50.0, 64.0), (30.0, 60.0, 90.0), 6.5, None),
    ((50.0, 58.0, 70.0), 10.0, (50.0, 58.0, 70.0), (90.0, 40.0, 30.0), 20.0, [[0, 0, 1], [0, 1, 0], [1, 0, 0]]),
]


@pytest.mark.parametrize("center,radius,source,dest,period,transform", SWIRLS)
def test_swirl_copy_brute_force(center, radius, source, dest, period, transform):
    """The copy swirl (hash_dag_editors.h:364-379) against a brute force that turns each
    destination voxel's offset from dest about the y axis by angle y / period in FP64
    (x' = x cos - z sin, z' = x sin + z cos), applies T^-1, adds source and floors. The oracle
    turns in float32 with cosf / sinf; the two agree on every voxel whose source coordinate
    lies more than 1e-3 from an integer (a decision boundary), and those are checked: voxel
    set and the colours of newly set voxels."""
    from tests.test_oracle_edits import BOX_ADD, per_voxel_colours, touched
    p, sc, g = scene("d7")
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    before, cols_before = per_voxel_colours(od, oc, g)
    T = np.eye(3) if transform is None else np.array(transform, np.float64)
    od.edit_copy(center, radius, source, dest, oc, transform=None if transform is None else T,
                 swirl_period=period)
    after, cols_after = per_voxel_colours(od, oc, g)
    E = 1 << g.levels
    box_c = np.array(dest) + T @ (np.array(center) - np.array(source))
    box_r = radius * np.abs(T @ np.ones(3)).max()
    t = touched(BOX_ADD, tuple(box_c), box_r, E)
    Tinv = np.linalg.inv(T)
    src, boundary = {}, np.zeros_like(t)
    for v in zip(*np.nonzero(t)):
        x, y, z = (float(a) for a in np.array(v, np.float64) - np.array(dest))
        a = y / period
        rel = np.array([x * math.cos(a) - z * math.sin(a), y, x * math.sin(a) + z * math.cos(a)])
        s_ = np.array(source) + Tinv @ rel
        assert (s_ >= 0).all() and (s_ < E).all()
        near = np.abs(s_ - np.round(s_))
        if ((near < 1e-3) & (near != 0)).any():
            boundary[v] = 1
            continue
        src[v] = tuple(int(c) for c in np.floor(s_))
    expect = before.copy()
    for v, s_ in src.items():
        if before[s_]:
            expect[v] = 1
    keep = boundary == 0
    assert np.array_equal(after[keep], expect[keep])
    assert boundary.sum() <= 0.01 * t.sum()
    turned = sum(1 for v, s_ in src.items() if before[s_] and not before[v])
    assert turned > 50
    for v, col in cols_after.items():
        if v in src and not before[v]:
            assert col == cols_before[src[v]], v
        elif not t[v]:
            assert col == cols_before[v], v


# ---------------------------------------------------------------------------------------------
# fog: cases whose result the mathematics fixes
# ---------------------------------------------------------------------------------------------


def test_fog_special_cases():
    """applyFog (tracer.cu:551-572): fogAmount = 1 - exp(-distance * density * 1e-5), the fog
    colour lerp((187, 242, 250) / 255, white, (1.01 max(dir . sun, 0))^30), the pixel
    lerp(colour, fog colour, clamp(fogAmount)). Pinned by two cases:
    * a density so high that fogAmount is 1 for every pixel: each pixel is its fog colour,
      exactly the sky colour (187, 242, 250) where the direction faces away from the sun;
    * for a chosen hit pixel, the density with distance * density * 1e-5 = ln 2, so that
      exp gives 1/2: the pixel is the midpoint of its fog-free colour and its fog colour (to
      the RGB8 truncation of both)."""
    f = np.float32
    p, sc, g = scene("d7")
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    W, H = 48, 32
    # looking slightly down along the sun's azimuth: terrain below, sky toward the sun above
    cam = CameraView.look((10.0, 90.0, 10.0), math.atan2(0.3, 0.5), -0.25)
    prm = Oracle.ray_params(cam, DAGInfo.unit(g.levels), g.levels, W, H)
    paths = Oracle.trace_paths(od, prm, W, H)
    cols = Oracle.trace_colors(od, oc, paths, W, H)
    clear = Oracle.trace_shadows(od, prm, paths, cols, W, H, 1.0, 0.0)
    thick = Oracle.trace_shadows(od, prm, paths, cols, W, H, 1.0, 1e12)
    s = 1.0 / math.sqrt(0.3 * 0.3 + 1.0 + 0.25)
    sun = np.array([0.3 * s, s, 0.5 * s])
    camp, rmin, ddx, ddy = prm[0:3], prm[3:6], prm[6:9], prm[9:12]

    def direction(px, py):  # toward the hit voxel's corner, or the camera ray for the sky
        pp = paths[py, px, :3]
        if not pp.any():
            v = rmin + px * ddx + (H - 1 - py) * ddy - camp
        else:
            v = pp.astype(np.float64) - camp
        return v / np.linalg.norm(v), float(np.linalg.norm(v))

    def fog_colour(dirn):
        sa = (1.01 * max(float(dirn @ sun), 0.0)) ** 30
        return np.array([187.0, 242.0, 250.0]) / 255.0 * (1 - sa) + sa, sa

    def rgb(c):
        return np.array([(int(c) >> (8 * k)) & 0xFF for k in range(3)])

    away = sunward = 0
    hits = []
    for py in range(H):
        for px in range(W):
            dirn, dist = direction(px, py)
            fc, sa = fog_colour(dirn)
            got = rgb(thick[py, px])
            if sa == 0.0:
                assert tuple(got) == (187, 242, 250), (px, py)
                away += 1
            else:
                assert np.abs(got - np.minimum(fc, 1.0) * 255).max() <= 1.0, (px, py)
                sunward += 1
            if paths[py, px, :3].any():
                hits.append((px, py, dist, fc))
    assert away > 0 and sunward > 0 and len(hits) > 100
    rng = np.random.default_rng(0)
    for k in rng.choice(len(hits), 6, replace=False):
        px, py, dist, fc = hits[int(k)]
        density = math.log(2.0) / (dist * 1e-5)
        half = Oracle.trace_shadows(od, prm, paths, cols, W, H, 1.0, density)
        want = (rgb(clear[py, px]) + np.minimum(fc, 1.0) * 255) / 2
        assert np.abs(rgb(half[py, px]) - want).max() <= 1.0, (px, py, rgb(half[py, px]), want)
