"""Pins the oracle's restatement of the noise SphereNoiseEditor draws from FastNoise
(hash_dag_editors.h:762-809; FastNoise.cpp:195-237, 328-333, 760-797, 841-888). The reference's
FastNoise cannot be built here (FastNoise.h includes typedefs.h, which includes
cuda_runtime.h), and the reference holds no noise vectors, so the checks are what the
mathematics fixes:

* the permutation: std::mt19937_64, written here from the published 64-bit Mersenne Twister
  and pinned by the C++ standard's value of its 10000th output (9981545732273789042), drives
  FastNoise's seeded Fisher-Yates pass; the oracle's tables must equal it, with m_perm a
  permutation repeated twice and m_perm12 = m_perm % 12;
* gradient noise vanishes at every lattice point, and next to a lattice point it is the dot
  product of the offset with that point's gradient (quintic weights are flat at 0), the
  gradient being the table entry the permutation selects;
* it is continuous across cell faces;
* the fractal is the bounded octave sum of single noises with lacunarity 2 and gain 0.5;
* the editor on a dense grid: a buffer of 2 ceil(1.5 r) voxels per axis from center - 1.5 r,
  voxels within r (1 + noise) of its centre set to isAdding, the others loaded from the DAG at
  the world voxel truncate(min + p); a voxel w in the box becomes set when its buffer voxel is
  set or loaded set, and a voxel whose value changes gets the editor's colour (rgb565 min
  (94, 30, 0) / 255, max (125, 100, 89) / 255, 3-bit weight |w - c| / 1.5 r - 0.3).
"""
import numpy as np
import pytest

from oracle import OracleColors, OracleDAG
from oracle.binding import perlin_fractal
from tests.scenes import POOL_PAGES, scene

NOISE_ADD, NOISE_CARVE = 8, 9
GRAD = np.array([[1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1],
                 [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1]], np.float64)


class MT19937_64:
    """The 64-bit Mersenne Twister (Matsumoto and Nishimura, 2000), as std::mt19937_64."""
    N, M = 312, 156
    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.mt = [seed & self.MASK]
        for i in range(1, self.N):
            prev = self.mt[-1]
            self.mt.append((6364136223846793005 * (prev ^ (prev >> 62)) + i) & self.MASK)
        self.i = self.N

    def __call__(self) -> int:
        if self.i >= self.N:
            for k in range(self.N):
                x = (self.mt[k] & 0xFFFFFFFF80000000) | (self.mt[(k + 1) % self.N] & 0x7FFFFFFF)
                xa = x >> 1
                if x & 1:
                    xa ^= 0xB5026F5AA96619E9
                self.mt[k] = self.mt[(k + self.M) % self.N] ^ xa
            self.i = 0
        y = self.mt[self.i]
        self.i += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & self.MASK


def test_mt19937_64_standard_value():
    g = MT19937_64(5489)  # default seed
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


@pytest.mark.parametrize("seed", [1337, 1923, 40161, 2**31 - 1])
def test_permutation_tables(seed):
    _, perm, perm12 = perlin_fractal(seed, 0.01, 3, np.zeros((1, 3)), with_perm=True)
    g = MT19937_64(seed)
    want = list(range(256))
    for j in range(256):
        k = g() % (256 - j) + j
        want[j], want[k] = want[k], want[j]
    assert sorted(perm[:256].tolist()) == list(range(256))
    assert perm[:256].tolist() == want
    assert np.array_equal(perm[256:], perm[:256])
    assert np.array_equal(perm12, perm % 12)


def _corner_gradient(perm, perm12, x, y, z, offset):
    p = [int(v) for v in perm]
    return GRAD[int(perm12[(x & 255) + p[(y & 255) + p[(z & 255) + offset]]])]


@pytest.mark.parametrize("seed", [7, 1923])
def test_single_noise_lattice_and_gradients(seed):
    rng = np.random.default_rng(seed)
    lattice = rng.integers(-300, 300, size=(200, 3)).astype(np.float32)
    v, perm, perm12 = perlin_fractal(seed, 1.0, 1, lattice, with_perm=True)
    assert np.all(v == 0.0)
    eps = np.float32(2.0 ** -12)
    for axis in range(3):
        pts = lattice.copy()
        pts[:, axis] += eps
        got = perlin_fractal(seed, 1.0, 1, pts).astype(np.float64) / float(eps)
        for q, gv in zip(lattice.astype(int), got):
            g = _corner_gradient(perm, perm12, *q.tolist(), int(perm[0]))
            assert abs(gv - g[axis]) < 1e-3, (q, axis, gv, g)


def test_single_noise_continuous_across_cell_faces():
    rng = np.random.default_rng(3)
    base = rng.uniform(-50, 50, size=(300, 3)).astype(np.float32)
    for axis in range(3):
        face = base.copy()
        face[:, axis] = np.floor(face[:, axis]) + 1
        lo, hi = face.copy(), face.copy()
        lo[:, axis] = np.nextafter(face[:, axis], np.float32(-1e9))
        hi[:, axis] = np.nextafter(face[:, axis], np.float32(1e9))
        a = perlin_fractal(1337, 1.0, 1, lo)
        b = perlin_fractal(1337, 1.0, 1, hi)
        assert np.abs(a - b).max() < 1e-4
    v = perlin_fractal(1337, 1.0, 1, base)
    assert np.abs(v).max() <= 1.0 and np.abs(v).max() > 0.3


def test_fractal_is_bounded_octave_sum():
    """Octave i is single noise at 2^i p with offset perm[i]; an offset o shifts the z lattice
    by o (m_perm is stored twice), so it is the one-octave noise (offset perm[0]) at
    z + perm[i] - perm[0]. On points of a 1/64 grid these shifts are exact in float32, and the
    fractal must equal (s0 + s1 * 0.5 + s2 * 0.25) * (1 / 1.75) evaluated in float32."""
    f32 = np.float32
    rng = np.random.default_rng(5)
    pts = (rng.integers(0, 64 * 200, size=(400, 3)) / 64.0).astype(np.float32)
    got, perm, _ = perlin_fractal(911, 1.0, 3, pts, with_perm=True)
    octs = []
    for i in range(3):
        q = pts * f32(2 ** i)
        q[:, 2] = q[:, 2] + f32(int(perm[i]) - int(perm[0]))
        octs.append(perlin_fractal(911, 1.0, 1, q))
    bounding = f32(1.0) / f32(f32(1.0) + f32(0.5) + f32(0.25))
    want = ((octs[0] + octs[1] * f32(0.5)) + octs[2] * f32(0.25)) * bounding
    assert np.array_equal(got, want.astype(np.float32))


def brute_force_noise(before, cols_before, kind, center, radius):
    """SphereNoiseEditor on a dense grid; returns the grid after and {voxel: colour} of the
    voxels whose value changed."""
    f32 = np.float32
    r = f32(radius)
    br = f32(f32(1.5) * r)
    fmin = np.array(center, np.float32) - br
    size = 2 * int(np.ceil(br))
    seed = abs(1923 + int(center[0]) + 23 * int(center[1]) + 2938 * int(center[2]))
    idx = np.arange(size)
    X, Y, Z = np.meshgrid(idx, idx, idx, indexing="ij")
    pts = np.stack([X, Y, Z], -1).reshape(-1, 3).astype(np.float32)
    noise = perlin_fractal(seed, f32(1) / r, 3, pts).astype(np.float64).reshape(size, size, size)
    c0 = size // 2
    dist = np.sqrt(((c0 - X) ** 2 + (c0 - Y) ** 2 + (c0 - Z) ** 2).astype(np.float64))
    setmask = dist < float(r) * (1.0 + noise)
    E = before.shape[0]

    def world(q):
        return tuple(int(f32(fmin[a] + f32(q[a]))) for a in range(3))

    bmin = fmin
    bmax = np.array(center, np.float32) + br
    after, new_cols = before.copy(), {}
    lo = [max(0, int(np.floor(bmin[a])) - 1) for a in range(3)]
    hi = [min(E, int(np.ceil(bmax[a])) + 1) for a in range(3)]
    mn565 = lambda c: (int(f32(c[0]) * f32(31)) | (int(f32(c[1]) * f32(63)) << 5) | (int(f32(c[2]) * f32(31)) << 11))  # noqa: E731
    bits = mn565([f32(94) / f32(255), f32(30) / f32(255), f32(0) / f32(255)]) | \
        (mn565([f32(125) / f32(255), f32(100) / f32(255), f32(89) / f32(255)]) << 16)
    for x in range(lo[0], hi[0]):
        for y in range(lo[1], hi[1]):
            for z in range(lo[2], hi[2]):
                w = (x, y, z)
                if not all(f32(w[a] + 1) > bmin[a] and f32(w[a]) < bmax[a] for a in range(3)):
                    continue
                q = tuple(int(f32(f32(w[a]) - fmin[a])) for a in range(3))
                if not all(0 <= v < size for v in q):
                    continue
                if setmask[q]:
                    val = kind == NOISE_ADD
                else:
                    ww = world(q)
                    val = all(0 <= ww[a] < E for a in range(3)) and bool(before[ww])
                new = bool(before[w]) or val
                if new != bool(before[w]):
                    d = [f32(f32(w[a]) - f32(center[a])) for a in range(3)]
                    ln = f32(np.sqrt(f32(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])))
                    wt = int(f32(f32(ln / br - f32(0.3)) * f32(7)))
                    new_cols[w] = (bits, min(max(wt, 0), 7) | (3 << 8))
                    after[w] = 1
    return after, new_cols


@pytest.mark.parametrize("kind,where,radius", [(NOISE_ADD, "above", 5.0), (NOISE_ADD, "surface", 7.5),
                                               (NOISE_CARVE, "surface", 6.0)])
def test_noise_editor_brute_force(kind, where, radius):
    from hashdag_amd.scene import terrain_height
    from tests.test_oracle_edits import per_voxel_colours
    p, sc, g = scene("d7")
    od = OracleDAG.from_basic(g, sc.dag, POOL_PAGES)
    oc = OracleColors.from_scene(sc)
    before, cols_before = per_voxel_colours(od, oc, g)
    x, z = 61, 67
    h = terrain_height(p, x, z)
    center = (float(x), float(h + (6 if where == "above" else 0)), float(z))
    want, want_cols = brute_force_noise(before, cols_before, kind, center, radius)
    od.edit(kind, center, radius, oc)
    after, cols_after = per_voxel_colours(od, oc, g)
    assert np.array_equal(after, want)
    if kind == NOISE_ADD:
        assert int(after.sum()) > int(before.sum())
    for v, col in cols_after.items():
        if v in want_cols:
            assert col == want_cols[v], v
        else:
            assert col == cols_before[v], v
