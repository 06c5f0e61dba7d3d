"""Pins the oracle's hash table (hash_table.h:155-560, 749-809), its BasicDAG import
(hash_dag_factory.cpp:5-198), Path/leaf bit layout and GC (hash_dag.h:407-547) with checks
that do not restate the code under test:

- the imported DAG decodes to exactly the voxel grid the terrain heights define;
- get_value agrees with that grid at random points;
- every node sits in the bucket its hash selects (hashes are pinned by published vectors),
  never straddles a page, and every child pointer lies in the child level's address range;
- the full node of every level is where full_nodes says and holds the full content;
- GC keeps the canonical form, is idempotent, and leaves only reachable nodes in use.
"""
import numpy as np
import pytest

from tests.scenes import POOL_PAGES, scene


def vaddr(g):
    top = g.top()

    def make_ptr(level, bucket, pos):
        if level < top:
            return (level * (1 << g.bucket_bits_top) + bucket) * g.bucket_size_top + pos
        return (top * (1 << g.bucket_bits_top) * g.bucket_size_top +
                ((level - top) * (1 << g.bucket_bits_low) + bucket) * g.bucket_size_low + pos)
    return make_ptr


def level_of(g, va):
    mp = vaddr(g)
    for lvl in range(g.levels - 1, -1, -1):
        if va >= mp(lvl, 0, 0):
            return lvl
    return 0


class Table:
    def __init__(self, od):
        self.g = od.geometry
        self.pool, self.pt, self.bs, self.full, self.pool_top = od.export()
        self.root = od.root

    def word(self, va):
        ps = self.g.page_size
        return int(self.pool[int(self.pt[va // ps]) * ps + va % ps])

    def node(self, level, va):
        if level == self.g.leaf_level:
            return [self.word(va), self.word(va + 1)]
        h = self.word(va)
        n = 1 + bin(h & 0xFF).count("1")
        return [self.word(va + i) for i in range(n)]

    def reachable(self):
        """(level, address) of every node reachable from the root."""
        seen = {(0, self.root)}
        stack = [(0, self.root)]
        while stack:
            lvl, va = stack.pop()
            if lvl == self.g.leaf_level:
                continue
            for c in self.node(lvl, va)[1:]:
                if (lvl + 1, c) not in seen:
                    seen.add((lvl + 1, c))
                    stack.append((lvl + 1, c))
        return seen


def bucket_of(g, level, va):
    mp = vaddr(g)
    size = g.bucket_size_top if level < g.top() else g.bucket_size_low
    b = (va - mp(level, 0, 0)) // size
    return b, va - mp(level, b, 0)


def node_hash(g, level, words):
    from oracle.binding import _p, orc
    if level == g.leaf_level:
        return orc.orc_murmurhash64(words[0] | (words[1] << 32)) & 0xFFFFFFFF
    arr = np.array(words, np.uint32)
    return orc.orc_murmurhash32xN(_p(arr), arr.size, 0)


def load(name):
    from oracle import OracleDAG
    p, sc, g = scene(name)
    return p, sc, g, OracleDAG.from_basic(g, sc.dag, POOL_PAGES)


@pytest.mark.parametrize("name", ["d7", "d8s"])
def test_import_decodes_to_the_height_grid(name):
    from hashdag_amd.scene import dense_from_heights
    p, sc, g, od = load(name)
    assert np.array_equal(od.dense(), dense_from_heights(p))


def test_get_value_matches_grid():
    from hashdag_amd.scene import dense_from_heights
    p, sc, g, od = load("d7")
    grid = dense_from_heights(p)
    rng = np.random.default_rng(3)
    occupied = np.argwhere(grid)
    pts = list(occupied[rng.choice(len(occupied), 200)]) + list(rng.integers(0, 128, size=(200, 3)))
    for x, y, z in pts:
        assert od.get_value(int(x), int(y), int(z)) == bool(grid[x, y, z])


@pytest.mark.parametrize("name", ["d8", "d13"])
def test_nodes_live_in_their_hash_bucket(name):
    p, sc, g, od = load(name)
    t = Table(od)
    mp = vaddr(g)
    for lvl, va in t.reachable():
        words = t.node(lvl, va)
        b, pos = bucket_of(g, lvl, va)
        nb = 1 << (g.bucket_bits_top if lvl < g.top() else g.bucket_bits_low)
        assert node_hash(g, lvl, words) & (nb - 1) == b, (lvl, va)
        assert pos + len(words) <= int(t.bs[bucket_index(g, lvl, b)])
        assert pos // g.page_size == (pos + len(words) - 1) // g.page_size, "node straddles a page"
        if lvl < g.leaf_level:
            for c in words[1:]:
                assert level_of(g, c) == lvl + 1
                assert c >= mp(lvl + 1, 0, 0)


def bucket_index(g, level, b):
    top = g.top()
    if level < top:
        return level * (1 << g.bucket_bits_top) + b
    return top * (1 << g.bucket_bits_top) + (level - top) * (1 << g.bucket_bits_low) + b


def test_full_nodes():
    p, sc, g, od = load("d13")
    t = Table(od)
    L = g.leaf_level
    assert t.node(L, int(t.full[L])) == [0xFFFFFFFF, 0xFFFFFFFF]
    for lvl in range(L - 1, 0, -1):
        words = t.node(lvl, int(t.full[lvl]))
        count = 0 if lvl < g.color_tree_levels else 8 ** (g.levels - lvl)
        assert words == [(count << 8) | 0xFF] + [int(t.full[lvl + 1])] * 8


def test_gc_is_idempotent_and_keeps_only_reachable():
    from tests.test_edit_gpu import random_edits
    p, sc, g, od = load("d8")
    before = od.canonical()
    for kind, center, r in random_edits(p, 10, seed=9):
        od.edit(kind, center, r)
    edited = od.canonical()
    assert not np.array_equal(before, edited)
    od.gc(g.leaf_level)
    assert np.array_equal(od.canonical(), edited)
    t1 = Table(od)
    used1 = int(t1.bs.astype(np.uint64).sum())
    # every level holds its reachable nodes and its full node, plus at most page padding
    words = {}
    for lvl, va in t1.reachable():
        words[lvl] = words.get(lvl, 0) + len(t1.node(lvl, va))
    for lvl in range(g.levels - 1):
        full = 2 if lvl == g.leaf_level else 9
        lo = words.get(lvl, 0)
        b0 = bucket_index(g, lvl, 0)
        nb = 1 << (g.bucket_bits_top if lvl < g.top() else g.bucket_bits_low)
        used = int(t1.bs[b0:b0 + nb].astype(np.uint64).sum())
        assert lo <= used <= lo + full + 8 * (used // g.page_size + nb), lvl
    od.gc(g.leaf_level)
    t2 = Table(od)
    assert np.array_equal(od.canonical(), edited)
    assert int(t2.bs.astype(np.uint64).sum()) == used1
