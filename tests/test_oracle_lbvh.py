"""The oracle's restatement of the GPU-built BVH (oracle/lbvh.c), pinned
before it is used to check rt_bvh_build_device:

* structure: every triangle in exactly one leaf, stored with the caller's
  exact vertices, every box on a path encloses its triangles (padded), the
  reported node counts and depths are what a walk finds;
* Morton order: the leaf order is the order of (30-bit Morton code of the
  triangle centre in the centre bounds, index), recomputed here in numpy;
* radix splits: at every kept node, the triangles of its two children are
  separated by the highest bit in which the node's first and last 62-bit
  keys (code, sorted position) differ -- the definition of a binary radix
  tree, checked directly;
* the four-wide collapse: breadth-first numbering, the traversal stack bound
  is the deepest sum of (kids - 1), and the restated traversal never needs
  more;
* traversal of both layouts against brute force on every unambiguous ray."""
import numpy as np
import pytest

from test_oracle_bvh import walk
from test_oracle_bvh4 import walk4


def soup(rt, n, seed, dup=False):
    if n == 0:
        return np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32)
    xyz, idx = rt.gen_soup(n, seed=seed)
    if dup:  # repeated triangles: equal Morton codes, split by position
        xyz = xyz.reshape(n, 9)
        xyz[n // 3:] = xyz[n // 3]
        xyz = np.ascontiguousarray(xyz.reshape(-1, 3))
    return xyz, idx


def spread(v):
    r = np.zeros_like(v, dtype=np.uint64)
    for b in range(10):
        r |= ((v.astype(np.uint64) >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b)
    return r


def morton_order(xyz, idx):
    v = xyz[idx.astype(np.int64)]  # (nt, 3, 3)
    lo, hi = v.min(axis=1), v.max(axis=1)
    c = (np.float32(0.5) * lo + np.float32(0.5) * hi).astype(np.float32)
    cmin, cmax = c.min(axis=0), c.max(axis=0)
    ext = (cmax - cmin).astype(np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        u = np.where(ext > 0, (c - cmin) / np.where(ext > 0, ext, 1) * np.float32(1024), 0).astype(np.float32)
    q = np.clip(u, 0, 1023).astype(np.uint32)
    code = spread(q[:, 0]) << np.uint64(2) | spread(q[:, 1]) << np.uint64(1) | spread(q[:, 2])
    order = np.lexsort((np.arange(len(code)), code))
    return order, code[order]


def leaf_range(link):
    enc = (~int(link)) & 0xFFFFFFFF
    return enc >> 3, (enc & 7) + 1


def subtree_range(nodes, link):
    """Sorted positions covered by a child link of the two-wide tree."""
    if link < 0:
        f, c = leaf_range(link)
        return f, f + c - 1
    ci = nodes[int(link), 12:14].copy().view(np.int32)
    return subtree_range(nodes, ci[0])[0], subtree_range(nodes, ci[1])[1]


def area(lo, hi):
    d = np.maximum(0.0, hi.astype(np.float64) - lo.astype(np.float64))
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])


class SahRule:
    """The SAH leaf rule recomputed from the sorted keys: the radix subtree
    over positions [lo, hi] splits at the highest bit in which its first and
    last keys differ; a subtree over at most max_leaf triangles (not the
    root) is a leaf when triangles x area <= trav x area + its children's
    costs, a single triangle costing its box's area."""

    def __init__(self, key, tlo, thi, max_leaf, trav=float(np.float32(1.2))):
        self.key, self.tlo, self.thi, self.max_leaf, self.trav = key, tlo, thi, max_leaf, trav
        self.memo = {}

    def cost(self, lo, hi, root=False):
        if lo == hi:
            return area(self.tlo[lo], self.thi[lo]), False
        if (lo, hi, root) in self.memo:
            return self.memo[(lo, hi, root)]
        bit = (int(self.key[lo]) ^ int(self.key[hi])).bit_length() - 1
        s = lo
        while (int(self.key[s + 1]) >> bit) & 1 == (int(self.key[lo]) >> bit) & 1:
            s += 1
        c0, _ = self.cost(lo, s)
        c1, _ = self.cost(s + 1, hi)
        a = area(self.tlo[lo:hi + 1].min(axis=0), self.thi[lo:hi + 1].max(axis=0))
        split = self.trav * a + c0 + c1
        size = hi - lo + 1
        leaf = not root and size <= self.max_leaf and size * a <= split
        r = (size * a if leaf else split), leaf
        self.memo[(lo, hi, root)] = r
        return r


def stack_need(nodes4, n=0):
    refs = nodes4[n, 4:8].copy().view(np.int32)
    valid = (int(nodes4[n, 3:4].copy().view(np.uint32)[0]) >> 24) & 0xF
    nk = bin(valid).count("1")
    sub = [stack_need(nodes4, int(refs[i])) for i in range(nk) if refs[i] >= 0]
    return nk - 1 + max(sub, default=0)


CASES = [(0, 4, False), (1, 4, False), (2, 4, False), (3, 4, False), (5, 4, False), (9, 1, False),
         (100, 4, False), (1000, 8, False), (3000, 1, True), (5000, 4, False)]


@pytest.mark.parametrize("n,max_leaf,dup", CASES)
def test_structure_and_order(rt, orc, n, max_leaf, dup):
    xyz, idx = soup(rt, n, seed=n + 3, dup=dup)
    inf, nodes, tris, nodes4 = orc.lbvh_build(xyz, idx, max_leaf=max_leaf)
    reached, nn, depth = walk(nodes, tris)
    ids = tris[:, 3].copy().view(np.int32)
    real = sorted(int(ids[k]) for k in reached if ids[k] >= 0)
    assert real == list(range(n)), "every triangle in exactly one leaf"
    assert nn == inf["num_nodes"] and depth == inf["max_depth"]
    assert inf["num_leaves"] == (inf["num_nodes"] + 1 if n > 1 else n)
    assert ids[-1] == -1
    for k in range(n):
        v = tris[k, [0, 1, 2, 4, 5, 6, 8, 9, 10]].reshape(3, 3)
        assert np.array_equal(v, xyz[idx[ids[k]]])
    if n:
        order, _ = morton_order(xyz, idx)
        assert np.array_equal(ids[:n], order)
    # leaves hold at most max_leaf triangles; every kept node and every leaf
    # is what the SAH rule chooses
    ci = nodes[:, 12:14].copy().view(np.int32)
    if n > 1:
        _, code = morton_order(xyz, idx)
        key = (code << np.uint64(32)) | np.arange(n, dtype=np.uint64)
        v = xyz[idx[ids[:n]].astype(np.int64)]
        rule = SahRule(key, v.min(axis=1), v.max(axis=1), max_leaf)
        for i in range(nn):
            lo, hi = subtree_range(nodes, i)
            assert not rule.cost(lo, hi, root=i == 0)[1], "a kept node the rule makes a leaf"
            for c in ci[i]:
                if c < 0:
                    f, cnt = leaf_range(c)
                    assert cnt <= max_leaf
                    assert cnt == 1 or rule.cost(f, f + cnt - 1)[1], "a leaf the rule splits"
    reached4, nn4, depth4 = walk4(nodes4, tris)
    assert nn4 == inf["num_nodes4"] and depth4 == inf["depth4"]
    assert sorted(int(ids[k]) for k in reached4 if ids[k] >= 0) == list(range(n))
    assert stack_need(nodes4) == inf["max_stack4"]


@pytest.mark.parametrize("n,dup", [(200, False), (300, True), (2000, False)])
def test_radix_splits(rt, orc, n, dup):
    xyz, idx = soup(rt, n, seed=n, dup=dup)
    inf, nodes, tris, _ = orc.lbvh_build(xyz, idx, max_leaf=1)
    _, code = morton_order(xyz, idx)
    key = (code << np.uint64(32)) | np.arange(n, dtype=np.uint64)
    ci = nodes[:, 12:14].copy().view(np.int32)
    for i in range(inf["num_nodes"]):
        a0, a1 = subtree_range(nodes, ci[i, 0])
        b0, b1 = subtree_range(nodes, ci[i, 1])
        assert a1 + 1 == b0
        x = int(key[a0]) ^ int(key[b1])
        bit = x.bit_length() - 1
        assert all((int(k) >> bit) & 1 == 0 for k in key[a0:a1 + 1])
        assert all((int(k) >> bit) & 1 == 1 for k in key[b0:b1 + 1])


@pytest.mark.parametrize("n", [7, 2000, 30000])
def test_traversal_matches_brute_force(rt, orc, n):
    xyz, idx = soup(rt, n, seed=n + 1)
    rays = rt.gen_rays(48, 32, n_incoherent=3000, seed=n)
    inf, nodes, tris, nodes4 = orc.lbvh_build(xyz, idx)
    bp, bt, fl = orc.brute_trace(xyz, idx, rays)
    clean = fl == 0
    assert clean.mean() > 0.99
    p2, t2, _, _ = orc.bvh2_trace(nodes, tris, rays)
    assert np.array_equal(p2[clean], bp[clean]) and np.array_equal(t2[clean], bt[clean])
    p4, t4, _, _, max_sp = orc.bvh4_trace(nodes4, tris, rays)
    assert np.array_equal(p4[clean], bp[clean]) and np.array_equal(t4[clean], bt[clean])
    assert max_sp <= inf["max_stack4"]


def test_bad_input_is_refused(rt, orc):
    xyz, idx = soup(rt, 10, seed=1)
    bad = idx.copy()
    bad[3, 1] = xyz.shape[0]
    with pytest.raises(ValueError):
        orc.lbvh_build(xyz, bad)
    x2 = xyz.copy()
    x2[5, 0] = np.nan
    with pytest.raises(ValueError):
        orc.lbvh_build(x2, idx)
