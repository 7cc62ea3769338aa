"""Pins the oracle's BVH code: the reference's Morton-pairing tree
(cuda_libraries/BVH.cu:153-448, shapes/Triangle.cu:703-776,
geometricMath.cu:16-44) and the counting traversal of the library's exported
two-wide layout, both against brute force over every triangle; plus the
structural invariants of the library's host-built tree (CPU only)."""
import ctypes as C

import numpy as np
import pytest


def morton_ref(u):
    """Interleave 10-bit quantised unit-cube coordinates bit by bit."""
    q = [int(min(max(x * 1024.0, 0.0), 1023.0)) for x in u]
    code = 0
    for bit in range(10):
        for axis, v in enumerate(q):  # x highest of each triple
            code |= ((v >> bit) & 1) << (3 * bit + (2 - axis))
    return code


def soup(rt, n, seed=7):
    return rt.gen_soup(n, seed=seed)


def test_morton_codes_and_sort(rt, orc):
    xyz, idx = soup(rt, 2000, seed=3)
    b = orc.RefBvh(xyz, idx)
    tri = xyz.reshape(-1, 3, 3).astype(np.float32)
    lo = np.minimum(tri.min(axis=1).min(axis=0), 0)  # the zero-box reduce seed
    hi = np.maximum(tri.max(axis=1).max(axis=0), 0)
    codes = []
    for i in range(idx.shape[0]):
        c = (np.float32(1.0) / np.float32(3.0)) * ((tri[i, 0] + tri[i, 1]) + tri[i, 2])
        u = (c - lo) / (hi - lo)
        want = morton_ref(u.astype(np.float64))
        got = b.morton(i)
        assert got == want, (i, got, want)
        codes.append(got)
    order = b.sorted_ids()
    key = [(codes[i], i) for i in range(len(codes))]
    assert [k[1] for k in sorted(key)] == list(order)


@pytest.mark.parametrize("n", [2, 3, 17, 1000, 20000])
def test_reference_bvh_matches_brute_force(rt, orc, n):
    xyz, idx = soup(rt, n, seed=n)
    rays = rt.gen_rays(48, 24, n_incoherent=3000, seed=n + 1)
    bp, bt, fl = orc.brute_trace(xyz, idx, rays)
    rp, rtt = orc.RefBvh(xyz, idx).trace(rays)
    clean = fl == 0
    assert clean.mean() > 0.99
    # the reference's slab test is not conservative, so a handful of its
    # results may be misses where brute force hits; everything else agrees
    agree = rp[clean] == bp[clean]
    assert agree.mean() > 0.999
    assert np.array_equal(rtt[clean][agree], bt[clean][agree])


def test_ambiguity_flags(rt, orc):
    # two copies of one triangle: every ray hitting it is an exact tie
    tri = np.array([[0, 0, 2], [1, 0, 2], [0, 1, 2]], np.float32)
    xyz = np.concatenate([tri, tri]).astype(np.float32)
    idx = np.array([[0, 1, 2], [3, 4, 5]], np.uint32)
    rays = [np.array([0.2], np.float32), np.array([0.2], np.float32), np.array([0.0], np.float32),
            np.array([0.0], np.float32), np.array([0.0], np.float32), np.array([1.0], np.float32)]
    p, t, f = orc.brute_trace(xyz, idx, rays)
    assert p[0] == 0 and (f[0] & 1)
    # a deltaT-masked triangle in front of an accepted one (Triangle.cu:573, 592)
    steep = np.array([[-1, -1, 1e-7 - 1], [1, -1, 1e-7 - 1], [0, 1, 1e-7 + 1]], np.float32)
    xyz = np.concatenate([steep, tri]).astype(np.float32)
    p, t, f = orc.brute_trace(xyz, idx, [np.array([0.0], np.float32)] * 5 + [np.array([1.0], np.float32)])
    assert p[0] == 1 and (f[0] & 2)


def host_bvh(rt, xyz, idx, max_leaf=4):
    L = rt.lib()
    h = C.c_void_p()
    o = rt.BvhOpts(max_leaf, 32, 1.2)
    xyz = np.ascontiguousarray(xyz, np.float32)
    idx = np.ascontiguousarray(idx, np.uint32)
    rt.check(L.rt_bvh_build(None, xyz.ctypes.data, xyz.shape[0], idx.ctypes.data, idx.shape[0],
                            C.byref(o), C.byref(h)))
    inf = rt.BvhInfo()
    rt.check(L.rt_bvh_get_info(h, C.byref(inf)))
    nodes = np.zeros((inf.num_nodes, 16), np.float32)
    tris = np.zeros((inf.num_tris + 1, 12), np.float32)
    rt.check(L.rt_bvh_export(h, nodes.ctypes.data, tris.ctypes.data))
    L.rt_bvh_destroy(h)
    return inf, nodes, tris


def walk(nodes, tris):
    """Yields (box_lo, box_hi, child_ref) for every child slot and checks the
    tree structure; returns the list of triangle record indices reached."""
    reached = []
    seen_nodes = set()
    stack = [0]
    ci = nodes[:, 12:14].copy().view(np.int32)
    depth_max = 0
    stack = [(0, 1)]
    while stack:
        n, dep = stack.pop()
        assert n not in seen_nodes
        seen_nodes.add(n)
        depth_max = max(depth_max, dep)
        f = nodes[n]
        boxes = [((f[0], f[2], f[8]), (f[1], f[3], f[9])), ((f[4], f[6], f[10]), (f[5], f[7], f[11]))]
        for s in range(2):
            c = int(ci[n, s])
            lo, hi = np.array(boxes[s][0]), np.array(boxes[s][1])
            if c >= 0:
                g = nodes[c]
                for lo2, hi2 in (((g[0], g[2], g[8]), (g[1], g[3], g[9])), ((g[4], g[6], g[10]), (g[5], g[7], g[11]))):
                    assert np.all(lo <= np.array(lo2)) and np.all(np.array(hi2) <= hi), "child box escapes parent"
                stack.append((c, dep + 1))
            else:
                enc = (~c) & 0xFFFFFFFF
                first, cnt = enc >> 3, (enc & 7) + 1
                for k in range(first, first + cnt):
                    v = tris[k, [0, 1, 2, 4, 5, 6, 8, 9, 10]].reshape(3, 3)
                    reached.append(k)
                    if tris[k, 3:4].view(np.int32)[0] < 0:
                        continue  # degenerate sentinel: never accepted, box irrelevant
                    assert np.all(lo < v.min(axis=0)) and np.all(v.max(axis=0) < hi), "leaf box not padded around tri"
    return reached, len(seen_nodes), depth_max


@pytest.mark.parametrize("n,max_leaf", [(1, 4), (3, 4), (4, 4), (5, 4), (1000, 4), (5000, 1), (5000, 8)])
def test_host_bvh_structure(rt, n, max_leaf):
    xyz, idx = soup(rt, n, seed=n)
    inf, nodes, tris = host_bvh(rt, xyz, idx, max_leaf)
    reached, nnodes, depth = walk(nodes, tris)
    ids = tris[:, 3].copy().view(np.int32)
    real = sorted(int(ids[k]) for k in reached if ids[k] >= 0)
    assert real == list(range(n)), "every triangle in exactly one leaf"
    assert nnodes == inf.num_nodes
    assert depth == inf.max_depth
    assert ids[-1] == -1  # sentinel record
    # the stored vertices are the caller's exact FP32 values
    for k in reached:
        if ids[k] >= 0:
            v = tris[k, [0, 1, 2, 4, 5, 6, 8, 9, 10]].reshape(3, 3)
            assert np.array_equal(v, xyz[idx[ids[k]]])


def test_host_bvh_empty(rt, orc):
    xyz = np.zeros((0, 3), np.float32)
    idx = np.zeros((0, 3), np.uint32)
    inf, nodes, tris = host_bvh(rt, xyz, idx)
    assert inf.num_nodes == 1 and inf.num_tris == 0
    rays = rt.gen_rays(8, 8, n_incoherent=64, seed=1)
    p, t, nn, nt = orc.bvh2_trace(nodes, tris, rays)
    assert np.all(p == -1)


@pytest.mark.parametrize("n", [5, 2000, 30000])
def test_bvh2_traversal_matches_brute_force(rt, orc, n):
    xyz, idx = soup(rt, n, seed=n + 5)
    inf, nodes, tris = host_bvh(rt, xyz, idx)
    rays = rt.gen_rays(64, 32, n_incoherent=4096, seed=n + 6)
    # include axis-parallel rays (zero direction components)
    rng = np.random.default_rng(n)
    m = 512
    ax = [rng.uniform(0, 1, m).astype(np.float32) for _ in range(3)]
    dirs = np.zeros((3, m), np.float32)
    dirs[rng.integers(0, 3, m), np.arange(m)] = rng.choice([-1.0, 1.0], m)
    rays = [np.concatenate([rays[k], ax[k] if k < 3 else dirs[k - 3]]) for k in range(6)]
    bp, bt, fl = orc.brute_trace(xyz, idx, rays)
    p2, t2, nn, ntr = orc.bvh2_trace(nodes, tris, rays)
    clean = fl == 0
    assert np.array_equal(p2[clean], bp[clean])
    assert np.array_equal(t2[clean], bt[clean])
    assert nn > 0
