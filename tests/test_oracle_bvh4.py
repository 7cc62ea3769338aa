"""The library's four-wide quantized tree (rt_bvh_export4), built on the
host without a GPU: structure (every triangle in exactly one leaf, every
link in range, each dequantized child box encloses every triangle below it,
padded), the stack bound the kernel sizes its LDS stack with, and the
oracle's operation-for-operation restatement of the kernel's four-wide
traversal against brute force over every triangle."""
import ctypes as C

import numpy as np
import pytest


def host_bvh4(rt, xyz, idx, max_leaf=4):
    L = rt.lib()
    h = C.c_void_p()
    o = rt.BvhOpts(max_leaf, 32, 1.2, 4)
    xyz = np.ascontiguousarray(xyz, np.float32)
    idx = np.ascontiguousarray(idx, np.uint32)
    rt.check(L.rt_bvh_build(None, xyz.ctypes.data, xyz.shape[0], idx.ctypes.data, idx.shape[0],
                            C.byref(o), C.byref(h)))
    inf = rt.BvhInfo()
    rt.check(L.rt_bvh_get_info(h, C.byref(inf)))
    nodes = np.zeros((inf.num_nodes, 16), np.float32)
    tris = np.zeros((inf.num_tris + 1, 12), np.float32)
    nodes4 = np.zeros((inf.num_nodes4, 16), np.float32)
    rt.check(L.rt_bvh_export(h, nodes.ctypes.data, tris.ctypes.data))
    rt.check(L.rt_bvh_export4(h, nodes4.ctypes.data))
    L.rt_bvh_destroy(h)
    return inf, nodes, nodes4, tris


def child_boxes(node):
    """Dequantized child boxes (exact, float64) and the valid mask."""
    u = node.copy().view(np.uint32)
    meta = int(u[3])
    lo = np.zeros((4, 3))
    hi = np.zeros((4, 3))
    for a in range(3):
        sc = 2.0 ** (((meta >> (8 * a)) & 0xFF) - 127)
        for i in range(4):
            lo[i, a] = float(node[a]) + ((int(u[8 + 2 * a]) >> (8 * i)) & 0xFF) * sc
            hi[i, a] = float(node[a]) + ((int(u[9 + 2 * a]) >> (8 * i)) & 0xFF) * sc
    return lo, hi, (meta >> 24) & 0xF


def walk4(nodes4, tris):
    refs = nodes4[:, 4:8].copy().view(np.int32)
    seen = set()
    reached = []
    depth_max = 0
    stack = [(0, 1, [])]  # node, depth, ancestor boxes enclosing this subtree
    while stack:
        n, dep, anc = stack.pop()
        assert n not in seen, "node reached twice"
        seen.add(n)
        depth_max = max(depth_max, dep)
        lo, hi, valid = child_boxes(nodes4[n])
        assert valid in (0b11, 0b111, 0b1111), "slots are filled from 0"
        u = nodes4[n].copy().view(np.uint32)
        for i in range(4):
            if not (valid >> i) & 1:
                # unused slot: an empty box (lo byte 255, hi byte 0 on every
                # axis) around the degenerate sentinel leaf
                for a in range(3):
                    assert (int(u[8 + 2 * a]) >> (8 * i)) & 0xFF == 255
                    assert (int(u[9 + 2 * a]) >> (8 * i)) & 0xFF == 0
                enc = (~int(refs[n, i])) & 0xFFFFFFFF
                assert int(refs[n, i]) < 0 and enc >> 3 == tris.shape[0] - 1 and enc & 7 == 0
                continue
            c = int(refs[n, i])
            boxes = anc + [(lo[i], hi[i])]
            if c >= 0:
                assert c < nodes4.shape[0], "link out of range"
                stack.append((c, dep + 1, boxes))
            else:
                enc = (~c) & 0xFFFFFFFF
                first, cnt = enc >> 3, (enc & 7) + 1
                assert first + cnt <= tris.shape[0], "leaf out of range"
                for k in range(first, first + cnt):
                    reached.append(k)
                    if tris[k, 3:4].view(np.int32)[0] < 0:
                        continue  # degenerate sentinel
                    v = tris[k, [0, 1, 2, 4, 5, 6, 8, 9, 10]].reshape(3, 3).astype(np.float64)
                    for blo, bhi in boxes:
                        assert np.all(blo < v.min(axis=0)) and np.all(v.max(axis=0) < bhi), \
                            "a box on the path does not enclose the triangle (padded)"
    return reached, len(seen), depth_max


@pytest.mark.parametrize("n,max_leaf", [(0, 4), (1, 4), (3, 4), (4, 4), (5, 4), (1000, 4), (5000, 1), (5000, 8)])
def test_wide_tree_structure(rt, n, max_leaf):
    if n:
        xyz, idx = rt.gen_soup(n, seed=n + 11)
    else:
        xyz, idx = np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32)
    inf, nodes, nodes4, tris = host_bvh4(rt, xyz, idx, max_leaf)
    assert inf.width == 4
    reached, nn, depth = walk4(nodes4, tris)
    assert nn == inf.num_nodes4
    assert depth == inf.depth4
    ids = tris[:, 3].copy().view(np.int32)
    real = sorted(int(ids[k]) for k in reached if ids[k] >= 0)
    assert real == list(range(n)), "every triangle in exactly one leaf"
    if n > 100:
        assert inf.num_nodes4 < 0.6 * inf.num_nodes  # a wide node stands for about three two-wide nodes


@pytest.mark.parametrize("n", [1, 5, 2000, 30000])
def test_wide_traversal_matches_brute_force(rt, orc, n):
    xyz, idx = rt.gen_soup(n, seed=n + 5)
    inf, nodes, nodes4, tris = host_bvh4(rt, xyz, idx)
    rays = rt.gen_rays(64, 32, n_incoherent=4096, seed=n + 6)
    rng = np.random.default_rng(n)
    m = 512
    ax = [rng.uniform(0, 1, m).astype(np.float32) for _ in range(3)]
    dirs = np.zeros((3, m), np.float32)
    dirs[rng.integers(0, 3, m), np.arange(m)] = rng.choice([-1.0, 1.0], m)
    rays = [np.concatenate([rays[k], ax[k] if k < 3 else dirs[k - 3]]) for k in range(6)]
    bp, bt, fl = orc.brute_trace(xyz, idx, rays)
    p4, t4, nn4, nt4, msp = orc.bvh4_trace(nodes4, tris, rays)
    p2, t2, nn2, nt2 = orc.bvh2_trace(nodes, tris, rays)
    clean = fl == 0
    assert np.array_equal(p4[clean], bp[clean])
    assert np.array_equal(t4[clean], bt[clean])
    assert msp <= inf.stack_cap - 1, "the stack bound the kernel allocates is too small"
    print(f"[wide] n={n}: visits/ray two-wide {nn2 / len(p2):.2f}, four-wide {nn4 / len(p4):.2f}; "
          f"tests/ray {nt2 / len(p2):.2f} vs {nt4 / len(p4):.2f}; stack {msp} of {inf.stack_cap}")
    if n >= 2000:
        assert nn4 < 0.75 * nn2


def test_wide_traversal_tmax_and_far_origins(rt, orc):
    """tMax clipping and origins well outside the scene (camera-style rays)."""
    xyz, idx = rt.gen_soup(3000, seed=99, edge_mu=0.05)
    inf, nodes, nodes4, tris = host_bvh4(rt, xyz, idx)
    rng = np.random.default_rng(5)
    m = 4096
    o = rng.uniform(-20, 20, (m, 3)).astype(np.float32)
    tgt = rng.uniform(0, 1, (m, 3)).astype(np.float32)
    d = (tgt - o).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = [o[:, 0].copy(), o[:, 1].copy(), o[:, 2].copy(), d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy()]
    tmax = rng.uniform(0, 40, m).astype(np.float32)
    bp, bt, fl = orc.brute_trace(xyz, idx, rays, tmax=tmax)
    p4, t4, _, _, _ = orc.bvh4_trace(nodes4, tris, rays, tmax=tmax)
    clean = fl == 0
    assert np.array_equal(p4[clean], bp[clean])
    assert np.array_equal(t4[clean], bt[clean])
    assert (bp[clean] >= 0).sum() > 500


def test_self_candidates_do_not_hide_hits(rt, orc):
    """Rays spawned from a large wall re-meet it at t ~ 0 as a candidate that
    fails the deltaT rejection.  The library (and its restated traversal)
    applies deltaT before lowering the hit distance, so such rays still get
    the closest passing hit, equal to brute force on every ray.  The
    reference's order (hitOffset lowered first, Triangle.cu:569-593, kept in
    the oracle's literal test and reference tree) loses the hit whenever the
    wall is tested first; its Morton tree does that for some of these rays."""
    import material_scenes as ms
    W = 64
    d = ms.probe("lambert", width=W, height=W, spp=1)
    d.max_depth = 2
    s = orc.OrcScene(d)
    segs = []
    for y in range(W):
        for x in range(W):
            ok, L, seg = s.path_segments(x, y, 0)
            assert ok and len(seg) == 2
            assert np.array_equal(L, np.float32(0.5) * np.array(ms.KE, np.float32))
            segs.append(seg[1])
    segs = np.array(segs, np.float32)
    rays = [segs[:, k] for k in range(6)]
    inf, nodes, nodes4, tris = host_bvh4(rt, d.xyz, d.idx)
    bp, bt, fl = orc.brute_trace(d.xyz, d.idx, rays)
    masked = (fl & 2) != 0
    assert masked.sum() > 100, "the scene must produce self candidates"
    assert (bp >= 0).all()
    p4, t4, *_ = orc.bvh4_trace(nodes4, tris, rays)
    p2, t2, *_ = orc.bvh2_trace(nodes, tris, rays)
    assert np.array_equal(p4, bp) and np.array_equal(t4, bt)
    assert np.array_equal(p2, bp) and np.array_equal(t2, bt)
    ref, _ = orc.RefBvh(d.xyz, d.idx).trace(rays)
    assert (ref[masked] < 0).sum() > 0, "the reference's order loses some of these hits"
    assert np.array_equal(ref[~masked], bp[~masked])
