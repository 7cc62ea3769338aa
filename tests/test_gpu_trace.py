"""GPU parity of rt_trace_closest / rt_trace_any (called through the C ABI)
against the oracle: brute force over every triangle with the reference's
Triangle::intersects arithmetic (shapes/Triangle.cu:433-643), the restated
reference BVH (cuda_libraries/BVH.cu:355-448) and the restated kernel
traversal on the same exported tree.

Bar: hit ids bit-exact and t bitwise equal on every ray that is not flagged
ambiguous (exact-t ties, deltaT masking: Triangle.cu:569-593)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def to_dev(arrs):
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]


def edge_rays(n, seed):
    """Axis-parallel rays (zero direction components) and rays starting on
    the unit-cube faces."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(0, 1, (3, n)).astype(np.float32)
    d = np.zeros((3, n), np.float32)
    d[rng.integers(0, 3, n), np.arange(n)] = rng.choice([-1.0, 1.0], n)
    return [o[0], o[1], o[2], d[0], d[1], d[2]]


def cat(a, b):
    return [np.concatenate([x, y]) for x, y in zip(a, b)]


@pytest.mark.parametrize("width", [4, 2])
@pytest.mark.parametrize("nt", [0, 1, 3, 4, 5, 100, 4096, 32768])
def test_closest_hit_matches_brute_force(rt, orc, gpu_ctx, nt, width):
    print(f"[closest nt={nt} width={width}]", flush=True)
    if nt:
        xyz, idx = rt.gen_soup(nt, seed=100 + nt)
    else:
        xyz, idx = np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32)
    rays = cat(rt.gen_rays(64, 48, n_incoherent=4096, seed=nt), edge_rays(512, nt))
    bvh = rt.Bvh(gpu_ctx, xyz, idx, width=width)
    prim, t = bvh.trace_closest(*to_dev(rays))
    torch.cuda.synchronize()
    prim, t = prim.cpu().numpy(), t.cpu().numpy()
    bp, bt, fl = orc.brute_trace(xyz, idx, rays)
    clean = fl == 0
    assert clean.mean() > 0.99
    assert np.array_equal(prim[clean], bp[clean])
    assert np.array_equal(t[clean], bt[clean])
    assert np.all(np.isinf(t[prim < 0]))
    bvh.close()


def test_tmax_clips_closest_hit(rt, orc, gpu_ctx):
    xyz, idx = rt.gen_soup(4096, seed=5)
    rays = rt.gen_rays(32, 32, n_incoherent=4096, seed=6)
    bp, bt, fl = orc.brute_trace(xyz, idx, rays)
    rng = np.random.default_rng(1)
    tmax = np.where(bp >= 0, bt * rng.uniform(0.5, 1.5, bp.shape), 3.0).astype(np.float32)
    cp, ct, cfl = orc.brute_trace(xyz, idx, rays, tmax=tmax)
    bvh = rt.Bvh(gpu_ctx, xyz, idx)
    prim, t = bvh.trace_closest(*to_dev(rays), tmax=torch.from_numpy(tmax).cuda())
    torch.cuda.synchronize()
    clean = cfl == 0
    assert np.array_equal(prim.cpu().numpy()[clean], cp[clean])
    assert np.array_equal(t.cpu().numpy()[clean], ct[clean])


def test_any_hit_with_transparency(rt, orc, gpu_ctx):
    xyz, idx = rt.gen_soup(8192, seed=9, edge_mu=0.03)  # larger triangles: many occluders
    rays = rt.gen_rays(32, 32, n_incoherent=6000, seed=10)
    rng = np.random.default_rng(2)
    transparent = (rng.uniform(size=idx.shape[0]) < 0.3).astype(np.uint8)
    bp, bt, fl = orc.brute_trace(xyz, idx, rays, transparent=transparent, check_shadow=True)
    tmax = np.where(bp >= 0, bt * rng.choice([0.5, 2.0], bp.shape), 5.0).astype(np.float32)
    bvh = rt.Bvh(gpu_ctx, xyz, idx)
    bvh.set_transparent(transparent)
    occ = bvh.trace_any(*to_dev(rays), torch.from_numpy(tmax).cuda())
    torch.cuda.synchronize()
    occ = occ.cpu().numpy().astype(bool)
    want = (bp >= 0) & (bt <= tmax)
    clean = fl == 0
    assert np.array_equal(occ[clean], want[clean])
    assert want.sum() > 100 and (~want).sum() > 100


def test_host_pointer_path_equals_device_path(rt, gpu_ctx):
    xyz, idx = rt.gen_soup(20000, seed=3)
    rays = rt.gen_rays(64, 64, n_incoherent=8192, seed=4)
    bvh = rt.Bvh(gpu_ctx, xyz, idx)
    hp, ht = bvh.trace_closest(*rays, host=True)
    dp, dt = bvh.trace_closest(*to_dev(rays))
    torch.cuda.synchronize()
    assert np.array_equal(hp, dp.cpu().numpy())
    assert np.array_equal(ht, dt.cpu().numpy())


@pytest.mark.parametrize("width", [2, 4])
def test_kernel_counts_match_oracle_traversal(rt, orc, gpu_ctx, width):
    """The counting kernel variant against the oracle's restatement of the
    kernel traversal on the same tree.  Two-wide: exactly the same node
    visits and triangle tests, and the same hit on every ray (same order).
    Four-wide: the kernel parks reached leaves until the wavefront's leaf
    phase, so its running hit distance culls a little later -- it visits
    every node and tests every triangle the immediate-leaf order does, plus
    a few more (bounded here), and hits differ only on rays whose result
    depends on the test order (ambiguous for brute force)."""
    xyz, idx = rt.gen_soup(50000, seed=12)
    rays = rt.gen_rays(64, 64, n_incoherent=8192, seed=13)
    bvh = rt.Bvh(gpu_ctx, xyz, idx, width=width)
    nodes, tris = bvh.export()
    if width == 2:
        p2, t2, nn, ntr = orc.bvh2_trace(nodes, tris, rays)
    else:
        p2, t2, nn, ntr, _ = orc.bvh4_trace(bvh.export4(), tris, rays)
    gn, gt = bvh.trace_stats(*to_dev(rays))
    prim, t = bvh.trace_closest(*to_dev(rays))
    torch.cuda.synchronize()
    prim, t = prim.cpu().numpy(), t.cpu().numpy()
    print(f"[counts width={width}] nodes {gn} vs {nn} ({gn / nn:.4f}), tris {gt} vs {ntr} ({gt / max(ntr, 1):.4f})",
          flush=True)
    if width == 2:
        assert (gn, gt) == (nn, ntr)
        assert np.array_equal(prim, p2)  # same tree, same order: every ray
        assert np.array_equal(t, t2)
    else:
        assert nn <= gn <= 1.15 * nn
        assert ntr <= gt <= 1.3 * ntr
        diff = np.nonzero((prim != p2) | (t != t2))[0]
        assert diff.size <= 1e-3 * prim.size
        if diff.size:
            _, _, fl = orc.brute_trace(xyz, idx, [a[diff] for a in rays])
            assert np.all(fl != 0), diff[fl == 0][:10]


def test_full_size_c1_against_reference_bvh(rt, orc, gpu_ctx):
    """1M-triangle soup (seed 7): a sample of coherent and incoherent rays
    through the kernel, the restated kernel traversal and the restated
    reference BVH; disagreements with the reference tree must be rays that
    brute force flags as ambiguous or that the reference's non-conservative
    slab test loses."""
    print("[c1 full size] building 1M soup", flush=True)
    xyz, idx = rt.gen_soup(1 << 20, seed=7)
    coh = rt.gen_rays(4096, 2048, seed=7)
    inc = rt.gen_rays(0, 0, n_incoherent=1 << 16, seed=7)
    sel = np.random.default_rng(0).choice(4096 * 2048, 1 << 16, replace=False)
    rays = cat([a[sel] for a in coh], inc)
    bvh = rt.Bvh(gpu_ctx, xyz, idx)
    prim, t = bvh.trace_closest(*to_dev(rays))
    torch.cuda.synchronize()
    prim, t = prim.cpu().numpy(), t.cpu().numpy()
    nodes, tris = bvh.export()
    p2, t2, _, _, _ = orc.bvh4_trace(bvh.export4(), tris, rays, nthreads=16)
    # the kernel's leaf order differs from the restated immediate order: only
    # order-dependent (ambiguous) rays may differ
    dk = np.nonzero((prim != p2) | (t != t2))[0]
    print(f"[c1 full size] kernel vs restated traversal: {dk.size} rays differ", flush=True)
    assert dk.size <= 1e-4 * prim.size
    if dk.size:
        _, _, fk = orc.brute_trace(xyz, idx, [a[dk] for a in rays], nthreads=16)
        assert np.all(fk != 0), dk[fk == 0][:10]
    rp, rtt = orc.RefBvh(xyz, idx).trace(rays, nthreads=16)
    diff = np.nonzero(prim != rp)[0]
    print(f"[c1 full size] hits {np.mean(prim >= 0):.3f}, differ from reference tree: {diff.size}", flush=True)
    assert diff.size < 1e-3 * prim.size
    if diff.size:
        sub = [a[diff] for a in rays]
        bp, bt, fl = orc.brute_trace(xyz, idx, sub, nthreads=16)
        explained = (fl != 0) | ((bp == prim[diff]) & (bt == t[diff]))
        assert explained.all(), diff[~explained][:10]


def test_shadow_query_replays_scene_is_visible_through_the_c_abi(rt, orc, gpu_ctx):
    """The reference's shadow test as its callers consume the boundary:
    Scene::intersects calls BVH::intersectRays(rays, ids, checkShadows=true)
    on host rays with maximumOffset = +inf (spawnRay), re-tests
    triangles[id].intersects(ray, hitOffset = inf, checkShadow) for t
    (Scene.cu:108-150), and Scene::isVisible says visible = !hit || t > dist
    (Scene.cu:238-276).  Here the boundary call is rt_trace_closest with
    RT_FLAG_HOST_PTRS | RT_FLAG_SHADOW through ctypes on numpy buffers, the
    re-test is the oracle's literal Triangle::intersects, and the result must
    equal the oracle's brute-force closest opaque hit, and rt_trace_any with
    tmax = the light distance, on every unambiguous ray."""
    import ctypes as C
    print("[shadow query] 8192 triangles, 30 % transparent", flush=True)
    xyz, idx = rt.gen_soup(8192, seed=21, edge_mu=0.03)
    rng = np.random.default_rng(22)
    transparent = (rng.uniform(size=idx.shape[0]) < 0.3).astype(np.uint8)
    n = 8192
    o = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    light = rng.uniform(-0.2, 1.2, (n, 3)).astype(np.float32)
    to_l = (light - o).astype(np.float32)
    dist = np.linalg.norm(to_l, axis=1).astype(np.float32)
    d = (to_l / dist[:, None]).astype(np.float32)
    rays = [np.ascontiguousarray(o[:, 0]), np.ascontiguousarray(o[:, 1]), np.ascontiguousarray(o[:, 2]),
            np.ascontiguousarray(d[:, 0]), np.ascontiguousarray(d[:, 1]), np.ascontiguousarray(d[:, 2])]
    bvh = rt.Bvh(gpu_ctx, xyz, idx)
    bvh.set_transparent(transparent)
    L = rt.lib()
    r = rt.Rays(*[a.ctypes.data for a in rays], None)  # maximumOffset = +inf
    ids = np.empty(n, np.int32)
    tt = np.empty(n, np.float32)
    rt.check(L.rt_trace_closest(bvh.h, C.byref(r), n, ids.ctypes.data, tt.ctypes.data,
                                rt.RT_FLAG_HOST_PTRS | rt.RT_FLAG_SHADOW, None))
    # Scene::intersects: re-test the returned triangle for its distance
    tri = xyz[idx]
    visible = np.ones(n, bool)
    for i in np.nonzero(ids >= 0)[0]:
        ok, t_re, _ = orc.tri_intersects(o[i], d[i], *tri[ids[i]])
        assert ok, i
        assert t_re == tt[i], (i, t_re, tt[i])
        visible[i] = t_re > dist[i]
    bp, bt, fl = orc.brute_trace(xyz, idx, rays, transparent=transparent, check_shadow=True)
    clean = fl == 0
    assert clean.mean() > 0.99
    assert np.array_equal(ids[clean], bp[clean]) and np.array_equal(tt[clean], bt[clean])
    assert not np.any(transparent[ids[ids >= 0]])
    want_visible = (bp < 0) | (bt > dist)
    assert np.array_equal(visible[clean], want_visible[clean])
    occ = bvh.trace_any(*rays, dist, host=True).astype(bool)
    assert np.array_equal(occ[clean], ~want_visible[clean])
    assert want_visible.sum() > 200 and (~want_visible).sum() > 200
    # without the flag, transparent triangles count
    ids2, _ = bvh.trace_closest(*rays, host=True)
    assert np.any(transparent[ids2[ids2 >= 0]])
    bvh.close()
