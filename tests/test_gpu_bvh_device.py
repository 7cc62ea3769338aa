"""rt_bvh_build_device (the BVH built on the GPU from device arrays, the
arrangement of the reference's BVH constructor, cuda_libraries/BVH.cu:
314-341) against the oracle's restatement of the same build
(oracle/lbvh.c): the exported two-wide nodes, triangles and four-wide nodes
are bitwise equal and the reported figures match; traces on the device-built
tree agree with brute force on every unambiguous ray; transparency and the
error paths work as on the host-built tree."""
import ctypes as C

import numpy as np
import pytest
import torch

from test_oracle_lbvh import soup

pytestmark = pytest.mark.gpu


def to_dev(arrs):
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]


def device_bvh(rt, ctx, xyz, idx, max_leaf=4, width=4):
    x = torch.from_numpy(np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)).cuda()
    i = torch.from_numpy(np.ascontiguousarray(idx, np.uint32).reshape(-1, 3).view(np.int32)).cuda()
    if x.shape[0] == 0:  # the entry point wants a vertex for any triangle; none here
        x = torch.zeros((1, 3), dtype=torch.float32, device="cuda")
    return rt.Bvh.from_device(ctx, x, i, max_leaf=max_leaf, width=width)


CASES = [(0, 4, 4, False), (1, 4, 4, False), (2, 4, 2, False), (3, 4, 4, False), (5, 1, 4, False),
         (100, 4, 2, False), (1000, 8, 4, False), (3000, 1, 4, True), (4096, 4, 4, False),
         (50000, 4, 4, False), (50000, 4, 2, False), (200000, 2, 4, True)]


@pytest.mark.parametrize("n,max_leaf,width,dup", CASES)
def test_device_build_equals_oracle(rt, orc, gpu_ctx, n, max_leaf, width, dup):
    print(f"[bvh device] n={n} max_leaf={max_leaf} width={width} dup={dup}", flush=True)
    xyz, idx = soup(rt, n, seed=n + 3, dup=dup)
    bvh = device_bvh(rt, gpu_ctx, xyz, idx, max_leaf, width)
    inf = bvh.info()
    nodes, tris = bvh.export()
    want, wn, wt, w4 = orc.lbvh_build(xyz, idx, max_leaf=max_leaf, wide=width == 4)
    assert (inf.num_nodes, inf.num_leaves, inf.max_depth) == (want["num_nodes"], want["num_leaves"],
                                                              want["max_depth"])
    assert inf.num_tris == n and inf.width == width
    assert np.array_equal(nodes.view(np.uint32), wn.view(np.uint32))
    assert np.array_equal(tris.view(np.uint32), wt.view(np.uint32))
    if width == 4:
        assert (inf.num_nodes4, inf.depth4, inf.stack_cap) == (want["num_nodes4"], want["depth4"],
                                                               want["max_stack4"] + 1)
        assert np.array_equal(bvh.export4().view(np.uint32), w4.view(np.uint32))
    else:
        assert inf.num_nodes4 == 0 and inf.stack_cap == want["max_depth"] + 1
    print(f"[bvh device] n={n}: {inf.num_nodes} nodes, depth {inf.max_depth}, sah {inf.sah_cost:.2f}, "
          f"{inf.build_ms:.2f} ms", flush=True)
    bvh.close()


@pytest.mark.parametrize("width", [4, 2])
@pytest.mark.parametrize("n", [1, 5, 4096, 32768])
def test_traces_on_device_tree_match_brute_force(rt, orc, gpu_ctx, n, width):
    print(f"[bvh device trace] n={n} width={width}", flush=True)
    xyz, idx = soup(rt, n, seed=100 + n)
    rays = rt.gen_rays(64, 48, n_incoherent=4096, seed=n)
    bvh = device_bvh(rt, gpu_ctx, xyz, idx, width=width)
    prim, t = bvh.trace_closest(*to_dev(rays))
    torch.cuda.synchronize()
    prim, t = prim.cpu().numpy(), t.cpu().numpy()
    bp, bt, fl = orc.brute_trace(xyz, idx, rays)
    clean = fl == 0
    assert clean.mean() > 0.99
    assert np.array_equal(prim[clean], bp[clean])
    assert np.array_equal(t[clean], bt[clean])
    bvh.close()


def test_any_hit_with_transparency_on_device_tree(rt, orc, gpu_ctx):
    print("[bvh device any-hit]", flush=True)
    xyz, idx = rt.gen_soup(8192, seed=9, edge_mu=0.03)
    rays = rt.gen_rays(32, 32, n_incoherent=6000, seed=10)
    rng = np.random.default_rng(2)
    transparent = (rng.uniform(size=idx.shape[0]) < 0.3).astype(np.uint8)
    bp, bt, fl = orc.brute_trace(xyz, idx, rays, transparent=transparent, check_shadow=True)
    tmax = np.where(bp >= 0, bt * rng.choice([0.5, 2.0], bp.shape), 5.0).astype(np.float32)
    bvh = device_bvh(rt, gpu_ctx, xyz, idx)
    bvh.set_transparent(transparent)
    occ = bvh.trace_any(*to_dev(rays), torch.from_numpy(tmax).cuda())
    torch.cuda.synchronize()
    occ = occ.cpu().numpy().astype(bool)
    want = (bp >= 0) & (bt <= tmax)
    clean = fl == 0
    assert np.array_equal(occ[clean], want[clean])
    assert want.sum() > 100 and (~want).sum() > 100
    # the flags land on the exported records by triangle id
    _, tris = bvh.export()
    ids = tris[:-1, 3].copy().view(np.int32)
    assert np.array_equal(tris[:-1, 7].copy().view(np.uint32), transparent[ids].astype(np.uint32))
    bvh.close()


def test_bad_input_is_refused(rt, gpu_ctx):
    print("[bvh device errors]", flush=True)
    xyz, idx = rt.gen_soup(64, seed=1)
    bad = idx.copy()
    bad[3, 1] = xyz.shape[0]  # one past the last vertex: the kernel reads vertex 0 instead and reports it
    with pytest.raises(RuntimeError, match="out of range"):
        device_bvh(rt, gpu_ctx, xyz, bad)
    x2 = xyz.copy()
    x2[5, 0] = np.inf
    with pytest.raises(RuntimeError, match="non-finite"):
        device_bvh(rt, gpu_ctx, x2, idx)
    h = C.c_void_p()
    opts = rt.BvhOpts(9, 32, 1.2, 4)
    assert rt.lib().rt_bvh_build_device(gpu_ctx.h, None, 0, None, 0, C.byref(opts), C.byref(h)) != 0
    # the context still works
    bvh = device_bvh(rt, gpu_ctx, xyz, idx)
    assert bvh.info().num_tris == 64
    bvh.close()


def test_full_size_device_build(rt, orc, gpu_ctx):
    """The 1M-triangle soup (seed 7): the device build equals the oracle's
    record for record, and the kernel's closest hits on it agree with the
    restated traversal of the same tree except on rays brute force flags as
    ambiguous."""
    print("[bvh device full size] 1M soup", flush=True)
    xyz, idx = rt.gen_soup(1 << 20, seed=7)
    bvh = device_bvh(rt, gpu_ctx, xyz, idx)
    inf = bvh.info()
    print(f"[bvh device full size] {inf.build_ms:.1f} ms, {inf.num_nodes} nodes, depth {inf.max_depth}, "
          f"wide {inf.num_nodes4} depth4 {inf.depth4} stack {inf.stack_cap}, sah {inf.sah_cost:.2f}", flush=True)
    nodes, tris = bvh.export()
    nodes4 = bvh.export4()
    want, wn, wt, w4 = orc.lbvh_build(xyz, idx)
    assert np.array_equal(nodes.view(np.uint32), wn.view(np.uint32))
    assert np.array_equal(tris.view(np.uint32), wt.view(np.uint32))
    assert np.array_equal(nodes4.view(np.uint32), w4.view(np.uint32))
    coh = rt.gen_rays(4096, 2048, seed=7)
    inc = rt.gen_rays(0, 0, n_incoherent=1 << 15, seed=7)
    sel = np.random.default_rng(0).choice(4096 * 2048, 1 << 15, replace=False)
    rays = [np.concatenate([a[sel], b]) for a, b in zip(coh, inc)]
    prim, t = bvh.trace_closest(*to_dev(rays))
    torch.cuda.synchronize()
    prim, t = prim.cpu().numpy(), t.cpu().numpy()
    p2, t2, _, _, _ = orc.bvh4_trace(nodes4, tris, rays, nthreads=16)
    dk = np.nonzero((prim != p2) | (t != t2))[0]
    print(f"[bvh device full size] kernel vs restated traversal: {dk.size} rays differ", flush=True)
    assert dk.size <= 1e-4 * prim.size
    if dk.size:
        _, _, fk = orc.brute_trace(xyz, idx, [a[dk] for a in rays], nthreads=16)
        assert np.all(fk != 0), dk[fk == 0][:10]
    bvh.close()
