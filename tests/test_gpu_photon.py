"""Photon map k-nearest search on the GPU (rt_photon_map_knn) against the
oracle's brute force and restated tree search (cuda_libraries/kdtree.h,
integration/Scene.cu:704-722).

Counts must be equal for every query.  On queries whose k-th and (k+1)-th
qualifying distances differ, the photon sets must be equal and the float
distances bit-identical (same formula, correctly rounded square root)."""
import numpy as np
import pytest
import torch

import oracle
from test_oracle_photon import as_sets, photons, unambiguous

pytestmark = pytest.mark.gpu


def gpu_knn(rt, ctx, x, q, k, radius, shape=None, qshape=None):
    dev = torch.device("cuda", ctx.device)
    pm = rt.PhotonMap(ctx, x, shape)
    qs = None if qshape is None else torch.from_numpy(qshape).to(dev)
    cnt, idx, dist = pm.nearest(torch.from_numpy(q).to(dev), k, radius, qs)
    torch.cuda.synchronize()
    out = cnt.cpu().numpy().astype(np.uint32), idx.cpu().numpy().astype(np.uint32), dist.cpu().numpy()
    pm.close()
    return out


def assert_same(got, want, ok):
    gc, gi, gd = got
    wc, wi, wd = want
    assert np.array_equal(gc, wc)
    for i in np.flatnonzero(ok):
        assert set(gi[i, :gc[i]].tolist()) == set(wi[i, :wc[i]].tolist()), i
        assert np.array_equal(np.sort(gd[i, :gc[i]]), np.sort(wd[i, :wc[i]])), i


@pytest.mark.parametrize("k,radius", [(1, np.inf), (8, np.inf), (100, 100.0), (20, 0.15), (64, 0.05)])
@pytest.mark.parametrize("clustered", [False, True])
def test_knn_matches_brute_force(rt, gpu_ctx, k, radius, clustered):
    print(f"[photon] k={k} r={radius} clustered={clustered}", flush=True)
    x = photons(20000, 41, clustered)
    q = photons(2048, 42, clustered)
    got = gpu_knn(rt, gpu_ctx, x, q, k, radius)
    want = oracle.knn(x, q, k, radius)
    ok = unambiguous(x, q, k, radius)
    assert ok.mean() > 0.9
    assert_same(got, want, ok)


def test_shape_filter(rt, gpu_ctx):
    print("[photon] shape filter", flush=True)
    x = photons(10000, 51)
    sh = np.random.default_rng(52).integers(0, 6, len(x)).astype(np.int32)
    q = photons(1024, 53)
    qs = np.random.default_rng(54).integers(-1, 6, len(q)).astype(np.int32)
    got = gpu_knn(rt, gpu_ctx, x, q, 32, 0.4, sh, qs)
    want = oracle.knn(x, q, 32, 0.4, sh, qs)
    assert_same(got, want, unambiguous(x, q, 32, 0.4, sh, qs))


def test_reference_constants_on_a_million_photons(rt, gpu_ctx):
    """N_PHOTON_SAMPLES = 1M photons, MAX_NEIGHBOURS = 100 within
    PHOTON_SEARCH_RADIUS = 100 (cuda_libraries/types.h:57-60): the GPU equals
    the oracle's tree search on every unambiguous query and the brute force
    on a sample."""
    print("[photon] 1M photons, 8192 queries, k=100, r=100", flush=True)
    x = photons(1 << 20, 61, clustered=True)
    q = photons(8192, 62, clustered=True)
    got = gpu_knn(rt, gpu_ctx, x, q, 100, 100.0)
    tree = oracle.kd_build(x)[:2]
    want = oracle.knn(x, q, 100, 100.0, tree=tree, nthreads=8)
    ok = unambiguous(x, q[:256], 100, 100.0)
    assert_same(tuple(a[:256] for a in got), oracle.knn(x, q[:256], 100, 100.0), ok)
    sets_g, sets_w = as_sets(got[0], got[1]), as_sets(want[0], want[1])
    same = np.array([a == b for a, b in zip(sets_g, sets_w)])
    assert np.array_equal(got[0], want[0])
    assert same.mean() > 0.98


def test_edge_cases(rt, gpu_ctx):
    x = photons(5, 71)
    q = photons(9, 72)
    cnt, idx, dist = gpu_knn(rt, gpu_ctx, x, q, 8, np.inf)
    assert np.all(cnt == 5)
    for i in range(len(q)):
        assert set(idx[i, :5].tolist()) == set(range(5))
    cnt, _, _ = gpu_knn(rt, gpu_ctx, x, q, 3, 1e-6)
    assert np.all(cnt == 0)
    pm = rt.PhotonMap(gpu_ctx, np.zeros((0, 3), np.float32))
    c, _, _ = pm.nearest(torch.from_numpy(q).cuda(), 4)
    torch.cuda.synchronize()
    assert int(c.sum()) == 0 and pm.info() == (0, 0)
    pm.close()


def test_bad_arguments(rt, gpu_ctx):
    with pytest.raises(rt.RtError):
        rt.PhotonMap(gpu_ctx, np.array([[0, 0, np.nan]], np.float32))
    pm = rt.PhotonMap(gpu_ctx, photons(10, 81))
    q = torch.zeros((4, 3), dtype=torch.float32, device="cuda")
    for k, r in [(0, 1.0), (2000, 1.0), (4, 0.0), (4, -1.0)]:
        with pytest.raises(rt.RtError):
            pm.nearest(q, k, r)
    pm.close()


@pytest.mark.parametrize("k,radius", [(1, np.inf), (20, 0.3), (100, 100.0)])
def test_gather_lambert_matches_oracle(rt, gpu_ctx, k, radius):
    """combinePhotonContributions on Lambertian surfaces (Scene.cu:607-634):
    equal to the oracle within 1e-12 relative on unambiguous queries (same
    photons; only the order of the double sum may differ)."""
    print(f"[photon] gather k={k} r={radius}", flush=True)
    rng = np.random.default_rng(91)
    x = photons(20000, 92, clustered=True)
    sh = rng.integers(0, 4, len(x)).astype(np.int32)
    rgb = rng.random((len(x), 3)) * 5
    d = rng.normal(size=(len(x), 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    q = photons(2048, 93, clustered=True)
    nrm = rng.normal(size=(len(q), 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    kd = rng.random((len(q), 3)).astype(np.float32)
    qs = rng.integers(-1, 4, len(q)).astype(np.int32)
    dev = torch.device("cuda", gpu_ctx.device)
    pm = rt.PhotonMap(gpu_ctx, x, sh)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    got = pm.gather_lambert(t(rgb), t(d), t(q), t(nrm), t(kd), k, radius, t(qs))
    torch.cuda.synchronize()
    got = got.cpu().numpy()
    pm.close()
    want = oracle.photon_gather_lambert(x, rgb, d, q, nrm, kd, k, radius, sh, qs)
    ok = unambiguous(x, q, k, radius, sh, qs)
    assert ok.mean() > 0.9
    assert np.allclose(got[ok], want[ok], rtol=1e-12, atol=0)


@pytest.mark.parametrize("k", [1, 2, 5, 16, 100, 129, 1024])
def test_result_order_equals_the_reference_heap_order(rt, gpu_ctx, k):
    """The search returns its vector as kdtree.h leaves it: insertion order
    until k photons were accepted, then the layout of libstdc++'s
    make_heap / push_heap / pop_heap (the oracle's order is checked against
    the standard library itself in test_oracle_photon_order.py).  With no
    coordinate ties the GPU's tree is the oracle's, so indices and
    distances must agree position by position on every query."""
    print(f"[photon] result order k={k}", flush=True)
    rng = np.random.default_rng(70 + k)
    # distinct coordinates on every axis (random float32s collide): one tree
    x = np.stack([(rng.permutation(1 << 20)[:30000] + 0.5) / (1 << 20) for _ in range(3)], 1).astype(np.float32)
    for a in range(3):
        assert np.unique(x[:, a]).size == x.shape[0]
    q = rng.random((4096, 3)).astype(np.float32)
    q[:64] = x[:64]  # queries on photons: a distance-0 candidate
    radius = 0.2 if k <= 100 else 0.5  # k > 128: the heap in the output slice
    got = gpu_knn(rt, gpu_ctx, x, q, k, radius)
    tree = oracle.kd_build(x)[:2]
    want = oracle.knn(x, q, k, radius, tree=tree)
    assert np.array_equal(got[0], want[0])
    for i in range(q.shape[0]):
        c = got[0][i]
        assert np.array_equal(got[1][i, :c], want[1][i, :c]), i
        assert np.array_equal(got[2][i, :c], want[2][i, :c]), i
    assert (got[0] == k).mean() > 0.5


@pytest.mark.parametrize("k,radius", [(16, 0.2), (100, 100.0)])
def test_shape_filtered_order_equals_the_reference_heap_order(rt, gpu_ctx, k, radius):
    """Shape-restricted searches, as the photon-mapping render makes them:
    the GPU passes over subtrees whose shape signature (a 32-bit OR of the
    shapes' bits) lacks the query's shape, which must leave the walk's
    insertions, and so the result order, exactly the oracle's.  Shapes are
    spatial cells (as surfaces are), 40 of them so that shapes 32 apart share
    a signature bit, plus queries with no shape (-1) and on shapes with no
    photon."""
    print(f"[photon] shape-filtered result order k={k} r={radius}", flush=True)
    rng = np.random.default_rng(90 + k)
    x = np.stack([(rng.permutation(1 << 20)[:30000] + 0.5) / (1 << 20) for _ in range(3)], 1).astype(np.float32)
    sh = (np.floor(x[:, 0] * 8) + 8 * np.floor(x[:, 1] * 5)).astype(np.int32)
    sh[rng.random(len(x)) < 0.02] = -1  # photons with no shape
    q = rng.random((4096, 3)).astype(np.float32)
    q[:64] = x[:64]
    qs = (np.floor(q[:, 0] * 8) + 8 * np.floor(q[:, 1] * 5)).astype(np.int32)
    qs[rng.random(len(q)) < 0.1] = -1
    qs[rng.random(len(q)) < 0.05] = 45  # no photon of this shape
    got = gpu_knn(rt, gpu_ctx, x, q, k, radius, sh, qs)
    tree = oracle.kd_build(x)[:2]
    want = oracle.knn(x, q, k, radius, sh, qs, tree=tree)
    assert np.array_equal(got[0], want[0])
    for i in range(q.shape[0]):
        c = got[0][i]
        assert np.array_equal(got[1][i, :c], want[1][i, :c]), i
        assert np.array_equal(got[2][i, :c], want[2][i, :c]), i
    assert np.all(got[0][qs == 45] == 0)
    assert (got[0] == k).mean() > 0.5
