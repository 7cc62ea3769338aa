"""Oracle pinning for the photon map (cuda_libraries/kdtree.h build_tree and
nearest_neighbors, as Scene::search_nearest calls them,
integration/Scene.cu:704-722).

The brute force states the result by definition: the k photons of smallest
float distance strictly within the radius (on the query's shape).  The
restated tree search must return the same set wherever the k-th and the
(k+1)-th qualifying distances differ (elsewhere the reference's own result
depends on its visit order).  The restated build must have the kd-tree's
properties at every node: the median splits its range on the axis of the
range's largest extent, with every element of the left part at or below the
split coordinate and every element of the right part at or above it."""
import numpy as np
import pytest

import oracle


def photons(n, seed, clustered=False):
    rng = np.random.default_rng(seed)
    if clustered:
        c = rng.random((16, 3)) * 10
        x = c[rng.integers(0, 16, n)] + rng.normal(0, 0.2, (n, 3))
    else:
        x = rng.random((n, 3)) * np.array([4.0, 1.0, 2.0])
    return x.astype(np.float32)


def unambiguous(xyz, queries, k, radius, shape=None, query_shape=None):
    """Queries whose k-th and (k+1)-th qualifying distances differ."""
    cnt, _, dist = oracle.knn(xyz, queries, k + 1, radius, shape, query_shape)
    ok = np.ones(len(queries), bool)
    full = cnt == k + 1
    ok[full] = dist[full, k - 1] != dist[full, k]
    return ok


def as_sets(cnt, idx):
    return [frozenset(idx[i, :cnt[i]].tolist()) for i in range(len(cnt))]


@pytest.mark.parametrize("n", [1, 2, 3, 7, 100, 5000])
def test_build_properties(n):
    x = photons(n, n)
    order, axis, depth = oracle.kd_build(x)
    assert sorted(order.tolist()) == list(range(n))
    assert depth == (int(np.floor(np.log2(n))) + 1 if n else 0)

    def check(l, r):
        if r - l <= 1:
            return
        m = (l + r) // 2
        pts = x[order[l:r]]
        ext = pts.max(0) - pts.min(0)
        assert axis[m] == int(np.argmax(ext))
        a = axis[m]
        split = x[order[m], a]
        assert np.all(x[order[l:m], a] <= split) and np.all(x[order[m + 1:r], a] >= split)
        check(l, m)
        check(m + 1, r)
    check(0, n)


@pytest.mark.parametrize("k,radius", [(1, np.inf), (8, np.inf), (100, np.inf), (20, 0.15), (100, 100.0)])
@pytest.mark.parametrize("clustered", [False, True])
def test_tree_search_equals_brute_force(k, radius, clustered):
    x = photons(4000, 11, clustered)
    q = photons(300, 12, clustered)
    tree = oracle.kd_build(x)[:2]
    bc, bi, bd = oracle.knn(x, q, k, radius)
    tc, ti, td = oracle.knn(x, q, k, radius, tree=tree)
    ok = unambiguous(x, q, k, radius)
    assert ok.mean() > 0.9
    assert np.array_equal(bc, tc)
    for i in np.flatnonzero(ok):
        assert set(bi[i, :bc[i]]) == set(ti[i, :tc[i]])
        assert sorted(bd[i, :bc[i]]) == sorted(td[i, :tc[i]])


def test_shape_filter_and_radius():
    x = photons(3000, 21)
    sh = np.random.default_rng(22).integers(0, 5, len(x)).astype(np.int32)
    q = photons(200, 23)
    qs = np.random.default_rng(24).integers(-1, 5, len(q)).astype(np.int32)
    tree = oracle.kd_build(x)[:2]
    bc, bi, bd = oracle.knn(x, q, 16, 0.3, shape=sh, query_shape=qs)
    tc, ti, _ = oracle.knn(x, q, 16, 0.3, shape=sh, query_shape=qs, tree=tree)
    ok = unambiguous(x, q, 16, 0.3, sh, qs)
    assert np.array_equal(bc, tc)
    for i in np.flatnonzero(ok):
        got = set(ti[i, :tc[i]].tolist())
        assert got == set(bi[i, :bc[i]].tolist())
        if qs[i] != -1:
            assert all(sh[j] == qs[i] for j in got)
        d = np.sqrt(((x[list(got)].astype(np.float64) - q[i]) ** 2).sum(1)) if got else np.zeros(0)
        assert np.all(d < 0.3 + 1e-6)


def test_brute_force_against_numpy():
    """The brute force equals a numpy sort of all distances (float64 for the
    order, the oracle's float distances within rounding)."""
    x = photons(2000, 31)
    q = photons(50, 32)
    cnt, idx, dist = oracle.knn(x, q, 10, 0.2)
    d = np.sqrt(((x[None, :, :].astype(np.float64) - q[:, None, :]) ** 2).sum(-1))
    for i in range(len(q)):
        inside = np.flatnonzero(d[i] < 0.2)
        want = inside[np.argsort(d[i, inside], kind="stable")][:10]
        assert cnt[i] == len(want)
        assert np.allclose(dist[i, :cnt[i]], d[i, want], rtol=1e-6)
        # equal as sets up to exchanges among (near-)equal distances
        assert len(set(idx[i, :cnt[i]]) ^ set(want)) <= 2


def test_gather_closed_form():
    """One photon of radiance L arriving along -n at distance d from the
    query on a surface of albedo kd: L * kd / pi * 1 / (pi d^2) (the density
    estimate's kernel 1 / (pi d_max^2), Scene.cu:607-634); two photons at
    distances d and 2d give (L1 + L2 cos) * kd / pi / (pi (2d)^2)."""
    x = np.array([[0.0, 0.0, 0.5], [1.0, 0.0, 0.0]], np.float32)
    rgb = np.array([[2.0, 1.0, 0.5], [4.0, 4.0, 4.0]])
    dirs = np.array([[0.0, 0.0, -1.0], [-0.6, 0.0, -0.8]], np.float32)
    q = np.zeros((1, 3), np.float32)
    n = np.array([[0.0, 0.0, 1.0]], np.float32)
    kd = np.array([[0.5, 0.25, 1.0]], np.float32)
    one = oracle.photon_gather_lambert(x, rgb, dirs, q, n, kd, 1, 100.0)
    want = rgb[0] * kd[0] / np.pi / (np.pi * 0.25)
    assert np.allclose(one[0], want, rtol=1e-6)
    two = oracle.photon_gather_lambert(x, rgb, dirs, q, n, kd, 2, 100.0)
    want2 = (rgb[0] * 1.0 + rgb[1] * 0.8) * kd[0] / np.pi / (np.pi * 1.0)
    assert np.allclose(two[0], want2, rtol=1e-6)
    none = oracle.photon_gather_lambert(x, rgb, dirs, q, n, kd, 2, 0.1)
    assert not none.any()
