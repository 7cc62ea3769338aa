"""Pins the oracle's photon pre-pass (generatePhotons + computePhotons,
integration/photonDispersion.cu:26-192, Scene.cu:427-455, 989-1021), its
general photon-mapping density estimate (combinePhotonContributions,
Scene.cu:607-634) and the subsurface term (combinePhotonsSubsurface,
Scene.cu:667-702, with the dipole of geometricMath.cu:451-522), against what
geometry and probability fix:

* a point light over a large Lambertian floor: every first-bounce photon
  carries color * I * 4 pi / N exactly (one light: p = 1, n = N), the
  fraction stored is the floor's solid angle over 4 pi, and the radii follow
  P(r < R) = 1 - h / sqrt(h^2 + R^2);
  the stored normal faces the light and the direction runs light -> point;
* a closed Lambertian cube: each bounce multiplies the carried light by kd
  exactly; with Russian roulette from depth 1 the survivors' throughput
  returns to 1 and the survival fractions are 1 - q;
* two lights of different power: the selection fractions follow the powers
  and every photon's light is recovered from its radiance with
  n_A + n_B = N;
* the dipole's bracket integrates in closed form over the plane
  (integral of z (s d + 1) e^(-s d) / d^3 * 2 pi r dr = 2 pi e^(-s z));
* the subsurface sum visits exactly the photons a brute-force radius search
  on the query's shape finds, and the general gather on a Lambertian
  surface equals the Lambertian-only estimate."""
import importlib
import math

import numpy as np
import pytest

scenes = importlib.import_module("2023_render_engine_amd.scenes")

LAMB, DIEL, TROW, TROWB, LIGHT, MIRROR, PLAIN = range(7)


def box_desc(kd=0.5, light=((0, 0, 0), (1, 1, 1), 2.0), lights=None, size=1.0, floor_only=False):
    a = size
    faces = [((-a, -a, -a), (a, -a, -a), (a, a, -a), (-a, a, -a))]
    if floor_only:  # a floor at z = -1 reaching to +-size
        faces = [((-a, -a, -1), (a, -a, -1), (a, a, -1), (-a, a, -1))]
    else:
        faces += [((-a, -a, a), (-a, a, a), (a, a, a), (a, -a, a)), ((a, -a, -a), (a, -a, a), (a, a, a), (a, a, -a)),
                  ((-a, -a, -a), (-a, a, -a), (-a, a, a), (-a, -a, a)),
                  ((-a, -a, -a), (-a, -a, a), (a, -a, a), (a, -a, -a)), ((-a, a, -a), (a, a, -a), (a, a, a), (-a, a, a))]
    xyz, idx = [], []
    for f in faces:
        v, t = scenes.quad(*f)
        idx.append(t + 4 * len(xyz))
        xyz.append(v)
    xyz = np.concatenate(xyz).astype(np.float32)
    idx = np.concatenate(idx).astype(np.uint32)
    mats = [dict(type=LAMB, kd=(kd, kd, kd))]
    cam = dict(pos=(0, 0, 0.5), lookat=(0, 0, -1), up=(0, 1, 0), cts=(0, 0, -1.0), frame_y=1.0, width=4, height=4)
    return scenes.SceneDesc(xyz, idx, np.zeros(idx.shape[0], np.uint32), mats, [], lights or [light], cam)


def test_first_bounce_on_a_floor(orc):
    h, N, I = 1.0, 40000, 2.0
    d = box_desc(light=((0, 0, -1 + h), (0.5, 1.0, 0.25), I), size=50.0, floor_only=True)
    ph, m = orc.OrcScene(d).trace_photons(N, seed=3, max_depth=1)
    assert m == ph["key"].size
    assert np.all((ph["key"] & 31) == 0)
    want = np.array([0.5, 1.0, 0.25]) * I * 4 * math.pi / N
    assert np.array_equal(ph["radiance"], np.tile(want, (m, 1)))
    # the square floor of half-width a subtends 4 asin(a^2 / (a^2 + h^2))
    a = 50.0
    want_frac = math.asin(a * a / (a * a + h * h)) / math.pi
    assert abs(m / N - want_frac) < 5 * math.sqrt(want_frac * (1 - want_frac) / N)
    p = ph["location"].astype(np.float64)
    assert np.allclose(p[:, 2], -1, atol=1e-5)
    r = np.hypot(p[:, 0], p[:, 1])
    for R in (0.5, 1.0, 2.0, 5.0):
        # downward fraction within R, over the stored (square-floor) fraction
        want_f = (1 - h / math.sqrt(h * h + R * R)) / 2 / want_frac
        assert abs(np.mean(r < R) - want_f) < 5 * math.sqrt(want_f * (1 - want_f) / m) + 1e-3, R
    assert np.allclose(ph["normal"], [0, 0, 1], atol=1e-6)
    # the direction runs from the light to the point
    near = r < 5
    dvec = p[near] - np.array([0, 0, -1 + h])
    dvec /= np.linalg.norm(dvec, axis=1, keepdims=True)
    assert np.allclose(ph["direction"][near], dvec, atol=1e-4)
    # shape ids: no analytic shapes, so triangle t -> t
    assert set(np.unique(ph["shape"])) <= {0, 1}


def test_bounces_multiply_by_kd_in_a_closed_box(orc):
    kd, N = 0.5, 4000
    d = box_desc(kd=kd)
    ph, m = orc.OrcScene(d).trace_photons(N, seed=7, max_depth=3, rr_depth=10)
    path, b = ph["key"] >> 5, ph["key"] & 31
    # closed box: every path hits at every bounce and nothing ends early
    assert m == 3 * N
    assert np.array_equal(np.bincount(b), [N, N, N])
    rgb0 = 1.0 * 2.0 * 4 * math.pi / N
    for k in range(3):
        sel = b == k
        assert np.allclose(ph["radiance"][sel], rgb0 * kd ** k, rtol=1e-6), k
    assert np.array_equal(path, np.repeat(np.arange(N), 3))


def test_russian_roulette_survival_and_weights(orc):
    kd, N = 0.5, 20000
    d = box_desc(kd=kd)
    ph, m = orc.OrcScene(d).trace_photons(N, seed=11, max_depth=6, rr_depth=0)
    b = ph["key"] & 31
    cnt = np.bincount(b, minlength=6)
    rgb0 = 2.0 * 4 * math.pi / N
    # depth 0 and 1 always stored; RR after depth 1 with beta = 0.25 (q = 0.75),
    # then beta = 0.5 each bounce (q = 0.5); survivors are rescaled to 1
    assert cnt[0] == N and cnt[1] == N
    for k, p in ((2, 0.25), (3, 0.5), (4, 0.5), (5, 0.5)):
        want = cnt[k - 1] * p
        assert abs(cnt[k] - want) < 5 * math.sqrt(cnt[k - 1] * p * (1 - p)) + 1, (k, cnt)
    assert np.allclose(ph["radiance"][b == 1], rgb0 * kd, rtol=1e-6)
    for k in (2, 3, 4, 5):
        assert np.allclose(ph["radiance"][b == k], rgb0, rtol=1e-5), k


def test_light_selection_by_power(orc):
    N = 30000
    lights = [((0.3, 0, 0), (1, 0, 0), 3.0), ((-0.3, 0, 0), (0, 1, 0), 1.0)]
    d = box_desc(lights=lights)
    o = orc.OrcScene(d)
    p, cdf = o.photon_light_table()
    assert np.allclose(p, [0.75, 0.25]) and cdf[-1] == pytest.approx(1.0)
    ph, m = o.trace_photons(N, seed=5, max_depth=1)
    first = ph["radiance"]
    red = first[:, 0] > 0
    assert np.all(first[red, 1] == 0) and np.all(first[~red, 0] == 0)
    # radiance = c * I * 4 pi / n_l / p_l: recover n_l
    n_a = 3.0 * 4 * math.pi / first[red, 0][0] / 0.75
    n_b = 1.0 * 4 * math.pi / first[~red, 1][0] / 0.25
    assert abs(n_a - round(n_a)) < 1e-6 and abs(n_b - round(n_b)) < 1e-6
    assert round(n_a) + round(n_b) == N
    assert abs(round(n_a) / N - 0.75) < 5 * math.sqrt(0.75 * 0.25 / N)
    # closed box: every photon's first bounce is stored
    assert m == N and red.sum() == round(n_a)


def test_no_lights_no_photons(orc):
    d = box_desc(lights=[])
    d.lights = []
    ph, m = orc.OrcScene(d).trace_photons(100, seed=1)
    assert m == 0


def test_dipole_bracket_integrates_in_closed_form(orc):
    """Rd / (C1 (1 - Fdr) alpha') is z_r (s d_r + 1) e^(-s d_r) / d_r^3 plus
    the same at z_v; each integrates to 2 pi e^(-s z) over the plane."""
    from scipy.integrate import quad
    for ss, sa, g, eta in ((1.0, 0.1, 0.0, 1.3), (2.5, 0.4, 0.3, 1.5), (0.7, 0.05, -0.2, 1.33)):
        ssp = (1 - g) * ss
        stp = ssp + sa
        alpha = ssp / stp
        s = math.sqrt(3 * sa * stp)
        fdr = np.float32(-1.440 / np.float32(eta * eta) + 0.710 / np.float32(eta) + 0.668 + 0.0636 * np.float32(eta))
        A = float((1 + fdr) / (1 - fdr))
        zr = math.sqrt(3 * (1 - alpha)) / s
        zv = zr * A

        def bracket(r):
            dr = math.sqrt(r * r + zr * zr)
            c1 = zr * (s + 1 / dr)
            return orc.dipole_rd(ss, sa, g, eta, r) / (c1 * (1 - float(fdr)) * alpha)

        val, err = quad(lambda r: bracket(r) * 2 * math.pi * r, 0, 60 / s, limit=400)
        want = 2 * math.pi * (math.exp(-s * zr) + math.exp(-s * zv))
        assert val == pytest.approx(want, rel=2e-4), (ss, sa, g, eta)


def test_subsurface_sum_covers_the_radius_search(orc):
    rng = np.random.default_rng(4)
    n = 3000
    xyz = rng.random((n, 3)).astype(np.float32)
    shape = rng.integers(0, 3, n).astype(np.int32)
    rgb = rng.random((n, 3))
    pdir = rng.normal(size=(n, 3)).astype(np.float32)
    pdir /= np.linalg.norm(pdir, axis=1, keepdims=True)
    pnrm = -pdir + 0.3 * rng.normal(size=(n, 3)).astype(np.float32)
    pnrm /= np.linalg.norm(pnrm, axis=1, keepdims=True)
    tree = orc.kd_build(xyz)[:2]
    q = rng.random((40, 3)).astype(np.float32)
    qs = rng.integers(-1, 3, 40).astype(np.int32)
    qn = np.tile(np.float32([0, 0, 1]), (40, 1))
    qv = np.tile(np.float32([0, 0.6, 0.8]), (40, 1))
    params = ((0.8, 0.6, 0.4), (1.0, 1.2, 1.5), (0.05, 0.1, 0.2), 0.1, 1.3)
    out = orc.photon_subsurface(xyz, tree, rgb, pdir, pnrm, q, qn, qv, params, radius=0.2, shape=shape,
                                query_shape=qs)
    # brute force: the same terms over every photon within the radius
    kd, ss, sa, g, eta = params
    ft = lambda c: 1 - orc.lib().orc_fresnel_dielectric(c, 1.0, eta)
    cos_o = float(np.float32(0.8))
    for i in range(40):
        dd = np.sqrt(((xyz.astype(np.float64) - q[i]) ** 2).sum(1)).astype(np.float32)
        sel = (dd < 0.2) & ((qs[i] == -1) | (shape == qs[i]))
        if not sel.any():
            assert np.all(np.isnan(out[i]))
            continue
        acc = np.zeros(3)
        for e in np.nonzero(sel)[0]:
            ci = float(np.dot(pnrm[e], -pdir[e]) / (np.linalg.norm(pnrm[e]) * np.linalg.norm(pdir[e])))
            for c in range(3):
                rd = orc.dipole_rd(ss[c], sa[c], g, eta, float(dd[e]))
                acc[c] += rgb[e, c] * ((1 / math.pi) * ft(ci) * rd * ft(cos_o)) * kd[c]
        want = acc / (2 * math.pi * float(dd[sel].max()) ** 2)
        assert np.allclose(out[i], want, rtol=2e-5), i


def test_general_gather_equals_the_lambertian_estimate(orc):
    rng = np.random.default_rng(8)
    n, m, k = 4000, 60, 16
    xyz = rng.random((n, 3)).astype(np.float32)
    rgb = rng.random((n, 3))
    pdir = rng.normal(size=(n, 3)).astype(np.float32)
    pdir /= np.linalg.norm(pdir, axis=1, keepdims=True)
    q = rng.random((m, 3)).astype(np.float32)
    nr = rng.normal(size=(m, 3)).astype(np.float32)
    nr /= np.linalg.norm(nr, axis=1, keepdims=True)
    kd = rng.random((m, 3)).astype(np.float32)
    views = nr.copy()
    cnt, idx, dist = orc.knn(xyz, q, k, 0.3)
    mats = [dict(type=LAMB, kd=tuple(kd[i])) for i in range(m)]
    got = orc.photon_gather_mat(rgb, pdir, cnt, idx, dist, nr, views, mats)
    want = orc.photon_gather_lambert(xyz, rgb, pdir, q, nr, kd, k, 0.3)
    assert np.allclose(got, want, rtol=1e-5, atol=0)
    # plain and the non-scattering materials reflect no photons
    for t in (PLAIN, DIEL, MIRROR, LIGHT):
        z = orc.photon_gather_mat(rgb, pdir, cnt, idx, dist, nr, views, [dict(type=t, kd=(1, 1, 1))] * m)
        assert np.all(z == 0)
