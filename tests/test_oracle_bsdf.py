"""Pins the oracle's BSDF, Russian-roulette and sphere-light functions.

Every test calls the oracle's own function (oracle/shade.c through
oracle/__init__.py) and checks it against something that does not share its
arithmetic:
  * closed forms derived by hand from the geometry (disk warps, the half
    vector at normal incidence, Snell's law, Fresnel at normal incidence and
    for k = 0, the irradiance of a sphere light);
  * quadrature of the oracle's own densities (GGX D, the visible-normal
    normalisation with its Lambda, the sampled pdf);
  * sampling statistics of the oracle's own samplers (chi-square of sampled
    directions against the oracle's pdf, E[weight] against the quadrature of
    f, survival rates, branch frequencies).
A dropped term, a wrong sign or a transposed operand in any of these
functions fails at least one test.

Reference lines: Sampler.cu:13-44 (warps), TrowbridgeBSDF.cu:56-246 (GGX),
geometricMath.cu:202-271 (Fresnel), DielectricMaterial.cu:59-77,
SpecularRefractionBSDF.cu:31-63, SpecularReflectionBSDF.cu:30-42,
LambertianBSDF.cu:22-74, pathTracing.cu:91-102 (Russian roulette),
Spectrum.cu:68-75 (brightness).
"""
import math

import numpy as np
import pytest
from scipy import stats

LAMBERT, DIELECTRIC, TROWBRIDGE, TROWBRIDGE_BASIC, LIGHT, MIRROR, PLAIN = range(7)


def dirs(theta, phi):
    st = np.sin(theta)
    return np.stack([st * np.cos(phi), st * np.sin(phi), np.cos(theta)], -1)


def upper_grid(n_mu, n_phi):
    """Midpoint grid over the upper hemisphere in (mu = cos theta, phi):
    directions and the solid angle of each cell (d mu d phi)."""
    mu = (np.arange(n_mu) + 0.5) / n_mu
    ph = (np.arange(n_phi) + 0.5) / n_phi * 2 * np.pi
    M, P = np.meshgrid(mu, ph, indexing="ij")
    w = dirs(np.arccos(M.ravel()), P.ravel())
    return w, np.full(w.shape[0], (1.0 / n_mu) * (2 * np.pi / n_phi))


# ---------------------------------------------------------------------------
# sampling warps
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("u,want", [
    ((0.5, 0.5), (0.0, 0.0)),
    ((1.0, 0.5), (1.0, 0.0)),
    ((0.5, 1.0), (0.0, 1.0)),
    ((0.0, 0.5), (-1.0, 0.0)),
    ((0.75, 0.5), (0.5, 0.0)),
    ((1.0, 1.0), (math.sqrt(0.5), math.sqrt(0.5))),
    ((0.0, 0.0), (-math.sqrt(0.5), -math.sqrt(0.5))),
    ((0.75, 1.0), (math.cos(math.pi / 2 - math.pi / 8), math.sin(math.pi / 2 - math.pi / 8))),
])
def test_concentric_disk_points(orc, u, want):
    """Shirley-Chiu concentric map (Sampler.cu:13-36): squares of side 2r go
    to circles of radius r; the diagonal to 45 degrees; (0.75, 1) lies on the
    top edge at offset x = 0.5 of r = 1, angle pi/2 - pi/4 * 0.5."""
    got = orc.concentric_disk(*u)
    assert np.allclose(got, want, atol=2e-7), (u, got, want)


def test_lambertian_cosine_warp_statistics(orc):
    """Cosine-hemisphere sampling: cos^2(theta) = 1 - r^2 is uniform on [0, 1]
    and phi uniform, so equal-probability bins in (cos^2 theta, phi) must hold
    equal counts (chi-square); E[cos theta] = 2/3; weight = kd exactly; the
    pdf is cos / pi and f * |cos| = kd / pi * cos."""
    m = orc.material(LAMBERT, kd=(0.25, 0.5, 0.75))
    rng = np.random.default_rng(11)
    n = 400_000
    u = rng.random((n, 2), dtype=np.float32)
    wo = np.array([0.3, -0.4, np.sqrt(0.75)], np.float32)
    wi, pdf, delta, wgt = orc.bsdf_sample(m, wo, u)
    assert orc.bsdf_draws(m) == 2 and not delta.any()
    assert np.all(wi[:, 2] >= 0)
    assert np.allclose(np.linalg.norm(wi, axis=1), 1, atol=3e-6)
    assert np.array_equal(wgt, np.broadcast_to(np.float32([0.25, 0.5, 0.75]), wgt.shape))
    assert np.allclose(pdf, wi[:, 2] / np.pi, rtol=2e-6, atol=1e-9)
    c = wi[:, 2].astype(np.float64)
    assert abs(c.mean() - 2 / 3) < 4 * math.sqrt(1 / 18 / n)
    # chi-square over 10 x 16 equal-probability cells
    a = np.clip((c * c * 10).astype(int), 0, 9)
    b = np.clip(((np.arctan2(wi[:, 1], wi[:, 0]) + np.pi) / (2 * np.pi) * 16).astype(int), 0, 15)
    counts = np.bincount(a * 16 + b, minlength=160)
    chi2 = ((counts - n / 160) ** 2 / (n / 160)).sum()
    assert stats.chi2.sf(chi2, 159) > 1e-4, chi2
    # f * |cos wi| and pdf on fixed directions, both hemispheres
    w = np.float32([[0, 0, 1], [0.6, 0, 0.8], [0, 0.8, -0.6]])
    f, p = orc.bsdf_eval(m, wo, w)
    cos = np.abs(w[:, 2])
    assert np.allclose(f, np.outer(cos, [0.25, 0.5, 0.75]) / np.pi, rtol=1e-6)
    assert np.allclose(p, [1 / np.pi, 0.8 / np.pi, 0.0], rtol=1e-6)


# ---------------------------------------------------------------------------
# GGX (Trowbridge-Reitz)
# ---------------------------------------------------------------------------

def ggx_hemisphere_nodes(alpha, n_x=4000, n_phi=24):
    """Quadrature nodes over the half-vector hemisphere concentrated where
    GGX has its mass: theta = atan(alpha tan x), x Gauss-Legendre on
    [0, pi/2); returns directions and solid-angle weights."""
    x, wx = np.polynomial.legendre.leggauss(n_x)
    x = (x + 1) * np.pi / 4
    wx = wx * np.pi / 4
    th = np.arctan(alpha * np.tan(x))
    dth = alpha / np.cos(x) ** 2 / (1 + (alpha * np.tan(x)) ** 2)
    ph = (np.arange(n_phi) + 0.5) / n_phi * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    W = (wx * dth * np.sin(th))[:, None] * np.full(n_phi, 2 * np.pi / n_phi)[None, :]
    return dirs(T.ravel(), P.ravel()), W.ravel()


@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.5, 0.9])
def test_ggx_D_projects_to_one(orc, alpha):
    """integral of D(h) cos(theta_h) over the hemisphere = 1, with the
    oracle's D (TrowbridgeBSDF.cu:71-84); D(n) = 1 / (pi alpha^2)."""
    h, w = ggx_hemisphere_nodes(alpha)
    D = orc.ggx_D(alpha, h).astype(np.float64)
    assert abs((D * h[:, 2] * w).sum() - 1) < 1e-3
    assert abs(orc.ggx_D(alpha, [0, 0, 1])[0] * np.pi * alpha * alpha - 1) < 1e-6


@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.5])
@pytest.mark.parametrize("theta_o", [0.0, 0.5, 1.2])
def test_ggx_visible_normals_normalise(orc, alpha, theta_o):
    """The Smith masking function with the oracle's Lambda
    (TrowbridgeBSDF.cu:86-113) satisfies the visible-normal identity
    integral G1(wo) max(0, wo.h) D(h) dh = cos(theta_o) (Heitz 2014, eq. 2)
    for every wo; Lambda(n) = 0 and Lambda follows the closed form
    (sqrt(1 + alpha^2 tan^2) - 1) / 2; G1 <= 1."""
    wo = dirs(np.float64(theta_o), np.float64(0.7))
    h, w = ggx_hemisphere_nodes(alpha)
    D = orc.ggx_D(alpha, h).astype(np.float64)
    lam = float(orc.ggx_lambda(alpha, wo)[0])
    want = (math.sqrt(1 + alpha * alpha * math.tan(theta_o) ** 2) - 1) / 2
    assert abs(lam - want) <= 1e-6 * max(1.0, want)
    g1 = 1 / (1 + lam)
    assert g1 <= 1
    integ = (g1 * np.maximum(0, h @ wo) * D * w).sum()
    assert abs(integ - math.cos(theta_o)) < 2e-3


@pytest.mark.parametrize("u0,u1", [(0.1, 0.1), (0.3, 0.37), (0.75, 0.8), (0.95, 0.5), (0.5, 0.25)])
@pytest.mark.parametrize("alpha", [0.1, 0.4])
def test_ggx_half_vector_at_normal_incidence(orc, alpha, u0, u1):
    """At wo = n the visible-normal sampler reduces to the stretched
    hemisphere: with the reference's uniformSampleDisk(2u - 1)
    (TrowbridgeBSDF.cu:138-139, Sampler.cu:39-44) the disk point has
    r^2 = |2 u0 - 1| and angle 2 pi (2 u1 - 1), so
    h ~ (alpha r cos, alpha r sin, sqrt(1 - r^2)).  A plain
    uniformSampleDisk(u) would put r^2 = u0 and the angle at 2 pi u1."""
    h = orc.ggx_sample_half(alpha, [0, 0, 1], u0, u1).astype(np.float64)
    r = math.sqrt(abs(2 * u0 - 1))
    ph = 2 * math.pi * (2 * u1 - 1)
    want = np.array([alpha * r * math.cos(ph), alpha * r * math.sin(ph), max(1e-6, math.sqrt(1 - r * r))])
    want /= np.linalg.norm(want)
    assert np.allclose(h, want, atol=3e-6), (h, want)


def _sample_bins(pilot_mu, pilot_ph, nb_mu, nb_ph):
    """Equal-count bins from a pilot sample: mu quantile bands, then phi
    quantiles inside each band."""
    mu_edges = np.quantile(pilot_mu, np.linspace(0, 1, nb_mu + 1))
    mu_edges[0], mu_edges[-1] = 0.0, 1.0
    ph_edges = []
    for k in range(nb_mu):
        sel = (pilot_mu >= mu_edges[k]) & (pilot_mu < mu_edges[k + 1])
        e = np.quantile(pilot_ph[sel], np.linspace(0, 1, nb_ph + 1))
        e[0], e[-1] = -np.pi, np.pi
        ph_edges.append(e)
    return mu_edges, ph_edges


def _bin_integrals(orc, m, wo, mu_edges, ph_edges, sub=24):
    """Integral of the oracle's pdf and f * |cos| over every bin (midpoint
    rule on a sub x sub grid in (mu, phi) per bin)."""
    nb_mu, nb_ph = len(mu_edges) - 1, len(ph_edges[0]) - 1
    pdf_int = np.zeros((nb_mu, nb_ph))
    f_int = np.zeros((nb_mu, nb_ph, 3))
    s = (np.arange(sub) + 0.5) / sub
    for k in range(nb_mu):
        m0, m1 = mu_edges[k], mu_edges[k + 1]
        mus = m0 + (m1 - m0) * s
        for j in range(nb_ph):
            p0, p1 = ph_edges[k][j], ph_edges[k][j + 1]
            phs = p0 + (p1 - p0) * s
            M, P = np.meshgrid(mus, phs, indexing="ij")
            w = dirs(np.arccos(M.ravel()), P.ravel())
            f, pdf = orc.bsdf_eval(m, wo, w)
            da = (m1 - m0) * (p1 - p0) / (sub * sub)
            pdf_int[k, j] = pdf.astype(np.float64).sum() * da
            f_int[k, j] = f.astype(np.float64).sum(0) * da
    return pdf_int, f_int


@pytest.mark.parametrize("alpha,theta_o,kind", [
    (0.2, 0.5, TROWBRIDGE_BASIC), (0.5, 1.1, TROWBRIDGE_BASIC), (0.3, 0.8, TROWBRIDGE)])
def test_ggx_samples_follow_pdf_and_weights_match_f(orc, alpha, theta_o, kind):
    """Directions drawn by the oracle's GGX sampleDirection follow the
    oracle's pdf (chi-square over equal-count bins plus one bin for
    directions below the horizon, whose mass is 1 - the pdf's integral), and
    the sampled weights average to the integral of f * |cos| (the
    directional albedo) by quadrature."""
    m = orc.material(kind, kd=(0.9, 0.7, 0.5), roughness=alpha * alpha, eta=(0.2, 0.9, 1.1), k=(3.9, 2.4, 2.2))
    wo = dirs(np.float64(theta_o), np.float64(0.3)).astype(np.float32)
    rng = np.random.default_rng(5)
    pilot = orc.bsdf_sample(m, wo, rng.random((200_000, 2), dtype=np.float32))[0]
    up = pilot[:, 2] > 0
    mu_edges, ph_edges = _sample_bins(pilot[up, 2], np.arctan2(pilot[up, 1], pilot[up, 0]), 16, 16)
    n = 1_000_000
    wi, pdf, delta, wgt = orc.bsdf_sample(m, wo, rng.random((n, 2), dtype=np.float32))
    assert orc.bsdf_draws(m) == 2 and not delta.any()
    pdf_int, f_int = _bin_integrals(orc, m, wo, mu_edges, ph_edges)
    total = pdf_int.sum()
    assert total <= 1 + 2e-3
    # counts per bin
    upper = wi[:, 2] > 0
    mu = wi[upper, 2]
    ph = np.arctan2(wi[upper, 1], wi[upper, 0])
    kb = np.clip(np.searchsorted(mu_edges, mu, side="right") - 1, 0, 15)
    counts = np.zeros((16, 16))
    for k in range(16):
        sel = kb == k
        jb = np.clip(np.searchsorted(ph_edges[k], ph[sel], side="right") - 1, 0, 15)
        counts[k] = np.bincount(jb, minlength=16)
    obs = np.append(counts.ravel(), n - upper.sum())
    exp = np.append(pdf_int.ravel(), max(0.0, 1 - total)) * n
    keep = exp > 20
    chi2 = ((obs[keep] - exp[keep]) ** 2 / exp[keep]).sum()
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-4, (chi2, keep.sum())
    # the pdf the sampler returns is the pdf of the direction it drew
    _, pdf_back = orc.bsdf_eval(m, wo, wi[upper][:2000])
    assert np.allclose(pdf[upper][:2000], pdf_back, rtol=1e-4)
    # E[weight] = integral of f |cos| (weights below the horizon are 0)
    albedo = f_int.sum((0, 1))
    mean = wgt.astype(np.float64).mean(0)
    se = wgt.astype(np.float64).std(0) / math.sqrt(n)
    assert np.all(np.abs(mean - albedo) < 5 * se + 2e-3 * albedo), (mean, albedo, se)


def test_ggx_smooth_limit_is_a_mirror(orc):
    """alpha < 1e-3 (roughness < 1e-6) samples the mirror direction with no
    draws; Trowbridge weight = F_conductor / cos, TrowbridgeBasic 1."""
    wo = np.float32([0.28, -0.3, math.sqrt(1 - 0.28 ** 2 - 0.09)])
    for kind in (TROWBRIDGE, TROWBRIDGE_BASIC):
        m = orc.material(kind, roughness=1e-8, eta=(0.2, 0.2, 0.2), k=(3.0, 3.0, 3.0))
        assert orc.bsdf_draws(m) == 0
        wi, pdf, delta, wgt = orc.bsdf_sample(m, wo, [[0.3, 0.6]])
        assert delta[0] == 1
        assert np.allclose(wi[0], [-wo[0], -wo[1], wo[2]], atol=1e-7)
        if kind == TROWBRIDGE:
            F = orc.fresnel_conductor(float(wo[2]), 0.2, 3.0)
            assert np.allclose(wgt[0], F / wo[2], rtol=1e-6)
        else:
            assert np.array_equal(wgt[0], [1, 1, 1])


# ---------------------------------------------------------------------------
# Fresnel
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("eta,k", [(0.2, 3.9), (1.1, 2.2), (0.05, 4.2), (2.5, 0.0)])
def test_fresnel_conductor_normal_incidence(orc, eta, k):
    """F(1) = ((n - 1)^2 + k^2) / ((n + 1)^2 + k^2) (geometricMath.cu:242-271)."""
    want = ((eta - 1) ** 2 + k * k) / ((eta + 1) ** 2 + k * k)
    assert abs(orc.fresnel_conductor(1.0, eta, k) - want) < 2e-6


@pytest.mark.parametrize("n", [1.33, 1.5, 2.4])
def test_fresnel_conductor_without_absorption_is_dielectric_fresnel(orc, n):
    """With k = 0 the conductor formula is the textbook unpolarised
    dielectric Fresnel from air into index n; at grazing incidence it is 1."""
    for ci in (0.05, 0.2, 0.5, 0.8, 0.99):
        si = math.sqrt(1 - ci * ci)
        ct = math.sqrt(1 - (si / n) ** 2)
        rs = ((ci - n * ct) / (ci + n * ct)) ** 2
        rp = ((n * ci - ct) / (n * ci + ct)) ** 2
        assert abs(orc.fresnel_conductor(ci, n, 0.0) - (rs + rp) / 2) < 3e-6
    assert abs(orc.fresnel_conductor(0.0, n, 0.0) - 1) < 1e-6


# ---------------------------------------------------------------------------
# dielectric (0.7 / 0.3 split, refraction, reflection)
# ---------------------------------------------------------------------------

def _lit_fresnel(ci, n_o, n_d):
    """The reference's fresnelDielectric as written (geometricMath.cu:
    202-240): r = (n_d cos_i - n_d cos_t) / (n_d cos_i + n_d cos_t),
    F = r^2 |cos_i|, 1 when sin_i >= 1."""
    ci = abs(ci)
    si = math.sqrt(max(0, 1 - ci * ci))
    sd = n_o / n_d * si
    cd = math.sqrt(max(0, 1 - sd * sd))
    r = (ci - cd) / (ci + cd)
    return r * r * ci


@pytest.mark.parametrize("theta", [0.1, 0.6, 1.2])
def test_dielectric_refraction_follows_snell(orc, theta):
    """u < 0.7 refracts (DielectricMaterial.cu:59-77): entering from outside
    (cos wo > 0) the direction leaves through the other side in the plane of
    incidence with sin(theta_t) = sin(theta_o) / ior, and the weight is
    kd (1 - F); u >= 0.7 mirrors wo about n with weight kd F(cos wi, 1, ior)."""
    m = orc.material(DIELECTRIC, kd=(1.0, 0.8, 0.6), ior=1.5)
    wo = dirs(np.float64(theta), np.float64(0.4)).astype(np.float32)
    wi, pdf, delta, wgt = orc.bsdf_sample(m, wo, [[0.3, 0.0], [0.6999999, 0.0], [0.7, 0.0], [0.95, 0.0]])
    assert orc.bsdf_draws(m) == 1 and delta.all()
    for i in (0, 1):
        t = wi[i].astype(np.float64)
        assert t[2] < 0 and abs(np.linalg.norm(t) - 1) < 1e-6
        st = math.hypot(t[0], t[1])
        assert abs(st - math.sin(theta) / 1.5) < 2e-6
        # opposite azimuth: the transmitted tangent is -wo's
        assert np.dot(t[:2], wo[:2]) <= 0 and abs(t[0] * wo[1] - t[1] * wo[0]) < 1e-6
        F = _lit_fresnel(math.cos(theta), 1.0, 1.5)
        assert np.allclose(wgt[i], np.float32([1.0, 0.8, 0.6]) * (1 - F), rtol=2e-5)
    for i in (2, 3):
        assert np.allclose(wi[i], [-wo[0], -wo[1], wo[2]], atol=0)
        F = _lit_fresnel(float(wo[2]), 1.0, 1.5)
        assert np.allclose(wgt[i], np.float32([1.0, 0.8, 0.6]) * F, rtol=2e-5, atol=1e-8)


def test_dielectric_total_internal_reflection_threshold(orc):
    """From inside (cos wo < 0) the indices swap, eta = 1.5: refraction
    exists iff sin(theta_o) < 1 / 1.5, with sin(theta_t) = 1.5 sin(theta_o);
    beyond the critical angle the refraction branch returns weight 0."""
    m = orc.material(DIELECTRIC, kd=(1, 1, 1), ior=1.5)
    crit = math.asin(1 / 1.5)
    for theta, tir in ((crit - 0.02, False), (crit - 1e-4, False), (crit + 1e-4, True), (crit + 0.3, True)):
        wo = dirs(np.float64(math.pi - theta), np.float64(1.0)).astype(np.float32)
        wi, pdf, delta, wgt = orc.bsdf_sample(m, wo, [[0.1, 0.0]])
        if tir:
            assert np.array_equal(wgt[0], [0, 0, 0]) and np.array_equal(wi[0], [0, 0, 0])
        else:
            t = wi[0].astype(np.float64)
            assert t[2] > 0 and wgt[0][0] > 0
            assert abs(math.hypot(t[0], t[1]) - min(1.0, 1.5 * math.sin(theta))) < 5e-4


def test_dielectric_branch_frequency_on_a_pcg_stream(orc):
    """Over a path's PCG32 stream the refraction branch is taken with
    probability 0.7 (the split is not compensated)."""
    n = 1_000_000
    k = orc.lib().orc_dielectric_refractions(1, n)
    assert abs(k / n - 0.7) < 4 * math.sqrt(0.21 / n)


def test_mirror_and_light_sampling(orc):
    """Mirror: wi = (-x, -y, z), weight = colour (PerfectReflectionBSDF);
    light and plain: weight 0 (NullBSDF, PlainMaterial::sampleDirection)."""
    wo = np.float32([0.1, 0.2, math.sqrt(0.95)])
    m = orc.material(MIRROR, kd=(0.5, 0.6, 0.7))
    wi, pdf, delta, wgt = orc.bsdf_sample(m, wo, [[0.5, 0.5]])
    assert delta[0] == 1 and np.allclose(wi[0], [-0.1, -0.2, wo[2]]) and np.allclose(wgt[0], [0.5, 0.6, 0.7])
    for kind in (LIGHT, PLAIN):
        wi, pdf, delta, wgt = orc.bsdf_sample(orc.material(kind), wo, [[0.5, 0.5]])
        assert not wgt.any()


# ---------------------------------------------------------------------------
# Russian roulette and brightness
# ---------------------------------------------------------------------------

def test_brightness_weights(orc):
    """Spectrum::brightness: Rec. 709 luminance 0.212671 r + 0.715160 g + 0.072169 b."""
    for rgb, want in (((1, 0, 0), 0.212671), ((0, 1, 0), 0.715160), ((0, 0, 1), 0.072169), ((1, 1, 1), 1.0),
                      ((0.5, 0.25, 2.0), 0.5 * 0.212671 + 0.25 * 0.715160 + 2 * 0.072169)):
        assert abs(orc.luminance(rgb) - want) < 1e-6


@pytest.mark.parametrize("beta", [(0.3, 0.3, 0.3), (0.9, 0.1, 0.05), (2.0, 2.0, 2.0), (0.01, 0.0, 0.0)])
def test_russian_roulette_is_unbiased(orc, beta):
    """q = max(0.05, 1 - brightness(beta)); the path ends when u < q, else
    beta / (1 - q): over stratified u the survival rate is 1 - q and
    E[beta'] = beta."""
    q = max(0.05, 1 - (0.212671 * beta[0] + 0.715160 * beta[1] + 0.072169 * beta[2]))
    n = 20000
    alive = 0
    acc = np.zeros(3)
    for i in range(n):
        ok, b = orc.russian_roulette(beta, (i + 0.5) / n)
        if ok:
            alive += 1
            acc += b
            # 1 - q is formed in float: relative rounding of eps / (1 - q)
            assert np.allclose(b, np.float32(beta) / np.float32(1 - q), rtol=1e-4)
    assert abs(alive / n - (1 - q)) <= 1.0 / n + 1e-6
    # stratified u: the survivor count is off by at most one, so E[beta'] by
    # at most beta / ((1 - q) n) beyond rounding
    assert np.all(np.abs(acc / n - beta) <= np.float64(beta) / ((1 - q) * n) + 2e-4 * np.float64(beta) + 1e-9)


# ---------------------------------------------------------------------------
# sphere area light (the library's own light type; no reference counterpart)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("r,d", [(0.2, 1.0), (0.5, 0.7), (0.05, 3.0)])
def test_sphere_light_cone_sampling(orc, r, d):
    """Uniform cone sampling of the visible cap: every direction hits the
    sphere at the returned distance, the pdf is 1 / (2 pi (1 - cos_max)) and
    equals sphere_light_pdf; E[cos to centre] = (1 + cos_max) / 2; and the
    estimator E[max(0, cos theta) / pdf] of a unit-radiance sphere straight
    above the point is its irradiance pi (r / d)^2."""
    c = np.float32([0.1, -0.2, 0.3])
    p = c - np.float32([0, 0, d])
    cmax = math.sqrt(1 - (r / d) ** 2)
    n = 128
    cos_c, est = [], []
    for i in range(n):
        for j in range(n):
            ok, wi, dist, pdf = orc.sphere_light_sample(c, r, p, (i + 0.5) / n, (j + 0.5) / n)
            assert ok
            hitp = p.astype(np.float64) + dist * wi
            assert abs(np.linalg.norm(hitp - c) - r) < 2e-5 * max(1.0, d)
            assert abs(pdf * 2 * math.pi * (1 - cmax) - 1) < 2e-3
            cos_c.append(wi[2])
            est.append(max(0.0, wi[2]) / pdf)
    assert abs(orc.sphere_light_pdf(c, r, p) - pdf) <= 1e-6 * pdf
    assert abs(np.mean(cos_c) - (1 + cmax) / 2) < 1e-4
    assert abs(np.mean(est) / (math.pi * (r / d) ** 2) - 1) < 3e-3
    ok, *_ = orc.sphere_light_sample(c, r, c, 0.5, 0.5)
    assert not ok and orc.sphere_light_pdf(c, r, c) == 0
