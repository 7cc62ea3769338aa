"""The oracle's photon-mapping render mode (orc_render_photon: RENDER_MODE
PHOTON_MAPPING, integration/Scene.cu:912-967, pathTracing.cu:84-88), pinned
by what the mathematics fixes:

* A Lambertian floor lit by one point light, photons stored at their first
  hit only: the estimate sum(photon flux * lightReflected) / (pi r_k^2)
  converges to kd/pi * I cos^2(theta) / d^2 at the point seen.  The photon
  flux color * I * 4 pi / N already carries the incidence cosine through the
  photons' density, and lightReflected = f * |cos wi| multiplies by it once
  more (Scene.cu:607-634 with LambertianBSDF::lightReflected), hence cos^2.
  The photon search is filtered by shape id, which is per triangle
  (Scene.cu:736-741), so the view stays far from the floor quad's diagonal,
  where half of each search disk lies on the other triangle.  A missing pi,
  a wrong normal or view, a search on the wrong shape (the scene carries a
  sphere, so triangle shape ids are offset by it), or next-event estimation
  added on top would each move the image by a constant factor.
* No next-event estimation: with no photons every pixel is 0.
* A path ends at its first scattering hit: photons placed only where
  second-bounce vertices could land change nothing.
* Specular bounces are followed: behind a glass sphere the floor is seen
  through it."""
import math

import numpy as np

from test_oracle_photon_pass import box_desc


CX, CY = 10.0, -10.0  # 14 units from the floor quad's diagonal y = x


def floor_scene(width=24, height=24, kd=0.5, I=2.0, h=1.0, sphere=True):
    d = box_desc(kd=kd, light=((CX, CY, -1 + h), (1, 1, 1), I), size=50.0, floor_only=True)
    d.camera = dict(d.camera, width=width, height=height, pos=(CX, CY, 0.5), lookat=(CX, CY, -1), frame_y=1.0)
    d.spp = 4
    d.max_depth = 8
    if sphere:  # far out of view and out of the light's reach of the floor seen
        d.materials = d.materials + [dict(type=0, kd=(0.5, 0.5, 0.5))]
        d.spheres = [((40.0, 40.0, 5.0), 1.0, 1, False)]
    return d


def analytic_floor(orc_scene, d, kd, I, h):
    """kd/pi * I cos^2 / d^2 at each pixel centre's floor point."""
    W, H = d.width, d.height
    out = np.zeros((H, W))
    lz = -1 + h
    for y in range(H):
        for x in range(W):
            o, dv = orc_scene.camera_ray(x, y, 0.5, 0.5)
            t = (-1 - o[2]) / dv[2]
            p = o + t * dv
            r2 = (p[0] - CX) ** 2 + (p[1] - CY) ** 2 + (lz + 1) ** 2
            cos = (lz + 1) / math.sqrt(r2)
            out[y, x] = kd / math.pi * I * cos * cos / r2
    return out


def test_floor_estimate_converges_to_direct_radiance(orc):
    kd, I, h = 0.5, 2.0, 1.0
    d = floor_scene(kd=kd, I=I, h=h)
    sc = orc.OrcScene(d)
    ph, m = sc.trace_photons(300_000, seed=5, max_depth=1)
    assert m > 100_000
    acc, nv = sc.render_photon(ph, k=100, radius=100.0, spp=4)
    img = (acc.reshape(-1, 3) / np.maximum(nv, 1)[:, None]).reshape(d.height, d.width, 3)
    assert np.all(nv == 4)
    want = analytic_floor(sc, d, kd, I, h)
    ratio = img[..., 0] / want
    # density estimation noise ~ 1/sqrt(k) per estimate, four per pixel
    assert abs(ratio.mean() - 1) < 0.02, ratio.mean()
    assert np.mean(np.abs(ratio - 1) < 0.25) > 0.95
    assert np.allclose(img[..., 0], img[..., 1]) and np.allclose(img[..., 0], img[..., 2])


def test_no_photons_no_light(orc):
    d = floor_scene()
    sc = orc.OrcScene(d)
    ph = dict(location=np.zeros((1, 3), np.float32) + 1e3, shape=np.array([-5], np.int32),
              radiance=np.ones((1, 3)), direction=np.array([[0, 0, -1]], np.float32))
    acc, nv = sc.render_photon(ph, k=10, radius=1.0, spp=2)
    assert np.all(acc == 0) and np.all(nv == 2)


def test_paths_end_at_the_first_scattering_hit(orc):
    """Photons far above the floor (where only bounced paths could reach
    them) on the floor's own shape ids change nothing."""
    kd, I, h = 0.5, 2.0, 1.0
    d = floor_scene(kd=kd, I=I, h=h)
    sc = orc.OrcScene(d)
    ph, m = sc.trace_photons(50_000, seed=7, max_depth=1)
    a, _ = sc.render_photon(ph, k=50, radius=0.5, spp=2)
    extra = dict(location=np.concatenate([ph["location"], ph["location"] + np.float32([0, 0, 3])]),
                 shape=np.concatenate([ph["shape"], ph["shape"]]),
                 radiance=np.concatenate([ph["radiance"], ph["radiance"] * 100]),
                 direction=np.concatenate([ph["direction"], ph["direction"]]))
    b, _ = sc.render_photon(extra, k=50, radius=0.5, spp=2)
    assert np.array_equal(a, b)


def test_glass_is_seen_through(orc):
    """A glass sphere between the camera and the floor: its pixels see the
    floor through it (specular bounces continue), so they carry light, and
    a diffuse sphere there would end the paths on itself instead."""
    kd, I, h = 0.5, 2.0, 1.0
    d = floor_scene(kd=kd, I=I, h=h, sphere=False)
    d.materials = d.materials + [dict(type=1, ior=1.5, kd=(1.0, 1.0, 1.0))]  # colour: the dielectric weight
    d.spheres = [((CX, CY, -0.6), 0.2, 1, False)]
    sc = orc.OrcScene(d)
    ph, m = sc.trace_photons(100_000, seed=9, max_depth=8)  # photons pass the glass to the floor below it
    acc, nv = sc.render_photon(ph, k=50, radius=100.0, spp=32)
    img = acc.reshape(d.height, d.width, 3)[..., 0]
    W = d.width
    centre = img[W // 2 - 2:W // 2 + 2, W // 2 - 2:W // 2 + 2]
    # a path refracts through both surfaces with probability ~0.3^2 (the
    # fixed split), so the block as a whole carries light
    assert centre.mean() > 0 and np.mean(centre > 0) > 0.5
    d.materials[-1] = dict(type=0, kd=(0.0, 0.0, 0.0))  # black diffuse sphere: no photons stored on it
    acc2, _ = orc.OrcScene(d).render_photon(ph, k=50, radius=100.0, spp=32)
    img2 = acc2.reshape(d.height, d.width, 3)[..., 0]
    assert np.all(img2[W // 2 - 2:W // 2 + 2, W // 2 - 2:W // 2 + 2] == 0)


def test_hybrid_is_the_mean_of_both_estimates_on_a_floor(orc):
    """RENDER_MODE HYBRID (Scene.cu:963-967): (photon + direct) / 2 at each
    scattering hit, and the path goes on.  On a floor under a point light a
    continued path never hits again, and the hybrid draws no extra samples,
    so every hybrid pixel is half the photon-mapping pixel plus half the
    path-tracing pixel (up to the FP64 film's summation order)."""
    d = floor_scene()
    sc = orc.OrcScene(d)
    ph, m = sc.trace_photons(50_000, seed=13, max_depth=1)
    pm, _ = sc.render_photon(ph, k=50, radius=100.0, spp=4)
    hy, nv = sc.render_photon(ph, k=50, radius=100.0, spp=4, hybrid=True)
    pt, _ = sc.render(spp=4)
    assert np.all(nv == 4)
    assert pm.mean() > 0 and pt.mean() > 0 and not np.allclose(pm, pt, rtol=1e-2)  # distinct estimates
    assert np.allclose(hy, 0.5 * (pm + pt), rtol=1e-6, atol=0)


def test_hybrid_continues_past_scattering_hits(orc):
    """In a closed Lambertian box the hybrid keeps bouncing (unlike the
    photon-mapping mode, which stops at the first hit): with photons
    removed it is exactly half of path tracing's next-event terms plus
    nothing, so it equals half the path-traced image."""
    from test_oracle_photon_pass import box_desc as box
    d = box(kd=0.6, light=((0.1, 0.2, 0.3), (1, 1, 1), 2.0))
    d.camera = dict(d.camera, width=16, height=16, pos=(0, 0, 0.5), lookat=(0, 0, -1), frame_y=1.0)
    sc = orc.OrcScene(d)
    none = dict(location=np.zeros((1, 3), np.float32) + 1e3, shape=np.array([-5], np.int32),
                radiance=np.ones((1, 3)), direction=np.array([[0, 0, -1]], np.float32))
    hy, nv = sc.render_photon(none, k=10, radius=1.0, spp=4, max_depth=6, hybrid=True)
    pt, _ = sc.render(spp=4, max_depth=6)
    assert pt.mean() > 0
    assert np.allclose(hy, 0.5 * pt, rtol=1e-6, atol=0)
