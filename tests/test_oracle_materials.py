"""The oracle's Mirror and Plain materials and its camera-ray rule, pinned
by scenes whose every path has a closed-form value (no lights, so no
next-event terms):

- camera -> Plain:  L = kd (PlainMaterial::getEmittedRadiance towards a
  camera ray, materials/PlainMaterial.cu:62-75);
- camera -> Mirror -> Plain:  L = mirror colour * kd (the reflected ray is
  still a camera ray: SurfaceInteraction.cu:252-254; PerfectReflectionBSDF
  weight = colour);
- camera -> Lambertian (kd 0.5) -> Plain:  L = 0.5 * ke (a Lambertian hit
  ends the camera ray; Lambertian weight = kd);
- camera -> Trowbridge -> Plain (kd red, ke green):  only red, since a
  Trowbridge hit is specular (TrowbridgeMaterial.cu:55-63)."""
import numpy as np

import material_scenes as ms


def paths(orc, d, n=64):
    s = orc.OrcScene(d)
    out = []
    for k in range(n):
        ok, L = s.path(k % d.width, (k * 7) % d.height, k)
        assert ok
        out.append(L)
    return np.array(out)


def f32(v):
    return np.array(v, np.float32)


def test_camera_sees_plain_kd(orc):
    L = paths(orc, ms.probe("plain"))
    assert np.array_equal(L, np.tile(f32(ms.KD), (len(L), 1)))


def test_mirror_keeps_the_camera_ray(orc):
    L = paths(orc, ms.probe("mirror"))
    want = f32(ms.MIRROR_COLOR) * f32(ms.KD)
    assert np.array_equal(L, np.tile(want, (len(L), 1)))


def test_lambertian_ends_the_camera_ray(orc):
    L = paths(orc, ms.probe("lambert"), n=1024)
    want = np.float32(0.5) * f32(ms.KE)
    assert np.array_equal(L, np.tile(want, (len(L), 1)))


def test_trowbridge_keeps_the_camera_ray(orc):
    L = paths(orc, ms.probe("ggx"), n=256)
    assert np.all(L[:, 1:] == 0)
    assert (L[:, 0] > 0).mean() > 0.5


def test_emissive_mesh_and_material_cornell_render(orc):
    for d in (ms.cornell_materials(16, 16, 4), ms.cornell_emissive_ceiling(16, 16, 4)):
        acc, nv = orc.OrcScene(d).render(spp=4, nthreads=8)
        img = acc.reshape(-1, 3) / np.maximum(nv, 1)[:, None]
        assert np.isfinite(img).all() and img.mean() > 0
