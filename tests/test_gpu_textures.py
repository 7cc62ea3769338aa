"""GPU parity of textures (rt_scene_add_texture, rt_scene_add_mesh_uv,
rt_material.texture; textures/Texture.cu:172-183) against the oracle:

1. the probes of a Plain material (kd 1, no lights) textured with a 2 x 2
   texture on a plane, on a two-triangle quad with corner uvs and without
   (the reference's default corners): every path's value is one texel, and
   the GPU film equals the oracle's film;
2. a textured sphere (surface point (theta, phi) in degrees) and a Cornell
   box with a checker-textured Lambertian floor and a textured Plain back
   wall match the oracle to the render parity bar."""
import numpy as np
import pytest

import material_scenes as ms
from test_gpu_materials import compare, gpu_mean

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["plane", "quad", "default"])
def test_texture_probes_equal_oracle(rt, orc, gpu_ctx, kind):
    d = ms.textured_probe(kind, width=32, height=32, spp=4)
    print(f"[textures] probe {kind}", flush=True)
    gi, nv = gpu_mean(rt, gpu_ctx, d)
    oa, on = orc.OrcScene(d).render(spp=d.spp, nthreads=16)
    oi = oa.reshape(-1, 3) / np.maximum(on, 1)[:, None]
    assert np.all(nv == 4)
    assert np.allclose(gi, oi, rtol=1e-12, atol=0)
    # the plane and the quad with corner uvs span the whole texture; with the
    # default corners the surface point is (u, v) or (v, v - u), which never
    # reaches texel (0, 1)
    colours = {tuple(np.round(c, 6)) for c in gi}
    want = [(0, 0), (1, 0), (0, 1), (1, 1)] if kind != "default" else [(0, 0), (1, 0), (1, 1)]
    for x, y in want:
        assert tuple(np.round(ms.texel(ms.TEX2, x, y).astype(np.float64), 6)) in colours
    if kind == "default":
        assert tuple(np.round(ms.texel(ms.TEX2, 0, 1).astype(np.float64), 6)) not in colours


def test_textured_sphere_matches_oracle(rt, orc, gpu_ctx):
    compare(rt, orc, gpu_ctx, ms.textured_probe("sphere", width=32, height=32, spp=8), "textured sphere")


def test_textured_cornell_matches_oracle(rt, orc, gpu_ctx):
    compare(rt, orc, gpu_ctx, ms.cornell_textured(64, 64, 16), "textured cornell")
