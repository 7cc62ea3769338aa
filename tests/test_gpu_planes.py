"""GPU parity of the Plane shape (rt_scene_add_plane, shapes/Plane.cu)
against the oracle:

1. a Plain plane that fills the view gives kd on every pixel, exactly;
2. the reference driver's Cornell box made of five planes
   (cornellBoxGuion.cu:60-92) with the workload's spheres and point light,
   and the same box with a triangle mesh in it, match the oracle's render to
   the render parity bar (RMSE below 5% of the image noise, 90% of pixels
   within 1e-4 relative): closest hits between planes, spheres and triangles,
   and shadow rays blocked by planes;
3. an axis shorter than 1 (a zero-size plane after the reference's integer
   truncation) is refused."""
import ctypes as C

import numpy as np
import pytest
import torch

import material_scenes as ms
from test_gpu_materials import compare, gpu_mean

pytestmark = pytest.mark.gpu


def test_plain_plane_fills_the_view(rt, gpu_ctx):
    d = ms.probe("plain", spp=4)
    d.xyz, d.idx, d.tri_mat = np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32), np.zeros(0, np.uint32)
    d.planes = [((0.5, -1.0, -1.0), (0, 2, 0), (0, 0, 2), (-1, 0, 0), 0)]
    print("[planes] plain plane probe", flush=True)
    img, nv = gpu_mean(rt, gpu_ctx, d)
    assert np.all(nv == 4)
    assert np.allclose(img, np.array(ms.KD, np.float32)[None, :].astype(np.float64), rtol=1e-6, atol=0)


@pytest.mark.parametrize("mesh", [False, True])
def test_plane_cornell_matches_oracle(rt, orc, gpu_ctx, mesh):
    compare(rt, orc, gpu_ctx, ms.cornell_planes(64, 64, 16, with_mesh=mesh), f"plane cornell (mesh {mesh})")


def test_short_axis_is_refused(rt, gpu_ctx):
    L = rt.lib()
    h = C.c_void_p()
    rt.check(L.rt_scene_create(gpu_ctx.h, C.byref(h)))
    m = rt.Material(0, rt.f3((1, 1, 1)), 1.5, 0.0, rt.f3((1, 1, 1)), rt.f3((0, 0, 0)), 0.0, rt.f3((0, 0, 0)))
    mid = C.c_uint32()
    rt.check(L.rt_scene_add_material(h, C.byref(m), C.byref(mid)))
    v = [np.array(x, np.float32) for x in ((0, 0, 0), (0.9, 0, 0), (0, 1, 0), (0, 0, 1))]
    st = L.rt_scene_add_plane(h, *[v_.ctypes.data for v_ in v], 0)
    assert st != 0 and b"shorter than 1" in L.rt_last_error()
    v[1] = np.array((1.0, 0, 0), np.float32)
    assert L.rt_scene_add_plane(h, *[v_.ctypes.data for v_ in v], 0) == 0
    L.rt_scene_destroy(h)
