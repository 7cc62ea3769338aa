"""The oracle's textures (textures/Texture.cu:172-183 and their callers),
pinned by scenes whose every path value is fixed: a Plain material (kd 1,
no lights) seen from the camera shows exactly the texel at the hit's
surface point.

- plane: surface point = (u, v) of the unit square: the four texels of a
  2 x 2 texture appear in the four quadrants of the view;
- triangles with corner uvs giving the same (u, v): the same image;
- triangles without uvs: the reference's default corners (0,0), (1,0),
  (1,1) (Triangle.cu:134-152) give surface point (b1 + b2, b2);
- sphere: surface point (theta, phi) in degrees (Sphere.cu:269-275), so the
  lookup clamps to the last column almost everywhere."""
import numpy as np

import material_scenes as ms


def paths(orc, d):
    s = orc.OrcScene(d)
    out = {}
    for y in range(d.height):
        for x in range(d.width):
            ok, L, seg = s.path_segments(x, y, 0)
            assert ok
            out[(x, y)] = (L, seg[0])
    return out


def hit_yz(o, dvec, x=0.5):
    t = (x - o[0]) / dvec[0]
    return o[1] + t * dvec[1], o[2] + t * dvec[2]


def test_plane_and_quad_show_the_texel_under_each_hit(orc):
    for kind in ("plane", "quad"):
        res = paths(orc, ms.textured_probe(kind))
        seen = set()
        for (x, y), (L, ray) in res.items():
            yy, zz = hit_yz(ray[:3], ray[3:])
            u, v = (yy + 1) / 2, (zz + 1) / 2
            if min(abs(u - 0.5), abs(v - 0.5)) < 1e-3:
                continue  # on a texel boundary
            want = ms.texel(ms.TEX2, int(u * 2), int(v * 2))
            assert np.array_equal(L, want), (kind, x, y, u, v)
            seen.add((int(u * 2), int(v * 2)))
        assert seen == {(0, 0), (1, 0), (0, 1), (1, 1)}


def test_default_triangle_uvs(orc):
    res = paths(orc, ms.textured_probe("default"))
    for (x, y), (L, ray) in res.items():
        yy, zz = hit_yz(ray[:3], ray[3:])
        u, v = (yy + 1) / 2, (zz + 1) / 2
        # surface point (b1 + b2, b2).  Triangle (0,1,2) covers u >= v with
        # corners (0,0), (1,0), (1,1) in (u, v): b1 = u - v, b2 = v -> (u, v).
        # Triangle (0,2,3) covers u < v with corners (0,0), (1,1), (0,1):
        # b1 = u, b2 = v - u -> (v, v - u).
        su, sv = (u, v) if u >= v else (v, v - u)
        if min(abs(su - 0.5), abs(sv - 0.5), abs(u - v)) < 2e-3:
            continue
        want = ms.texel(ms.TEX2, min(int(su * 2), 1), min(int(sv * 2), 1))
        assert np.array_equal(L, want), (x, y, su, sv)


def test_sphere_surface_point_is_degrees(orc):
    res = paths(orc, ms.textured_probe("sphere"))
    vals = [L for L, _ in res.values() if L.any()]
    last = ms.texel(ms.TEX2, 1, 1)
    frac = np.mean([np.array_equal(L, last) or np.array_equal(L, ms.texel(ms.TEX2, 1, 0)) for L in vals])
    assert len(vals) > 100 and frac == 1.0  # theta in degrees always reaches the last column
    assert np.mean([np.array_equal(L, last) for L in vals]) > 0.45


def test_textured_cornell_renders(orc):
    d = ms.cornell_textured(16, 16, 4)
    acc, nv = orc.OrcScene(d).render(spp=4, nthreads=8)
    img = acc.reshape(-1, 3) / np.maximum(nv, 1)[:, None]
    assert np.isfinite(img).all() and img.mean() > 0
