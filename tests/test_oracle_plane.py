"""The oracle's Plane (shapes/Plane.cu) pinned without the reference:
closed-form hits on the reference driver's back wall (cornellBoxGuion.cu:60),
its quirks (a local d.x of 0 gives a NaN distance and no hit; axis lengths
truncated to integers; strict < against the current hit), the equivalent
two-triangle quad, and the constructor's matrices against numpy."""
import numpy as np
import pytest

import material_scenes as ms


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


@pytest.fixture
def back(orc):
    return orc.Plane(*ms.REF_PLANES["plane1"])


def test_constructor_matrices(orc):
    for name, (c, ax, ay, n) in ms.REF_PLANES.items():
        p = orc.Plane(c, ax, ay, n)
        cr = np.cross(ax, ay)
        N = unit(cr if np.dot(cr, n) > 0 else -cr)
        want = np.eye(4)
        want[:3, 0] = unit(ax) * int(np.linalg.norm(ax))
        want[:3, 1] = unit(ay) * int(np.linalg.norm(ay))
        want[:3, 2] = N
        want[:3, 3] = c
        assert np.allclose(p.o2w, want, atol=1e-7), name
        assert np.allclose(p.w2o, np.linalg.inv(want), atol=1e-6), name
    # the axis lengths are truncated: 1.9 acts as 1, 2.0 as 2
    p = orc.Plane((0, 0, 0), (0, 1.9, 0), (0, 0, 1), (1, 0, 0))
    assert np.allclose(p.o2w[:3, 0], (0, 1, 0))
    p = orc.Plane((0, 0, 0), (0, 2.0, 0), (0, 0, 1), (1, 0, 0))
    assert np.allclose(p.o2w[:3, 0], (0, 2, 0))


def test_closed_form_hit(back):
    o = np.array([-1.0, 0.1, 0.2])
    d = unit([1.0, 0.05, 0.02])
    t = 1.5 / d[0]
    ok, off, p, n, e = back.intersect(o, d)
    assert ok
    assert abs(off - t) <= 2e-6 * t
    assert np.allclose(p, o + t * d, atol=2e-6)
    assert np.allclose(n, (-1, 0, 0))  # faces the incoming ray
    assert np.all(e >= 1e-4) and np.all(e < 1.1e-4)
    # from behind the normal turns the other way
    ok, _, _, n2, _ = back.intersect(np.array([2.0, 0.1, 0.2]), unit([-1.0, 0.05, 0.02]))
    assert ok and np.allclose(n2, (1, 0, 0))


def test_quirks_and_bounds(orc, back):
    o = np.array([-1.0, 0.1, 0.2])
    # local x is world y here: a ray with d.y == 0 gets a NaN distance and misses
    assert not back.intersect(o, unit([1.0, 0.0, 0.02]))[0]
    # outside the unit square (y = 0.7 -> u = 1.2)
    assert not back.intersect(np.array([-1.0, 0.7, 0.2]), unit([1.0, 0.01, 0.0]))[0]
    # parallel to the plane
    assert not back.intersect(o, unit([0.0, 1.0, 0.3]))[0]
    # strict < against the current hit distance
    d = unit([1.0, 0.05, 0.02])
    ok, off, *_ = back.intersect(o, d)
    assert not back.intersect(o, d, hit_offset=off)[0]
    assert back.intersect(o, d, hit_offset=np.nextafter(np.float32(off), np.float32(np.inf)))[0]
    # tMax: a hit beyond it misses
    assert not back.intersect(o, d, tmax=1.0)[0]
    # truncated axes: 1.9 covers u in [0, 1] only
    p19 = orc.Plane((0.5, -0.5, -0.5), (0, 1.9, 0), (0, 0, 1), (-0.5, 0, 0))
    p20 = orc.Plane((0.5, -0.5, -0.5), (0, 2.0, 0), (0, 0, 1), (-0.5, 0, 0))
    o2 = np.array([-1.0, 0.7, 0.2])
    d2 = unit([1.0, 0.01, 0.0])
    assert not p19.intersect(o2, d2)[0] and p20.intersect(o2, d2)[0]


def test_matches_the_two_triangle_quad(orc):
    """Random rays through the plane's parallelogram hit at the triangle
    test's t, within 4e-6 relative plus the rounding that getOffset's
    division by the local x direction amplifies (ulp(o.x) / |d.x|); rays that
    pass outside miss both."""
    rng = np.random.default_rng(3)
    for name, (c, ax, ay, n) in ms.REF_PLANES.items():
        p = orc.Plane(c, ax, ay, n)
        c, ax, ay = (np.array(v, np.float64) for v in (c, ax, ay))
        quad = [c, c + ax, c + ax + ay, c + ay]
        for k in range(200):
            uv = rng.uniform(-0.3, 1.3, 2)
            target = c + uv[0] * ax + uv[1] * ay
            o = rng.uniform(-0.4, 0.4, 3)
            d = unit(target - o)
            if abs(d).min() < 1e-3:
                continue
            inside = 0.01 < uv[0] < 0.99 and 0.01 < uv[1] < 0.99
            outside = uv[0] < -0.01 or uv[0] > 1.01 or uv[1] < -0.01 or uv[1] > 1.01
            ok, off, *_ = p.intersect(o, d)
            a1, t1, _ = orc.tri_intersects(o, d, quad[0], quad[1], quad[2])
            a2, t2, _ = orc.tri_intersects(o, d, quad[0], quad[2], quad[3])
            tt = t1 if a1 else (t2 if a2 else None)
            if inside:
                ol = p.w2o[:3, :3] @ o + p.w2o[:3, 3]
                dl = p.w2o[:3, :3] @ d
                slack = 4 * np.finfo(np.float32).eps * (abs(ol[0]) + 1) / abs(dl[0])
                assert ok and tt is not None and abs(off - tt) <= 4e-6 * tt + slack, (name, uv)
            if outside:
                assert not ok and tt is None, (name, uv)


def test_oracle_renders_the_plane_box(orc):
    for mesh in (False, True):
        d = ms.cornell_planes(16, 16, 4, with_mesh=mesh)
        acc, nv = orc.OrcScene(d).render(spp=4, nthreads=8)
        img = acc.reshape(-1, 3) / np.maximum(nv, 1)[:, None]
        assert np.isfinite(img).all() and (img.sum(1) > 0).mean() > 0.9
