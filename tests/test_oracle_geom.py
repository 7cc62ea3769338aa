"""Pins the oracle's restated geometry (shapes/Triangle.cu:433-643,
geometry/BoundingBox.cu:296-329, shapes/Sphere.cu:18-398,
cuda_libraries/math.cu:117-194) against closed forms, textbook algorithms in
float64, brute force and the special cases the reference's code has."""
import math

import numpy as np
import pytest

F32 = np.float32


def moller_trumbore(o, d, p0, p1, p2):
    """Textbook ray/triangle intersection in float64: (t, b1, b2) or None."""
    o, d, p0, p1, p2 = (np.asarray(x, np.float64) for x in (o, d, p0, p1, p2))
    e1, e2 = p1 - p0, p2 - p0
    pv = np.cross(d, e2)
    det = e1 @ pv
    if abs(det) < 1e-300:
        return None
    inv = 1.0 / det
    tv = o - p0
    u = (tv @ pv) * inv
    qv = np.cross(tv, e1)
    v = (d @ qv) * inv
    t = (e2 @ qv) * inv
    return t, u, v


def test_error_bound_closed_form(orc):
    me = np.float64(np.finfo(np.float32).eps) * 0.5
    for n in (1, 2, 3, 4, 5, 7):
        want = F32((n * me) / (1 - n * me))
        assert orc.lib().orc_error_bound(n) == want


@pytest.mark.parametrize("x", [0.0, -0.0, 1.0, -1.0, 1e-30, -3.5e12, 1.4e-45, np.inf, -np.inf])
def test_next_float_matches_nextafter(orc, x):
    L = orc.lib()
    x32 = F32(x)
    up, down = L.orc_next_float_up(x32), L.orc_next_float_down(x32)
    # The reference maps -0 to +0 before stepping up and +0 to -0 before stepping down.
    if np.isinf(x32) and x32 > 0:
        assert up == np.inf
    else:
        assert F32(up) == np.nextafter(F32(0.0) if x32 == 0 else x32, F32(np.inf))
    if np.isinf(x32) and x32 < 0:
        assert down == -np.inf
    else:
        assert F32(down) == np.nextafter(F32(-0.0) if x32 == 0 else x32, F32(-np.inf))


def test_triangle_axis_aligned_exact(orc):
    ok, t, info = orc.tri_intersects([0.2, 0.2, 0], [0, 0, 1], [0, 0, 2], [1, 0, 2], [0, 1, 2])
    assert ok and info.reached and not info.masked
    assert abs(t - 2.0) <= 2 * np.spacing(F32(2.0))
    # outside the triangle (x + y > 1)
    ok, _, info = orc.tri_intersects([0.7, 0.7, 0], [0, 0, 1], [0, 0, 2], [1, 0, 2], [0, 1, 2])
    assert not ok and not info.reached
    # behind the origin
    ok, _, _ = orc.tri_intersects([0.2, 0.2, 3], [0, 0, 1], [0, 0, 2], [1, 0, 2], [0, 1, 2])
    assert not ok


def test_triangle_back_face_and_tmax(orc):
    # reversed winding: det < 0 branch (Triangle.cu:547-552)
    ok, t, _ = orc.tri_intersects([0.2, 0.2, 0], [0, 0, 1], [0, 0, 2], [0, 1, 2], [1, 0, 2])
    assert ok and abs(t - 2.0) < 1e-6
    # tMax clipping in the scaled range test (:547-558)
    ok, _, _ = orc.tri_intersects([0.2, 0.2, 0], [0, 0, 1], [0, 0, 2], [1, 0, 2], [0, 1, 2], tmax=1.5)
    assert not ok
    ok, _, _ = orc.tri_intersects([0.2, 0.2, 0], [0, 0, 1], [0, 0, 2], [1, 0, 2], [0, 1, 2], tmax=2.5)
    assert ok
    # running hitOffset: a farther candidate is rejected (:569-571)
    ok, h, _ = orc.tri_intersects([0.2, 0.2, 0], [0, 0, 1], [0, 0, 2], [1, 0, 2], [0, 1, 2], hit_offset=1.0)
    assert not ok and h == 1.0


def test_triangle_against_textbook_float64(orc):
    rng = np.random.default_rng(3)
    n_checked = 0
    for _ in range(3000):
        p = rng.uniform(-1, 1, (3, 3)).astype(F32)
        o = rng.uniform(-3, 3, 3).astype(F32)
        target = (p[0] + p[1] + p[2]) / 3 + rng.normal(0, 0.4, 3)
        d = (target - o)
        d = (d / np.linalg.norm(d)).astype(F32)
        mt = moller_trumbore(o, d, *p)
        ok, t, info = orc.tri_intersects(o, d, *p)
        if mt is None:
            continue
        tt, u, v = mt
        margin = min(u, v, 1 - u - v)
        if margin > 1e-3 and tt > 1e-3:
            assert ok, (o, d, p)
            assert abs(t - tt) <= 1e-5 * abs(tt)
            n_checked += 1
        elif margin < -1e-3 or tt < -1e-3:
            assert not ok
    assert n_checked > 300


def test_triangle_watertight_shared_edge(orc):
    """Rays through points of a shared edge hit at least one of the two
    triangles (the property the shear/edge-function form guarantees)."""
    rng = np.random.default_rng(5)
    a = np.array([0.1, -0.3, 1.7], F32)
    b = np.array([0.9, 0.8, 2.1], F32)
    c = np.array([-0.6, 0.9, 1.9], F32)
    e = np.array([1.2, -0.5, 2.4], F32)
    misses = 0
    for _ in range(2000):
        s = rng.uniform(0.01, 0.99)
        q = (a + F32(s) * (b - a)).astype(F32)
        o = rng.uniform(-0.5, 0.5, 3).astype(F32)
        d = (q - o).astype(F32)
        hit1, _, _ = orc.tri_intersects(o, d, a, b, c)
        hit2, _, _ = orc.tri_intersects(o, d, b, a, e)
        misses += not (hit1 or hit2)
    assert misses == 0


def test_triangle_fp64_fallback_branch(orc):
    # vertex 0 on the ray axis: two edge functions are exactly 0 in FP32
    ok, t, info = orc.tri_intersects([0, 0, 0], [0, 0, 1], [0, 0, 1], [1, 0, 1], [0, 1, 1])
    assert info.fp64 == 1
    assert ok and t == 1.0


def test_triangle_hit_offset_lowered_before_deltaT(orc):
    """Triangle.cu:573 lowers hitOffset before the deltaT test at :592: a
    candidate at a distance below its error bound is rejected but still
    lowers hitOffset."""
    # a steep plane z = y + 1e-7 crossing the ray at t = 1e-7 while the
    # triangle spans |z| ~ 1: deltaT ~ 1e-6 > t
    p0, p1, p2 = [-1, -1, 1e-7 - 1], [1, -1, 1e-7 - 1], [0, 1, 1e-7 + 1]
    ok, h, info = orc.tri_intersects([0, 0, 0], [0, 0, 1], p0, p1, p2, hit_offset=10.0)
    assert info.reached and info.masked and not ok
    assert h == info.t and h < 10.0


def test_reference_slab_test(orc):
    L = orc.lib()
    import ctypes as C
    lo = np.zeros(3, F32)
    hi = np.ones(3, F32)
    h = C.c_float()
    r = orc.make_ray([-1, 0.5, 0.5], [1, 0, 0])
    assert L.orc_box_intersects(lo.ctypes.data, hi.ctypes.data, C.byref(r), C.byref(h)) and h.value == 1.0
    r = orc.make_ray([0.5, 0.5, 0.5], [0, 1, 0])  # inside: hitOffset = exit distance
    assert L.orc_box_intersects(lo.ctypes.data, hi.ctypes.data, C.byref(r), C.byref(h)) and h.value == 0.5
    r = orc.make_ray([-1, 2, 0.5], [1, 0, 0])  # parallel slab, outside
    assert not L.orc_box_intersects(lo.ctypes.data, hi.ctypes.data, C.byref(r), C.byref(h))
    r = orc.make_ray([-1, 0.5, 0.5], [1, 0, 0], tmax=0.5)  # beyond tMax
    assert not L.orc_box_intersects(lo.ctypes.data, hi.ctypes.data, C.byref(r), C.byref(h))


def sphere_closed_form(o, d, c, r):
    o, d, c = (np.asarray(x, np.float64) for x in (o, d, c))
    d = d / np.linalg.norm(d)
    oc = o - c
    b = oc @ d
    disc = b * b - (oc @ oc - r * r)
    if disc < 0:
        return None
    s = math.sqrt(disc)
    return -b - s, -b + s


def test_sphere_against_closed_form(orc):
    rng = np.random.default_rng(9)
    n_hit = 0
    for _ in range(2000):
        c = rng.uniform(-2, 2, 3).astype(F32)
        r = F32(rng.uniform(0.05, 1.0))
        s = orc.Sphere(c, r)
        o = rng.uniform(-4, 4, 3).astype(F32)
        d = (c + rng.normal(0, 0.8 * float(r), 3) - o).astype(F32)  # aimed near the sphere
        cf = sphere_closed_form(o, d, c, float(r))
        ok, t = s.intersects(o, d)
        dn = np.linalg.norm(d.astype(np.float64))
        if cf is None:
            assert not ok or True  # grazing rays may round either way
            continue
        t0, t1 = cf
        # The reference solves b^2 - 4ac in FP32 (math.cu:91-115), which
        # cancels when |o - c| >> r: the root moves by ~eps |o - c|^2 / half-chord.
        oc2 = float(np.sum((o.astype(np.float64) - c) ** 2))
        half_chord = max(0.5 * (t1 - t0), 1e-6)
        # plus the rounding of the FP32 world->object translation
        scale = float(np.abs(o).max() + np.abs(c).max() + r)
        tol = lambda tt: 1e-6 * abs(tt) + 16 * 6e-8 * (oc2 / half_chord + scale)
        if t0 > 1e-3 * r:
            assert ok
            # the object-space direction is renormalised (Sphere.cu:287), so t
            # is a distance along the unit direction, as in the closed form
            assert abs(t - t0) <= tol(t0)
            n_hit += 1
        elif t0 < -1e-3 * r and t1 > 1e-3 * r:
            assert ok and abs(t - t1) <= tol(t1)  # origin inside: far root
        elif t1 < -1e-3 * r:
            assert not ok
        assert dn > 0
    assert n_hit > 100


def test_sphere_strict_less_than(orc):
    s = orc.Sphere([0, 0, 5], 1.0)
    ok, t = s.intersects([0, 0, 0], [0, 0, 1])
    assert ok and abs(t - 4.0) < 1e-5
    ok2, t2 = s.intersects([0, 0, 0], [0, 0, 1], hit_offset=t)  # equal t is not a hit (Sphere.cu:390)
    assert not ok2 and t2 == t


def test_sphere_surface_point_and_normal(orc):
    s = orc.Sphere([1.0, -2.0, 0.5], 0.5)
    ok, t, p, n, e = s.intersect([1.0, -2.0, -3.0], [0, 0, 1])
    assert ok
    assert np.allclose(p, [1.0, -2.0, 0.0], atol=1e-5)
    assert np.allclose(n, [0, 0, -1], atol=1e-5)  # outward, not flipped (Sphere.cu:160)
    assert np.all(e >= 0) and np.all(e < 1e-5)


def test_mat4_inverse(orc):
    rng = np.random.default_rng(1)
    m = rng.normal(size=(4, 4)).astype(F32)
    out = np.zeros(16, F32)
    orc.lib().orc_mat4_inverse(np.ascontiguousarray(m).ctypes.data, out.ctypes.data)
    assert np.allclose(out.reshape(4, 4) @ m, np.eye(4), atol=1e-4)
