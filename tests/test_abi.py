"""The C ABI library loads and exports every entry point include/rt.h
declares; host-side argument checking works without a GPU."""
import ctypes as C
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(rt_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol(rt):
    L = rt.lib()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding declares a signature for each of them
    assert set(names) <= set(rt.exported_symbols())


def test_version_and_error_text(rt):
    L = rt.lib()
    assert L.rt_version() >> 16 == 1
    h = C.c_void_p()
    xyz = np.zeros((3, 3), np.float32)
    idx = np.array([[0, 1, 5]], np.uint32)  # index out of range
    st = L.rt_bvh_build(None, xyz.ctypes.data, 3, idx.ctypes.data, 1, None, C.byref(h))
    assert st == rt.RT_ERR_INVALID
    assert b"out of range" in L.rt_last_error()


def test_nonfinite_vertices_rejected(rt):
    L = rt.lib()
    h = C.c_void_p()
    xyz = np.array([[0, 0, 0], [1, 0, np.nan], [0, 1, 0]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    assert L.rt_bvh_build(None, xyz.ctypes.data, 3, idx.ctypes.data, 1, None, C.byref(h)) == rt.RT_ERR_INVALID


def test_host_only_bvh_refuses_traces(rt):
    L = rt.lib()
    xyz, idx = rt.gen_soup(10, seed=1)
    h = C.c_void_p()
    rt.check(L.rt_bvh_build(None, xyz.ctypes.data, xyz.shape[0], idx.ctypes.data, 10, None, C.byref(h)))
    rays = rt.gen_rays(2, 2, seed=1)
    r = rt.Rays(*[a.ctypes.data for a in rays], None)
    prim = np.empty(4, np.int32)
    assert L.rt_trace_closest(h, C.byref(r), 4, prim.ctypes.data, None, rt.RT_FLAG_HOST_PTRS, None) == rt.RT_ERR_INVALID
    assert b"host-only" in L.rt_last_error()
    L.rt_bvh_destroy(h)


def test_bad_options_rejected(rt):
    L = rt.lib()
    xyz, idx = rt.gen_soup(10, seed=1)
    h = C.c_void_p()
    o = rt.BvhOpts(9, 32, 1.2)
    assert L.rt_bvh_build(None, xyz.ctypes.data, 30, idx.ctypes.data, 10, C.byref(o), C.byref(h)) == rt.RT_ERR_INVALID


def test_device_build_needs_a_context(rt):
    """rt_bvh_build_device checks its arguments before any device call."""
    L = rt.lib()
    h = C.c_void_p()
    assert L.rt_bvh_build_device(None, None, 0, None, 0, None, C.byref(h)) == rt.RT_ERR_INVALID
    assert b"ctx" in L.rt_last_error()
    assert not h.value


def test_generators_deterministic(rt):
    a, _ = rt.gen_soup(1000, seed=7)
    b, _ = rt.gen_soup(1000, seed=7)
    c, _ = rt.gen_soup(1000, seed=8)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    tri = a.reshape(-1, 3, 3).astype(np.float64)
    edges = np.linalg.norm(tri - tri.mean(axis=1, keepdims=True), axis=2)
    # vertex distances from the centre are log-normal(ln 0.003, 0.5)
    assert 0.0025 < np.median(edges) < 0.0036
    cen = tri.mean(axis=1)
    assert cen.min() > -0.05 and cen.max() < 1.05
    r1 = rt.gen_rays(16, 8, n_incoherent=100, seed=3)
    r2 = rt.gen_rays(16, 8, n_incoherent=100, seed=3)
    for x, y in zip(r1, r2):
        assert np.array_equal(x, y)
    d = np.stack(r1[3:], 1).astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-6)
    assert np.all(r1[5][:128] > 0)  # primary rays look along +z
