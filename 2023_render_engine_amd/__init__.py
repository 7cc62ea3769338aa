"""Python binding of the C ABI in include/rt.h (librt_amd.so).

The shared library is the product: HIP kernels for gfx950 behind a C ABI.
This module only loads it with ctypes, declares the signatures and offers
small conveniences (device buffers are torch tensors on ``cuda:0``).  It
never falls back to a CPU implementation: if the library is missing, or no
GPU is visible when a GPU call is made, it raises.

The package directory name starts with a digit, so load it with
``importlib.import_module("2023_render_engine_amd")``.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librt_amd.so")

RT_OK, RT_ERR_INVALID, RT_ERR_HIP, RT_ERR_OOM, RT_ERR_UNSUPPORTED, RT_ERR_NO_DEVICE = range(6)
RT_FLAG_HOST_PTRS = 1


class RtError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"rt status {status}: {msg}")
        self.status = status


class BvhOpts(C.Structure):
    _fields_ = [("max_leaf", C.c_uint32), ("bins", C.c_uint32), ("trav_cost", C.c_float)]


class BvhInfo(C.Structure):
    _fields_ = [
        ("num_tris", C.c_uint32), ("num_nodes", C.c_uint32), ("num_leaves", C.c_uint32),
        ("max_depth", C.c_uint32), ("node_bytes", C.c_uint64), ("tri_bytes", C.c_uint64),
        ("build_ms", C.c_double), ("sah_cost", C.c_double),
    ]


class Rays(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("ox", "oy", "oz", "dx", "dy", "dz", "tmax")]


_P = C.c_void_p
_U32 = C.c_uint32
_SIGS = {
    "rt_last_error": (C.c_char_p, []),
    "rt_version": (C.c_uint32, []),
    "rt_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "rt_ctx_destroy": (C.c_int, [_P]),
    "rt_ctx_synchronize": (C.c_int, [_P]),
    "rt_ctx_stream": (_P, [_P]),
    "rt_malloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "rt_free": (C.c_int, [_P, _P]),
    "rt_memcpy_h2d": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "rt_memcpy_d2h": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "rt_bvh_build": (C.c_int, [_P, _P, _U32, _P, _U32, C.POINTER(BvhOpts), C.POINTER(_P)]),
    "rt_bvh_destroy": (C.c_int, [_P]),
    "rt_bvh_get_info": (C.c_int, [_P, C.POINTER(BvhInfo)]),
    "rt_bvh_set_transparent": (C.c_int, [_P, _P]),
    "rt_bvh_export": (C.c_int, [_P, _P, _P]),
    "rt_bvh_last_kernel_ms": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "rt_trace_closest": (C.c_int, [_P, C.POINTER(Rays), _U32, _P, _P, _U32, _P]),
    "rt_trace_any": (C.c_int, [_P, C.POINTER(Rays), _U32, _P, _U32, _P]),
    "rt_trace_stats": (C.c_int, [_P, C.POINTER(Rays), _U32, C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "rt_gen_soup": (C.c_int, [_U32, C.c_uint64, C.c_float, C.c_float, _P]),
    "rt_gen_rays": (C.c_int, [_U32, _U32, _P, C.c_float, _U32, C.c_uint64, _P, _P, _P, _P, _P, _P]),
}

_lib = None


def lib() -> C.CDLL:
    """Load librt_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def check(status: int) -> None:
    if status != RT_OK:
        msg = lib().rt_last_error()
        raise RtError(status, msg.decode() if msg else "")


def _ptr(a) -> int:
    """Address of a numpy array or torch tensor (contiguous)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        assert a.is_contiguous()
        return a.data_ptr()
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class Context:
    """An rt_ctx on one HIP device."""

    def __init__(self, device: int = 0):
        h = _P()
        check(lib().rt_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            lib().rt_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(lib().rt_ctx_synchronize(self.h))

    @property
    def stream(self) -> int:
        return lib().rt_ctx_stream(self.h)


class Bvh:
    """rt_bvh over a triangle soup given as host float32 vertices (nv,3) and
    uint32 indices (nt,3)."""

    def __init__(self, ctx: Context, xyz, idx, max_leaf=4, bins=32, trav_cost=1.2):
        import numpy as np
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1, 3)
        self.ctx = ctx
        opts = BvhOpts(max_leaf, bins, trav_cost)
        h = _P()
        check(lib().rt_bvh_build(ctx.h, _ptr(xyz), xyz.shape[0], _ptr(idx), idx.shape[0],
                                 C.byref(opts), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().rt_bvh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> BvhInfo:
        inf = BvhInfo()
        check(lib().rt_bvh_get_info(self.h, C.byref(inf)))
        return inf

    def export(self):
        import numpy as np
        inf = self.info()
        nodes = np.zeros((inf.num_nodes, 16), np.float32)
        tris = np.zeros((inf.num_tris + 1, 12), np.float32)
        check(lib().rt_bvh_export(self.h, _ptr(nodes), _ptr(tris)))
        return nodes, tris

    def set_transparent(self, flags):
        import numpy as np
        flags = np.ascontiguousarray(flags, dtype=np.uint8)
        check(lib().rt_bvh_set_transparent(self.h, _ptr(flags)))

    @staticmethod
    def _rays(o6, tmax):
        return Rays(*[_ptr(x) for x in o6], _ptr(tmax))

    def trace_closest(self, ox, oy, oz, dx, dy, dz, tmax=None, prim=None, t=None, host=False, stream=None):
        """Device tensors in, device tensors out (or numpy with host=True)."""
        n = int(ox.shape[0])
        r = self._rays((ox, oy, oz, dx, dy, dz), tmax)
        if host:
            import numpy as np
            prim = np.empty(n, np.int32) if prim is None else prim
            t = np.empty(n, np.float32) if t is None else t
            check(lib().rt_trace_closest(self.h, C.byref(r), n, _ptr(prim), _ptr(t), RT_FLAG_HOST_PTRS, None))
            return prim, t
        import torch
        dev = ox.device
        prim = torch.empty(n, dtype=torch.int32, device=dev) if prim is None else prim
        t = torch.empty(n, dtype=torch.float32, device=dev) if t is None else t
        s = stream if stream is not None else torch.cuda.current_stream(dev).cuda_stream
        check(lib().rt_trace_closest(self.h, C.byref(r), n, _ptr(prim), _ptr(t), 0, s))
        return prim, t

    def trace_any(self, ox, oy, oz, dx, dy, dz, tmax, occ=None, host=False, stream=None):
        n = int(ox.shape[0])
        r = self._rays((ox, oy, oz, dx, dy, dz), tmax)
        if host:
            import numpy as np
            occ = np.empty(n, np.uint8) if occ is None else occ
            check(lib().rt_trace_any(self.h, C.byref(r), n, _ptr(occ), RT_FLAG_HOST_PTRS, None))
            return occ
        import torch
        occ = torch.empty(n, dtype=torch.uint8, device=ox.device) if occ is None else occ
        s = stream if stream is not None else torch.cuda.current_stream(ox.device).cuda_stream
        check(lib().rt_trace_any(self.h, C.byref(r), n, _ptr(occ), 0, s))
        return occ

    def trace_stats(self, ox, oy, oz, dx, dy, dz, tmax=None, any_hit=False):
        """(node visits, triangle tests) summed over the rays (device tensors)."""
        n = int(ox.shape[0])
        r = self._rays((ox, oy, oz, dx, dy, dz), tmax)
        a = C.c_uint64()
        b = C.c_uint64()
        check(lib().rt_trace_stats(self.h, C.byref(r), n, int(any_hit), C.byref(a), C.byref(b)))
        return a.value, b.value

    def last_kernel_ms(self) -> float:
        ms = C.c_float()
        check(lib().rt_bvh_last_kernel_ms(self.h, C.byref(ms)))
        return ms.value


def gen_soup(nt: int, seed: int = 7, edge_mu: float = 0.003, edge_sigma: float = 0.5):
    """C1 triangle soup: returns (xyz (3nt,3) float32, idx (nt,3) uint32)."""
    import numpy as np
    tri = np.empty((nt, 9), np.float32)
    check(lib().rt_gen_soup(nt, seed, edge_mu, edge_sigma, _ptr(tri)))
    xyz = tri.reshape(-1, 3)
    idx = np.arange(3 * nt, dtype=np.uint32).reshape(-1, 3)
    return xyz, idx


def gen_rays(w: int, h: int, eye=(0.5, 0.5, -1.0), vfov=60.0, n_incoherent: int = 0, seed: int = 7):
    """C1 rays: w*h coherent pinhole rays then n_incoherent rays; returns six
    float32 numpy arrays (ox, oy, oz, dx, dy, dz)."""
    import numpy as np
    n = w * h + n_incoherent
    arrs = [np.empty(n, np.float32) for _ in range(6)]
    e = np.asarray(eye, np.float32)
    check(lib().rt_gen_rays(w, h, _ptr(e), vfov, n_incoherent, seed, *[_ptr(a) for a in arrs]))
    return arrs
