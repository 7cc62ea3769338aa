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
RT_FLAG_SHADOW = 4
RT_FLAG_PROFILE = 8
RT_FLAG_COUNT = 16


class RtError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"rt status {status}: {msg}")
        self.status = status


class BvhOpts(C.Structure):
    _fields_ = [("max_leaf", C.c_uint32), ("bins", C.c_uint32), ("trav_cost", C.c_float), ("width", C.c_uint32)]


class BvhInfo(C.Structure):
    _fields_ = [
        ("num_tris", C.c_uint32), ("num_nodes", C.c_uint32), ("num_leaves", C.c_uint32),
        ("max_depth", C.c_uint32), ("node_bytes", C.c_uint64), ("tri_bytes", C.c_uint64),
        ("build_ms", C.c_double), ("sah_cost", C.c_double), ("width", C.c_uint32), ("num_nodes4", C.c_uint32),
        ("depth4", C.c_uint32), ("stack_cap", C.c_uint32),
    ]


class Rays(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("ox", "oy", "oz", "dx", "dy", "dz", "tmax")]


MAT_LAMBERTIAN, MAT_DIELECTRIC, MAT_TROWBRIDGE, MAT_TROWBRIDGE_BASIC, MAT_LIGHT = range(5)


class Material(C.Structure):
    _fields_ = [("type", C.c_uint32), ("kd", C.c_float * 3), ("ior", C.c_float), ("roughness", C.c_float),
                ("eta", C.c_float * 3), ("k", C.c_float * 3), ("intensity", C.c_float), ("ke", C.c_float * 3),
                ("texture", C.c_uint32)]


class PhotonOpts(C.Structure):
    _fields_ = [("photons", C.c_uint32), ("max_depth", C.c_uint32), ("rr_depth", C.c_uint32), ("seed", C.c_uint32)]


class PhotonRenderOpts(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("k", C.c_uint32), ("radius", C.c_float)]


RT_RENDER_PHOTON_MAPPING = 1
RT_RENDER_HYBRID = 2


class SssParams(C.Structure):
    _fields_ = [("kd", C.c_float * 3), ("sigma_s", C.c_float * 3), ("sigma_a", C.c_float * 3), ("g", C.c_float),
                ("eta", C.c_float)]


class PointLight(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("color", C.c_float * 3), ("intensity", C.c_float)]


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("lookat", C.c_float * 3), ("up", C.c_float * 3), ("cts", C.c_float * 3),
                ("frame_y", C.c_float), ("width", C.c_uint32), ("height", C.c_uint32)]


class RenderOpts(C.Structure):
    _fields_ = [("spp", C.c_uint32), ("spp_offset", C.c_uint32), ("max_depth", C.c_uint32),
                ("rr_depth", C.c_uint32), ("seed", C.c_uint32), ("path_capacity", C.c_uint32),
                ("grid", C.c_uint32), ("chunk_paths", C.c_uint32)]


class CacheFrameOpts(C.Structure):
    _fields_ = [("cache_depth", C.c_uint32), ("harvest_shift", C.c_uint32), ("train_steps", C.c_uint32),
                ("plain", C.c_uint32)]


class CacheFrameStats(C.Structure):
    _fields_ = [("samples", C.c_uint32), ("steps", C.c_uint32), ("loss", C.c_float), ("fallback", C.c_uint32),
                ("reset", C.c_uint32)]


class RenderStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("ms", C.c_double), ("extend_ms", C.c_double), ("connect_ms", C.c_double),
                ("extend_launches", C.c_uint32), ("connect_launches", C.c_uint32),
                ("ext_node_visits", C.c_uint64), ("ext_tri_tests", C.c_uint64),
                ("shadow_node_visits", C.c_uint64), ("shadow_tri_tests", C.c_uint64),
                ("shadow_occluded", C.c_uint64)]


def f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


class BatchOpts(C.Structure):
    _fields_ = [("spp", C.c_uint32), ("spp_offset", C.c_uint32), ("max_depth", C.c_uint32),
                ("rr_depth", C.c_uint32), ("path_capacity", C.c_uint32), ("grid", C.c_uint32),
                ("chunk_paths", C.c_uint32)]


class TuneEntry(C.Structure):
    _fields_ = [("path_capacity", C.c_uint32), ("grid", C.c_uint32), ("ms", C.c_float)]


class CacheOpts(C.Structure):
    _fields_ = [("levels", C.c_uint32), ("log2_table", C.c_uint32), ("base_res", C.c_uint32),
                ("per_level_scale", C.c_float), ("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float),
                ("eps", C.c_float), ("batch", C.c_uint32), ("seed", C.c_uint32), ("aabb_lo", C.c_float * 3),
                ("aabb_hi", C.c_float * 3)]


class CacheInfo(C.Structure):
    _fields_ = [("levels", C.c_uint32), ("table_size", C.c_uint32), ("res", C.c_uint32 * 16),
                ("dense", C.c_uint32 * 16), ("mlp_params", C.c_uint64), ("grid_params", C.c_uint64),
                ("steps", C.c_uint32)]


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
    "rt_bvh_build_device": (C.c_int, [_P, _P, _U32, _P, _U32, C.POINTER(BvhOpts), C.POINTER(_P)]),
    "rt_scene_set_bvh": (C.c_int, [_P, C.POINTER(BvhOpts), _U32]),
    "rt_bvh_destroy": (C.c_int, [_P]),
    "rt_bvh_get_info": (C.c_int, [_P, C.POINTER(BvhInfo)]),
    "rt_bvh_set_transparent": (C.c_int, [_P, _P]),
    "rt_bvh_export": (C.c_int, [_P, _P, _P]),
    "rt_bvh_export4": (C.c_int, [_P, _P]),
    "rt_bvh_last_kernel_ms": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "rt_trace_closest": (C.c_int, [_P, C.POINTER(Rays), _U32, _P, _P, _U32, _P]),
    "rt_trace_any": (C.c_int, [_P, C.POINTER(Rays), _U32, _P, _U32, _P]),
    "rt_trace_stats": (C.c_int, [_P, C.POINTER(Rays), _U32, C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "rt_scene_create": (C.c_int, [_P, C.POINTER(_P)]),
    "rt_scene_destroy": (C.c_int, [_P]),
    "rt_scene_add_material": (C.c_int, [_P, C.POINTER(Material), C.POINTER(_U32)]),
    "rt_scene_add_mesh": (C.c_int, [_P, _P, _U32, _P, _U32, _U32, C.POINTER(_U32)]),
    "rt_scene_add_sphere": (C.c_int, [_P, _P, C.c_float, _U32, C.c_int]),
    "rt_scene_add_plane": (C.c_int, [_P, _P, _P, _P, _P, _U32]),
    "rt_scene_add_mesh_uv": (C.c_int, [_P, _P, _U32, _P, _U32, _P, _U32, _P]),
    "rt_scene_add_mesh_normals": (C.c_int, [_P, _P, _U32, _P, _U32, _P, _P, _U32, _P]),
    "rt_render_cache_frame": (C.c_int, [_P, _P, _P, C.POINTER(CacheFrameOpts), _P, _P, C.POINTER(CacheFrameStats),
                                        _U32, _P]),
    "rt_scene_add_texture": (C.c_int, [_P, _P, _U32, _U32, _P]),
    "rt_scene_add_point_light": (C.c_int, [_P, C.POINTER(PointLight)]),
    "rt_scene_set_camera": (C.c_int, [_P, C.POINTER(Camera)]),
    "rt_scene_commit": (C.c_int, [_P]),
    "rt_scene_bvh": (_P, [_P]),
    "rt_scene_camera_ray": (C.c_int, [_P, _U32, _U32, C.c_float, C.c_float, _P, _P]),
    "rt_render": (C.c_int, [_P, C.POINTER(RenderOpts), _P, _P, _U32, _P]),
    "rt_render_photon": (C.c_int, [_P, C.POINTER(RenderOpts), _P, _P, _P, C.POINTER(PhotonRenderOpts), _P, _P, _U32,
                                   _P]),
    "rt_render_get_stats": (C.c_int, [_P, C.POINTER(RenderStats)]),
    "rt_render_cached": (C.c_int, [_P, C.POINTER(RenderOpts), _P, _U32, _P, _P, _U32, _P]),
    "rt_batch_create": (C.c_int, [_P, C.POINTER(_P), _U32, C.POINTER(_P)]),
    "rt_batch_destroy": (C.c_int, [_P]),
    "rt_batch_render": (C.c_int, [_P, C.POINTER(BatchOpts), _P, C.POINTER(_P), C.POINTER(_P), _U32, _P]),
    "rt_batch_autotune": (C.c_int, [_P, C.POINTER(BatchOpts), C.POINTER(TuneEntry), C.POINTER(TuneEntry), _U32,
                                    C.POINTER(_U32)]),
    "rt_batch_get_stats": (C.c_int, [_P, C.POINTER(RenderStats)]),
    "rt_cache_train_scene": (C.c_int, [_P, _P, C.POINTER(RenderOpts), _U32, _U32, C.POINTER(C.c_float),
                                       C.POINTER(_U32), _P]),
    "rt_cache_create": (C.c_int, [_P, C.POINTER(CacheOpts), C.POINTER(_P)]),
    "rt_cache_destroy": (C.c_int, [_P]),
    "rt_cache_get_info": (C.c_int, [_P, C.POINTER(CacheInfo)]),
    "rt_cache_query": (C.c_int, [_P, _P, _U32, _P, _P]),
    "rt_cache_train_step": (C.c_int, [_P, _P, _P, _U32, C.POINTER(C.c_float), _P]),
    "rt_cache_gradients": (C.c_int, [_P, _P, _P, _U32, _P, _P, C.POINTER(C.c_float)]),
    "rt_cache_get_params": (C.c_int, [_P, _P, _P]),
    "rt_cache_set_params": (C.c_int, [_P, _P, _P]),
    "rt_image_pixel_bytes": (C.c_size_t, [C.c_int, _U32, _U32]),
    "rt_image_header": (C.c_size_t, [C.c_int, _U32, _U32, _P, C.c_size_t]),
    "rt_film_tonemap": (C.c_int, [_P, _P, _P, _U32, _U32, C.c_uint64, C.c_int, C.c_int, _P, C.c_size_t, _P]),
    "rt_film_tonemap_noalloc": (C.c_int, [_P, _P, _P, _U32, _U32, C.c_uint64, C.c_int, C.c_int, _P, C.c_size_t,
                                          _P]),
    "rt_photon_map_build": (C.c_int, [_P, _P, _P, _U32, C.POINTER(_P)]),
    "rt_photon_map_destroy": (C.c_int, [_P]),
    "rt_photon_map_get_info": (C.c_int, [_P, C.POINTER(_U32), C.POINTER(_U32)]),
    "rt_photon_map_knn": (C.c_int, [_P, _P, _P, _U32, _U32, C.c_float, _P, _P, _P, _P]),
    "rt_photon_gather_lambert": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _U32, _U32, C.c_float, _P, _P]),
    "rt_photon_gather": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _U32, _U32, C.c_float, _P, _P]),
    "rt_photon_subsurface": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _U32, C.c_float, _P, _P, _P]),
    "rt_scene_trace_photons": (C.c_int, [_P, _P, _U32, _P, _P, _P, _P, _P, _P, C.POINTER(_U32), _P]),
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
        self._ext = None

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

    def launch(self, fn, stream=None):
        """Calls fn(stream_handle) in order with torch's current stream.

        A created torch stream is passed through.  torch's default stream has
        handle 0, which the C ABI reads as "the context's own stream" (a
        non-blocking stream that does not synchronise with the default one),
        so the call then runs on the context's stream fenced by events on
        both sides: it sees all earlier torch work and later torch work sees
        its results."""
        if stream is not None:
            return fn(stream)
        import torch
        dev = torch.device("cuda", self.device)
        cur = torch.cuda.current_stream(dev)
        if cur.cuda_stream:
            return fn(cur.cuda_stream)
        if self._ext is None:
            self._ext = torch.cuda.ExternalStream(self.stream, device=dev)
        self._ext.wait_stream(cur)
        try:
            return fn(self._ext.cuda_stream)
        finally:
            cur.wait_stream(self._ext)

    def fence_torch(self):
        """Waits for torch's current stream (before calls that take no stream
        and run on the context's own)."""
        import torch
        torch.cuda.current_stream(torch.device("cuda", self.device)).synchronize()


class Bvh:
    """rt_bvh over a triangle soup given as host float32 vertices (nv,3) and
    uint32 indices (nt,3)."""

    def __init__(self, ctx: Context, xyz, idx, max_leaf=4, bins=32, trav_cost=1.2, width=4):
        import numpy as np
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1, 3)
        self.ctx = ctx
        opts = BvhOpts(max_leaf, bins, trav_cost, width)
        h = _P()
        check(lib().rt_bvh_build(ctx.h, _ptr(xyz), xyz.shape[0], _ptr(idx), idx.shape[0],
                                 C.byref(opts), C.byref(h)))
        self.h = h

    @classmethod
    def from_device(cls, ctx: Context, xyz, idx, max_leaf=4, width=4):
        """rt_bvh_build_device: the tree built on the GPU from device tensors
        (float32 vertices (nv,3), int32/uint32 indices (nt,3)) on ctx's
        device.  Waits for torch's current stream first (the build runs on
        the context's stream and returns when the tree is ready)."""
        import torch
        xyz = xyz.reshape(-1, 3).contiguous()
        idx = idx.reshape(-1, 3).contiguous()
        if xyz.dtype != torch.float32 or idx.dtype not in (torch.int32, torch.uint32):
            raise TypeError("xyz must be float32 and idx int32 or uint32")
        if not (xyz.is_cuda and idx.is_cuda):
            raise ValueError("xyz and idx must be device tensors")
        self = cls.__new__(cls)
        self.ctx = ctx
        opts = BvhOpts(max_leaf, 32, 1.2, width)
        h = _P()
        ctx.fence_torch()
        check(lib().rt_bvh_build_device(ctx.h, _ptr(xyz), xyz.shape[0], _ptr(idx), idx.shape[0],
                                        C.byref(opts), C.byref(h)))
        self.h = h
        return self

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

    def export4(self):
        import numpy as np
        inf = self.info()
        nodes4 = np.zeros((inf.num_nodes4, 16), np.float32)
        check(lib().rt_bvh_export4(self.h, _ptr(nodes4)))
        return nodes4

    def set_transparent(self, flags):
        import numpy as np
        flags = np.ascontiguousarray(flags, dtype=np.uint8)
        check(lib().rt_bvh_set_transparent(self.h, _ptr(flags)))

    @staticmethod
    def _rays(o6, tmax):
        return Rays(*[_ptr(x) for x in o6], _ptr(tmax))

    def trace_closest(self, ox, oy, oz, dx, dy, dz, tmax=None, prim=None, t=None, host=False, stream=None,
                      shadow=False):
        """Device tensors in, device tensors out (or numpy with host=True).
        shadow=True: the closest opaque hit (RT_FLAG_SHADOW, the reference's
        intersectRays with checkShadows)."""
        n = int(ox.shape[0])
        r = self._rays((ox, oy, oz, dx, dy, dz), tmax)
        fl = RT_FLAG_SHADOW if shadow else 0
        if host:
            import numpy as np
            prim = np.empty(n, np.int32) if prim is None else prim
            t = np.empty(n, np.float32) if t is None else t
            check(lib().rt_trace_closest(self.h, C.byref(r), n, _ptr(prim), _ptr(t), RT_FLAG_HOST_PTRS | fl, None))
            return prim, t
        import torch
        dev = ox.device
        prim = torch.empty(n, dtype=torch.int32, device=dev) if prim is None else prim
        t = torch.empty(n, dtype=torch.float32, device=dev) if t is None else t
        self.ctx.launch(lambda s: check(lib().rt_trace_closest(self.h, C.byref(r), n, _ptr(prim), _ptr(t), fl, s)),
                        stream)
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
        self.ctx.launch(lambda s: check(lib().rt_trace_any(self.h, C.byref(r), n, _ptr(occ), 0, s)), stream)
        return occ

    def trace_stats(self, ox, oy, oz, dx, dy, dz, tmax=None, any_hit=False):
        """(node visits, triangle tests) summed over the rays (device tensors)."""
        n = int(ox.shape[0])
        r = self._rays((ox, oy, oz, dx, dy, dz), tmax)
        a = C.c_uint64()
        b = C.c_uint64()
        self.ctx.fence_torch()
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


class Scene:
    """rt_scene built from a scenes.SceneDesc; film buffers are torch
    tensors on the context's device (FP64 sums + per-pixel valid counts)."""

    def __init__(self, ctx: Context, desc, bvh_on_device=False, bvh_width=4, bvh_max_leaf=4):
        """bvh_on_device: the scene's BVH is built on the GPU
        (rt_scene_set_bvh, rt_bvh_build_device) instead of the host's binned
        SAH."""
        import numpy as np
        self.ctx = ctx
        self.desc = desc
        L = lib()
        h = _P()
        check(L.rt_scene_create(ctx.h, C.byref(h)))
        self.h = h
        if bvh_on_device or bvh_width != 4 or bvh_max_leaf != 4:
            bo = BvhOpts(bvh_max_leaf, 32, 1.2, bvh_width)
            check(L.rt_scene_set_bvh(h, C.byref(bo), int(bool(bvh_on_device))))
        for m in desc.material_structs():
            mid = C.c_uint32()
            check(L.rt_scene_add_material(h, C.byref(m), C.byref(mid)))
        for rgb in getattr(desc, "textures", []):
            t = np.ascontiguousarray(rgb, np.uint8)
            check(L.rt_scene_add_texture(h, _ptr(t), t.shape[1], t.shape[0], None))
        uv = getattr(desc, "tri_uv", None)
        uv = None if uv is None else np.ascontiguousarray(uv, np.float32).reshape(-1, 6)
        nrm9 = getattr(desc, "tri_nrm", None)
        nrm9 = None if nrm9 is None else np.ascontiguousarray(nrm9, np.float32).reshape(-1, 9)
        xyz = np.ascontiguousarray(desc.xyz, np.float32)
        tm = np.asarray(desc.tri_mat, np.uint32)
        idx = np.ascontiguousarray(desc.idx, np.uint32)
        # one rt_scene_add_mesh per run of equal material ids (mesh order = id order)
        nt = idx.shape[0]
        start = 0
        while start < nt:
            m = tm[start]
            end = start + int(np.argmax(tm[start:] != m)) if np.any(tm[start:] != m) else nt
            sub = np.ascontiguousarray(idx[start:end])
            suv = None if uv is None else np.ascontiguousarray(uv[start:end])
            if nrm9 is not None:
                snr = np.ascontiguousarray(nrm9[start:end])
                check(L.rt_scene_add_mesh_normals(h, _ptr(xyz), xyz.shape[0], _ptr(sub), sub.shape[0], _ptr(suv),
                                                  _ptr(snr), int(m), None))
            elif uv is None:
                check(L.rt_scene_add_mesh(h, _ptr(xyz), xyz.shape[0], _ptr(sub), sub.shape[0], int(m), None))
            else:
                check(L.rt_scene_add_mesh_uv(h, _ptr(xyz), xyz.shape[0], _ptr(sub), sub.shape[0], _ptr(suv), int(m),
                                             None))
            start = end
        for c, r, m, is_light in desc.spheres:
            cc = np.asarray(c, np.float32)
            check(L.rt_scene_add_sphere(h, _ptr(cc), float(r), int(m), int(is_light)))
        for corner, ax, ay, nrm, m in getattr(desc, "planes", []):
            vs = [np.asarray(v, np.float32) for v in (corner, ax, ay, nrm)]
            check(L.rt_scene_add_plane(h, *[_ptr(v) for v in vs], int(m)))
        for p, col, inten in desc.lights:
            pl = PointLight(f3(p), f3(col), float(inten))
            check(L.rt_scene_add_point_light(h, C.byref(pl)))
        cam = desc.camera_struct()
        check(L.rt_scene_set_camera(h, C.byref(cam)))
        check(L.rt_scene_commit(h))

    def close(self):
        if getattr(self, "h", None):
            lib().rt_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def trace_photons(self, photons=1000000, max_depth=30, rr_depth=3, seed=1, capacity=None, stream=None):
        """rt_scene_trace_photons: the photon pre-pass (generatePhotons +
        computePhotons).  Returns a dict of device tensors location,
        direction, normal (n, 3 float32), radiance (n, 3 float64), shape
        (n, int32) and key (n, int32: path << 5 | bounce), ordered by photon
        path then bounce.  capacity (default 4 per path) grows once if the
        pass stores more."""
        import torch
        cap = capacity if capacity is not None else 4 * photons
        for _ in range(2):
            dev = self.ctx.device
            out = dict(location=torch.empty((cap, 3), dtype=torch.float32, device=dev),
                       direction=torch.empty((cap, 3), dtype=torch.float32, device=dev),
                       normal=torch.empty((cap, 3), dtype=torch.float32, device=dev),
                       radiance=torch.empty((cap, 3), dtype=torch.float64, device=dev),
                       shape=torch.empty(cap, dtype=torch.int32, device=dev),
                       key=torch.empty(cap, dtype=torch.int32, device=dev))
            o = PhotonOpts(photons, max_depth, rr_depth, seed)
            n = _U32()
            st = self.ctx.launch(lambda s: lib().rt_scene_trace_photons(
                self.h, C.byref(o), cap, _ptr(out["location"]), _ptr(out["direction"]), _ptr(out["normal"]),
                _ptr(out["radiance"]), _ptr(out["shape"]), _ptr(out["key"]), C.byref(n), s), stream)
            if st == 0:
                return {k: v[:n.value] for k, v in out.items()}
            if n.value <= cap:
                check(st)
            cap = n.value
        check(st)

    def new_film(self):
        import torch
        n = self.desc.width * self.desc.height
        dev = torch.device("cuda", self.ctx.device)
        return (torch.zeros(n * 3, dtype=torch.float64, device=dev), torch.zeros(n, dtype=torch.int32, device=dev))

    def render(self, film, spp=None, spp_offset=0, max_depth=None, rr_depth=None, seed=None, path_capacity=0,
               grid=0, chunk_paths=0, stream=None, profile=False, count=False):
        """rt_render.  profile: time every traversal launch with HIP events
        (RT_FLAG_PROFILE); count: counting traversal kernels (RT_FLAG_COUNT);
        both are read back with stats()."""
        o = self._opts(spp, spp_offset, max_depth, rr_depth, seed, path_capacity, grid, chunk_paths)
        accum, nvalid = film
        fl = (RT_FLAG_PROFILE if profile else 0) | (RT_FLAG_COUNT if count else 0)
        self.ctx.launch(lambda s: check(lib().rt_render(self.h, C.byref(o), _ptr(accum), _ptr(nvalid), fl, s)),
                        stream)
        return film

    def render_photon(self, film, pmap, radiance, direction, k=100, radius=100.0, spp=None, spp_offset=0,
                      max_depth=None, rr_depth=None, seed=None, path_capacity=0, stream=None, hybrid=False):
        """rt_render_photon.  RENDER_MODE PHOTON_MAPPING: paths end at their
        first scattering hit with the photon map's estimate there instead of
        next-event estimation; hybrid=True (RENDER_MODE HYBRID): every
        scattering hit takes half of each and the path goes on.  pmap: a
        PhotonMap built on the photons' locations and shapes; radiance (n,3)
        float64 and direction (n,3) float32 device tensors in the map's input
        order."""
        o = self._opts(spp, spp_offset, max_depth, rr_depth, seed, path_capacity)
        po = PhotonRenderOpts(RT_RENDER_HYBRID if hybrid else RT_RENDER_PHOTON_MAPPING, k, radius)
        accum, nvalid = film
        rgb = radiance.contiguous()
        dr = direction.contiguous()
        self.ctx.launch(lambda s: check(lib().rt_render_photon(self.h, C.byref(o), pmap.h, _ptr(rgb), _ptr(dr),
                                                               C.byref(po), _ptr(accum), _ptr(nvalid), 0, s)), stream)
        return film

    def write_image(self, film, name, samples, mode="gammaClamp", fmt="bmp"):
        """Film::writeToBMP / writeToPPM of a render film (accum, nvalid)
        holding `samples` samples per pixel."""
        return write_image(self.ctx, name, film[0], self.desc.width, self.desc.height, samples, mode, fmt,
                           nvalid=film[1])

    def _opts(self, spp, spp_offset, max_depth, rr_depth, seed, path_capacity, grid=0, chunk_paths=0):
        d = self.desc
        return RenderOpts(spp if spp is not None else d.spp, spp_offset, max_depth or d.max_depth,
                          d.rr_depth if rr_depth is None else rr_depth, d.seed if seed is None else seed,
                          path_capacity, grid, chunk_paths)

    def render_cached(self, film, cache, cache_depth=2, spp=None, spp_offset=0, max_depth=None, rr_depth=None,
                      seed=None, path_capacity=0, stream=None, profile=False, count=False):
        """rt_render_cached: paths end at their first cacheable vertex at
        depth >= cache_depth and take the radiance cache's answer."""
        o = self._opts(spp, spp_offset, max_depth, rr_depth, seed, path_capacity)
        accum, nvalid = film
        fl = (RT_FLAG_PROFILE if profile else 0) | (RT_FLAG_COUNT if count else 0)
        self.ctx.launch(lambda s: check(lib().rt_render_cached(self.h, C.byref(o), cache.h, cache_depth, _ptr(accum),
                                                               _ptr(nvalid), fl, s)), stream)
        return film

    def render_cache_frame(self, film, cache, cache_depth=2, harvest_shift=4, train_steps=1, plain=False, spp=None,
                           spp_offset=0, max_depth=None, rr_depth=None, seed=None, path_capacity=0, stream=None,
                           profile=False, count=False):
        """rt_render_cache_frame: one frame whose paths end in the cache,
        except one in 2^harvest_shift that goes full length and yields a
        training sample; then train_steps Adam steps on the frame's samples.
        plain=True: no path uses the cache (all of them harvest).  Returns
        CacheFrameStats (the call synchronises)."""
        o = self._opts(spp, spp_offset, max_depth, rr_depth, seed, path_capacity)
        co = CacheFrameOpts(cache_depth, harvest_shift, train_steps, 1 if plain else 0)
        st = CacheFrameStats()
        accum, nvalid = film
        fl = (RT_FLAG_PROFILE if profile else 0) | (RT_FLAG_COUNT if count else 0)
        self.ctx.launch(lambda s: check(lib().rt_render_cache_frame(self.h, C.byref(o), cache.h, C.byref(co),
                                                                    _ptr(accum), _ptr(nvalid), C.byref(st), fl, s)),
                        stream)
        return st

    def train_cache(self, cache, frame, n_paths=65536, cache_depth=2, max_depth=None, rr_depth=None, seed=None,
                    want_loss=True, stream=None):
        """rt_cache_train_scene: one online training step of `cache` on
        n_paths training paths of frame `frame`; returns (loss, samples), or
        (None, None) without waiting when want_loss is False."""
        o = self._opts(1, frame, max_depth, rr_depth, seed, 0)
        loss = C.c_float(0.0)
        n = C.c_uint32(0)
        lp, np_ = (C.byref(loss), C.byref(n)) if want_loss else (None, None)
        self.ctx.launch(lambda s: check(lib().rt_cache_train_scene(cache.h, self.h, C.byref(o), cache_depth, n_paths,
                                                                   lp, np_, s)), stream)
        return (loss.value, n.value) if want_loss else (None, None)

    def stats(self) -> RenderStats:
        st = RenderStats()
        check(lib().rt_render_get_stats(self.h, C.byref(st)))
        return st

    def camera_ray(self, px, py, u0, u1):
        import numpy as np
        o = np.zeros(3, np.float32)
        d = np.zeros(3, np.float32)
        check(lib().rt_scene_camera_ray(self.h, px, py, u0, u1, _ptr(o), _ptr(d)))
        return o, d


RT_FLAG_GRAPH = 2


class Batch:
    """rt_batch: several committed Scenes rendered in one launch sequence
    with shared ray queues.  Films are the scenes' own (Scene.new_film)."""

    def __init__(self, ctx: Context, scenes):
        self.ctx = ctx
        self.scenes = list(scenes)
        arr = (_P * len(self.scenes))(*[sc.h for sc in self.scenes])
        h = _P()
        check(lib().rt_batch_create(ctx.h, arr, len(self.scenes), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().rt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, films, spp, spp_offset=0, max_depth=12, rr_depth=3, seeds=None, path_capacity=0, grid=0,
               graph=False, chunk_paths=0, stream=None, profile=False, count=False):
        import numpy as np
        o = BatchOpts(spp, spp_offset, max_depth, rr_depth, path_capacity, grid, chunk_paths)
        n = len(self.scenes)
        acc = (_P * n)(*[_ptr(f[0]) for f in films])
        nv = (_P * n)(*[_ptr(f[1]) for f in films])
        sd = None if seeds is None else np.ascontiguousarray(seeds, np.uint32)
        flags = ((RT_FLAG_GRAPH if graph else 0) | (RT_FLAG_PROFILE if profile else 0) |
                 (RT_FLAG_COUNT if count else 0))
        self.ctx.launch(lambda s: check(lib().rt_batch_render(self.h, C.byref(o), None if sd is None else _ptr(sd),
                                                              acc, nv, flags, s)), stream)
        return films

    def autotune(self, spp=1, max_depth=12, rr_depth=3):
        """Returns (best TuneEntry, [TuneEntry ...]) and makes best the default."""
        o = BatchOpts(spp, 0, max_depth, rr_depth, 0, 0, 0)
        best = TuneEntry()
        ents = (TuneEntry * 64)()
        cnt = C.c_uint32(0)
        self.ctx.fence_torch()
        check(lib().rt_batch_autotune(self.h, C.byref(o), C.byref(best), ents, 64, C.byref(cnt)))
        return best, [ents[i] for i in range(cnt.value)]

    def stats(self) -> RenderStats:
        st = RenderStats()
        check(lib().rt_batch_get_stats(self.h, C.byref(st)))
        return st


def reference_sums(film, spp_total):
    """Per-pixel sums as the reference's Film holds them (H*W*3 float64):
    the valid samples' sum plus the valid mean for every invalid sample,
    i.e. sum * spp / nvalid (integration/pathTracing.cu:474-486)."""
    import numpy as np
    accum, nvalid = film
    a = accum.double().cpu().numpy().reshape(-1, 3) if hasattr(accum, "cpu") else np.asarray(accum).reshape(-1, 3)
    n = (nvalid.cpu().numpy() if hasattr(nvalid, "cpu") else np.asarray(nvalid)).astype(np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(n[:, None] > 0, a * (spp_total / n[:, None]), 0.0)


def write_hdrbmp(path, sums, width, height, samples):
    """Film::storeContributionsToFile (camera/Film.cu:366-394): writes
    path + ".hdrbmp" as "P6\\n<w> <h>F\\n<samples>\\n" then "r g b " per pixel
    with 6 significant digits."""
    import numpy as np
    s = np.asarray(sums, np.float64).reshape(-1, 3)
    if s.shape[0] != width * height:
        raise ValueError("sums do not match the resolution")
    with open(path + ".hdrbmp", "w") as f:
        f.write(f"P6\n{width} {height}F\n{int(samples)}\n")
        f.write("".join(f"{r:.6g} {g:.6g} {b:.6g} " for r, g, b in s))


def read_hdrbmp(path, width=None, height=None):
    """Film::loadContributionsFromFile (camera/Film.cu:396-449): returns
    (sums (H*W, 3) float64, samples, width, height).  Raises ValueError on a
    bad header, a resolution other than the one asked for, or truncated
    pixel data (the reference prints an error for a mismatch)."""
    import numpy as np
    with open(path + ".hdrbmp") as f:
        tok = f.read().split()
    if len(tok) < 5 or tok[0] != "P6" or not tok[2].endswith(("F", "f", "d")):
        raise ValueError(f"{path}.hdrbmp: bad header")
    w, h, n = int(tok[1]), int(tok[2][:-1]), int(tok[3])
    if (width is not None and w != width) or (height is not None and h != height):
        raise ValueError(f"{path}.hdrbmp: resolution {w}x{h} does not match the film")
    vals = tok[4:]
    if len(vals) < 3 * w * h:
        raise ValueError(f"{path}.hdrbmp: truncated pixel data")
    return np.array(vals[:3 * w * h], np.float64).reshape(-1, 3), n, w, h


# ToneMappingType (camera/Film.cu:13-21), in the reference's order
TONE_MAPPINGS = ("equalize", "reinhard", "reinhardJodie", "ACESfilmic", "uncharted2", "gammaClamp")
IMAGE_FORMATS = ("bmp", "ppm")


def _tone_id(mode):
    if isinstance(mode, str):
        if mode not in TONE_MAPPINGS:
            raise ValueError(f"unknown tone mapping {mode!r}; one of {TONE_MAPPINGS}")
        return TONE_MAPPINGS.index(mode)
    return int(mode)


def _format_id(fmt):
    if fmt not in IMAGE_FORMATS:
        raise ValueError(f"unknown image format {fmt!r}; one of {IMAGE_FORMATS}")
    return IMAGE_FORMATS.index(fmt)


def image_header(fmt, width, height) -> bytes:
    """The bytes before the pixel array: BitmapImage's 54-byte header
    (camera/BMP_Format.cu) or "P6\\n<w> <h>\\n255\\n" (Film.cu:339-350)."""
    buf = (C.c_uint8 * 64)()
    n = lib().rt_image_header(_format_id(fmt), width, height, buf, 64)
    if n == 0:
        raise ValueError(f"bad image size {width}x{height}")
    return bytes(buf[:n])


def tone_map(ctx, sums, width, height, samples, mode="gammaClamp", fmt="bmp", nvalid=None, out=None, stream=None):
    """Film::toneMapBMP / toneMapPPM (camera/Film.cu:285-337) on the GPU:
    returns the image's pixel array as a uint8 device tensor.

    sums: float64 device tensor of width*height*3 pixel sums (Film::pixelLight),
    or, with nvalid (rt_render's valid counts), the valid-sample sums of a
    render film; samples: the film's sample count; mode: a ToneMappingType
    name or id."""
    import torch
    nbytes = lib().rt_image_pixel_bytes(_format_id(fmt), width, height)
    if nbytes == 0:
        raise ValueError(f"bad image size {width}x{height}")
    if sums.dtype != torch.float64 or sums.device.type != "cuda" or sums.numel() != width * height * 3:
        raise ValueError("sums must be a float64 device tensor of width*height*3 values")
    if nvalid is not None and (nvalid.dtype not in (torch.int32, torch.uint32) or nvalid.numel() != width * height):
        raise ValueError("nvalid must be an int32 tensor of width*height counts")
    if out is None:
        out = torch.empty(nbytes, dtype=torch.uint8, device=sums.device)
    elif out.dtype != torch.uint8 or out.numel() < nbytes:
        raise ValueError("out must be a uint8 tensor of at least the pixel array's size")
    ctx.launch(lambda s: check(lib().rt_film_tonemap(ctx.h, _ptr(sums), _ptr(nvalid), width, height, int(samples),
                                                     _tone_id(mode), _format_id(fmt), _ptr(out), out.numel(), s)),
               stream)
    return out


def write_image(ctx, name, sums, width, height, samples, mode="gammaClamp", fmt="bmp", nvalid=None):
    """Film::writeToBMP / writeToPPM (camera/Film.cu:339-362): tone-maps on
    the GPU and writes name + ".bmp" (or ".ppm").  Returns the path."""
    import torch
    pixels = tone_map(ctx, sums, width, height, samples, mode, fmt, nvalid)
    torch.cuda.current_stream(pixels.device).synchronize()
    path = f"{name}.{fmt}"
    with open(path, "wb") as f:
        f.write(image_header(fmt, width, height))
        f.write(pixels.cpu().numpy().tobytes())
    return path


class PhotonMap:
    """rt_photon_map: the kd-tree photon map (Scene::buildPhotonStructures,
    integration/Scene.cu:1102-1114) and its k-nearest search
    (Scene::search_nearest, Scene.cu:704-722)."""

    def __init__(self, ctx, xyz, shape=None):
        import numpy as np
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        sh = None if shape is None else np.ascontiguousarray(shape, np.int32).reshape(-1)
        if sh is not None and sh.size != xyz.shape[0]:
            raise ValueError("one shape id per photon")
        h = _P()
        check(lib().rt_photon_map_build(ctx.h, _ptr(xyz), _ptr(sh), xyz.shape[0], C.byref(h)))
        self.ctx, self.h, self.n = ctx, h, xyz.shape[0]

    def info(self):
        n, d = _U32(), _U32()
        check(lib().rt_photon_map_get_info(self.h, C.byref(n), C.byref(d)))
        return n.value, d.value

    def nearest(self, queries, k, max_distance=float("inf"), query_shape=None, stream=None):
        """KDTree::nearest_neighbors for each query (a float32 device tensor
        of shape (m, 3)): returns (count (m,), index (m, k), distance (m, k))
        device tensors."""
        import torch
        q = queries.contiguous()
        if q.dtype != torch.float32 or q.dim() != 2 or q.shape[1] != 3:
            raise ValueError("queries must be a float32 (m, 3) device tensor")
        m = q.shape[0]
        qs = None if query_shape is None else query_shape.to(device=q.device, dtype=torch.int32).contiguous()
        if qs is not None and qs.numel() != m:
            raise ValueError("one query shape id per query")
        cnt = torch.empty(m, dtype=torch.int32, device=q.device)
        idx = torch.empty((m, k), dtype=torch.int32, device=q.device)
        dist = torch.empty((m, k), dtype=torch.float32, device=q.device)
        self.ctx.launch(lambda s: check(lib().rt_photon_map_knn(self.h, _ptr(q), _ptr(qs), m, k, max_distance,
                                                                _ptr(cnt), _ptr(idx), _ptr(dist), s)), stream)
        return cnt, idx, dist

    def gather_lambert(self, photon_rgb, photon_dir, points, normals, kd, k=100, max_distance=100.0,
                       query_shape=None, stream=None):
        """combinePhotonContributions on Lambertian surfaces (Scene.cu:607-634):
        device tensors photon_rgb (n, 3) float64 and photon_dir (n, 3)
        float32 indexed like the map's photons; points, normals, kd (m, 3)
        float32.  Returns the (m, 3) float64 radiance estimates."""
        import torch
        m = points.shape[0]
        for name, t, dt, rows in (("photon_rgb", photon_rgb, torch.float64, self.n),
                                  ("photon_dir", photon_dir, torch.float32, self.n),
                                  ("points", points, torch.float32, m), ("normals", normals, torch.float32, m),
                                  ("kd", kd, torch.float32, m)):
            if t.dtype != dt or tuple(t.shape) != (rows, 3) or not t.is_contiguous():
                raise ValueError(f"{name} must be a contiguous {dt} ({rows}, 3) tensor")
        qs = None if query_shape is None else query_shape.to(device=points.device, dtype=torch.int32).contiguous()
        if qs is not None and qs.numel() != m:
            raise ValueError("one query shape id per query")
        out = torch.empty((m, 3), dtype=torch.float64, device=points.device)
        self.ctx.launch(lambda s: check(lib().rt_photon_gather_lambert(
            self.h, _ptr(photon_rgb), _ptr(photon_dir), _ptr(points), _ptr(normals), _ptr(kd), _ptr(qs), m, k,
            max_distance, _ptr(out), s)), stream)
        return out

    def gather(self, photon_rgb, photon_dir, points, normals, views, materials, k=100, max_distance=100.0,
               query_shape=None, stream=None):
        """combinePhotonContributions for any material (rt_photon_gather):
        views (m, 3) float32 towards the viewer, materials a sequence of m
        Material structs (or one for all).  Returns (m, 3) float64."""
        import torch
        m = points.shape[0]
        if isinstance(materials, Material):
            materials = [materials] * m
        if len(materials) != m:
            raise ValueError("one material per query")
        mats = (Material * max(1, m))(*materials)
        mt = torch.frombuffer(bytearray(bytes(mats)), dtype=torch.uint8).to(points.device)
        qs = None if query_shape is None else query_shape.to(device=points.device, dtype=torch.int32).contiguous()
        out = torch.empty((m, 3), dtype=torch.float64, device=points.device)
        self.ctx.launch(lambda s: check(lib().rt_photon_gather(
            self.h, _ptr(photon_rgb), _ptr(photon_dir), _ptr(points), _ptr(normals), _ptr(views), _ptr(mt), _ptr(qs),
            m, k, max_distance, _ptr(out), s)), stream)
        return out

    def subsurface(self, photon_rgb, photon_dir, photon_nrm, points, normals, views, params, radius=0.3,
                   query_shape=None, stream=None):
        """Scene::combinePhotonsSubsurface (rt_photon_subsurface); params =
        (kd, sigma_s, sigma_a, g, eta).  Returns (m, 3) float64 (NaN where
        no photon lies within the radius, as in the reference)."""
        import torch
        m = points.shape[0]
        kd, ss, sa, g, eta = params
        prm = SssParams((C.c_float * 3)(*kd), (C.c_float * 3)(*ss), (C.c_float * 3)(*sa), g, eta)
        qs = None if query_shape is None else query_shape.to(device=points.device, dtype=torch.int32).contiguous()
        out = torch.empty((m, 3), dtype=torch.float64, device=points.device)
        self.ctx.launch(lambda s: check(lib().rt_photon_subsurface(
            self.h, _ptr(photon_rgb), _ptr(photon_dir), _ptr(photon_nrm), _ptr(points), _ptr(normals), _ptr(views),
            _ptr(qs), m, radius, C.byref(prm), _ptr(out), s)), stream)
        return out

    def close(self):
        if self.h:
            lib().rt_photon_map_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def film_image(film, spp_total):
    """Mean radiance per pixel (H*W*3 float64 numpy) from FP64 sums and valid
    counts: the reference's pixel sum is sum * (spp / nvalid)
    (integration/pathTracing.cu:474-486), divided by spp for the mean."""
    import numpy as np
    accum, nvalid = film
    a = accum.double().cpu().numpy().reshape(-1, 3)
    n = nvalid.cpu().numpy().astype(np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        img = np.where(n[:, None] > 0, a / n[:, None], 0.0)
    return img


def compute_rmse(img_a, img_b):
    """computeRMSE (camera/Film.cu:611-632): sqrt of the mean over pixels of
    the per-pixel squared norm of the difference of mean radiances (the RGB
    channel SUM of squares), accumulated here in float64."""
    import numpy as np
    d = np.asarray(img_a, np.float64).reshape(-1, 3) - np.asarray(img_b, np.float64).reshape(-1, 3)
    return float(np.sqrt(np.mean(np.sum(d * d, axis=1))))


class RadianceCache:
    """rt_cache: hash-grid + MLP radiance cache trained on the GPU."""

    def __init__(self, ctx: Context, aabb_lo, aabb_hi, levels=16, log2_table=19, base_res=16, per_level_scale=1.38,
                 lr=1e-2, beta1=0.9, beta2=0.99, eps=1e-8, batch=65536, seed=1):
        self.ctx = ctx
        o = CacheOpts(levels, log2_table, base_res, per_level_scale, lr, beta1, beta2, eps, batch, seed,
                      f3(aabb_lo), f3(aabb_hi))
        h = _P()
        check(lib().rt_cache_create(ctx.h, C.byref(o), C.byref(h)))
        self.h = h
        self.opts = o

    def close(self):
        if getattr(self, "h", None):
            lib().rt_cache_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> CacheInfo:
        i = CacheInfo()
        check(lib().rt_cache_get_info(self.h, C.byref(i)))
        return i

    def query(self, inputs, rgb=None, stream=None):
        import torch
        n = int(inputs.shape[0])
        rgb = torch.empty((n, 3), dtype=torch.float32, device=inputs.device) if rgb is None else rgb
        self.ctx.launch(lambda s: check(lib().rt_cache_query(self.h, _ptr(inputs), n, _ptr(rgb), s)), stream)
        return rgb

    def train_step(self, inputs, targets, want_loss=True, stream=None):
        import torch
        loss = C.c_float(0.0)
        self.ctx.launch(lambda s: check(lib().rt_cache_train_step(self.h, _ptr(inputs), _ptr(targets),
                                                                  int(inputs.shape[0]),
                                                                  C.byref(loss) if want_loss else None, s)), stream)
        return loss.value if want_loss else None

    def gradients(self, inputs, targets):
        import numpy as np
        inf = self.info()
        mg = np.zeros(inf.mlp_params, np.float32)
        gg = np.zeros(inf.grid_params, np.float32)
        loss = C.c_float(0.0)
        self.ctx.fence_torch()
        check(lib().rt_cache_gradients(self.h, _ptr(inputs), _ptr(targets), int(inputs.shape[0]), _ptr(mg),
                                       _ptr(gg), C.byref(loss)))
        return loss.value, mg, gg

    def get_params(self):
        import numpy as np
        inf = self.info()
        mg = np.zeros(inf.mlp_params, np.float32)
        gg = np.zeros(inf.grid_params, np.float32)
        check(lib().rt_cache_get_params(self.h, _ptr(mg), _ptr(gg)))
        return mg, gg

    def set_params(self, mlp, grid):
        import numpy as np
        mlp = np.ascontiguousarray(mlp, np.float32)
        grid = np.ascontiguousarray(grid, np.float32)
        check(lib().rt_cache_set_params(self.h, _ptr(mlp), _ptr(grid)))
