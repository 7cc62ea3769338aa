"""ctypes loader of liboracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle is the CPU restatement of the reference's hot path (see
oracle.h).  It is imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, and only as the checker or the timed CPU
baseline; the shipped library never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_P = C.c_void_p
_U32 = C.c_uint32
_F = C.c_float


class OrcPhotonRender(C.Structure):
    _fields_ = [("xyz", _P), ("shape", _P), ("order", _P), ("axis", _P), ("n", _U32), ("rgb", _P), ("dir", _P),
                ("k", _U32), ("radius", _F), ("mode", _U32)]


class OrcRay(C.Structure):
    _fields_ = [("o", _F * 3), ("d", _F * 3), ("inv", _F * 3), ("tmax", _F)]


class OrcTriInfo(C.Structure):
    _fields_ = [("reached", C.c_int), ("t", _F), ("masked", C.c_int), ("fp64", C.c_int)]


class OrcSphere(C.Structure):
    _fields_ = [("o2w", _F * 16), ("w2o", _F * 16), ("radius", _F), ("center", _F * 3)]


_SIGS = {
    "orc_error_bound": (_F, [C.c_int]),
    "orc_next_float_up": (_F, [_F]),
    "orc_next_float_down": (_F, [_F]),
    "orc_ray_make": (None, [C.POINTER(OrcRay), _P, _P, _F]),
    "orc_render_photon": (None, [_P, _U32, _U32, _U32, _U32, _U32, C.c_uint64, _U32, _U32, _U32, _U32, _P, _P, C.c_int,
                                 C.POINTER(OrcPhotonRender)]),
    "orc_lbvh_build": (C.c_int, [_P, _U32, _P, _U32, _U32, _F, C.c_int, _P, _P, _P, _P]),
    "orc_tri_intersects": (C.c_int, [C.POINTER(OrcRay), _P, _P, _P, C.POINTER(_F), C.POINTER(OrcTriInfo)]),
    "orc_box_intersects": (C.c_int, [_P, _P, C.POINTER(OrcRay), C.POINTER(_F)]),
    "orc_refbvh_build": (_P, [_P, _P, _U32]),
    "orc_refbvh_free": (None, [_P]),
    "orc_refbvh_root": (_U32, [_P]),
    "orc_refbvh_sorted_ids": (None, [_P, _P]),
    "orc_refbvh_morton": (_U32, [_P, _U32]),
    "orc_refbvh_intersect": (C.c_int32, [_P, C.POINTER(OrcRay), _P, C.c_int, C.POINTER(_F), _P, _P]),
    "orc_refbvh_trace": (None, [_P] + [_P] * 7 + [_U32, _P, _P, C.c_int]),
    "orc_brute_trace": (None, [_P, _P, _U32, _P, C.c_int] + [_P] * 7 + [_U32, _P, _P, _P, C.c_int]),
    "orc_bvh2_trace": (None, [_P, _P] + [_P] * 7 + [_U32, _P, _P, _P, _P, C.c_int]),
    "orc_sphere_init": (None, [C.POINTER(OrcSphere), _P, _F]),
    "orc_sphere_intersects": (C.c_int, [C.POINTER(OrcSphere), C.POINTER(OrcRay), C.POINTER(_F)]),
    "orc_sphere_intersect": (C.c_int, [C.POINTER(OrcSphere), C.POINTER(OrcRay), C.POINTER(_F), _P, _P, _P]),
    "orc_mat4_inverse": (None, [_P, _P]),
    "orc_tri_barycentrics": (None, [C.POINTER(OrcRay), _P, _P, _P, _P]),
    "orc_camera_ray": (None, [_P, _P, _U32, _U32, _U32, _U32, _F, _F, _P, _P]),
    "orc_fresnel_dielectric": (_F, [_F, _F, _F]),
    "orc_path_initstate": (C.c_uint64, [_U32, _U32]),
    "orc_path_segments": (C.c_int, [_P, _U32, _U32, _U32, _U32, _U32, _U32, C.c_uint64, _P, _P, _P]),
    "orc_scene_create": (_P, [_P, _P, _U32, _P, _P, _P, _P, _U32, _P, _U32, _P, _U32, _P]),
    "orc_scene_free": (None, [_P]),
    "orc_scene_add_planes": (None, [_P, _P, _P, _U32]),
    "orc_scene_add_texture": (_U32, [_P, _P, _U32, _U32]),
    "orc_scene_set_tri_uv": (None, [_P, _P]),
    "orc_scene_set_tri_normals": (None, [_P, _P]),
    "orc_plane_init": (None, [_P, _P, _P, _P, _P]),
    "orc_plane_intersect": (C.c_int, [_P, C.POINTER(OrcRay), C.POINTER(_F), _P, _P, _P]),
    "orc_scene_camera": (None, [_P, _P, _P]),
    "orc_path": (C.c_int, [_P, _U32, _U32, _U32, _U32, _U32, _U32, C.c_uint64, _P]),
    "orc_bvh4_trace": (None, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _U32, _P, _P, C.POINTER(C.c_uint64),
                              C.POINTER(C.c_uint64), C.POINTER(_U32), C.c_int]),
    "orc_render": (None, [_P, _U32, _U32, _U32, _U32, _U32, C.c_uint64, _U32, _U32, _U32, _U32, _P, _P, C.c_int]),
    "orc_render_const_cache": (None, [_P, _U32, _U32, _U32, _U32, _U32, C.c_uint64, _U32, _U32, _U32, _U32, _P, _P,
                                      C.c_int, _U32, _P]),
    "orc_compute_rmse": (_F, [_P, C.c_double, _P, C.c_double, _U32]),
    "orc_cache_loss_grad": (C.c_double, [_P, _P, _P, _P, _P, _U32, C.c_int, _P, _P, _P, _P]),
    "orc_film_pixel": (None, [_P, _P, C.c_uint64, C.c_size_t, _P]),
    "orc_film_maxima": (None, [_P, _P, C.c_size_t, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "orc_brightness": (C.c_double, [_P]),
    "orc_tonemap_pixel": (None, [_P, C.c_double, C.c_double, C.c_uint64, C.c_int, _P]),
    "orc_to_byte": (C.c_uint8, [C.c_double]),
    "orc_image_pixel_bytes": (C.c_size_t, [C.c_int, _U32, _U32]),
    "orc_image_header": (C.c_size_t, [C.c_int, _U32, _U32, _P, C.c_size_t]),
    "orc_film_image": (None, [_P, _P, _U32, _U32, C.c_uint64, C.c_int, C.c_int, _P, _P]),
    "orc_kd_build": (_U32, [_P, _U32, _P, _P]),
    "orc_knn_batch": (None, [_P, _P, _P, _P, _U32, _P, _P, _U32, _U32, _F, C.c_int, _P, _P, _P, C.c_int]),
    "orc_photon_gather_lambert": (None, [_P, _P, _U32, _P, _P, _P, _P, _P, _P, _U32, _U32, _F, _P]),
    "orc_render_f64": (None, [_P, _U32, _U32, _U32, _U32, _U32, _U32, _U32, _U32, _P, _P, C.c_int]),
    "orc_photon_light_table": (None, [_P, _P, _P]),
    "orc_trace_photons": (_U32, [_P, _U32, _U32, _U32, _U32, _U32, _P, _P, _P, _P, _P, _P]),
    "orc_photon_gather_mat": (None, [_P, _P, _U32, _P, _P, _P, _U32, _P, _P, _P, _P]),
    "orc_photon_subsurface": (None, [_P, _P, _P, _P, _U32, _P, _P, _P, _P, _P, _P, _P, _U32, _F, _P, _P]),
    "orc_dipole_rd": (C.c_double, [C.c_double, C.c_double, _F, _F, _F]),
    "orc_heap_sequence": (None, [_P, _U32, _U32, _P, _P, _P]),
    "orc_concentric_disk": (None, [_F, _F, _P]),
    "orc_ggx_sample_half": (None, [_F, _P, _F, _F, _P]),
    "orc_ggx_D_n": (None, [_F, _P, _U32, _P]),
    "orc_ggx_lambda_n": (None, [_F, _P, _U32, _P]),
    "orc_fresnel_conductor": (_F, [_F, _F, _F]),
    "orc_bsdf_draws": (C.c_int, [_P]),
    "orc_bsdf_eval_n": (None, [_P, _P, _P, _U32, _P, _P]),
    "orc_bsdf_sample_n": (None, [_P, _P, _P, _U32, _P, _P, _P, _P]),
    "orc_dielectric_refractions": (_U32, [_U32, _U32]),
    "orc_luminance": (_F, [_P]),
    "orc_russian_roulette": (C.c_int, [_P, _F, _P]),
    "orc_sphere_light_sample": (C.c_int, [_P, _F, _P, _F, _F, _P, C.POINTER(_F), C.POINTER(_F)]),
    "orc_sphere_light_pdf": (_F, [_P, _F, _P]),
}


# ---- the shading functions one at a time (shading frame, z = normal) ----

def material(type, kd=(1, 1, 1), ior=1.5, roughness=0.0, eta=(1, 1, 1), k=(0, 0, 0), intensity=0.0):
    """An orc_material (same fields as the library's rt_material)."""
    return OrcMaterial(type, _f3(kd), ior, roughness, _f3(eta), _f3(k), intensity, _f3((0, 0, 0)), 0)


def concentric_disk(u0, u1):
    out = np.zeros(2, np.float32)
    lib().orc_concentric_disk(u0, u1, _p(out))
    return out


def ggx_sample_half(alpha, wo, u0, u1):
    w = f32(wo)
    h = np.zeros(3, np.float32)
    lib().orc_ggx_sample_half(alpha, _p(w), u0, u1, _p(h))
    return h


def ggx_D(alpha, h):
    h = f32(h).reshape(-1, 3)
    out = np.zeros(h.shape[0], np.float32)
    lib().orc_ggx_D_n(alpha, _p(h), h.shape[0], _p(out))
    return out


def ggx_lambda(alpha, w):
    w = f32(w).reshape(-1, 3)
    out = np.zeros(w.shape[0], np.float32)
    lib().orc_ggx_lambda_n(alpha, _p(w), w.shape[0], _p(out))
    return out


def fresnel_conductor(cosi, eta, k):
    return lib().orc_fresnel_conductor(cosi, eta, k)


def bsdf_draws(m):
    return lib().orc_bsdf_draws(C.byref(m))


def bsdf_eval(m, wo, wi):
    """(f * |cos wi| (n, 3), pdf (n,)) for direction pairs (n, 3)."""
    wi = f32(wi).reshape(-1, 3)
    wo = f32(np.broadcast_to(f32(wo).reshape(-1, 3), wi.shape))
    n = wi.shape[0]
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    lib().orc_bsdf_eval_n(C.byref(m), _p(wo), _p(wi), n, _p(f), _p(pdf))
    return f, pdf


def bsdf_sample(m, wo, u):
    """sampleDirection with explicit draws u (n, 2): (wi, pdf, delta, weight)."""
    u = f32(u).reshape(-1, 2)
    n = u.shape[0]
    wo = f32(np.broadcast_to(f32(wo).reshape(-1, 3), (n, 3)))
    wi = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    delta = np.zeros(n, np.int32)
    wgt = np.zeros((n, 3), np.float32)
    lib().orc_bsdf_sample_n(C.byref(m), _p(wo), _p(u), n, _p(wi), _p(pdf), _p(delta), _p(wgt))
    return wi, pdf, delta, wgt


def luminance(rgb):
    c = f32(rgb)
    return lib().orc_luminance(_p(c))


def russian_roulette(beta, u):
    """(survives, beta after the step)."""
    b = f32(beta)
    out = np.zeros(3, np.float32)
    alive = lib().orc_russian_roulette(_p(b), u, _p(out))
    return bool(alive), out


def sphere_light_sample(center, radius, p, u0, u1):
    """(ok, wi, distance to the near surface point, pdf per solid angle)."""
    c, q = f32(center), f32(p)
    wi = np.zeros(3, np.float32)
    dist, pdf = _F(), _F()
    ok = lib().orc_sphere_light_sample(_p(c), radius, _p(q), u0, u1, _p(wi), C.byref(dist), C.byref(pdf))
    return bool(ok), wi, dist.value, pdf.value


def sphere_light_pdf(center, radius, p):
    c, q = f32(center), f32(p)
    return lib().orc_sphere_light_pdf(_p(c), radius, _p(q))


def photon_gather_lambert(xyz, rgb, dirs, points, normals, kd, k, radius, shape=None, query_shape=None):
    """combinePhotonContributions on Lambertian surfaces (Scene.cu:607-634)."""
    x = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    c = np.ascontiguousarray(rgb, np.float64).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    q = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    nr = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    kd_ = np.ascontiguousarray(kd, np.float32).reshape(-1, 3)
    sh = None if shape is None else np.ascontiguousarray(shape, np.int32)
    qs = None if query_shape is None else np.ascontiguousarray(query_shape, np.int32)
    out = np.zeros((q.shape[0], 3), np.float64)
    lib().orc_photon_gather_lambert(_p(x), _p(sh), x.shape[0], _p(c), _p(d), _p(q), _p(nr), _p(kd_), _p(qs),
                                    q.shape[0], k, float(radius), _p(out))
    return out


def photon_gather_mat(rgb, dirs, count, index, dist, normals, views, materials):
    """combinePhotonContributions for any material over given search results
    (count (m,), index / dist (m, k)); materials: a sequence of material dicts
    (the scene description's form), one per query."""
    c = np.ascontiguousarray(rgb, np.float64).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    cnt = np.ascontiguousarray(count, np.uint32)
    idx = np.ascontiguousarray(index, np.uint32)
    ds = np.ascontiguousarray(dist, np.float32)
    m, k = idx.shape
    nr = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    vw = np.ascontiguousarray(views, np.float32).reshape(-1, 3)
    mats = (OrcMaterial * max(1, m))()
    for i, mt in enumerate(materials):
        mats[i] = _desc_material(mt)
    out = np.zeros((m, 3), np.float64)
    lib().orc_photon_gather_mat(_p(c), _p(d), m, _p(cnt), _p(idx), _p(ds), k, _p(nr), _p(vw), C.cast(mats, _P),
                                _p(out))
    return out


def photon_subsurface(xyz, tree, rgb, dirs, nrms, points, normals, views, params, radius=0.3, shape=None,
                      query_shape=None):
    """Scene::combinePhotonsSubsurface per query over the restated tree
    (order, axis); params = (kd[3], sigma_s[3], sigma_a[3], g, eta)."""
    x = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    order, axis = np.ascontiguousarray(tree[0], np.uint32), np.ascontiguousarray(tree[1], np.uint8)
    c = np.ascontiguousarray(rgb, np.float64).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    pn = np.ascontiguousarray(nrms, np.float32).reshape(-1, 3)
    q = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    qn = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    qv = np.ascontiguousarray(views, np.float32).reshape(-1, 3)
    kd, ss, sa, g, eta = params
    prm = np.array([*kd, *ss, *sa, g, eta], np.float32)
    sh = None if shape is None else np.ascontiguousarray(shape, np.int32)
    qs = None if query_shape is None else np.ascontiguousarray(query_shape, np.int32)
    out = np.zeros((q.shape[0], 3), np.float64)
    lib().orc_photon_subsurface(_p(x), _p(sh), _p(order), _p(axis), x.shape[0], _p(c), _p(d), _p(pn), _p(q), _p(qn),
                                _p(qv), _p(qs), q.shape[0], float(radius), _p(prm), _p(out))
    return out


def dipole_rd(sigma_s, sigma_a, g, eta, r):
    """impPaper_dipoleDiffusionAproximation of one channel at distance r."""
    return lib().orc_dipole_rd(float(sigma_s), float(sigma_a), float(g), float(eta), float(r))


def heap_sequence(dist, k):
    """nearest_neighbors_impl's vector after accepting candidates 0..n-1 with
    the given distances: (ids, distances) in vector order."""
    d = f32(dist)
    ids = np.zeros(max(1, min(d.size, k)), np.uint32)
    out = np.zeros(max(1, min(d.size, k)), np.float32)
    cnt = C.c_uint32()
    lib().orc_heap_sequence(_p(d), d.size, k, _p(ids), _p(out), C.byref(cnt))
    return ids[:cnt.value], out[:cnt.value]


def kd_build(xyz):
    """kdtree.h build_tree: (order, axis, depth)."""
    x = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    order = np.zeros(x.shape[0], np.uint32)
    axis = np.zeros(x.shape[0], np.uint8)
    depth = lib().orc_kd_build(_p(x), x.shape[0], _p(order), _p(axis))
    return order, axis, depth


def knn(xyz, queries, k, radius=np.inf, shape=None, query_shape=None, tree=None, nthreads=8):
    """k nearest photons per query: brute force (tree None) or the restated
    kd-tree search over tree = (order, axis).  Returns (count, index, distance)
    with index / distance of shape (m, k) (entries past count are undefined)."""
    x = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    q = np.ascontiguousarray(queries, np.float32).reshape(-1, 3)
    sh = None if shape is None else np.ascontiguousarray(shape, np.int32)
    qs = None if query_shape is None else np.ascontiguousarray(query_shape, np.int32)
    m = q.shape[0]
    cnt = np.zeros(m, np.uint32)
    idx = np.zeros((m, k), np.uint32)
    dist = np.zeros((m, k), np.float32)
    order, axis = (None, None) if tree is None else (np.ascontiguousarray(tree[0], np.uint32),
                                                     np.ascontiguousarray(tree[1], np.uint8))
    lib().orc_knn_batch(_p(x), _p(sh), _p(order), _p(axis), x.shape[0], _p(q), _p(qs), m, k, float(radius),
                        0 if tree is None else 1, _p(cnt), _p(idx), _p(dist), nthreads)
    return cnt, idx, dist

# ToneMappingType (camera/Film.cu:13-21) and the two image formats
TONE_MAPPINGS = ("equalize", "reinhard", "reinhardJodie", "ACESfilmic", "uncharted2", "gammaClamp")
IMAGE_FORMATS = ("bmp", "ppm")


def _tone(mode):
    return TONE_MAPPINGS.index(mode) if isinstance(mode, str) else int(mode)


def film_maxima(sums, samples, nvalid=None):
    """(max channel, max luminance) of Film.cu:120-154, each / samples."""
    s = np.ascontiguousarray(sums, np.float64).reshape(-1)
    nv = None if nvalid is None else np.ascontiguousarray(nvalid, np.uint32)
    m, ml = C.c_double(), C.c_double()
    lib().orc_film_maxima(_p(s), _p(nv), s.size // 3, int(samples), C.byref(m), C.byref(ml))
    return m.value, ml.value


def tonemap_pixel(rgb, max_ch, max_lum, samples, mode):
    """toneMaxPixel (Film.cu:266-283) on one pixel sum."""
    i = np.ascontiguousarray(rgb, np.float64)
    o = np.zeros(3, np.float64)
    lib().orc_tonemap_pixel(_p(i), max_ch, max_lum, int(samples), _tone(mode), _p(o))
    return o


def to_byte(v):
    return int(lib().orc_to_byte(float(v)))


def image_header(fmt, width, height):
    buf = np.zeros(64, np.uint8)
    n = lib().orc_image_header(IMAGE_FORMATS.index(fmt), width, height, _p(buf), 64)
    return buf[:n].tobytes()


def film_image(sums, width, height, samples, mode, fmt="bmp", nvalid=None):
    """Pixel array of Film::writeToBMP / writeToPPM (Film.cu:285-362) and the
    tone-mapped values before the byte conversion (H*W*3, film order)."""
    s = np.ascontiguousarray(sums, np.float64).reshape(-1)
    assert s.size == width * height * 3
    nv = None if nvalid is None else np.ascontiguousarray(nvalid, np.uint32).reshape(-1)
    f = IMAGE_FORMATS.index(fmt)
    px = np.zeros(lib().orc_image_pixel_bytes(f, width, height), np.uint8)
    mapped = np.zeros(width * height * 3, np.float64)
    lib().orc_film_image(_p(s), _p(nv), width, height, int(samples), _tone(mode), f, _p(px), _p(mapped))
    return px, mapped


class OrcCacheCfg(C.Structure):
    _fields_ = [("levels", _U32), ("table_size", _U32), ("res", _U32 * 16), ("dense", _U32 * 16),
                ("lo", _F * 3), ("inv_ext", _F * 3)]


CACHE_PARAMS = 17408


def cache_cfg(levels, log2_table, base_res, scale, lo, hi):
    """Level table exactly as rt_cache_create derives it."""
    import math
    c = OrcCacheCfg()
    c.levels = levels
    c.table_size = 1 << log2_table
    for l in range(levels):
        r = math.floor(base_res * math.pow(scale, l))
        c.res[l] = r
        c.dense[l] = 1 if (r + 1.0) ** 3 <= c.table_size else 0
    for k in range(3):
        c.lo[k] = lo[k]
        c.inv_ext[k] = np.float32(1.0) / (np.float32(hi[k]) - np.float32(lo[k]))
    return c


def cache_loss_grad(cfg, mlp, grid, inputs, targets, emulate_bf16=True, want_grad=True, den_fixed=None):
    """(loss, mlp_grad float64[17408], grid_grad float64[L*T*2], y float64[n,3])."""
    mlp = f32(mlp)
    grid = f32(grid)
    inputs = f32(inputs).reshape(-1, 12)
    targets = f32(targets).reshape(-1, 3)
    n = inputs.shape[0]
    mg = np.zeros(CACHE_PARAMS) if want_grad else None
    gg = np.zeros(cfg.levels * cfg.table_size * 2) if want_grad else None
    y = np.zeros((n, 3))
    den = None if den_fixed is None else np.ascontiguousarray(den_fixed, np.float64)
    loss = lib().orc_cache_loss_grad(C.byref(cfg), _p(mlp), _p(grid), _p(inputs), _p(targets), n, int(emulate_bf16),
                                     _p(mg), _p(gg), _p(y), _p(den))
    return loss, mg, gg, y


class OrcMaterial(C.Structure):
    _fields_ = [("type", C.c_uint32), ("kd", _F * 3), ("ior", _F), ("roughness", _F), ("eta", _F * 3),
                ("k", _F * 3), ("intensity", _F), ("ke", _F * 3), ("texture", _U32)]


class OrcPointLight(C.Structure):
    _fields_ = [("pos", _F * 3), ("color", _F * 3), ("intensity", _F)]


class OrcCamera(C.Structure):
    _fields_ = [("pos", _F * 3), ("lookat", _F * 3), ("up", _F * 3), ("cts", _F * 3), ("frame_y", _F),
                ("width", _U32), ("height", _U32)]


def _f3(v):
    return (_F * 3)(*[float(x) for x in v])


def _desc_material(m):
    """An orc_material from a scene description's material dict."""
    return OrcMaterial(m["type"], _f3(m.get("kd", (0, 0, 0))), m.get("ior", 1.5), m.get("roughness", 0.0),
                       _f3(m.get("eta", (1, 1, 1))), _f3(m.get("k", (0, 0, 0))), m.get("intensity", 0.0),
                       _f3(m.get("ke", (0, 0, 0))), int(m.get("texture", 0)))


class OrcScene:
    """The oracle's copy of a scene description (same arrays the library
    gets); renders with the restated estimator on the host."""

    def __init__(self, desc):
        self.desc = desc
        self.xyz = f32(desc.xyz).reshape(-1, 3)
        self.idx = np.ascontiguousarray(desc.idx, np.uint32)
        self.tri_mat = np.ascontiguousarray(desc.tri_mat, np.uint32)
        sph = np.array([[*c, r] for c, r, m, l in desc.spheres], np.float32).reshape(-1, 4)
        self.sph = np.ascontiguousarray(sph)
        self.sph_mat = np.array([m for c, r, m, l in desc.spheres], np.uint32)
        self.sph_light = np.array([l for c, r, m, l in desc.spheres], np.uint8)
        mats = (OrcMaterial * max(1, len(desc.materials)))()
        for i, m in enumerate(desc.materials):
            mats[i] = _desc_material(m)
        self.mats = mats
        lights = (OrcPointLight * max(1, len(desc.lights)))()
        for i, (p, c, inten) in enumerate(desc.lights):
            lights[i] = OrcPointLight(_f3(p), _f3(c), inten)
        self.lights = lights
        c = desc.camera
        self.cam = OrcCamera(_f3(c["pos"]), _f3(c["lookat"]), _f3(c["up"]), _f3(c["cts"]), c["frame_y"],
                             c["width"], c["height"])
        self.h = lib().orc_scene_create(_p(self.xyz), _p(self.idx), self.idx.shape[0], _p(self.tri_mat),
                                        _p(self.sph), _p(self.sph_mat), _p(self.sph_light), self.sph.shape[0],
                                        C.cast(self.mats, _P), len(desc.materials), C.cast(self.lights, _P),
                                        len(desc.lights), C.byref(self.cam))
        self.tex = []
        for rgb in getattr(desc, "textures", []):
            t = np.ascontiguousarray(rgb, np.uint8)
            self.tex.append(t)
            lib().orc_scene_add_texture(self.h, _p(t), t.shape[1], t.shape[0])
        uv = getattr(desc, "tri_uv", None)
        if uv is not None and self.idx.shape[0]:
            self.tri_uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 6)
            lib().orc_scene_set_tri_uv(self.h, _p(self.tri_uv))
        nrm9 = getattr(desc, "tri_nrm", None)
        if nrm9 is not None and self.idx.shape[0]:
            self.tri_nrm = np.ascontiguousarray(nrm9, np.float32).reshape(-1, 9)
            lib().orc_scene_set_tri_normals(self.h, _p(self.tri_nrm))
        planes = getattr(desc, "planes", [])
        if planes:
            self.planes = np.array([[*c, *ax, *ay, *n] for c, ax, ay, n, m in planes], np.float32)
            self.plane_mat = np.array([m for c, ax, ay, n, m in planes], np.uint32)
            lib().orc_scene_add_planes(self.h, _p(self.planes), _p(self.plane_mat), len(planes))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_free(self.h)
            self.h = None

    def camera(self):
        r2c = np.zeros(16, np.float32)
        c2w = np.zeros(16, np.float32)
        lib().orc_scene_camera(self.h, _p(r2c), _p(c2w))
        return r2c, c2w

    def camera_ray(self, px, py, u0, u1):
        r2c, c2w = self.camera()
        o = np.zeros(3, np.float32)
        d = np.zeros(3, np.float32)
        lib().orc_camera_ray(_p(r2c), _p(c2w), self.cam.width, self.cam.height, px, py, u0, u1, _p(o), _p(d))
        return o, d

    def path(self, px, py, sample, seed=None, max_depth=None, rr_depth=None, stream_base=0):
        d = self.desc
        out = np.zeros(3, np.float32)
        ok = lib().orc_path(self.h, px, py, sample, d.seed if seed is None else seed, max_depth or d.max_depth,
                            d.rr_depth if rr_depth is None else rr_depth, stream_base, _p(out))
        return bool(ok), out

    def path_segments(self, px, py, sample, seed=None, max_depth=None, rr_depth=None, stream_base=0):
        """(ok, L, rays (nseg, 6)): the path and the ray of each of its segments."""
        d = self.desc
        md = max_depth or d.max_depth
        out = np.zeros(3, np.float32)
        seg = np.zeros((md, 6), np.float32)
        ns = C.c_uint32()
        ok = lib().orc_path_segments(self.h, px, py, sample, d.seed if seed is None else seed, md,
                                     d.rr_depth if rr_depth is None else rr_depth, stream_base, _p(out), _p(seg),
                                     C.byref(ns))
        return bool(ok), out, seg[:ns.value]

    def render_f64(self, spp=None, window=None, seed=None, nthreads=8, max_depth=None):
        """orc_render with the reference's double-precision Spectrum for
        throughput, radiance and light terms (orc_render_f64)."""
        d = self.desc
        W, H = d.width, d.height
        x0, y0, x1, y1 = window or (0, 0, W, H)
        accum = np.zeros(W * H * 3, np.float64)
        nvalid = np.zeros(W * H, np.uint32)
        lib().orc_render_f64(self.h, spp if spp is not None else d.spp, d.seed if seed is None else seed,
                             max_depth or d.max_depth, d.rr_depth, x0, y0, x1, y1, _p(accum), _p(nvalid), nthreads)
        return accum, nvalid

    def photon_light_table(self):
        n = len(self.desc.lights)
        p = np.zeros(max(1, n), np.float64)
        cdf = np.zeros(max(1, n), np.float64)
        lib().orc_photon_light_table(self.h, _p(p), _p(cdf))
        return p[:n], cdf[:n]

    def trace_photons(self, photons, seed=1, max_depth=30, rr_depth=3, cap=None):
        """The photon pre-pass: dict of location, direction, normal (n, 3
        float32), radiance (n, 3 float64), shape (n,) and key (path << 5 |
        bounce), and the number the pass produced (may exceed cap)."""
        cap = cap if cap is not None else photons * 4
        out = dict(location=np.zeros((cap, 3), np.float32), direction=np.zeros((cap, 3), np.float32),
                   normal=np.zeros((cap, 3), np.float32), radiance=np.zeros((cap, 3), np.float64),
                   shape=np.zeros(cap, np.int32), key=np.zeros(cap, np.uint32))
        m = lib().orc_trace_photons(self.h, photons, seed, max_depth, rr_depth, cap, _p(out["location"]),
                                    _p(out["direction"]), _p(out["normal"]), _p(out["radiance"]), _p(out["shape"]),
                                    _p(out["key"]))
        n = min(m, cap)
        return {k: v[:n] for k, v in out.items()}, m

    def render(self, spp=None, spp_offset=0, window=None, seed=None, nthreads=8, stream_base=0,
               max_depth=None, cache_depth=None, cache_rgb=None):
        """Returns (accum (H*W*3 float64), nvalid (H*W uint32)) for the window
        (x0, y0, x1, y1) (default: whole image).  With cache_rgb, paths end
        at their first cacheable vertex at depth >= cache_depth and take
        beta * cache_rgb (orc_render_const_cache)."""
        d = self.desc
        W, H = d.width, d.height
        x0, y0, x1, y1 = window or (0, 0, W, H)
        accum = np.zeros(W * H * 3, np.float64)
        nvalid = np.zeros(W * H, np.uint32)
        args = (self.h, spp if spp is not None else d.spp, d.seed if seed is None else seed,
                max_depth or d.max_depth, d.rr_depth, spp_offset, stream_base, x0, y0, x1, y1,
                _p(accum), _p(nvalid), nthreads)
        if cache_rgb is None:
            lib().orc_render(*args)
        else:
            c = np.ascontiguousarray(cache_rgb, np.float32)
            lib().orc_render_const_cache(*args, cache_depth, _p(c))
        return accum, nvalid

    def render_photon(self, photons, k=100, radius=100.0, spp=None, spp_offset=0, window=None, seed=None,
                      nthreads=8, stream_base=0, max_depth=None, hybrid=False):
        """orc_render_photon (RENDER_MODE PHOTON_MAPPING, or HYBRID) over photons =
        {location, shape, radiance, direction} (numpy, as trace_photons
        returns them); the kd-tree is built here.  Returns (accum, nvalid)
        like render."""
        d = self.desc
        W, H = d.width, d.height
        x0, y0, x1, y1 = window or (0, 0, W, H)
        loc = np.ascontiguousarray(photons["location"], np.float32).reshape(-1, 3)
        shp = np.ascontiguousarray(photons["shape"], np.int32).reshape(-1)
        rgb = np.ascontiguousarray(photons["radiance"], np.float64).reshape(-1, 3)
        dr = np.ascontiguousarray(photons["direction"], np.float32).reshape(-1, 3)
        order, axis, _ = kd_build(loc)
        ph = OrcPhotonRender(_p(loc), _p(shp), _p(order), _p(axis), loc.shape[0], _p(rgb), _p(dr), k, radius,
                             2 if hybrid else 1)
        accum = np.zeros(W * H * 3, np.float64)
        nvalid = np.zeros(W * H, np.uint32)
        lib().orc_render_photon(self.h, spp if spp is not None else d.spp, d.seed if seed is None else seed,
                                max_depth or d.max_depth, d.rr_depth, spp_offset, stream_base, x0, y0, x1, y1,
                                _p(accum), _p(nvalid), nthreads, C.byref(ph))
        return accum, nvalid

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            try:
                f = getattr(L, name)
            except AttributeError:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be contiguous"
    return a.ctypes.data


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def make_ray(o, d, tmax=np.inf) -> OrcRay:
    r = OrcRay()
    o = f32(o)
    d = f32(d)
    lib().orc_ray_make(C.byref(r), _p(o), _p(d), float(tmax))
    return r


def tri_intersects(o, d, p0, p1, p2, tmax=np.inf, hit_offset=np.inf):
    """One Triangle::intersects call; returns (accepted, new_hitOffset, info)."""
    r = make_ray(o, d, tmax)
    h = _F(hit_offset)
    info = OrcTriInfo()
    p0, p1, p2 = f32(p0), f32(p1), f32(p2)  # keep the arrays alive across the call
    ok = lib().orc_tri_intersects(C.byref(r), _p(p0), _p(p1), _p(p2), C.byref(h), C.byref(info))
    return bool(ok), h.value, info


class RefBvh:
    """The reference's Morton-pairing BVH restated."""

    def __init__(self, xyz, idx):
        self.xyz = f32(xyz).reshape(-1, 3)
        self.idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1, 3)
        self.h = lib().orc_refbvh_build(_p(self.xyz), _p(self.idx), self.idx.shape[0])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_refbvh_free(self.h)
            self.h = None

    def sorted_ids(self):
        out = np.empty(self.idx.shape[0], np.uint32)
        lib().orc_refbvh_sorted_ids(self.h, _p(out))
        return out

    def morton(self, i):
        return lib().orc_refbvh_morton(self.h, i)

    def trace(self, rays, tmax=None, nthreads=8):
        rays = [f32(a) for a in rays]
        n = rays[0].shape[0]
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        lib().orc_refbvh_trace(self.h, *[_p(a) for a in rays], _p(tmax), n, _p(prim), _p(t), nthreads)
        return prim, t


def lbvh_build(xyz, idx, max_leaf=4, wide=True, trav_cost=1.2):
    """The GPU-built BVH restated (lbvh.c): returns (info dict, nodes2
    (m, 16) float32, tris (nt + 1, 12) float32, nodes4 (m4, 16) float32 or
    None), in the library's export layouts."""
    x = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    ix = np.ascontiguousarray(idx, np.uint32).reshape(-1, 3)
    nt = ix.shape[0]
    cap = max(1, nt - 1)
    nodes2 = np.zeros((cap, 16), np.float32)
    tris = np.zeros((nt + 1, 12), np.float32)
    nodes4 = np.zeros((cap, 16), np.float32)
    info = np.zeros(6, np.uint32)
    if lib().orc_lbvh_build(_p(x), x.shape[0], _p(ix), nt, max_leaf, trav_cost, int(bool(wide)), _p(nodes2),
                            _p(tris), _p(nodes4), _p(info)) != 0:
        raise ValueError("orc_lbvh_build: bad input")
    keys = ("num_nodes", "num_leaves", "max_depth", "num_nodes4", "depth4", "max_stack4")
    inf = {k: int(v) for k, v in zip(keys, info)}
    return inf, nodes2[:inf["num_nodes"]], tris, (nodes4[:inf["num_nodes4"]] if wide else None)


def brute_trace(xyz, idx, rays, tmax=None, transparent=None, check_shadow=False, nthreads=8):
    """Brute-force closest hit + ambiguity flags (bit0 tie, bit1 deltaT mask)."""
    xyz = f32(xyz).reshape(-1, 3)
    idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1, 3)
    rays = [f32(a) for a in rays]
    n = rays[0].shape[0]
    prim = np.empty(n, np.int32)
    t = np.empty(n, np.float32)
    flags = np.empty(n, np.uint8)
    tr = None if transparent is None else np.ascontiguousarray(transparent, dtype=np.uint8)
    lib().orc_brute_trace(_p(xyz), _p(idx), idx.shape[0], _p(tr), int(check_shadow), *[_p(a) for a in rays],
                          _p(tmax), n, _p(prim), _p(t), _p(flags), nthreads)
    return prim, t, flags


def bvh2_trace(nodes, tris, rays, tmax=None, nthreads=8):
    """Counting traversal of the library's exported BVH2 layout."""
    rays = [f32(a) for a in rays]
    n = rays[0].shape[0]
    prim = np.empty(n, np.int32)
    t = np.empty(n, np.float32)
    nn = C.c_uint64()
    nt = C.c_uint64()
    nodes, tris = f32(nodes), f32(tris)
    lib().orc_bvh2_trace(_p(nodes), _p(tris), *[_p(a) for a in rays], _p(tmax), n, _p(prim), _p(t),
                         C.byref(nn), C.byref(nt), nthreads)
    return prim, t, nn.value, nt.value


def bvh4_trace(nodes4, tris, rays, tmax=None, nthreads=8):
    """Counting traversal of the library's exported four-wide layout:
    (prim, t, node visits, triangle tests, deepest stack)."""
    rays = [f32(a) for a in rays]
    n = rays[0].shape[0]
    prim = np.empty(n, np.int32)
    t = np.empty(n, np.float32)
    nn = C.c_uint64()
    nt = C.c_uint64()
    msp = C.c_uint32()
    nodes4, tris = f32(nodes4), f32(tris)
    lib().orc_bvh4_trace(_p(nodes4), _p(tris), *[_p(a) for a in rays], _p(tmax), n, _p(prim), _p(t),
                         C.byref(nn), C.byref(nt), C.byref(msp), nthreads)
    return prim, t, nn.value, nt.value, msp.value


class Plane:
    """Plane(corner, axis_x, axis_y, normal) restated (shapes/Plane.cu)."""

    def __init__(self, corner, axis_x, axis_y, normal):
        self.buf = np.zeros(32, np.float32)
        args = [f32(v) for v in (corner, axis_x, axis_y, normal)]
        lib().orc_plane_init(_p(self.buf), *[_p(a) for a in args])

    @property
    def o2w(self):
        return self.buf[:16].reshape(4, 4)

    @property
    def w2o(self):
        return self.buf[16:].reshape(4, 4)

    def intersect(self, o, d, tmax=np.inf, hit_offset=np.inf):
        """(hit, offset, point, normal, perr)."""
        r = make_ray(o, d, tmax)
        h = _F(hit_offset)
        p, n, e = np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(3, np.float32)
        ok = lib().orc_plane_intersect(_p(self.buf), C.byref(r), C.byref(h), _p(p), _p(n), _p(e))
        return bool(ok), h.value, p, n, e


class Sphere:
    def __init__(self, center, radius):
        self.s = OrcSphere()
        c = f32(center)
        lib().orc_sphere_init(C.byref(self.s), _p(c), float(radius))

    def intersects(self, o, d, tmax=np.inf, hit_offset=np.inf):
        r = make_ray(o, d, tmax)
        h = _F(hit_offset)
        ok = lib().orc_sphere_intersects(C.byref(self.s), C.byref(r), C.byref(h))
        return bool(ok), h.value

    def intersect(self, o, d, tmax=np.inf, hit_offset=np.inf):
        r = make_ray(o, d, tmax)
        h = _F(hit_offset)
        p = np.zeros(3, np.float32)
        n = np.zeros(3, np.float32)
        e = np.zeros(3, np.float32)
        ok = lib().orc_sphere_intersect(C.byref(self.s), C.byref(r), C.byref(h), _p(p), _p(n), _p(e))
        return bool(ok), h.value, p, n, e
