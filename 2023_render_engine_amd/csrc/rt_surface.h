// rt_surface.h -- the surface a ray hits (SurfaceInteraction as the
// renderer's kernels need it): hit point, shading normal, error bounds,
// material and texture coordinates of a triangle, sphere or plane hit, and
// the texture lookup.  Shared by the path tracer (render.hip) and the photon
// tracer (photon_trace.hip).
#pragma once

#include "rt_device.h"
#include "rt_render.h"
#include "rt_shade.h"

namespace rt {

struct Surface {
    V3 p, n, perr;
    uint32_t mat;
    float su, sv;  // SurfaceInteraction::surfacePoint, filled when the scene has textures
};

// Texture::getColor (textures/Texture.cu:172-183): texel
// ((unsigned)(u * width), (unsigned)(v * height)) clamped to the last
// row / column, channels / 255.  The float-to-unsigned conversion is taken
// as the reference's x86-64 host build performs it (truncation to a 64-bit
// integer, low 32 bits; NaN to 0), so a coordinate in (-1, 0) reads texel 0
// and one below -1 wraps to the last.
__device__ __forceinline__ uint32_t texel_index(float u, uint32_t n)
{
    float f = u * (float)n;
    if (f != f) f = 0.0f;
    f = fminf(fmaxf(f, -4.6e18f), 4.6e18f);
    const uint32_t x = (uint32_t)(unsigned long long)(long long)f;
    return x < n - 1 ? x : n - 1;
}

__device__ __forceinline__ V3 texture_color(const SceneDev &sc, uint32_t t, float u, float v)
{
    const TexDev &tx = sc.tex[t];
    const uint32_t x = texel_index(u, tx.width), y = texel_index(v, tx.height);
    const uint32_t c = sc.texels[tx.offset + (size_t)y * tx.width + x];
    return mk((float)(c & 0xffu) / 255.0f, (float)((c >> 8) & 0xffu) / 255.0f, (float)((c >> 16) & 0xffu) / 255.0f);
}

// A textured Lambertian's kd, or a textured plain material's camera-ray kd,
// times the texel at the hit (LambertianBSDF.cu:22-74, PlainMaterial.cu:62-75).
__device__ __forceinline__ void apply_texture(const SceneDev &sc, const Surface &s, rt_material &m)
{
    if (m.texture == 0 || !(m.type == RT_MAT_LAMBERTIAN || m.type == RT_MAT_PLAIN)) return;
    const V3 c = texture_color(sc, m.texture - 1, s.su, s.sv);
    m.kd[0] = m.kd[0] * c.x;
    m.kd[1] = m.kd[1] * c.y;
    m.kd[2] = m.kd[2] * c.z;
}

__device__ __forceinline__ void triangle_surface(const SceneDev &sc, int32_t id, V3 o, V3 d, Surface &s)
{
    const float *t = sc.tri_xyz + 9 * (size_t)id;
    const float4 a = make_float4(t[0], t[1], t[2], 0), b = make_float4(t[3], t[4], t[5], 0),
                 c = make_float4(t[6], t[7], t[8], 0);
    // barycentrics of the watertight test (Triangle.cu:161-290)
    const TriRay tr = make_tri_ray(o.x, o.y, o.z, d.x, d.y, d.z, INFINITY);
    float p0x = sel3(tr.kx, a.x, a.y, a.z) - tr.ox, p0y = sel3(tr.ky, a.x, a.y, a.z) - tr.oy,
          p0z = sel3(tr.kz, a.x, a.y, a.z) - tr.oz;
    float p1x = sel3(tr.kx, b.x, b.y, b.z) - tr.ox, p1y = sel3(tr.ky, b.x, b.y, b.z) - tr.oy,
          p1z = sel3(tr.kz, b.x, b.y, b.z) - tr.oz;
    float p2x = sel3(tr.kx, c.x, c.y, c.z) - tr.ox, p2y = sel3(tr.ky, c.x, c.y, c.z) - tr.oy,
          p2z = sel3(tr.kz, c.x, c.y, c.z) - tr.oz;
    p0x += tr.Sx * p0z; p0y += tr.Sy * p0z;
    p1x += tr.Sx * p1z; p1y += tr.Sy * p1z;
    p2x += tr.Sx * p2z; p2y += tr.Sy * p2z;
    float e0 = p1x * p2y - p1y * p2x, e1 = p2x * p0y - p2y * p0x, e2 = p0x * p1y - p0y * p1x;
    if (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f) {
        e0 = (float)((double)p2y * (double)p1x - (double)p2x * (double)p1y);
        e1 = (float)((double)p0y * (double)p2x - (double)p0x * (double)p2y);
        e2 = (float)((double)p1y * (double)p0x - (double)p1x * (double)p0y);
    }
    const float det = e0 + e1 + e2;
    const float invDet = 1 / det;
    const float b0 = e0 * invDet, b1 = e1 * invDet, b2 = e2 * invDet;
    const V3 P0 = mk(t[0], t[1], t[2]), P1 = mk(t[3], t[4], t[5]), P2 = mk(t[6], t[7], t[8]);
    s.p = add(add(mul(P0, b0), mul(P1, b1)), mul(P2, b2));
    const float xs = fabsf(b0 * P0.x) + fabsf(b1 * P1.x) + fabsf(b2 * P2.x);
    const float ys = fabsf(b0 * P0.y) + fabsf(b1 * P1.y) + fabsf(b2 * P2.y);
    const float zs = fabsf(b0 * P0.z) + fabsf(b1 * P1.z) + fabsf(b2 * P2.z);
    s.perr = mul(mk(xs, ys, zs), error_bound(7));
    V3 n = nrm(cross(sub(P0, P2), sub(P1, P2)));
    if (dot(n, d) > 0) n = neg(n);  // normal faces the ray (Triangle.cu:368-373)
    s.n = n;
    if (sc.tri_nrm) {
        // a mesh with vertex normals (Triangle.cu:379-418): the shading normal
        // is the interpolated one, not turned towards the ray, and the hit
        // point moves 0.15 along it
        const float *q = sc.tri_nrm + 9 * (size_t)id;
        if (q[0] == q[0]) {
            const V3 ns = nrm(add(add(mul(mk(q[0], q[1], q[2]), b0), mul(mk(q[3], q[4], q[5]), b1)),
                                  mul(mk(q[6], q[7], q[8]), b2)));
            s.n = ns;
            s.p = add(s.p, mul(ns, 0.15f));
        }
    }
    s.mat = sc.tri_mat[id];
    if (sc.ntex) {
        // surfacePoint = b0 uv0 + b1 uv1 + b2 uv2 (Triangle.cu:356-358), the
        // default corners (0,0), (1,0), (1,1) without texture coordinates
        // (:134-152)
        float u0 = 0, v0 = 0, u1 = 1, v1 = 0, u2 = 1, v2 = 1;
        if (sc.tri_uv) {
            const float *q = sc.tri_uv + 6 * (size_t)id;
            u0 = q[0]; v0 = q[1]; u1 = q[2]; v1 = q[3]; u2 = q[4]; v2 = q[5];
        }
        s.su = b0 * u0 + b1 * u1 + b2 * u2;
        s.sv = b0 * v0 + b1 * v1 + b2 * v2;
    }
}

// The hit surface of a path from its kPqRay and kPqHit quads; d = the ray
// direction.
__device__ __forceinline__ void surface_of(const SceneDev &sc, float4 qr, float4 qh, Surface &s, V3 &d)
{
    const V3 o = mk(qr.x, qr.y, qr.z);
    d = mk(qr.w, qh.x, qh.y);
    const int32_t hid = (int32_t)__float_as_uint(qh.w);
    const uint32_t k = (uint32_t)(-hid - 2);
    if (hid >= 0) {
        triangle_surface(sc, hid, o, d, s);
    } else if (k < sc.nsph) {
        const SphereDev &sp = sc.sph[k];
        V3 so, sd;
        const float ts = sphere_core(sp, o, d, INFINITY, so, sd);
        sphere_surface(sp, so, sd, ts, s.p, s.n, s.perr);
        s.mat = sp.material;
        if (sc.ntex) {
            // Sphere::localToSurface (Sphere.cu:269-275): (theta, phi) of the
            // object-space hit point, in degrees (geometricMath.cu:49-57)
            V3 hp = add(so, mul(sd, ts));
            const float dist = sqrtf(hp.x * hp.x + hp.y * hp.y + hp.z * hp.z);
            const float scl = sp.radius / dist;
            hp = mk(hp.x * scl, hp.y * scl, hp.z * scl);
            if (hp.x == 0 && hp.y == 0) hp.x = 1e-5f * sp.radius;
            const float th = atan2f(sqrtf(hp.x * hp.x + hp.y * hp.y), hp.z);
            const float ph = atan2f(hp.y, hp.x);
            s.su = (float)((double)th * 180.0 / 3.14159265358979323846);
            s.sv = (float)((double)ph * 180.0 / 3.14159265358979323846);
        }
    } else {
        const PlaneDev &pl = sc.plane[k - sc.nsph];
        float u, v, dz;
        plane_core(pl, o, d, INFINITY, u, v, dz);
        plane_surface(pl, u, v, dz, s.p, s.n, s.perr);
        s.mat = pl.material;
        s.su = u;
        s.sv = v;
    }
}

}  // namespace rt
