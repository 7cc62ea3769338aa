// rt_shade.h -- device-side shading of the wavefront path tracer: camera
// rays, sampling warps, the reference's BSDFs (Lambertian, dielectric with
// its literal Fresnel, Trowbridge-Reitz GGX), sphere intersection with the
// reference's error-bounded quadratic, and the sphere area light.
//
// FP32 throughout (compiled with -ffp-contract=off); operation order follows
// the reference files cited per function.
#pragma once

#include <hip/hip_runtime.h>

#include "rt_device.h"
#include "rt_pcg.h"
#include "rt_scene_types.h"

namespace rt {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 1.0f / kPi;

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 mk(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 mul(V3 a, float s) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ V3 hmul(V3 a, V3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ V3 neg(V3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
__device__ __forceinline__ V3 cross(V3 a, V3 b)
{
    return mk((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
__device__ __forceinline__ float len(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 nrm(V3 a)
{
    float inv = 1.0f / len(a);
    return mk(a.x * inv, a.y * inv, a.z * inv);
}
__device__ __forceinline__ V3 vabs(V3 a) { return mk(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
__device__ __forceinline__ bool black(V3 a) { return a.x == 0 && a.y == 0 && a.z == 0; }

// ---------------------------------------------------------------------
// next float up/down (cuda_libraries/math.cu:147-194)
// ---------------------------------------------------------------------
__device__ __forceinline__ float next_up(float x)
{
    if (isinf(x) && x > 0.0f) return x;
    if (x == -0.0f) x = 0.0f;
    uint32_t u = __float_as_uint(x);
    u = (x >= 0.0f) ? u + 1 : u - 1;
    return __uint_as_float(u);
}
__device__ __forceinline__ float next_down(float x)
{
    if (isinf(x) && x < 0.0f) return x;
    if (x == 0.0f) x = -0.0f;
    uint32_t u = __float_as_uint(x);
    u = (x > 0.0f) ? u - 1 : u + 1;
    return __uint_as_float(u);
}

// ---------------------------------------------------------------------
// Camera (camera/ProjectiveCamera.cu:71-104, Transform.cu:216-237)
// ---------------------------------------------------------------------
__device__ __forceinline__ V3 xform_point(const float *m, V3 p)
{
    float xp = m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3];
    float yp = m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7];
    float zp = m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11];
    float wp = m[12] * p.x + m[13] * p.y + m[14] * p.z + m[15];
    float inv = 1.0f / wp;
    return mk(inv * xp, inv * yp, inv * zp);
}
__device__ __forceinline__ V3 xform_vec(const float *m, V3 v)
{
    return mk(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
              m[8] * v.x + m[9] * v.y + m[10] * v.z);
}

__device__ __forceinline__ void camera_ray(const CameraDev &c, uint32_t px, uint32_t py, float u0, float u1, V3 &o,
                                           V3 &d)
{
    float fx = (float)px + u0, fy = (float)py + u1;
    fx = fminf(fmaxf(fx, 0.0f), (float)c.width);
    fy = fminf(fmaxf(fy, 0.0f), (float)c.height);
    V3 pc = xform_point(c.r2c, mk(fx, fy, 0));
    V3 dir = nrm(pc);
    const float eb3 = error_bound(3);
    V3 oerr = mul(mk(fabsf(c.c2w[3]), fabsf(c.c2w[7]), fabsf(c.c2w[11])), eb3);
    V3 org = xform_point(c.c2w, mk(0, 0, 0));
    V3 wd = xform_vec(c.c2w, dir);
    float ls = dot(wd, wd);
    if (ls > 0) {
        float dt = dot(vabs(wd), oerr) / ls;
        org = add(org, mul(wd, dt));
    }
    o = org;
    d = wd;
}

// ---------------------------------------------------------------------
// Sampling warps (sampling/Sampler.cu:13-96)
// ---------------------------------------------------------------------
__device__ __forceinline__ V3 cosine_hemisphere(float ux, float uy)
{
    float ox = 2.0f * ux - 1.0f, oy = 2.0f * uy - 1.0f;
    float x = 0, y = 0;
    if (!(ox == 0 && oy == 0)) {
        float theta, r;
        if (fabsf(ox) > fabsf(oy)) {
            r = ox;
            theta = (kPi / 4) * (oy / ox);
        } else {
            r = oy;
            theta = (kPi / 2) - (kPi / 4) * (ox / oy);
        }
        x = r * cosf(theta);
        y = r * sinf(theta);
    }
    float z = sqrtf(fmaxf(0.0f, 1 - x * x - y * y));
    return mk(x, y, z);
}

// Orthonormal basis about n (Duff et al., JCGT 2017).  The reference's
// frame (SurfaceInteraction.cu:54-107) differs only by a rotation about n,
// which no isotropic BSDF can see.
__device__ __forceinline__ void onb(V3 n, V3 &t, V3 &b)
{
    float s = copysignf(1.0f, n.z);
    float a = -1.0f / (s + n.z);
    float c = n.x * n.y * a;
    t = mk(1.0f + s * n.x * n.x * a, s * c, -s * n.x);
    b = mk(c, s + n.y * n.y * a, -n.y);
}

// ---------------------------------------------------------------------
// BSDFs
// ---------------------------------------------------------------------
__device__ __forceinline__ float clampf(float x, float a, float b) { return fmaxf(a, fminf(b, x)); }
__device__ __forceinline__ float cos2th(V3 w) { return w.z * w.z; }
__device__ __forceinline__ float sin2th(V3 w) { return fmaxf(0.0f, 1 - cos2th(w)); }
__device__ __forceinline__ float sinth(V3 w) { return sqrtf(sin2th(w)); }
__device__ __forceinline__ float cosphi(V3 w) { float s = sinth(w); return (s == 0) ? 1 : clampf(w.x / s, -1, 1); }
__device__ __forceinline__ float sinphi(V3 w) { float s = sinth(w); return (s == 0) ? 0 : clampf(w.y / s, -1, 1); }
__device__ __forceinline__ float tan2th(V3 w) { return sin2th(w) / cos2th(w); }
__device__ __forceinline__ float abscos(V3 w) { return fabsf(w.z / len(w)); }
__device__ __forceinline__ float abscos2(V3 a, V3 b) { return fabsf(dot(a, b)) / (len(a) * len(b)); }
__device__ __forceinline__ float cosine2(V3 a, V3 b) { return dot(a, b) / (len(a) * len(b)); }

// TrowbridgeBSDF::distribution (materials/TrowbridgeBSDF.cu:71-84)
__device__ __forceinline__ float ggx_D(float ax, float ay, V3 h)
{
    float t2 = tan2th(h);
    if (isinf(t2)) return 0;
    float c4 = cos2th(h) * cos2th(h);
    float cp = cosphi(h) / ax, sp = sinphi(h) / ay;
    float e = t2 * (cp * cp + sp * sp);
    float q = 1 + e;
    return 1 / (kPi * ax * ay * c4 * (q * q));
}
// lambda (TrowbridgeBSDF.cu:101-113)
__device__ __forceinline__ float ggx_lambda(float ax, float ay, V3 w)
{
    float t2 = tan2th(w);
    if (t2 == INFINITY) return 0;
    float a = cosphi(w) * ax, b = sinphi(w) * ay;
    float a2 = a * a + b * b;
    return (sqrtf(1 + a2 * t2) - 1) / 2;
}
// visible-normal distribution (TrowbridgeBSDF.cu:95-99)
__device__ __forceinline__ float ggx_Dv(float ax, float ay, V3 wo, V3 h)
{
    float g1 = 1 / (1 + ggx_lambda(ax, ay, wo));
    return g1 / abscos(wo) * ggx_D(ax, ay, h) * abscos2(wo, h);
}
// sampleHalf with the reference's uniformSampleDisk(2u - 1) (TrowbridgeBSDF.cu:120-151)
__device__ __forceinline__ V3 ggx_sample_half(float ax, float ay, V3 wo, float u0, float u1)
{
    V3 half = nrm(mk(ax * wo.x, ay * wo.y, wo.z));
    if (half.z < 0) half = neg(half);
    V3 T1 = (half.z < 0.99999f) ? nrm(cross(mk(0, 0, 1), half)) : mk(1, 0, 0);
    V3 T2 = cross(half, T1);
    float sx = 2 * u0 - 1, sy = 2 * u1 - 1;
    float r = sqrtf(fabsf(sx));
    float th = 2 * kPi * sy;
    float px = r * cosf(th), py = r * sinf(th);
    float hh = sqrtf(1 - px * px);
    float tl = (1 + half.z) / 2;
    py = (1 - tl) * hh + tl * py;
    float pz = sqrtf(fmaxf(0.0f, 1 - (px * px + py * py)));
    V3 nh = add(add(mul(T1, px), mul(T2, py)), mul(half, pz));
    return nrm(mk(ax * nh.x, ay * nh.y, fmaxf(1e-6f, nh.z)));
}

// fresnelConductor on eta + ik (geometricMath.cu:242-271)
struct Cpx {
    float re, im;
};
__device__ __forceinline__ Cpx cx(float r, float i) { return Cpx{r, i}; }
__device__ __forceinline__ Cpx cadd(Cpx a, Cpx b) { return cx(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ Cpx csub(Cpx a, Cpx b) { return cx(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ Cpx cmul(Cpx a, Cpx b) { return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
__device__ __forceinline__ Cpx cdiv(Cpx a, Cpx b)
{
    float s = 1 / (b.re * b.re + b.im * b.im);
    return cx(s * (a.re * b.re + a.im * b.im), s * (a.im * b.re - a.re * b.im));
}
__device__ __forceinline__ float cnorm(Cpx a) { return a.re * a.re + a.im * a.im; }
__device__ __forceinline__ Cpx c_sqrt(Cpx z)
{
    float n = sqrtf(sqrtf(cnorm(z)));
    if (n == 0) return cx(0, 0);
    float t = atan2f(z.im, z.re) / 2;
    return cx(n * cosf(t), n * sinf(t));
}
__device__ __forceinline__ float fresnel_conductor(float cosi, float eta, float k)
{
    cosi = clampf(cosi, 0, 1);
    Cpx e = cx(eta, k);
    float s2i = 1 - cosi * cosi;
    Cpx s2t = cdiv(cx(s2i, 0), cmul(e, e));
    Cpx ct = c_sqrt(csub(cx(1, 0), s2t));
    Cpx ec = cmul(e, cx(cosi, 0));
    Cpx rpl = cdiv(csub(ec, ct), cadd(ec, ct));
    Cpx ect = cmul(e, ct);
    Cpx rpp = cdiv(csub(cx(cosi, 0), ect), cadd(cx(cosi, 0), ect));
    return (cnorm(rpl) + cnorm(rpp)) / 2;
}

// fresnelDielectric, literally (geometricMath.cu:202-240)
__device__ __forceinline__ float fresnel_dielectric(float cosO, float nO, float nD)
{
    cosO = clampf(cosO, -1, 1);
    if (!(cosO > 0)) cosO = fabsf(cosO);
    float sinO = sqrtf(fmaxf(0.0f, 1 - cosO * cosO));
    float sinD = nO / nD * sinO;
    if (sinO >= 1) return 1;
    float cosD = sqrtf(fmaxf(0.0f, 1 - sinD * sinD));
    float r = ((nD * cosO) - (nD * cosD)) / ((nD * cosO) + (nD * cosD));
    return (r * r + r * r) / 2 * fabsf(cosO);
}

__device__ __forceinline__ bool is_scattering(uint32_t type)
{
    return type == RT_MAT_LAMBERTIAN || type == RT_MAT_TROWBRIDGE || type == RT_MAT_TROWBRIDGE_BASIC;
}

// lightReflected (f * |cos wi|) for the scattering materials, shading space
// BSDF value and pdf of one material family: LAMBERT (LambertianBSDF.cu)
// or the Trowbridge-Reitz family (TrowbridgeBSDF.cu; conductor Fresnel for
// RT_MAT_TROWBRIDGE, none for RT_MAT_TROWBRIDGE_BASIC).  A shading kernel
// that serves one family instantiates only its own branch.
template <bool LAMBERT>
__device__ __forceinline__ V3 bsdf_f_of(const rt_material &m, V3 wo, V3 wi)
{
    if constexpr (LAMBERT) {
        float c = abscos(wi);  // LambertianBSDF.cu:22-38
        return mk(m.kd[0] * kInvPi * c, m.kd[1] * kInvPi * c, m.kd[2] * kInvPi * c);
    } else {
        // TrowbridgeBSDF::lightReflected (TrowbridgeBSDF.cu:156-200)
        float a = sqrtf(m.roughness);
        V3 o = nrm(wo), i = nrm(wi);
        if (!(o.z * i.z > 0.0f)) return mk(0, 0, 0);
        if (a < 1e-3f) return mk(0, 0, 0);
        float co = abscos(o), ci = abscos(i);
        if (ci == 0 || co == 0) return mk(0, 0, 0);
        V3 h = add(i, o);
        if (dot(h, h) == 0) return mk(0, 0, 0);
        h = nrm(h);
        float D = ggx_D(a, a, h);
        float G = 1 / (1 + ggx_lambda(a, a, o) + ggx_lambda(a, a, i));
        float F0 = 1, F1 = 1, F2 = 1;
        if (m.type == RT_MAT_TROWBRIDGE) {
            float c = abscos2(o, h);
            F0 = fresnel_conductor(c, m.eta[0], m.k[0]);
            F1 = fresnel_conductor(c, m.eta[1], m.k[1]);
            F2 = fresnel_conductor(c, m.eta[2], m.k[2]);
        }
        float den = 4 * ci * co;
        return mk(D * F0 * G / den * m.kd[0] * ci, D * F1 * G / den * m.kd[1] * ci, D * F2 * G / den * m.kd[2] * ci);
    }
}

template <bool LAMBERT>
__device__ __forceinline__ float bsdf_pdf_of(const rt_material &m, V3 wo, V3 wi)
{
    if constexpr (LAMBERT) {
        if (wo.z * wi.z > 0.0f) return (wi.z / len(wi)) * kInvPi;  // LambertianBSDF.cu:40-50
        return 0;
    } else {
        // TrowbridgeBSDF::pdf (TrowbridgeBSDF.cu:202-214)
        float a = sqrtf(m.roughness);
        if (!(wo.z * wi.z > 0.0f)) return 0;
        if (a < 1e-3f) return 0;
        V3 h = add(wo, wi);
        if (dot(h, h) == 0) return 0;
        h = nrm(h);
        if (h.z < 0.f) h = neg(h);
        return ggx_Dv(a, a, wo, h) / (4 * abscos2(wo, h));
    }
}

// Any scattering material (Lambertian or Trowbridge).
__device__ __forceinline__ V3 bsdf_f(const rt_material &m, V3 wo, V3 wi)
{
    return m.type == RT_MAT_LAMBERTIAN ? bsdf_f_of<true>(m, wo, wi) : bsdf_f_of<false>(m, wo, wi);
}

__device__ __forceinline__ float bsdf_pdf(const rt_material &m, V3 wo, V3 wi)
{
    return m.type == RT_MAT_LAMBERTIAN ? bsdf_pdf_of<true>(m, wo, wi) : bsdf_pdf_of<false>(m, wo, wi);
}

// Material::sampleDirection.  Draws: Lambertian 2, dielectric 1, Trowbridge
// 2 (0 when effectively smooth), light, mirror and plain 0.  The dielectric
// instantiation also serves mirrors (both are the specular, non-scattering
// class).
template <uint32_t TYPE>
__device__ __forceinline__ V3 bsdf_sample(const rt_material &m, V3 wo, Pcg32 &g, V3 &wi, float &pdf, bool &delta)
{
    pdf = 0;
    delta = false;
    if (TYPE == RT_MAT_LAMBERTIAN) {
        float u0 = pcg32_float(g), u1 = pcg32_float(g);
        wi = cosine_hemisphere(u0, u1);
        pdf = (wi.z / len(wi)) * kInvPi;
        return mk(m.kd[0], m.kd[1], m.kd[2]);  // LambertianBSDF.cu:52-74: weight = kd
    }
    if (TYPE == RT_MAT_DIELECTRIC) {
        delta = true;
        if (m.type == RT_MAT_MIRROR) {
            // PerfectReflectionBSDF::sampleDirection: weight = colour, no draw
            wi = mk(-wo.x, -wo.y, wo.z);
            return mk(m.kd[0], m.kd[1], m.kd[2]);
        }
        float u = pcg32_float(g);
        if (u < 0.7f) {  // DielectricMaterial.cu:69-76, no 1/0.7 compensation
            float nO = 1, nD = m.ior;
            if (wo.z / len(wo) < 0) { float t = nO; nO = nD; nD = t; }
            float eta = nO / nD;
            V3 n = (dot(mk(0, 0, 1), wo) < 0.f) ? mk(0, 0, -1) : mk(0, 0, 1);
            float cosO = cosine2(n, wo);
            float s2o = fmaxf(0.f, 1.f - cosO * cosO);
            float s2t = eta * eta * s2o;
            if (s2t >= 1) { wi = mk(0, 0, 0); return mk(0, 0, 0); }
            float ct = sqrtf(1 - s2t);
            wi = add(mul(neg(wo), eta), mul(n, eta * cosO - ct));
            float tr = 1 - fresnel_dielectric(cosO, nO, nD);
            return mk(m.kd[0] * tr, m.kd[1] * tr, m.kd[2] * tr);
        }
        wi = mk(-wo.x, -wo.y, wo.z);  // SpecularReflectionBSDF.cu:30-42
        float fr = fresnel_dielectric(wi.z / len(wi), 1, m.ior);
        return mk(m.kd[0] * fr, m.kd[1] * fr, m.kd[2] * fr);
    }
    if (TYPE == RT_MAT_TROWBRIDGE || TYPE == RT_MAT_TROWBRIDGE_BASIC) {
        float a = sqrtf(m.roughness);  // TrowbridgeBSDF.cu:216-246
        V3 o = nrm(wo);
        if (a < 1e-3f) {
            delta = true;
            wi = mk(-o.x, -o.y, o.z);
            if (m.type == RT_MAT_TROWBRIDGE_BASIC) return mk(1, 1, 1);
            float c = abscos(wi);
            return mk(fresnel_conductor(c, m.eta[0], m.k[0]) / c, fresnel_conductor(c, m.eta[1], m.k[1]) / c,
                      fresnel_conductor(c, m.eta[2], m.k[2]) / c);
        }
        float u0 = pcg32_float(g), u1 = pcg32_float(g);
        V3 h = ggx_sample_half(a, a, o, u0, u1);
        wi = add(neg(o), mul(h, 2 * dot(o, h)));
        float p = ggx_Dv(a, a, o, h) / (4 * abscos2(o, h));
        pdf = p;
        V3 f = bsdf_f(m, o, wi);
        return mk(f.x / p, f.y / p, f.z / p);
    }
    wi = mk(0, 0, 0);  // light: NullBSDF
    return mk(0, 0, 0);
}

// Spectrum::brightness (Spectrum.cu:68-75): float weights widened to double
__device__ __forceinline__ float luminance(V3 c)
{
    return (float)(0.212671f * (double)c.x + 0.715160f * (double)c.y + 0.072169f * (double)c.z);
}

// SurfaceInteraction::spawnRay origin (SurfaceInteraction.cu:163-185)
__device__ __forceinline__ V3 spawn_origin(V3 p, V3 perr, V3 n, V3 w)
{
    float dd = dot(vabs(n), perr);
    V3 off = mul(n, dd);
    if (dot(w, n) < 0) off = neg(off);
    V3 po = add(p, off);
    po.x = off.x > 0 ? next_up(po.x) : (off.x < 0 ? next_down(po.x) : po.x);
    po.y = off.y > 0 ? next_up(po.y) : (off.y < 0 ? next_down(po.y) : po.y);
    po.z = off.z > 0 ? next_up(po.z) : (off.z < 0 ? next_down(po.z) : po.z);
    return po;
}

// ---------------------------------------------------------------------
// Sphere (shapes/Sphere.cu:18-398), object-space error-bounded quadratic
// ---------------------------------------------------------------------
struct EF {
    float v, e;
};
__device__ __forceinline__ EF ef(float v, float e) { return EF{v, e}; }
__device__ __forceinline__ EF ef_add(EF a, EF b) { return ef(a.v + b.v, a.e + b.e); }
__device__ __forceinline__ EF ef_sub(EF a, EF b) { return ef(a.v - b.v, a.e + b.e); }
__device__ __forceinline__ EF ef_mul(EF a, EF b)
{
    EF r;
    r.v = a.v * b.v;
    r.e = fabsf(r.v) * b.e + fabsf(b.v) * a.e;
    return r;
}
__device__ __forceinline__ EF ef_smul(float s, EF a)
{
    EF r;
    r.v = s * a.v;
    r.e = fabsf(r.v) * a.e;
    return r;
}
__device__ __forceinline__ EF ef_div(EF a, EF b)
{
    EF r;
    r.v = a.v / b.v;
    r.e = (fabsf(r.v) * b.e + fabsf(b.v) * a.e) / (b.v * b.v);
    return r;
}
__device__ __forceinline__ EF ef_sqrt(EF a)
{
    EF r;
    r.v = sqrtf(a.v);
    r.e = (float)(0.5 * (double)a.e / (double)r.v);
    return r;
}

// Returns the hit distance (object-space unit direction) or -1 on a miss,
// and the object-space origin/direction for the surface data.
__device__ __forceinline__ float sphere_core(const SphereDev &s, V3 wo_, V3 wd_, float tmax, V3 &o, V3 &d)
{
    const float *m = s.w2o;
    const float eb3 = error_bound(3);
    V3 oErr, dErr;
    oErr.x = (fabsf(m[0] * wo_.x) + fabsf(m[1] * wo_.y) + fabsf(m[2] * wo_.z) + fabsf(m[3])) * eb3;
    oErr.y = (fabsf(m[4] * wo_.x) + fabsf(m[5] * wo_.y) + fabsf(m[6] * wo_.z) + fabsf(m[7])) * eb3;
    oErr.z = (fabsf(m[8] * wo_.x) + fabsf(m[9] * wo_.y) + fabsf(m[10] * wo_.z) + fabsf(m[11])) * eb3;
    o = xform_point(m, wo_);
    dErr.x = (fabsf(m[0] * wd_.x) + fabsf(m[1] * wd_.y) + fabsf(m[2] * wd_.z)) * eb3;
    dErr.y = (fabsf(m[4] * wd_.x) + fabsf(m[5] * wd_.y) + fabsf(m[6] * wd_.z)) * eb3;
    dErr.z = (fabsf(m[8] * wd_.x) + fabsf(m[9] * wd_.y) + fabsf(m[10] * wd_.z)) * eb3;
    d = xform_vec(m, wd_);
    float tm = tmax;
    float ls = d.x * d.x + d.y * d.y + d.z * d.z;
    if (ls > 0) {
        float dt = (fabsf(d.x) * oErr.x + fabsf(d.y) * oErr.y + fabsf(d.z) * oErr.z) / ls;
        o.x += d.x * dt; o.y += d.y * dt; o.z += d.z * dt;
        tm -= dt;
    }
    float l = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
    float il = 1.0f / l;
    d = mk(d.x * il, d.y * il, d.z * il);
    EF ox = ef(o.x, oErr.x), oy = ef(o.y, oErr.y), oz = ef(o.z, oErr.z);
    EF dx = ef(d.x, dErr.x), dy = ef(d.y, dErr.y), dz = ef(d.z, dErr.z);
    EF a = ef_add(ef_add(ef_mul(dx, dx), ef_mul(dy, dy)), ef_mul(dz, dz));
    EF b = ef_smul(2.0f, ef_add(ef_add(ef_mul(dx, ox), ef_mul(dy, oy)), ef_mul(dz, oz)));
    EF rr = ef(s.radius, 0);
    EF c = ef_sub(ef_add(ef_add(ef_mul(ox, ox), ef_mul(oy, oy)), ef_mul(oz, oz)), ef_mul(rr, rr));
    // quadratic (math.cu:91-115)
    EF disc = ef_sub(ef_mul(b, b), ef_mul(ef_smul(4.0f, a), c));
    if (disc.v < 0.0f) return -1.0f;
    EF rd = ef_sqrt(disc);
    EF q = (b.v < 0.0f) ? ef_smul(-0.5f, ef_sub(b, rd)) : ef_smul(-0.5f, ef_add(b, rd));
    EF t0 = ef_div(q, a), t1 = ef_div(c, q);
    if (t0.v > t1.v) { EF t = t0; t0 = t1; t1 = t; }
    if (t0.v != t0.v || t1.v != t1.v) return -1.0f;
    // checkRayBounds (Sphere.cu:41-81)
    if (t0.v + t0.e > tm || t1.v - t1.e <= 0) return -1.0f;
    EF th = t0;
    if (t0.v - t0.e <= 0) {
        th = t1;
        if (t1.v + t1.e > tm) return -1.0f;
    }
    return th.v;
}

// World-space surface data of a sphere hit at object distance t
// (Sphere.cu:64-69, SurfaceInteraction.cu:339-357).
__device__ __forceinline__ void sphere_surface(const SphereDev &s, V3 o, V3 d, float t, V3 &p, V3 &n, V3 &perr)
{
    V3 hp = add(o, mul(d, t));
    float dist = sqrtf(hp.x * hp.x + hp.y * hp.y + hp.z * hp.z);
    float sc = s.radius / dist;
    hp = mk(hp.x * sc, hp.y * sc, hp.z * sc);
    if (hp.x == 0 && hp.y == 0) hp.x = 1e-5f * s.radius;
    const float *m = s.o2w;
    const float eb3 = error_bound(3);
    perr.x = (fabsf(m[0] * hp.x) + fabsf(m[1] * hp.y) + fabsf(m[2] * hp.z) + fabsf(m[3])) * eb3;
    perr.y = (fabsf(m[4] * hp.x) + fabsf(m[5] * hp.y) + fabsf(m[6] * hp.z) + fabsf(m[7])) * eb3;
    perr.z = (fabsf(m[8] * hp.x) + fabsf(m[9] * hp.y) + fabsf(m[10] * hp.z) + fabsf(m[11])) * eb3;
    p = xform_point(m, hp);
    V3 on = hp;
    float l = sqrtf(on.x * on.x + on.y * on.y + on.z * on.z);
    float il = 1.0f / l;
    on = mk(on.x * il, on.y * il, on.z * il);
    const float *mi = s.w2o;
    n = mk(mi[0] * on.x + mi[4] * on.y + mi[8] * on.z, mi[1] * on.x + mi[5] * on.y + mi[9] * on.z,
           mi[2] * on.x + mi[6] * on.y + mi[10] * on.z);
}

// Plane::intersect / intersects (shapes/Plane.cu:58-164).  The ray goes to
// object space with Transform::operator()(Ray) (Transform.cu:216-237: the
// origin error, errorBound(4) of the row sums, moves the origin along the
// unnormalised direction and shrinks tMax); |d.z| must exceed
// FLOAT_ERROR_MARGIN (1e-7); the distance is Ray::getOffset,
// (hit.x - o.x) / d.x (Ray.cu:69-72), so a local d.x of 0 gives NaN and no
// hit; the hit must lie in [0,1]^2 with 0 < offset <= tMax.  Returns the
// offset (the caller applies the strict offset < hitOffset), or -1; u, v and
// the local d.z are kept for the surface.
__device__ __forceinline__ float plane_core(const PlaneDev &pl, V3 wo_, V3 wd_, float tmax, float &u, float &v,
                                            float &dz)
{
    const float *m = pl.w2o;
    const float eb4 = error_bound(4);
    const float xp = m[0] * wo_.x + m[1] * wo_.y + m[2] * wo_.z + m[3];
    const float yp = m[4] * wo_.x + m[5] * wo_.y + m[6] * wo_.z + m[7];
    const float zp = m[8] * wo_.x + m[9] * wo_.y + m[10] * wo_.z + m[11];
    const float wp = m[12] * wo_.x + m[13] * wo_.y + m[14] * wo_.z + m[15];
    V3 oErr;
    oErr.x = (fabsf(m[0] * wo_.x) + fabsf(m[1] * wo_.y) + fabsf(m[2] * wo_.z) + fabsf(m[3])) * eb4 * fabsf(wp);
    oErr.y = (fabsf(m[4] * wo_.x) + fabsf(m[5] * wo_.y) + fabsf(m[6] * wo_.z) + fabsf(m[7])) * eb4 * fabsf(wp);
    oErr.z = (fabsf(m[8] * wo_.x) + fabsf(m[9] * wo_.y) + fabsf(m[10] * wo_.z) + fabsf(m[11])) * eb4 * fabsf(wp);
    const float inv = 1.0f / wp;
    V3 o = mk(inv * xp, inv * yp, inv * zp);
    const V3 d = xform_vec(m, wd_);
    float tm = tmax;
    const float ls = d.x * d.x + d.y * d.y + d.z * d.z;
    if (ls > 0) {
        const float dt = (fabsf(d.x) * oErr.x + fabsf(d.y) * oErr.y + fabsf(d.z) * oErr.z) / ls;
        o.x += d.x * dt; o.y += d.y * dt; o.z += d.z * dt;
        tm -= dt;
    }
    const float invx = 1.0f / d.x;
    dz = d.z;
    if (!((double)fabsf(d.z) > 0.0000001)) return -1.0f;  // the margin is a double literal
    const float sc = o.z / d.z;
    const float hx = o.x - sc * d.x, hy = o.y - sc * d.y;
    const float off = (hx - o.x) * invx;
    u = hx;
    v = hy;
    if (off <= 0 || off > tm || u < 0 || u > 1 || v < 0 || v > 1) return -1.0f;
    return off;  // NaN when the local d.x is 0: never below a hit distance
}

// The plane's surface: world point with pError = 1e-4 per axis plus the
// transform's error (errorBound(4)), normal (0,0,1) turned against the
// local direction and taken to world space through the inverse transpose
// (worldToObject's third row), normalised.
__device__ __forceinline__ void plane_surface(const PlaneDev &pl, float u, float v, float dz, V3 &p, V3 &n, V3 &perr)
{
    const float *m = pl.o2w;
    const float eb4 = error_bound(4);
    const float wp = m[12] * u + m[13] * v + m[14] * 0.0f + m[15];
    perr.x = 0.0001f + (fabsf(m[0] * u) + fabsf(m[1] * v) + fabsf(m[2] * 0.0f) + fabsf(m[3])) * eb4 * fabsf(wp);
    perr.y = 0.0001f + (fabsf(m[4] * u) + fabsf(m[5] * v) + fabsf(m[6] * 0.0f) + fabsf(m[7])) * eb4 * fabsf(wp);
    perr.z = 0.0001f + (fabsf(m[8] * u) + fabsf(m[9] * v) + fabsf(m[10] * 0.0f) + fabsf(m[11])) * eb4 * fabsf(wp);
    p = xform_point(m, mk(u, v, 0.0f));
    const float on = dz > 0.0f ? -1.0f : 1.0f;
    const float *mi = pl.w2o;
    n = nrm(mk(mi[8] * on, mi[9] * on, mi[10] * on));
}

// Solid-angle cone sampling of a sphere light (the reference has no area
// lights; this is the renderer's definition, shared with the oracle).
__device__ __forceinline__ bool sphere_light_sample(const SphereDev &sp, V3 p, float u0, float u1, V3 &wi,
                                                    float &dist, float &pdf)
{
    V3 c = mk(sp.center[0], sp.center[1], sp.center[2]);
    V3 pc = sub(c, p);
    float d2 = dot(pc, pc);
    float r = sp.radius;
    if (d2 <= r * r) return false;
    float dc = sqrtf(d2);
    V3 wc = mul(pc, 1.0f / dc);
    float s2 = r * r / d2;
    float cmax = sqrtf(fmaxf(0.0f, 1 - s2));
    float ct = 1 - u0 * (1 - cmax);
    float st = sqrtf(fmaxf(0.0f, 1 - ct * ct));
    float ph = 2 * kPi * u1;
    V3 t, b;
    onb(wc, t, b);
    wi = nrm(add(add(mul(t, st * cosf(ph)), mul(b, st * sinf(ph))), mul(wc, ct)));
    float bb = dot(pc, wi);
    float disc = bb * bb - (d2 - r * r);
    dist = bb - sqrtf(fmaxf(0.0f, disc));
    pdf = 1.0f / (2 * kPi * (1 - cmax));
    return true;
}

__device__ __forceinline__ float sphere_light_pdf(const SphereDev &sp, V3 p)
{
    V3 c = mk(sp.center[0], sp.center[1], sp.center[2]);
    V3 pc = sub(c, p);
    float d2 = dot(pc, pc);
    float r = sp.radius;
    if (d2 <= r * r) return 0;
    float cmax = sqrtf(fmaxf(0.0f, 1 - r * r / d2));
    return 1.0f / (2 * kPi * (1 - cmax));
}

}  // namespace rt
