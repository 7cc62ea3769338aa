/*
 * shade.c -- oracle (test infrastructure only): camera, BSDFs, lights and
 * the path-tracing estimator of the reference, restated in C with FP32
 * shading and FP64 pixel sums.  The GPU wavefront renderer implements the
 * same estimator with the same random-number draw order, so with equal
 * seeds the two agree path for path except where transcendental functions
 * (sinf/cosf/sqrtf-based warps) round differently on the two sides.
 *
 * Draw order per path (the reference's own sampler is unseeded and shared
 * between threads, sampling/Sampler.cu:139-145, so its values cannot be
 * reproduced; this order is this project's definition):
 *   camera: 2 (film jitter)
 *   per bounce at a scattering material: 2 per sphere area light (NEE)
 *   BSDF: Lambertian 2, Trowbridge 2 (0 if effectively smooth),
 *         dielectric 1, light, mirror and plain 0
 *   Russian roulette: 1, only when it applies.
 */
#include "oracle.h"
#include "oracle_shade.h"
#include "oracle_photon.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define PIf 3.14159265358979323846f
#define INV_PIf (1.0f / PIf)

/* ------------------------------------------------------------------ */
/* PCG32 (O'Neill 2014, pcg32 XSH-RR with pcg32_srandom seeding)        */
/* ------------------------------------------------------------------ */
typedef struct { uint64_t state, inc; } pcg32;

static uint32_t pcg_next(pcg32 *g)
{
    uint64_t old = g->state;
    g->state = old * 6364136223846793005ULL + g->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

static pcg32 pcg_seed(uint64_t initstate, uint64_t initseq)
{
    pcg32 g = {0u, (initseq << 1u) | 1u};
    pcg_next(&g);
    g.state += initstate;
    pcg_next(&g);
    return g;
}

static float pcg_float(pcg32 *g) { return (float)(pcg_next(g) >> 8) * (1.0f / 16777216.0f); }

uint64_t orc_path_initstate(uint32_t seed, uint32_t sample)
{
    return ((uint64_t)seed << 32) ^ (uint64_t)sample ^ 0x853c49e6748fea9bULL;
}

/* ------------------------------------------------------------------ */
/* small vector helpers (FP32, no contraction)                          */
/* ------------------------------------------------------------------ */
typedef struct { float x, y, z; } v3;
static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, float s) { return mk(s * a.x, s * a.y, s * a.z); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float dot(v3 a, v3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
static inline v3 cross(v3 a, v3 b)
{
    return mk((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
static inline float len(v3 a) { return sqrtf(dot(a, a)); }
/* normalize(): v / length() with operator/ multiplying by 1/len (Vector.cu:454-457, 163-169) */
static inline v3 nrm(v3 a) { float inv = (float)1 / len(a); return mk(a.x * inv, a.y * inv, a.z * inv); }
static inline v3 vabs(v3 a) { return mk(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }

/* ------------------------------------------------------------------ */
/* Camera (camera/ProjectiveCamera.cu, FileImporting.cu:404-436,        */
/* Transform.cu:475-504,611-640)                                        */
/* ------------------------------------------------------------------ */
static void mat_mul(const float *a, const float *b, float *out)
{
    float r[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float s = 0;
            for (int k = 0; k < 4; ++k) s += a[i * 4 + k] * b[k * 4 + j];
            r[i * 4 + j] = s;
        }
    memcpy(out, r, sizeof r);
}

void orc_camera_setup(const orc_camera *c, float raster_to_camera[16], float camera_to_world[16])
{
    /* lookAt (Transform.cu:611-640): the matrix built is cameraToWorld, the
     * Transform returned holds its inverse (the mInv argument is dropped by
     * Transform(m, mInv), Transform.cu:34-37); the camera is constructed with
     * inverse(lookAt) (FileImporting.cu:431), i.e. a double inversion. */
    v3 pos = mk(c->pos[0], c->pos[1], c->pos[2]);
    v3 look = mk(c->lookat[0], c->lookat[1], c->lookat[2]);
    v3 up = mk(c->up[0], c->up[1], c->up[2]);
    v3 dir = nrm(sub(look, pos));
    v3 right = nrm(cross(nrm(up), dir));
    v3 nup = cross(dir, right);
    float c2w[16] = {right.x, nup.x, dir.x, pos.x,
                     right.y, nup.y, dir.y, pos.y,
                     right.z, nup.z, dir.z, pos.z,
                     0, 0, 0, 1};
    float la[16];
    orc_mat4_inverse(c2w, la);
    orc_mat4_inverse(la, camera_to_world);

    /* screen window: half height = frameY / 2, half width scaled by the
     * aspect ratio (FileImporting.cu:425-433) */
    float vfs = c->frame_y;
    vfs /= 2;
    float hfs = (vfs / (float)c->height) * (float)c->width;
    float pminx = -hfs, pminy = -vfs, pmaxx = hfs, pmaxy = vfs;
    /* screenToRaster = scale(res) * scale(1/w, 1/(pMin.y - pMax.y)) *
     * translate(-pMin.x, -pMax.y) (ProjectiveCamera.cu:37-53) */
    float T[16] = {1, 0, 0, -pminx, 0, 1, 0, -pmaxy, 0, 0, 1, 0, 0, 0, 0, 1};
    float S1[16] = {1 / (pmaxx - pminx), 0, 0, 0, 0, 1 / (pminy - pmaxy), 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    float S2[16] = {(float)c->width, 0, 0, 0, 0, (float)c->height, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    float tmp[16], s2r[16], r2s[16];
    mat_mul(S1, T, tmp);       /* scale_to_0_1(move_to_origin) */
    mat_mul(S2, tmp, s2r);     /* scale_to_resolution(...) */
    orc_mat4_inverse(s2r, r2s);
    /* rasterToCamera = inverse(cameraToScreen) * rasterToScreen, with
     * cameraToScreen = translate(cts) */
    float cts[16] = {1, 0, 0, c->cts[0], 0, 1, 0, c->cts[1], 0, 0, 1, c->cts[2], 0, 0, 0, 1};
    float icts[16];
    orc_mat4_inverse(cts, icts);
    mat_mul(icts, r2s, raster_to_camera);
}

static v3 xform_point(const float *m, v3 p)
{
    float xp = m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3];
    float yp = m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7];
    float zp = m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11];
    float wp = m[12] * p.x + m[13] * p.y + m[14] * p.z + m[15];
    float inv = (float)1 / wp;
    return mk(inv * xp, inv * yp, inv * zp);
}

static v3 xform_vec(const float *m, v3 v)
{
    return mk(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
              m[8] * v.x + m[9] * v.y + m[10] * v.z);
}

/* ProjectiveCamera::generateRay (ProjectiveCamera.cu:71-104) with the
 * camera sample of Sampler::getCameraSample (Sampler.cu:166-181) and the
 * ray transform with origin error offset (Transform.cu:216-237). */
void orc_camera_ray(const float *r2c, const float *c2w, uint32_t w, uint32_t h, uint32_t px, uint32_t py,
                    float u0, float u1, float o[3], float d[3])
{
    float fx = (float)px + u0, fy = (float)py + u1;
    fx = fminf(fmaxf(fx, 0.0f), (float)w);
    fy = fminf(fmaxf(fy, 0.0f), (float)h);
    v3 pc = xform_point(r2c, mk(fx, fy, 0));
    v3 dir = nrm(pc);
    /* origin (0,0,0) through cameraToWorld with error bounds */
    float eb3 = orc_error_bound(3);
    v3 oerr = mul(mk(fabsf(c2w[3]), fabsf(c2w[7]), fabsf(c2w[11])), eb3);
    /* abs(m*0) terms are zero; the sums reduce to |m[i][3]| */
    v3 org = xform_point(c2w, mk(0, 0, 0));
    v3 wd = xform_vec(c2w, dir);
    float ls = dot(wd, wd);
    if (ls > 0) {
        float dt = dot(vabs(wd), oerr) / ls;
        org = add(org, mul(wd, dt));
    }
    o[0] = org.x; o[1] = org.y; o[2] = org.z;
    d[0] = wd.x; d[1] = wd.y; d[2] = wd.z;
}

/* ------------------------------------------------------------------ */
/* Sampling warps (sampling/Sampler.cu:13-96)                           */
/* ------------------------------------------------------------------ */
static void concentric_disk(float ux, float uy, float *dx, float *dy)
{
    float ox = 2.0f * ux - 1.0f, oy = 2.0f * uy - 1.0f;
    if (ox == 0 && oy == 0) { *dx = 0; *dy = 0; return; }
    float theta, r;
    if (fabsf(ox) > fabsf(oy)) {
        r = ox;
        theta = (PIf / 4) * (oy / ox);
    } else {
        r = oy;
        theta = (PIf / 2) - (PIf / 4) * (ox / oy);
    }
    *dx = r * cosf(theta);
    *dy = r * sinf(theta);
}

static v3 cosine_hemisphere(float ux, float uy)
{
    float x, y;
    concentric_disk(ux, uy, &x, &y);
    float z = sqrtf(fmaxf(0.0f, 1 - x * x - y * y));
    return mk(x, y, z);
}

static void uniform_disk(float ux, float uy, float *x, float *y)
{
    float r = sqrtf(fabsf(ux));
    float th = 2 * PIf * uy;
    *x = r * cosf(th);
    *y = r * sinf(th);
}

/* ------------------------------------------------------------------ */
/* Shading frame: orthonormal basis about the shading normal (Duff et    */
/* al., JCGT 2017).  The reference builds its frame from the normal and  */
/* dpdv (SurfaceInteraction.cu:54-107); every BSDF here is isotropic, so  */
/* the frame's rotation about n does not change any distribution.       */
/* ------------------------------------------------------------------ */
static void onb(v3 n, v3 *t, v3 *b)
{
    float s = copysignf(1.0f, n.z);
    float a = -1.0f / (s + n.z);
    float c = n.x * n.y * a;
    *t = mk(1.0f + s * n.x * n.x * a, s * c, -s * n.x);
    *b = mk(c, s + n.y * n.y * a, -n.y);
}

/* ------------------------------------------------------------------ */
/* BSDFs                                                                */
/* ------------------------------------------------------------------ */
static float cos2th(v3 w) { return w.z * w.z; }
static float sin2th(v3 w) { return fmaxf(0.0f, 1 - cos2th(w)); }
static float sinth(v3 w) { return sqrtf(sin2th(w)); }
static float clampf(float x, float a, float b) { return fmaxf(a, fminf(b, x)); }
static float cosphi(v3 w) { float s = sinth(w); return (s == 0) ? 1 : clampf(w.x / s, -1, 1); }
static float sinphi(v3 w) { float s = sinth(w); return (s == 0) ? 0 : clampf(w.y / s, -1, 1); }
static float tan2th(v3 w) { return sin2th(w) / cos2th(w); }
static float abscos(v3 w) { return fabsf(w.z / len(w)); }
static float abscos2(v3 a, v3 b) { return fabsf(dot(a, b)) / (len(a) * len(b)); }
static float cosine2(v3 a, v3 b) { return dot(a, b) / (len(a) * len(b)); }

/* TrowbridgeBSDF::distribution (materials/TrowbridgeBSDF.cu:71-84) */
static float ggx_D(float ax, float ay, v3 h)
{
    float t2 = tan2th(h);
    if (isinf(t2)) return 0;
    float c4 = cos2th(h) * cos2th(h);
    float cp = cosphi(h) / ax, sp = sinphi(h) / ay;
    float e = t2 * (cp * cp + sp * sp);
    float q = 1 + e;
    return 1 / (PIf * ax * ay * c4 * (q * q));
}

/* lambda (TrowbridgeBSDF.cu:101-113); isInf is an == +inf test (math.cu:83-88) */
static float ggx_lambda(float ax, float ay, v3 w)
{
    float t2 = tan2th(w);
    if (t2 == INFINITY) return 0;
    float a = cosphi(w) * ax, b = sinphi(w) * ay;
    float a2 = a * a + b * b;
    return (sqrtf(1 + a2 * t2) - 1) / 2;
}

/* distribution(wo, h): visible normals (TrowbridgeBSDF.cu:95-99) */
static float ggx_Dv(float ax, float ay, v3 wo, v3 h)
{
    float g1 = 1 / (1 + ggx_lambda(ax, ay, wo));
    return g1 / abscos(wo) * ggx_D(ax, ay, h) * abscos2(wo, h);
}

/* sampleHalf (TrowbridgeBSDF.cu:120-151), including the reference's
 * uniformSampleDisk(2u - 1) mapping */
static v3 ggx_sample_half(float ax, float ay, v3 wo, float u0, float u1)
{
    v3 half = nrm(mk(ax * wo.x, ay * wo.y, wo.z));
    if (half.z < 0) half = neg(half);
    v3 T1 = (half.z < 0.99999f) ? nrm(cross(mk(0, 0, 1), half)) : mk(1, 0, 0);
    v3 T2 = cross(half, T1);
    float sx = 2 * u0 - 1, sy = 2 * u1 - 1;
    float px, py;
    uniform_disk(sx, sy, &px, &py);
    float hh = sqrtf(1 - px * px);
    float tl = (1 + half.z) / 2;
    py = (1 - tl) * hh + tl * py; /* lerp(x, a, b) = (1 - x) a + x b (math.cu:11-16) */
    float pz = sqrtf(fmaxf(0.0f, 1 - (px * px + py * py)));
    v3 nh = add(add(mul(T1, px), mul(T2, py)), mul(half, pz));
    return nrm(mk(ax * nh.x, ay * nh.y, fmaxf(1e-6f, nh.z)));
}

/* fresnelConductor (geometricMath.cu:242-271) on complex eta = n + ik */
typedef struct { float re, im; } cpx;
static cpx cx(float r, float i) { cpx c = {r, i}; return c; }
static cpx cadd(cpx a, cpx b) { return cx(a.re + b.re, a.im + b.im); }
static cpx csub(cpx a, cpx b) { return cx(a.re - b.re, a.im - b.im); }
static cpx cmul(cpx a, cpx b) { return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static cpx cdiv(cpx a, cpx b)
{
    float s = 1 / (b.re * b.re + b.im * b.im);
    return cx(s * (a.re * b.re + a.im * b.im), s * (a.im * b.re - a.re * b.im));
}
static float cnorm(cpx a) { return a.re * a.re + a.im * a.im; }
static cpx c_sqrt(cpx z)
{
    float n = sqrtf(sqrtf(cnorm(z)));
    if (n == 0) return cx(0, 0);
    float t = atan2f(z.im, z.re) / 2;
    return cx(n * cosf(t), n * sinf(t));
}

static float fresnel_conductor(float cosi, float eta, float k)
{
    cosi = clampf(cosi, 0, 1);
    cpx e = cx(eta, k);
    float s2i = 1 - cosi * cosi;
    cpx s2t = cdiv(cx(s2i, 0), cmul(e, e));
    cpx ct = c_sqrt(csub(cx(1, 0), s2t));
    cpx ec = cmul(e, cx(cosi, 0));
    cpx rpl = cdiv(csub(ec, ct), cadd(ec, ct));
    cpx ect = cmul(e, ct);
    cpx rpp = cdiv(csub(cx(cosi, 0), ect), cadd(cx(cosi, 0), ect));
    return (cnorm(rpl) + cnorm(rpp)) / 2;
}

/* fresnelDielectric, literally (geometricMath.cu:202-240): both
 * polarisations use eta_t in numerator and denominator, the result is
 * (r_par^2 + r_perp^2) / 2 * |cos|, TIR is tested on sin(theta_i). */
float orc_fresnel_dielectric(float cosO, float nO, float nD)
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

/* Material::isScattering: Lambertian, Trowbridge and Plain
 * (PlainMaterial.cu:77-80; its next-event terms are 0: lightReflected = 0, :30-41) */
static int is_scattering(const orc_material *m)
{
    return m->type == ORC_MAT_LAMBERTIAN || m->type == ORC_MAT_TROWBRIDGE || m->type == ORC_MAT_TROWBRIDGE_BASIC ||
           m->type == ORC_MAT_PLAIN;
}

/* Material::isSpecular: a camera ray stays one through these hits
 * (SurfaceInteraction.cu:252-254); Trowbridge counts (TrowbridgeMaterial.cu:55-63) */
static int is_specular(const orc_material *m)
{
    return m->type == ORC_MAT_DIELECTRIC || m->type == ORC_MAT_MIRROR || m->type == ORC_MAT_TROWBRIDGE ||
           m->type == ORC_MAT_TROWBRIDGE_BASIC;
}

static float ggx_alpha(const orc_material *m) { return sqrtf(m->roughness); } /* roughnessToAlpha */

/* lightReflected for the scattering materials (f * |cos wi|), shading space */
static v3 bsdf_f(const orc_material *m, v3 wo, v3 wi)
{
    if (m->type == ORC_MAT_PLAIN) return mk(0, 0, 0); /* PlainMaterial::lightReflected */
    if (m->type == ORC_MAT_LAMBERTIAN) {
        /* LambertianBRDF: kd/pi * absCosine(wi) (LambertianBSDF.cu:22-38) */
        float c = abscos(wi);
        return mk(m->kd[0] * INV_PIf * c, m->kd[1] * INV_PIf * c, m->kd[2] * INV_PIf * c);
    }
    /* TrowbridgeBSDF::lightReflected (TrowbridgeBSDF.cu:156-200) */
    float a = ggx_alpha(m);
    v3 o = nrm(wo), i = nrm(wi);
    if (!(o.z * i.z > 0.0f)) return mk(0, 0, 0);
    if (a < 1e-3f) return mk(0, 0, 0);
    float co = abscos(o), ci = abscos(i);
    if (ci == 0 || co == 0) return mk(0, 0, 0);
    v3 h = add(i, o);
    if (dot(h, h) == 0) return mk(0, 0, 0);
    h = nrm(h);
    float D = ggx_D(a, a, h);
    float G = 1 / (1 + ggx_lambda(a, a, o) + ggx_lambda(a, a, i));
    float F[3] = {1, 1, 1};
    if (m->type == ORC_MAT_TROWBRIDGE) {
        float c = abscos2(o, h);
        for (int k = 0; k < 3; ++k) F[k] = fresnel_conductor(c, m->eta[k], m->k[k]);
    }
    float den = 4 * ci * co;
    v3 r;
    r.x = D * F[0] * G / den * m->kd[0] * ci;
    r.y = D * F[1] * G / den * m->kd[1] * ci;
    r.z = D * F[2] * G / den * m->kd[2] * ci;
    return r;
}

static float bsdf_pdf(const orc_material *m, v3 wo, v3 wi)
{
    if (m->type == ORC_MAT_LAMBERTIAN) {
        /* cosineHemispherePdf if same hemisphere (LambertianBSDF.cu:40-50) */
        if (wo.z * wi.z > 0.0f) return (wi.z / len(wi)) * INV_PIf;
        return 0;
    }
    if (m->type == ORC_MAT_TROWBRIDGE || m->type == ORC_MAT_TROWBRIDGE_BASIC) {
        /* TrowbridgeBSDF::pdf (TrowbridgeBSDF.cu:202-214) */
        float a = ggx_alpha(m);
        if (!(wo.z * wi.z > 0.0f)) return 0;
        if (a < 1e-3f) return 0;
        v3 h = add(wo, wi);
        if (dot(h, h) == 0) return 0;
        h = nrm(h);
        if (h.z < 0.f) h = neg(h); /* faceForward(h, (0,0,1)) */
        return ggx_Dv(a, a, wo, h) / (4 * abscos2(wo, h));
    }
    return 0;
}

/* Uniform draws Material::sampleDirection consumes for material m: the
 * Lambertian and rough Trowbridge warps take 2, the dielectric's branch
 * choice 1, the rest 0 (the draw order listed at the top of this file). */
static int bsdf_draws(const orc_material *m)
{
    if (m->type == ORC_MAT_LAMBERTIAN) return 2;
    if (m->type == ORC_MAT_DIELECTRIC) return 1;
    if (m->type == ORC_MAT_TROWBRIDGE || m->type == ORC_MAT_TROWBRIDGE_BASIC) return ggx_alpha(m) < 1e-3f ? 0 : 2;
    return 0;
}

/* Material::sampleDirection for every type with its uniform draws given in
 * u[0..bsdf_draws(m)-1]; returns the weight and sets wi. */
static v3 bsdf_sample_u(const orc_material *m, v3 wo, const float *u, v3 *wi, float *pdf_out, int *delta)
{
    *pdf_out = 0;
    *delta = 0;
    if (m->type == ORC_MAT_LAMBERTIAN) {
        float u0 = u[0], u1 = u[1];
        *wi = cosine_hemisphere(u0, u1);
        *pdf_out = (wi->z / len(*wi)) * INV_PIf;
        return mk(m->kd[0], m->kd[1], m->kd[2]); /* weight = kd (LambertianBSDF.cu:52-74) */
    }
    if (m->type == ORC_MAT_DIELECTRIC) {
        *delta = 1;
        if (u[0] < 0.7f) {
            /* SpecularRefractionBSDF::sampleDirection (SpecularRefractionBSDF.cu:31-63) */
            float nO = 1, nD = m->ior;
            if (wo.z / len(wo) < 0) { float t = nO; nO = nD; nD = t; }
            float eta = nO / nD;
            v3 n = (dot(mk(0, 0, 1), wo) < 0.f) ? mk(0, 0, -1) : mk(0, 0, 1);
            float cosO = cosine2(n, wo);
            float s2o = fmaxf(0.f, 1.f - cosO * cosO);
            float s2t = eta * eta * s2o;
            if (s2t >= 1) { *wi = mk(0, 0, 0); return mk(0, 0, 0); }
            float ct = sqrtf(1 - s2t);
            *wi = add(mul(neg(wo), eta), mul(n, eta * cosO - ct));
            float tr = 1 - orc_fresnel_dielectric(cosO, nO, nD);
            return mk(m->kd[0] * tr, m->kd[1] * tr, m->kd[2] * tr);
        }
        /* SpecularReflectionBSDF::sampleDirection (SpecularReflectionBSDF.cu:30-42) */
        *wi = mk(-wo.x, -wo.y, wo.z);
        float fr = orc_fresnel_dielectric(wi->z / len(*wi), 1, m->ior);
        return mk(m->kd[0] * fr, m->kd[1] * fr, m->kd[2] * fr);
    }
    if (m->type == ORC_MAT_TROWBRIDGE || m->type == ORC_MAT_TROWBRIDGE_BASIC) {
        /* TrowbridgeBSDF::sampleDirection (TrowbridgeBSDF.cu:216-246) */
        float a = ggx_alpha(m);
        v3 o = nrm(wo);
        if (a < 1e-3f) {
            *delta = 1;
            *wi = mk(-o.x, -o.y, o.z);
            if (m->type == ORC_MAT_TROWBRIDGE_BASIC) return mk(1, 1, 1);
            float c = abscos(*wi);
            return mk(fresnel_conductor(c, m->eta[0], m->k[0]) / c, fresnel_conductor(c, m->eta[1], m->k[1]) / c,
                      fresnel_conductor(c, m->eta[2], m->k[2]) / c);
        }
        v3 h = ggx_sample_half(a, a, o, u[0], u[1]);
        *wi = add(neg(o), mul(h, 2 * dot(o, h))); /* reflect (geometricMath.cu:196-199) */
        float pdf = ggx_Dv(a, a, o, h) / (4 * abscos2(o, h));
        *pdf_out = pdf;
        v3 f = bsdf_f(m, o, *wi);
        return mk(f.x / pdf, f.y / pdf, f.z / pdf);
    }
    if (m->type == ORC_MAT_MIRROR) {
        /* PerfectReflectionBSDF::sampleDirection (PerfectReflectionBSDF.cu): weight = colour */
        *delta = 1;
        *wi = mk(-wo.x, -wo.y, wo.z);
        return mk(m->kd[0], m->kd[1], m->kd[2]);
    }
    /* light: NullBSDF (NullBSDF.cu); plain: PlainMaterial::sampleDirection returns 0 */
    *wi = mk(0, 0, 0);
    return mk(0, 0, 0);
}

/* bsdf_sample_u with its draws taken from the path's PCG32 stream. */
static v3 bsdf_sample(const orc_material *m, v3 wo, pcg32 *g, v3 *wi, float *pdf_out, int *delta)
{
    float u[2] = {0, 0};
    const int n = bsdf_draws(m);
    for (int k = 0; k < n; ++k) u[k] = pcg_float(g);
    return bsdf_sample_u(m, wo, u, wi, pdf_out, delta);
}

/* Russian roulette (pathTracing.cu:91-102): q = max(0.05, 1 - brightness(beta))
 * with beta after this bounce's weight; the path ends if u < q, else beta is
 * divided by 1 - q.  Returns 1 if the path survives. */
static inline float lum(v3 c);
static int russian_roulette(v3 *beta, float u)
{
    float q = fmaxf(0.05f, 1 - lum(*beta));
    if (u < q) return 0;
    float sc = 1 - q;
    *beta = mk(beta->x / sc, beta->y / sc, beta->z / sc);
    return 1;
}

/* ------------------------------------------------------------------ */
/* Scene                                                                */
/* ------------------------------------------------------------------ */
struct orc_scene {
    uint32_t nt, nsph, nmat, nlight;
    float *xyz;        /* 9 floats per triangle */
    uint32_t *tri_mat;
    orc_sphere *sph;
    uint32_t *sph_mat;
    uint8_t *sph_light;
    uint8_t *tri_transparent;
    orc_material *mat;
    orc_point_light *pl;
    orc_camera cam;
    float r2c[16], c2w[16];
    orc_refbvh *bvh;
    uint32_t *idx_seq;
    orc_plane *plane;
    uint32_t *plane_mat;
    uint32_t nplane;
    float *tri_uv;        /* 6 per triangle, or NULL: (0,0), (1,0), (1,1) */
    float *tri_nrm;       /* 9 per triangle (corner normals, NaN first: none), or NULL */
    uint8_t **tex;        /* ntex textures, 3 bytes per texel */
    uint32_t *tex_w, *tex_h;
    uint32_t ntex;
};

orc_scene *orc_scene_create(const float *xyz, const uint32_t *idx, uint32_t nt, const uint32_t *tri_mat,
                            const float *spheres, const uint32_t *sph_mat, const uint8_t *sph_light,
                            uint32_t nsph, const orc_material *mats, uint32_t nmat,
                            const orc_point_light *lights, uint32_t nlight, const orc_camera *cam)
{
    orc_scene *s = calloc(1, sizeof *s);
    s->nt = nt; s->nsph = nsph; s->nmat = nmat; s->nlight = nlight;
    s->xyz = malloc(sizeof(float) * 9 * (nt ? nt : 1));
    s->idx_seq = malloc(sizeof(uint32_t) * 3 * (nt ? nt : 1));
    for (uint32_t i = 0; i < nt; ++i)
        for (int v = 0; v < 3; ++v) {
            for (int k = 0; k < 3; ++k) s->xyz[9 * i + 3 * v + k] = xyz[3 * idx[3 * i + v] + k];
            s->idx_seq[3 * i + v] = 3 * i + v;
        }
    s->tri_mat = malloc(sizeof(uint32_t) * (nt ? nt : 1));
    memcpy(s->tri_mat, tri_mat, sizeof(uint32_t) * nt);
    s->mat = malloc(sizeof(orc_material) * (nmat ? nmat : 1));
    memcpy(s->mat, mats, sizeof(orc_material) * nmat);
    s->tri_transparent = malloc(nt ? nt : 1);
    for (uint32_t i = 0; i < nt; ++i) s->tri_transparent[i] = mats[tri_mat[i]].type == ORC_MAT_DIELECTRIC;
    s->sph = malloc(sizeof(orc_sphere) * (nsph ? nsph : 1));
    s->sph_mat = malloc(sizeof(uint32_t) * (nsph ? nsph : 1));
    s->sph_light = malloc(nsph ? nsph : 1);
    for (uint32_t j = 0; j < nsph; ++j) {
        orc_sphere_init(&s->sph[j], spheres + 4 * j, spheres[4 * j + 3]);
        s->sph_mat[j] = sph_mat[j];
        s->sph_light[j] = sph_light ? sph_light[j] : 0;
    }
    s->pl = malloc(sizeof(orc_point_light) * (nlight ? nlight : 1));
    memcpy(s->pl, lights, sizeof(orc_point_light) * nlight);
    s->cam = *cam;
    orc_camera_setup(cam, s->r2c, s->c2w);
    s->bvh = orc_refbvh_build(s->xyz, s->idx_seq, nt);
    return s;
}

/* Plane shapes, tested after the spheres (Scene.cu:221-227 tests the
 * analytic shapes in insertion order): 12 floats each (corner, axis_x,
 * axis_y, normal) as Plane(corner, axisX, axisY, normal) takes them. */
void orc_scene_add_planes(orc_scene *s, const float *planes, const uint32_t *mats, uint32_t n)
{
    s->plane = realloc(s->plane, sizeof(orc_plane) * (s->nplane + n));
    s->plane_mat = realloc(s->plane_mat, sizeof(uint32_t) * (s->nplane + n));
    for (uint32_t k = 0; k < n; ++k) {
        const float *q = planes + 12 * k;
        orc_plane_init(&s->plane[s->nplane + k], q, q + 3, q + 6, q + 9);
        s->plane_mat[s->nplane + k] = mats[k];
    }
    s->nplane += n;
}

uint32_t orc_scene_add_texture(orc_scene *s, const uint8_t *rgb, uint32_t width, uint32_t height)
{
    s->tex = realloc(s->tex, sizeof(uint8_t *) * (s->ntex + 1));
    s->tex_w = realloc(s->tex_w, sizeof(uint32_t) * (s->ntex + 1));
    s->tex_h = realloc(s->tex_h, sizeof(uint32_t) * (s->ntex + 1));
    s->tex[s->ntex] = malloc((size_t)width * height * 3);
    memcpy(s->tex[s->ntex], rgb, (size_t)width * height * 3);
    s->tex_w[s->ntex] = width;
    s->tex_h[s->ntex] = height;
    return s->ntex++;
}

void orc_scene_set_tri_normals(orc_scene *s, const float *nrm9)
{
    free(s->tri_nrm);
    s->tri_nrm = malloc(sizeof(float) * 9 * (s->nt ? s->nt : 1));
    memcpy(s->tri_nrm, nrm9, sizeof(float) * 9 * s->nt);
}

void orc_scene_set_tri_uv(orc_scene *s, const float *uv)
{
    free(s->tri_uv);
    s->tri_uv = malloc(sizeof(float) * 6 * (s->nt ? s->nt : 1));
    memcpy(s->tri_uv, uv, sizeof(float) * 6 * s->nt);
}

/* Texture::getColor (textures/Texture.cu:172-183): texel
 * ((unsigned)(u * width), (unsigned)(v * height)) clamped to the last
 * column / row, channels / 255; the float-to-unsigned conversion as the
 * reference's x86-64 host build performs it (truncation to a 64-bit integer,
 * low 32 bits; NaN to 0). */
static uint32_t texel_index(float u, uint32_t n)
{
    float f = u * (float)n;
    if (f != f) f = 0.0f;
    f = fminf(fmaxf(f, -4.6e18f), 4.6e18f);
    uint32_t x = (uint32_t)(unsigned long long)(long long)f;
    return x < n - 1 ? x : n - 1;
}

static v3 texture_color(const orc_scene *s, uint32_t t, float u, float v)
{
    uint32_t x = texel_index(u, s->tex_w[t]), y = texel_index(v, s->tex_h[t]);
    const uint8_t *c = s->tex[t] + 3 * ((size_t)y * s->tex_w[t] + x);
    return mk((float)c[0] / 255.0f, (float)c[1] / 255.0f, (float)c[2] / 255.0f);
}

void orc_scene_free(orc_scene *s)
{
    if (!s) return;
    free(s->plane); free(s->plane_mat);
    for (uint32_t t = 0; t < s->ntex; ++t) free(s->tex[t]);
    free(s->tex); free(s->tex_w); free(s->tex_h); free(s->tri_uv); free(s->tri_nrm);
    orc_refbvh_free(s->bvh);
    free(s->xyz); free(s->idx_seq); free(s->tri_mat); free(s->mat); free(s->tri_transparent);
    free(s->sph); free(s->sph_mat); free(s->sph_light); free(s->pl); free(s);
}

void orc_scene_camera(const orc_scene *s, float r2c[16], float c2w[16])
{
    memcpy(r2c, s->r2c, sizeof s->r2c);
    memcpy(c2w, s->c2w, sizeof s->c2w);
}

typedef struct {
    int kind;       /* 0 miss, 1 triangle, 2 sphere */
    int32_t id;
    float t;
    v3 p, n, perr;  /* world hit point, shading normal, error */
    float su, sv;   /* SurfaceInteraction::surfacePoint */
} hit_rec;

/* Scene::intersectTriangles + intersectShapes (integration/Scene.cu:165-236):
 * the BVH's triangle, re-tested with Triangle::intersect from hitOffset =
 * +inf, then every sphere in insertion order with a strict < test.  The
 * triangle tests apply the deltaT rejection before lowering the hit distance
 * (orc_tri_intersects_guarded), as the library does: in the reference's
 * order a ray that re-meets the triangle it was spawned from (a candidate at
 * t ~ 0 that fails deltaT) loses every later hit, and whether that happens
 * depends on the tree's visit order, which the library does not share. */
static hit_rec scene_intersect(const orc_scene *s, v3 o, v3 d)
{
    hit_rec h;
    memset(&h, 0, sizeof h);
    h.t = INFINITY;
    orc_ray r;
    float of[3] = {o.x, o.y, o.z}, df[3] = {d.x, d.y, d.z};
    orc_ray_make(&r, of, df, INFINITY);
    if (s->nt) {
        float tt;
        int32_t id = orc_refbvh_intersect_guarded(s->bvh, &r, NULL, 0, &tt);
        if (id >= 0) {
            const float *t = s->xyz + 9 * id;
            float hh = INFINITY;
            orc_tri_info info;
            if (orc_tri_intersects_guarded(&r, t, t + 3, t + 6, &hh, &info)) {
                h.kind = 1; h.id = id; h.t = hh;
            }
        }
    }
    for (uint32_t j = 0; j < s->nsph; ++j) {
        float p[3], n[3], e[3];
        float hh = h.t;
        float uv[2];
        if (orc_sphere_intersect_uv(&s->sph[j], &r, &hh, p, n, e, uv)) {
            h.kind = 2; h.id = (int32_t)j; h.t = hh; h.su = uv[0]; h.sv = uv[1];
            h.p = mk(p[0], p[1], p[2]); h.n = mk(n[0], n[1], n[2]); h.perr = mk(e[0], e[1], e[2]);
        }
    }
    for (uint32_t k = 0; k < s->nplane; ++k) {
        float p[3], n[3], e[3];
        float hh = h.t;
        float uv[2];
        if (orc_plane_intersect_uv(&s->plane[k], &r, &hh, p, n, e, uv)) {
            h.kind = 3; h.id = (int32_t)k; h.t = hh; h.su = uv[0]; h.sv = uv[1];
            h.p = mk(p[0], p[1], p[2]); h.n = nrm(mk(n[0], n[1], n[2])); h.perr = mk(e[0], e[1], e[2]);
        }
    }
    if (h.kind == 1) {
        /* surface data of Triangle::intersect (Triangle.cu:161-431) */
        const float *t = s->xyz + 9 * h.id;
        float bc[3];
        orc_tri_barycentrics(&r, t, t + 3, t + 6, bc);
        v3 p0 = mk(t[0], t[1], t[2]), p1 = mk(t[3], t[4], t[5]), p2 = mk(t[6], t[7], t[8]);
        h.p = add(add(mul(p0, bc[0]), mul(p1, bc[1])), mul(p2, bc[2]));
        float xs = fabsf(bc[0] * p0.x) + fabsf(bc[1] * p1.x) + fabsf(bc[2] * p2.x);
        float ys = fabsf(bc[0] * p0.y) + fabsf(bc[1] * p1.y) + fabsf(bc[2] * p2.y);
        float zs = fabsf(bc[0] * p0.z) + fabsf(bc[1] * p1.z) + fabsf(bc[2] * p2.z);
        h.perr = mul(mk(xs, ys, zs), orc_error_bound(7));
        v3 n = nrm(cross(sub(p0, p2), sub(p1, p2)));
        if (dot(n, d) > 0) n = neg(n);
        h.n = n;
        if (s->tri_nrm && s->tri_nrm[9 * (size_t)h.id] == s->tri_nrm[9 * (size_t)h.id]) {
            /* mesh normals (Triangle.cu:379-418): ns = normalize(b0 n0 + b1 n1 +
             * b2 n2), not turned towards the ray, shades the hit, and the hit
             * point moves 0.15 along it */
            const float *q = s->tri_nrm + 9 * (size_t)h.id;
            v3 ns = nrm(add(add(mul(mk(q[0], q[1], q[2]), bc[0]), mul(mk(q[3], q[4], q[5]), bc[1])),
                            mul(mk(q[6], q[7], q[8]), bc[2])));
            h.n = ns;
            h.p = add(h.p, mul(ns, 0.15f));
        }
        /* surfacePoint = b0 uv0 + b1 uv1 + b2 uv2 (Triangle.cu:356-358, 134-152) */
        float uvd[6] = {0, 0, 1, 0, 1, 1};
        const float *q = s->tri_uv ? s->tri_uv + 6 * (size_t)h.id : uvd;
        h.su = bc[0] * q[0] + bc[1] * q[2] + bc[2] * q[4];
        h.sv = bc[0] * q[1] + bc[1] * q[3] + bc[2] * q[5];
    }
    return h;
}

/* Scene::intersects with checkShadow (Scene.cu:108-150) + isVisible
 * (:262-273): occluded iff an opaque triangle or sphere is hit at t <= dist.
 * skip_sphere excludes an area light's own sphere from its shadow test. */
static int occluded(const orc_scene *s, v3 o, v3 d, float dist, int skip_sphere)
{
    orc_ray r;
    float of[3] = {o.x, o.y, o.z}, df[3] = {d.x, d.y, d.z};
    orc_ray_make(&r, of, df, INFINITY);
    float t = INFINITY;
    int hit = 0;
    if (s->nt) {
        float tt;
        int32_t id = orc_refbvh_intersect_guarded(s->bvh, &r, s->tri_transparent, 1, &tt);
        if (id >= 0) { hit = 1; t = tt; }
    }
    for (uint32_t j = 0; j < s->nsph; ++j) {
        if ((int)j == skip_sphere) continue;
        if (s->mat[s->sph_mat[j]].type == ORC_MAT_DIELECTRIC) continue; /* Sphere.cu:362-363 */
        if (orc_sphere_intersects(&s->sph[j], &r, &t)) hit = 1;
    }
    for (uint32_t k = 0; k < s->nplane; ++k) {
        if (s->mat[s->plane_mat[k]].type == ORC_MAT_DIELECTRIC) continue; /* Plane.cu:128-129 */
        if (orc_plane_intersect(&s->plane[k], &r, &t, NULL, NULL, NULL)) hit = 1;
    }
    return hit && !(t > dist);
}

/* SurfaceInteraction::spawnRay (SurfaceInteraction.cu:163-185) */
static v3 spawn_origin(v3 p, v3 perr, v3 n, v3 w)
{
    float dd = dot(vabs(n), perr);
    v3 off = mul(n, dd);
    if (dot(w, n) < 0) off = neg(off);
    v3 po = add(p, off);
    float *pp = &po.x;
    const float *oo = &off.x;
    for (int i = 0; i < 3; ++i) {
        if (oo[i] > 0) pp[i] = orc_next_float_up(pp[i]);
        else if (oo[i] < 0) pp[i] = orc_next_float_down(pp[i]);
    }
    return po;
}

static inline float lum(v3 c) /* Spectrum::brightness, Spectrum.cu:68-75 */
{
    return (float)(0.212671f * (double)c.x + 0.715160f * (double)c.y + 0.072169f * (double)c.z);
}

/* Solid-angle cone sampling of a sphere light seen from p (the reference
 * has no area lights; this is the definition used by the renderer). */
static int sphere_light_sample(const orc_sphere *sp, v3 p, float u0, float u1, v3 *wi, float *dist, float *pdf)
{
    v3 c = mk(sp->center[0], sp->center[1], sp->center[2]);
    v3 pc = sub(c, p);
    float d2 = dot(pc, pc);
    float r = sp->radius;
    if (d2 <= r * r) return 0;
    float dc = sqrtf(d2);
    v3 wc = mul(pc, 1.0f / dc);
    float s2 = r * r / d2;
    float cmax = sqrtf(fmaxf(0.0f, 1 - s2));
    float ct = 1 - u0 * (1 - cmax);
    float st = sqrtf(fmaxf(0.0f, 1 - ct * ct));
    float ph = 2 * PIf * u1;
    v3 t, b;
    onb(wc, &t, &b);
    *wi = nrm(add(add(mul(t, st * cosf(ph)), mul(b, st * sinf(ph))), mul(wc, ct)));
    /* distance to the near intersection with the sphere along wi */
    float bb = dot(pc, *wi);
    float disc = bb * bb - (d2 - r * r);
    *dist = bb - sqrtf(fmaxf(0.0f, disc));
    *pdf = 1.0f / (2 * PIf * (1 - cmax));
    return 1;
}

static float sphere_light_pdf(const orc_sphere *sp, v3 p)
{
    v3 c = mk(sp->center[0], sp->center[1], sp->center[2]);
    v3 pc = sub(c, p);
    float d2 = dot(pc, pc);
    float r = sp->radius;
    if (d2 <= r * r) return 0;
    float cmax = sqrtf(fmaxf(0.0f, 1 - r * r / d2));
    return 1.0f / (2 * PIf * (1 - cmax));
}

/* Vertices a radiance cache may stand for: diffuse, or GGX with roughness
 * >= 0.04 (alpha >= 0.2).  The reference has no cache; this is the library's
 * rule, restated for the termination tests. */
static int cacheable(const orc_material *m)
{
    return m->type == ORC_MAT_LAMBERTIAN ||
           ((m->type == ORC_MAT_TROWBRIDGE || m->type == ORC_MAT_TROWBRIDGE_BASIC) && m->roughness >= 0.04f);
}

/* One path of the estimator (integration/pathTracing.cu:33-123,173-205,
 * Scene.cu:280-335,468-525).  Returns L in out[3] and 1 if finite.  With
 * cache_rgb non-NULL the path ends at its first cacheable vertex at depth >=
 * cache_depth and adds beta * cache_rgb there (a cache that answers one
 * constant). */
static int path_impl(const orc_scene *s, uint32_t px, uint32_t py, uint32_t sample, uint32_t seed,
                     uint32_t max_depth, uint32_t rr_depth, uint64_t stream_base, uint32_t cache_depth,
                     const float *cache_rgb, float out[3], float *seg, uint32_t *nseg,
                     const orc_photon_render *ph)
{
    if (nseg) *nseg = 0;
    const uint32_t W = s->cam.width;
    pcg32 g = pcg_seed(orc_path_initstate(seed, sample), stream_base + (uint64_t)py * W + px);
    float u0 = pcg_float(&g), u1 = pcg_float(&g);
    float of[3], df[3];
    orc_camera_ray(s->r2c, s->c2w, s->cam.width, s->cam.height, px, py, u0, u1, of, df);
    v3 o = mk(of[0], of[1], of[2]), d = mk(df[0], df[1], df[2]);
    v3 beta = mk(1, 1, 1), L = mk(0, 0, 0);
    float prev_pdf = 0;   /* BSDF pdf of the direction that produced this ray */
    int prev_delta = 1;   /* camera rays count as delta: no MIS on a light hit */
    int camera = 1;       /* Ray::cameraRay: primary, then kept through specular hits */
    v3 prev_p = o;
    for (uint32_t depth = 0; depth < max_depth; ++depth) {
        if (seg) {
            float *q = seg + 6 * (size_t)depth;
            q[0] = o.x; q[1] = o.y; q[2] = o.z; q[3] = d.x; q[4] = d.y; q[5] = d.z;
            *nseg = depth + 1;
        }
        hit_rec h = scene_intersect(s, o, d);
        if (h.kind == 0) break;
        const orc_material *m =
            &s->mat[h.kind == 1 ? s->tri_mat[h.id] : h.kind == 2 ? s->sph_mat[h.id] : s->plane_mat[h.id]];
        orc_material textured;
        if (m->texture && (m->type == ORC_MAT_LAMBERTIAN || m->type == ORC_MAT_PLAIN)) {
            /* LambertianBSDF.cu:22-74, PlainMaterial.cu:62-75: kd times the texel */
            textured = *m;
            v3 c = texture_color(s, m->texture - 1, h.su, h.sv);
            textured.kd[0] = m->kd[0] * c.x;
            textured.kd[1] = m->kd[1] * c.y;
            textured.kd[2] = m->kd[2] * c.z;
            m = &textured;
        }
        if (cache_rgb && depth >= cache_depth && cacheable(m)) {
            L = add(L, mk(beta.x * cache_rgb[0], beta.y * cache_rgb[1], beta.z * cache_rgb[2]));
            break;
        }
        float nee_scale = 1.0f;
        if (ph && is_scattering(m)) {
            /* RENDER_MODE PHOTON_MAPPING: no direct-light sampling; the
             * scattering hit takes getPhotonMappingContribution (Scene.cu:
             * 724-746: the k nearest photons within the radius on the hit's
             * shape, combinePhotonContributions), added times the throughput
             * (pathTracing.cu:186-204), then its emission, and the path ends
             * (finishStep, pathTracing.cu:84-88).  Photon shape ids: spheres,
             * planes, then triangles. */
            uint32_t idx[1024];
            float dist[1024];
            const int32_t qshape = h.kind == 1 ? (int32_t)(s->nsph + s->nplane + h.id)
                                   : h.kind == 2 ? (int32_t)h.id : (int32_t)(s->nsph + h.id);
            const float q[3] = {h.p.x, h.p.y, h.p.z};
            uint32_t cnt = orc_kd_knn(ph->xyz, ph->shape, ph->order, ph->axis, ph->n, q, qshape, ph->k, ph->radius,
                                      idx, dist);
            const float nq[3] = {h.n.x, h.n.y, h.n.z}, view[3] = {-d.x, -d.y, -d.z};
            double est[3];
            orc_photon_gather_mat(ph->rgb, ph->dir, 1, &cnt, idx, dist, ph->k, nq, view, m, est);
            if (ph->mode == 2) {
                /* HYBRID (Scene.cu:963-967): (photons + direct) / 2, added as
                 * the library adds them (the photon half, then the halved
                 * next-event terms); the path goes on */
                L = add(L, mk(beta.x * ((float)est[0] * 0.5f), beta.y * ((float)est[1] * 0.5f),
                              beta.z * ((float)est[2] * 0.5f)));
                nee_scale = 0.5f;
            } else {
                L = add(L, mk(beta.x * (float)est[0], beta.y * (float)est[1], beta.z * (float)est[2]));
                if (m->type == ORC_MAT_PLAIN) {
                    const float *e = camera ? m->kd : m->ke;
                    L = add(L, mk(beta.x * e[0], beta.y * e[1], beta.z * e[2]));
                }
                break;
            }
        }
        v3 n = h.n, t, b;
        onb(n, &t, &b);
        v3 wo_w = neg(d);
        v3 wo = mk(dot(wo_w, t), dot(wo_w, b), dot(wo_w, n));
        if (is_scattering(m)) {
            v3 Ld = mk(0, 0, 0);
            for (uint32_t l = 0; l < s->nlight; ++l) {
                /* PointLight::sampleLight (PointLight.cu:56-66) */
                v3 lp = mk(s->pl[l].pos[0], s->pl[l].pos[1], s->pl[l].pos[2]);
                v3 tol = sub(lp, h.p);
                v3 Lw = nrm(tol);
                float d2 = dot(tol, tol);
                float I = s->pl[l].intensity;
                v3 Li = mk(s->pl[l].color[0] * I * (1 / d2), s->pl[l].color[1] * I * (1 / d2),
                           s->pl[l].color[2] * I * (1 / d2));
                if (!(dot(Lw, n) > 0)) continue; /* isOrientedTo (SurfaceInteraction.cu:283) */
                v3 wi = mk(dot(Lw, t), dot(Lw, b), dot(Lw, n));
                v3 f = bsdf_f(m, wo, wi);
                float pf = bsdf_pdf(m, wo, wi);
                if ((f.x == 0 && f.y == 0 && f.z == 0) || (Li.x == 0 && Li.y == 0 && Li.z == 0)) continue;
                v3 so = spawn_origin(h.p, h.perr, n, Lw);
                if (occluded(s, so, Lw, len(tol), -1)) continue;
                float w = 1.0f / (1.0f + pf * pf); /* powerHeuristic(1, 1, 1, pf), Scene.cu:331 */
                Ld = add(Ld, mk(f.x * Li.x * w * nee_scale, f.y * Li.y * w * nee_scale, f.z * Li.z * w * nee_scale));
            }
            for (uint32_t j = 0; j < s->nsph; ++j) {
                if (!s->sph_light[j]) continue;
                float a0 = pcg_float(&g), a1 = pcg_float(&g);
                v3 Lw;
                float dist, pl;
                if (!sphere_light_sample(&s->sph[j], h.p, a0, a1, &Lw, &dist, &pl)) continue;
                if (!(dot(Lw, n) > 0)) continue;
                v3 wi = mk(dot(Lw, t), dot(Lw, b), dot(Lw, n));
                v3 f = bsdf_f(m, wo, wi);
                if (f.x == 0 && f.y == 0 && f.z == 0) continue;
                float pf = bsdf_pdf(m, wo, wi);
                v3 so = spawn_origin(h.p, h.perr, n, Lw);
                if (occluded(s, so, Lw, dist * 0.9999f, (int)j)) continue;
                const orc_material *lm = &s->mat[s->sph_mat[j]];
                float w = pl * pl / (pl * pl + pf * pf);
                float k = w / pl;
                Ld = add(Ld, mk(f.x * lm->kd[0] * lm->intensity * k * nee_scale,
                                f.y * lm->kd[1] * lm->intensity * k * nee_scale,
                                f.z * lm->kd[2] * lm->intensity * k * nee_scale));
            }
            L = add(L, mk(beta.x * Ld.x, beta.y * Ld.y, beta.z * Ld.z));
        }
        if (m->type == ORC_MAT_LIGHT) {
            /* getEmittedRadiance (LightMaterial.cu:20-23); MIS against the
             * light's own NEE when it is a sampled sphere light */
            float w = 1;
            if (h.kind == 2 && s->sph_light[h.id] && !prev_delta) {
                float pl = sphere_light_pdf(&s->sph[h.id], prev_p);
                w = prev_pdf * prev_pdf / (prev_pdf * prev_pdf + pl * pl);
            }
            float I = m->intensity;
            L = add(L, mk(beta.x * m->kd[0] * I * w, beta.y * m->kd[1] * I * w, beta.z * m->kd[2] * I * w));
        }
        if (m->type == ORC_MAT_PLAIN) {
            /* PlainMaterial::getEmittedRadiance (PlainMaterial.cu:62-75): kd to camera rays, ke otherwise */
            const float *e = camera ? m->kd : m->ke;
            L = add(L, mk(beta.x * e[0], beta.y * e[1], beta.z * e[2]));
        }
        v3 wi;
        float pdf;
        int delta;
        v3 wgt = bsdf_sample(m, wo, &g, &wi, &pdf, &delta);
        beta = mk(beta.x * wgt.x, beta.y * wgt.y, beta.z * wgt.z);
        if (wgt.x == 0 && wgt.y == 0 && wgt.z == 0) break;
        if (is_scattering(m) && depth > rr_depth && !russian_roulette(&beta, pcg_float(&g))) break;
        v3 ww = add(add(mul(t, wi.x), mul(b, wi.y)), mul(n, wi.z));
        o = spawn_origin(h.p, h.perr, n, ww);
        d = ww;
        prev_pdf = pdf;
        prev_delta = delta;
        prev_p = h.p;
        camera = camera && is_specular(m);
    }
    out[0] = L.x; out[1] = L.y; out[2] = L.z;
    return isfinite(L.x) && isfinite(L.y) && isfinite(L.z);
}

int orc_path(const orc_scene *s, uint32_t px, uint32_t py, uint32_t sample, uint32_t seed, uint32_t max_depth,
             uint32_t rr_depth, uint64_t stream_base, float out[3])
{
    return path_impl(s, px, py, sample, seed, max_depth, rr_depth, stream_base, 0, NULL, out, NULL, NULL, NULL);
}

int orc_path_segments(const orc_scene *s, uint32_t px, uint32_t py, uint32_t sample, uint32_t seed,
                      uint32_t max_depth, uint32_t rr_depth, uint64_t stream_base, float out[3], float *seg,
                      uint32_t *nseg)
{
    return path_impl(s, px, py, sample, seed, max_depth, rr_depth, stream_base, 0, NULL, out, seg, nseg, NULL);
}

typedef struct {
    const orc_scene *s;
    uint32_t spp, seed, max_depth, rr_depth, spp_offset;
    uint32_t x0, y0, x1, y1;
    double *accum;
    uint32_t *nvalid;
    uint64_t stream_base;
    uint32_t cache_depth;
    const float *cache_rgb;
    const orc_photon_render *ph;
} render_ctx;

static void *render_rows(void *p)
{
    render_ctx *c = p;
    const uint32_t W = c->s->cam.width;
    for (uint32_t y = c->y0; y < c->y1; ++y)
        for (uint32_t x = c->x0; x < c->x1; ++x) {
            double acc[3] = {0, 0, 0};
            uint32_t nv = 0;
            for (uint32_t k = 0; k < c->spp; ++k) {
                float L[3];
                if (path_impl(c->s, x, y, c->spp_offset + k, c->seed, c->max_depth, c->rr_depth, c->stream_base,
                              c->cache_depth, c->cache_rgb, L, NULL, NULL, c->ph)) {
                    acc[0] += L[0]; acc[1] += L[1]; acc[2] += L[2];
                    ++nv;
                }
            }
            size_t i = (size_t)y * W + x;
            c->accum[3 * i] += acc[0];
            c->accum[3 * i + 1] += acc[1];
            c->accum[3 * i + 2] += acc[2];
            c->nvalid[i] += nv;
        }
    return NULL;
}

static void render_impl(const orc_scene *s, uint32_t spp, uint32_t seed, uint32_t max_depth, uint32_t rr_depth,
                        uint32_t spp_offset, uint64_t stream_base, uint32_t x0, uint32_t y0, uint32_t x1,
                        uint32_t y1, double *accum, uint32_t *nvalid, int nthreads, uint32_t cache_depth,
                        const float *cache_rgb, const orc_photon_render *ph)
{
    if (nthreads < 1) nthreads = 1;
    uint32_t rows = y1 > y0 ? y1 - y0 : 0;
    if ((uint32_t)nthreads > rows) nthreads = rows ? (int)rows : 1;
    pthread_t *th = malloc(sizeof(pthread_t) * nthreads);
    render_ctx *cs = malloc(sizeof(render_ctx) * nthreads);
    /* interleave row blocks so threads get similar work */
    uint32_t chunk = (rows + nthreads - 1) / nthreads;
    for (int i = 0; i < nthreads; ++i) {
        render_ctx c = {s, spp, seed, max_depth, rr_depth, spp_offset, x0, y0 + i * chunk, x1, y0 + (i + 1) * chunk,
                        accum, nvalid, stream_base, cache_depth, cache_rgb, ph};
        if (c.y0 > y1) c.y0 = y1;
        if (c.y1 > y1) c.y1 = y1;
        cs[i] = c;
        pthread_create(&th[i], NULL, render_rows, &cs[i]);
    }
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th); free(cs);
}

void orc_render(const orc_scene *s, uint32_t spp, uint32_t seed, uint32_t max_depth, uint32_t rr_depth,
                uint32_t spp_offset, uint64_t stream_base, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1,
                double *accum, uint32_t *nvalid, int nthreads)
{
    render_impl(s, spp, seed, max_depth, rr_depth, spp_offset, stream_base, x0, y0, x1, y1, accum, nvalid,
                nthreads, 0, NULL, NULL);
}

void orc_render_photon(const orc_scene *s, uint32_t spp, uint32_t seed, uint32_t max_depth, uint32_t rr_depth,
                       uint32_t spp_offset, uint64_t stream_base, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1,
                       double *accum, uint32_t *nvalid, int nthreads, const orc_photon_render *ph)
{
    render_impl(s, spp, seed, max_depth, rr_depth, spp_offset, stream_base, x0, y0, x1, y1, accum, nvalid,
                nthreads, 0, NULL, ph);
}

void orc_render_const_cache(const orc_scene *s, uint32_t spp, uint32_t seed, uint32_t max_depth,
                            uint32_t rr_depth, uint32_t spp_offset, uint64_t stream_base, uint32_t x0, uint32_t y0,
                            uint32_t x1, uint32_t y1, double *accum, uint32_t *nvalid, int nthreads,
                            uint32_t cache_depth, const float cache_rgb[3])
{
    render_impl(s, spp, seed, max_depth, rr_depth, spp_offset, stream_base, x0, y0, x1, y1, accum, nvalid,
                nthreads, cache_depth, cache_rgb, NULL);
}

/* computeRMSE (camera/Film.cu:611-632): per pixel the squared norm of the
 * difference of mean radiances, as a float (Spectrum.cu:204-207),
 * accumulated in float, divided by the pixel count, square root. */
float orc_compute_rmse(const double *sum_a, double n_a, const double *sum_b, double n_b, uint32_t npix)
{
    float rmse = 0;
    for (uint32_t i = 0; i < npix; ++i) {
        double dr = sum_a[3 * i] / n_a - sum_b[3 * i] / n_b;
        double dg = sum_a[3 * i + 1] / n_a - sum_b[3 * i + 1] / n_b;
        double db = sum_a[3 * i + 2] / n_a - sum_b[3 * i + 2] / n_b;
        rmse += (float)(dr * dr + dg * dg + db * db);
    }
    rmse /= (float)npix;
    return sqrtf(rmse);
}

/* ------------------------------------------------------------------ */
/* The shading functions one at a time, for the tests that pin them     */
/* against closed forms, quadrature and sampling statistics.  Vectors   */
/* are in the shading frame (z = normal) unless stated.                 */
/* ------------------------------------------------------------------ */
void orc_concentric_disk(float u0, float u1, float out[2]) { concentric_disk(u0, u1, &out[0], &out[1]); }

void orc_ggx_sample_half(float alpha, const float wo[3], float u0, float u1, float h[3])
{
    v3 r = ggx_sample_half(alpha, alpha, mk(wo[0], wo[1], wo[2]), u0, u1);
    h[0] = r.x; h[1] = r.y; h[2] = r.z;
}

void orc_ggx_D_n(float alpha, const float *h, uint32_t n, float *out)
{
    for (uint32_t i = 0; i < n; ++i) out[i] = ggx_D(alpha, alpha, mk(h[3 * i], h[3 * i + 1], h[3 * i + 2]));
}

void orc_ggx_lambda_n(float alpha, const float *w, uint32_t n, float *out)
{
    for (uint32_t i = 0; i < n; ++i) out[i] = ggx_lambda(alpha, alpha, mk(w[3 * i], w[3 * i + 1], w[3 * i + 2]));
}

float orc_fresnel_conductor(float cosi, float eta, float k) { return fresnel_conductor(cosi, eta, k); }

int orc_bsdf_draws(const orc_material *m) { return bsdf_draws(m); }

/* lightReflected (f * |cos wi|) and pdf of n direction pairs. */
void orc_bsdf_eval_n(const orc_material *m, const float *wo, const float *wi, uint32_t n, float *f, float *pdf)
{
    for (uint32_t i = 0; i < n; ++i) {
        v3 o = mk(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), w = mk(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
        v3 r = bsdf_f(m, o, w);
        f[3 * i] = r.x; f[3 * i + 1] = r.y; f[3 * i + 2] = r.z;
        pdf[i] = bsdf_pdf(m, o, w);
    }
}

/* sampleDirection of n outgoing directions with explicit draws (2 per
 * sample; the first bsdf_draws(m) are used). */
void orc_bsdf_sample_n(const orc_material *m, const float *wo, const float *u, uint32_t n, float *wi, float *pdf,
                       int32_t *delta, float *weight)
{
    for (uint32_t i = 0; i < n; ++i) {
        v3 w;
        int dl;
        v3 r = bsdf_sample_u(m, mk(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), u + 2 * i, &w, &pdf[i], &dl);
        wi[3 * i] = w.x; wi[3 * i + 1] = w.y; wi[3 * i + 2] = w.z;
        weight[3 * i] = r.x; weight[3 * i + 1] = r.y; weight[3 * i + 2] = r.z;
        delta[i] = dl;
    }
}

/* The dielectric's branch over a PCG32 stream: how many of n draws (stream
 * seeded like a path's) take the refraction branch. */
uint32_t orc_dielectric_refractions(uint32_t seed, uint32_t n)
{
    pcg32 g = pcg_seed(orc_path_initstate(seed, 0), 0);
    orc_material m;
    memset(&m, 0, sizeof m);
    m.type = ORC_MAT_DIELECTRIC;
    m.kd[0] = m.kd[1] = m.kd[2] = 1;
    m.ior = 1.5f;
    uint32_t k = 0;
    for (uint32_t i = 0; i < n; ++i) {
        v3 w;
        float pdf;
        int dl;
        bsdf_sample(&m, mk(0.3f, 0.2f, 0.9327379f), &g, &w, &pdf, &dl);
        k += w.z < 0;  /* refracted directions leave through the other side */
    }
    return k;
}

float orc_luminance(const float rgb[3]) { return lum(mk(rgb[0], rgb[1], rgb[2])); }

int orc_russian_roulette(const float beta[3], float u, float out[3])
{
    v3 b = mk(beta[0], beta[1], beta[2]);
    int alive = russian_roulette(&b, u);
    out[0] = b.x; out[1] = b.y; out[2] = b.z;
    return alive;
}

int orc_sphere_light_sample(const float center[3], float radius, const float p[3], float u0, float u1, float wi[3],
                            float *dist, float *pdf)
{
    orc_sphere sp;
    orc_sphere_init(&sp, center, radius);
    v3 w;
    int ok = sphere_light_sample(&sp, mk(p[0], p[1], p[2]), u0, u1, &w, dist, pdf);
    wi[0] = w.x; wi[1] = w.y; wi[2] = w.z;
    return ok;
}

float orc_sphere_light_pdf(const float center[3], float radius, const float p[3])
{
    orc_sphere sp;
    orc_sphere_init(&sp, center, radius);
    return sphere_light_pdf(&sp, mk(p[0], p[1], p[2]));
}

/* ------------------------------------------------------------------ */
/* Photon pre-pass: generatePhotons + computePhotons                    */
/* (integration/photonDispersion.cu:26-192, Scene.cu:427-455, 989-1021) */
/* ------------------------------------------------------------------ */
/* Selection probability of every point light (sum of its power over the
 * total, PointLight::power = color * intensity * 4 pi, light/PointLight.cu:
 * 40-43) and their running sum, in double. */
void orc_photon_light_table(const orc_scene *s, double *p, double *cdf)
{
    double total = 0;
    for (uint32_t l = 0; l < s->nlight; ++l) {
        const orc_point_light *L = &s->pl[l];
        p[l] = ((double)L->color[0] + (double)L->color[1] + (double)L->color[2]) * (double)L->intensity *
               (4.0 * 3.14159265358979323846);
        total += p[l];
    }
    double run = 0;
    for (uint32_t l = 0; l < s->nlight; ++l) {
        p[l] = total > 0 ? p[l] / total : 0.0;
        run += p[l];
        cdf[l] = run;
    }
}

/* selectOneRandomLight: the first light whose running probability reaches u
 * (the reference subtracts each light's share from u until it is <= 0);
 * rounding past the end falls back to the last light with a share. */
static uint32_t photon_select(const double *cdf, uint32_t nl, float u)
{
    for (uint32_t l = 0; l < nl; ++l)
        if ((double)u <= cdf[l]) return l;
    uint32_t last = nl - 1;
    while (last > 0 && cdf[last] == cdf[last - 1]) --last;
    return last;
}

/* The photon path's generator: stream i of the seed's state. */
static pcg32 photon_rng(uint32_t seed, uint32_t i)
{
    return pcg_seed(((uint64_t)seed << 32) ^ 0x9E3779B97F4A7C15ull, (uint64_t)i);
}

uint32_t orc_trace_photons(const orc_scene *s, uint32_t n, uint32_t seed, uint32_t max_depth, uint32_t rr_depth,
                           uint32_t cap, float *pos, float *dir, float *nrm_out, double *rgb, int32_t *shape,
                           uint32_t *key)
{
    const uint32_t nl = s->nlight;
    if (nl == 0 || n == 0) return 0;
    double *p = malloc(sizeof(double) * nl), *cdf = malloc(sizeof(double) * nl), *rgbl = calloc(3 * nl, sizeof(double));
    uint32_t *cnt = calloc(nl, sizeof(uint32_t));
    orc_photon_light_table(s, p, cdf);
    /* generatePhotons: every photon's light first, then per light
     * color * intensity / uniformSpherePdf / n_l, divided by
     * lightSelectionPdf(l) (Scene.cu:996-1012, PointLight.cu:68-95) */
    for (uint32_t i = 0; i < n; ++i) {
        pcg32 g = photon_rng(seed, i);
        ++cnt[photon_select(cdf, nl, pcg_float(&g))];
    }
    for (uint32_t l = 0; l < nl; ++l)
        for (int c = 0; c < 3; ++c)
            if (cnt[l])
                rgbl[3 * l + c] = (double)s->pl[l].color[c] * (double)s->pl[l].intensity *
                                  (4.0 * 3.14159265358979323846) / (double)cnt[l] / p[l];
    const float kPi = 3.14159265358979323846f;
    const uint32_t nshapes = s->nsph + s->nplane;
    uint32_t m = 0;
    for (uint32_t i = 0; i < n; ++i) {
        pcg32 g = photon_rng(seed, i);
        const uint32_t l = photon_select(cdf, nl, pcg_float(&g));
        /* uniformSampleSphere (sampling/Sampler.cu:47-52) */
        const float u0 = pcg_float(&g), u1 = pcg_float(&g);
        const float phi = 2.0f * kPi * u0, theta = acosf(2.0f * u1 - 1.0f);
        v3 o = mk(s->pl[l].pos[0], s->pl[l].pos[1], s->pl[l].pos[2]);
        v3 d = mk(sinf(theta) * cosf(phi), sinf(theta) * sinf(phi), cosf(theta));
        double tl[3] = {rgbl[3 * l], rgbl[3 * l + 1], rgbl[3 * l + 2]};
        v3 beta = mk(1, 1, 1);
        for (uint32_t depth = 0; depth < max_depth; ++depth) {
            hit_rec h = scene_intersect(s, o, d);
            if (h.kind == 0) break; /* no interaction: finished */
            const orc_material *mt =
                &s->mat[h.kind == 1 ? s->tri_mat[h.id] : h.kind == 2 ? s->sph_mat[h.id] : s->plane_mat[h.id]];
            orc_material textured;
            if (mt->texture && (mt->type == ORC_MAT_LAMBERTIAN || mt->type == ORC_MAT_PLAIN)) {
                textured = *mt;
                v3 c = texture_color(s, mt->texture - 1, h.su, h.sv);
                textured.kd[0] = mt->kd[0] * c.x;
                textured.kd[1] = mt->kd[1] * c.y;
                textured.kd[2] = mt->kd[2] * c.z;
                mt = &textured;
            }
            const int scat = is_scattering(mt);
            if (scat) {
                /* samplePhotonIntersections: Photon(pathRay, worldPoint,
                 * shading normal, totalLight * pathLightFactor, shapeId)
                 * (photonDispersion.cu:40-55) */
                if (m < cap) {
                    pos[3 * (size_t)m] = h.p.x; pos[3 * (size_t)m + 1] = h.p.y; pos[3 * (size_t)m + 2] = h.p.z;
                    dir[3 * (size_t)m] = d.x; dir[3 * (size_t)m + 1] = d.y; dir[3 * (size_t)m + 2] = d.z;
                    nrm_out[3 * (size_t)m] = h.n.x; nrm_out[3 * (size_t)m + 1] = h.n.y; nrm_out[3 * (size_t)m + 2] = h.n.z;
                    rgb[3 * (size_t)m] = tl[0] * (double)beta.x;
                    rgb[3 * (size_t)m + 1] = tl[1] * (double)beta.y;
                    rgb[3 * (size_t)m + 2] = tl[2] * (double)beta.z;
                    shape[m] = h.kind == 1 ? (int32_t)(nshapes + (uint32_t)h.id)
                                           : h.kind == 2 ? h.id : (int32_t)(s->nsph + (uint32_t)h.id);
                    key[m] = (i << 5) | depth;
                }
                ++m;
            }
            /* sampleRay, emission into totalLight, pathLightFactor *= weight
             * (photonDispersion.cu:57-72) */
            v3 nn = h.n, t, b;
            onb(nn, &t, &b);
            v3 wo_w = neg(d);
            v3 wo = mk(dot(wo_w, t), dot(wo_w, b), dot(wo_w, nn));
            v3 wi;
            float pdf;
            int delta;
            v3 wgt = bsdf_sample(mt, wo, &g, &wi, &pdf, &delta);
            if (mt->type == ORC_MAT_LIGHT) {
                const float I = mt->intensity;
                tl[0] += (double)beta.x * (double)(mt->kd[0] * I);
                tl[1] += (double)beta.y * (double)(mt->kd[1] * I);
                tl[2] += (double)beta.z * (double)(mt->kd[2] * I);
            } else if (mt->type == ORC_MAT_PLAIN) {
                /* photon rays are not camera rays: ke */
                tl[0] += (double)beta.x * (double)mt->ke[0];
                tl[1] += (double)beta.y * (double)mt->ke[1];
                tl[2] += (double)beta.z * (double)mt->ke[2];
            }
            beta = mk(beta.x * wgt.x, beta.y * wgt.y, beta.z * wgt.z);
            if (wgt.x == 0 && wgt.y == 0 && wgt.z == 0) break;
            /* finishStep with photon_first_pass (pathTracing.cu:62-123) */
            if (scat && depth > rr_depth && !russian_roulette(&beta, pcg_float(&g))) break;
            v3 ww = add(add(mul(t, wi.x), mul(b, wi.y)), mul(nn, wi.z));
            o = spawn_origin(h.p, h.perr, nn, ww);
            d = ww;
        }
    }
    free(p); free(cdf); free(rgbl); free(cnt);
    return m;
}

/* combinePhotonContributions for any material over given search results
 * (count[q] photons idx[q*k ..] at distances dist[q*k ..]): the photon's
 * radiance times the material's lightReflected(wo, -photon dir) in the
 * shading frame of nrm[q] (Scene.cu:607-634, SurfaceInteraction.cu:278-292);
 * Lambertian and Trowbridge only.  Sum in double, times 1 / (PI * dmax^2). */
void orc_photon_gather_mat(const double *rgb, const float *dir, uint32_t m, const uint32_t *count, const uint32_t *idx,
                           const float *dist, uint32_t k, const float *nrm_q, const float *view,
                           const orc_material *mats, double *out)
{
    for (uint32_t q = 0; q < m; ++q) {
        double acc[3] = {0, 0, 0};
        const orc_material *mt = &mats[q];
        const int scat = mt->type == ORC_MAT_LAMBERTIAN || mt->type == ORC_MAT_TROWBRIDGE ||
                         mt->type == ORC_MAT_TROWBRIDGE_BASIC;
        if (count[q] > 0 && scat) {
            v3 n = mk(nrm_q[3 * (size_t)q], nrm_q[3 * (size_t)q + 1], nrm_q[3 * (size_t)q + 2]), t, b;
            onb(n, &t, &b);
            v3 vw = mk(view[3 * (size_t)q], view[3 * (size_t)q + 1], view[3 * (size_t)q + 2]);
            v3 wo = mk(dot(vw, t), dot(vw, b), dot(vw, n));
            float maxd = 0;
            for (uint32_t i = 0; i < count[q]; ++i) {
                const uint32_t e = idx[(size_t)q * k + i];
                if (dist[(size_t)q * k + i] > maxd) maxd = dist[(size_t)q * k + i];
                v3 w = mk(-dir[3 * (size_t)e], -dir[3 * (size_t)e + 1], -dir[3 * (size_t)e + 2]);
                v3 wi = mk(dot(w, t), dot(w, b), dot(w, n));
                v3 f = bsdf_f(mt, wo, wi);
                acc[0] += rgb[3 * (size_t)e] * (double)f.x;
                acc[1] += rgb[3 * (size_t)e + 1] * (double)f.y;
                acc[2] += rgb[3 * (size_t)e + 2] * (double)f.z;
            }
            const float kernel = 1.0f / (3.14159265358979323846f * maxd * maxd);
            for (int c = 0; c < 3; ++c) acc[c] *= (double)kernel;
        }
        out[3 * (size_t)q] = acc[0]; out[3 * (size_t)q + 1] = acc[1]; out[3 * (size_t)q + 2] = acc[2];
    }
}

/* ------------------------------------------------------------------ */
/* The estimator with the reference's Spectrum precision                */
/* ------------------------------------------------------------------ */
/* path_impl with throughput, radiance and every light term in double, as
 * the reference's Spectrum stores RGB (radiometry/Spectrum.cu:15); geometry,
 * sampling and the BSDF stay float (the reference's Float).  It measures how
 * far the float estimator of the library (and of the oracle) is from the
 * reference's own arithmetic; it is not a parity target. */
typedef struct { double x, y, z; } d3;
static inline d3 dmk(double x, double y, double z) { d3 r = {x, y, z}; return r; }

static int path_f64(const orc_scene *s, uint32_t px, uint32_t py, uint32_t sample, uint32_t seed,
                    uint32_t max_depth, uint32_t rr_depth, uint64_t stream_base, double out[3])
{
    const uint32_t W = s->cam.width;
    pcg32 g = pcg_seed(orc_path_initstate(seed, sample), stream_base + (uint64_t)py * W + px);
    float u0 = pcg_float(&g), u1 = pcg_float(&g);
    float of[3], df[3];
    orc_camera_ray(s->r2c, s->c2w, s->cam.width, s->cam.height, px, py, u0, u1, of, df);
    v3 o = mk(of[0], of[1], of[2]), d = mk(df[0], df[1], df[2]);
    d3 beta = dmk(1, 1, 1), L = dmk(0, 0, 0);
    float prev_pdf = 0;
    int prev_delta = 1, camera = 1;
    v3 prev_p = o;
    for (uint32_t depth = 0; depth < max_depth; ++depth) {
        hit_rec h = scene_intersect(s, o, d);
        if (h.kind == 0) break;
        const orc_material *m =
            &s->mat[h.kind == 1 ? s->tri_mat[h.id] : h.kind == 2 ? s->sph_mat[h.id] : s->plane_mat[h.id]];
        orc_material textured;
        if (m->texture && (m->type == ORC_MAT_LAMBERTIAN || m->type == ORC_MAT_PLAIN)) {
            textured = *m;
            v3 c = texture_color(s, m->texture - 1, h.su, h.sv);
            textured.kd[0] = m->kd[0] * c.x;
            textured.kd[1] = m->kd[1] * c.y;
            textured.kd[2] = m->kd[2] * c.z;
            m = &textured;
        }
        v3 n = h.n, t, b;
        onb(n, &t, &b);
        v3 wo_w = neg(d);
        v3 wo = mk(dot(wo_w, t), dot(wo_w, b), dot(wo_w, n));
        if (is_scattering(m)) {
            d3 Ld = dmk(0, 0, 0);
            for (uint32_t l = 0; l < s->nlight; ++l) {
                v3 lp = mk(s->pl[l].pos[0], s->pl[l].pos[1], s->pl[l].pos[2]);
                v3 tol = sub(lp, h.p);
                v3 Lw = nrm(tol);
                float d2 = dot(tol, tol);
                /* color * intensity * lightDecay (Float 1 / d^2), as a Spectrum */
                const double I = s->pl[l].intensity, dec = (double)(1 / d2);
                d3 Li = dmk(s->pl[l].color[0] * I * dec, s->pl[l].color[1] * I * dec, s->pl[l].color[2] * I * dec);
                if (!(dot(Lw, n) > 0)) continue;
                v3 wi = mk(dot(Lw, t), dot(Lw, b), dot(Lw, n));
                v3 f = bsdf_f(m, wo, wi);
                float pf = bsdf_pdf(m, wo, wi);
                if ((f.x == 0 && f.y == 0 && f.z == 0) || (Li.x == 0 && Li.y == 0 && Li.z == 0)) continue;
                v3 so = spawn_origin(h.p, h.perr, n, Lw);
                if (occluded(s, so, Lw, len(tol), -1)) continue;
                const double w = 1.0f / (1.0f + pf * pf);
                Ld = dmk(Ld.x + f.x * Li.x * w, Ld.y + f.y * Li.y * w, Ld.z + f.z * Li.z * w);
            }
            for (uint32_t j = 0; j < s->nsph; ++j) {
                if (!s->sph_light[j]) continue;
                float a0 = pcg_float(&g), a1 = pcg_float(&g);
                v3 Lw;
                float dist, pl;
                if (!sphere_light_sample(&s->sph[j], h.p, a0, a1, &Lw, &dist, &pl)) continue;
                if (!(dot(Lw, n) > 0)) continue;
                v3 wi = mk(dot(Lw, t), dot(Lw, b), dot(Lw, n));
                v3 f = bsdf_f(m, wo, wi);
                if (f.x == 0 && f.y == 0 && f.z == 0) continue;
                float pf = bsdf_pdf(m, wo, wi);
                v3 so = spawn_origin(h.p, h.perr, n, Lw);
                if (occluded(s, so, Lw, dist * 0.9999f, (int)j)) continue;
                const orc_material *lm = &s->mat[s->sph_mat[j]];
                const double w = pl * pl / (pl * pl + pf * pf), k = w / pl, I = lm->intensity;
                Ld = dmk(Ld.x + f.x * (lm->kd[0] * I) * k, Ld.y + f.y * (lm->kd[1] * I) * k,
                         Ld.z + f.z * (lm->kd[2] * I) * k);
            }
            L = dmk(L.x + beta.x * Ld.x, L.y + beta.y * Ld.y, L.z + beta.z * Ld.z);
        }
        if (m->type == ORC_MAT_LIGHT) {
            double w = 1;
            if (h.kind == 2 && s->sph_light[h.id] && !prev_delta) {
                float pl = sphere_light_pdf(&s->sph[h.id], prev_p);
                w = (double)prev_pdf * prev_pdf / ((double)prev_pdf * prev_pdf + (double)pl * pl);
            }
            const double I = m->intensity;
            L = dmk(L.x + beta.x * m->kd[0] * I * w, L.y + beta.y * m->kd[1] * I * w, L.z + beta.z * m->kd[2] * I * w);
        }
        if (m->type == ORC_MAT_PLAIN) {
            const float *e = camera ? m->kd : m->ke;
            L = dmk(L.x + beta.x * e[0], L.y + beta.y * e[1], L.z + beta.z * e[2]);
        }
        v3 wi;
        float pdf;
        int delta;
        v3 wgt = bsdf_sample(m, wo, &g, &wi, &pdf, &delta);
        beta = dmk(beta.x * wgt.x, beta.y * wgt.y, beta.z * wgt.z);
        if (wgt.x == 0 && wgt.y == 0 && wgt.z == 0) break;
        if (is_scattering(m) && depth > rr_depth) {
            /* Spectrum::brightness in double, then q = max(0.05, 1 - it) */
            const double q = fmax(0.05, 1 - (0.212671f * beta.x + 0.715160f * beta.y + 0.072169f * beta.z));
            if (pcg_float(&g) < q) break;
            beta = dmk(beta.x / (1 - q), beta.y / (1 - q), beta.z / (1 - q));
        }
        v3 ww = add(add(mul(t, wi.x), mul(b, wi.y)), mul(n, wi.z));
        o = spawn_origin(h.p, h.perr, n, ww);
        d = ww;
        prev_pdf = pdf;
        prev_delta = delta;
        prev_p = h.p;
        camera = camera && is_specular(m);
    }
    out[0] = L.x; out[1] = L.y; out[2] = L.z;
    return isfinite(L.x) && isfinite(L.y) && isfinite(L.z);
}

/* orc_render over the window with path_f64 (one thread per row block). */
typedef struct {
    const orc_scene *s;
    uint32_t spp, seed, max_depth, rr_depth, x0, y0, x1, y1;
    double *accum;
    uint32_t *nvalid;
} render64_ctx;

static void *render64_rows(void *p)
{
    render64_ctx *c = p;
    const uint32_t W = c->s->cam.width;
    for (uint32_t y = c->y0; y < c->y1; ++y)
        for (uint32_t x = c->x0; x < c->x1; ++x) {
            double acc[3] = {0, 0, 0};
            uint32_t nv = 0;
            for (uint32_t k = 0; k < c->spp; ++k) {
                double L[3];
                if (path_f64(c->s, x, y, k, c->seed, c->max_depth, c->rr_depth, 0, L)) {
                    acc[0] += L[0]; acc[1] += L[1]; acc[2] += L[2];
                    ++nv;
                }
            }
            size_t i = (size_t)y * W + x;
            c->accum[3 * i] += acc[0];
            c->accum[3 * i + 1] += acc[1];
            c->accum[3 * i + 2] += acc[2];
            c->nvalid[i] += nv;
        }
    return NULL;
}

void orc_render_f64(const orc_scene *s, uint32_t spp, uint32_t seed, uint32_t max_depth, uint32_t rr_depth,
                    uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1, double *accum, uint32_t *nvalid, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    uint32_t rows = y1 > y0 ? y1 - y0 : 0;
    if ((uint32_t)nthreads > rows) nthreads = rows ? (int)rows : 1;
    pthread_t *th = malloc(sizeof(pthread_t) * nthreads);
    render64_ctx *cs = malloc(sizeof(render64_ctx) * nthreads);
    uint32_t chunk = (rows + nthreads - 1) / nthreads;
    for (int i = 0; i < nthreads; ++i) {
        render64_ctx c = {s, spp, seed, max_depth, rr_depth, x0, y0 + i * chunk, x1, y0 + (i + 1) * chunk, accum, nvalid};
        if (c.y0 > y1) c.y0 = y1;
        if (c.y1 > y1) c.y1 = y1;
        cs[i] = c;
        pthread_create(&th[i], NULL, render64_rows, &cs[i]);
    }
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th); free(cs);
}
