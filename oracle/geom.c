/*
 * geom.c -- oracle (test infrastructure only): ray/triangle, ray/box and
 * ray/sphere tests restated from the reference in FP32 with the reference's
 * operation order.  Built with -ffp-contract=off and no fast-math.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <string.h>

/* cuda_libraries/types.h:23: MachineEpsilon = FLT_EPSILON * 0.5, a double
 * expression; errorBound is evaluated in double and returned as float
 * (math.cu:117-120). */
float orc_error_bound(int n)
{
    const double me = (double)FLT_EPSILON * 0.5;
    return (float)((n * me) / (1 - n * me));
}

/* math.cu:147-170 */
float orc_next_float_up(float x)
{
    if (isinf(x) && x > 0.0f) return x;
    if (x == -0.0f) x = 0.0f;
    uint32_t ui;
    memcpy(&ui, &x, 4);
    if (x >= 0.0f) ++ui; else --ui;
    memcpy(&x, &ui, 4);
    return x;
}

/* math.cu:172-194 */
float orc_next_float_down(float x)
{
    if (isinf(x) && x < 0.0f) return x;
    if (x == 0.0f) x = -0.0f;
    uint32_t ui;
    memcpy(&ui, &x, 4);
    if (x > 0.0f) --ui; else ++ui;
    memcpy(&x, &ui, 4);
    return x;
}

/* geometry/Ray.cu:32-39: invDirection = (1/d.x, 1/d.y, 1/d.z). */
void orc_ray_make(orc_ray *r, const float o[3], const float d[3], float tmax)
{
    for (int k = 0; k < 3; ++k) {
        r->o[k] = o[k];
        r->d[k] = d[k];
        r->inv[k] = 1.0f / d[k];
    }
    r->tmax = tmax;
}

static inline float max3f(float a, float b, float c) { return fmaxf(a, fmaxf(b, c)); }

/*
 * Triangle::intersects, shapes/Triangle.cu:433-643.
 *  - translate to the ray origin (:452-454), permute by the largest |d|
 *    component (:457-484), shear (:487-495);
 *  - FP32 edge functions with an FP64 recompute when any is exactly zero
 *    (:500-522);
 *  - mixed-sign rejection (:527-532), det == 0 rejection (:535-537);
 *  - scaled distance range test against tMax (:540-558);
 *  - t = scaledOffset * (1/det); reject if t > hitOffset, else LOWER
 *    hitOffset (:563-573) -- before the deltaT test;
 *  - conservative deltaT rejection (:575-593).
 * The tail (dpdu/dpdv, alpha stub) never rejects: the alpha stub returns
 * true (:154-158) and the derivative code has no rejecting branch.
 *
 * `guarded` applies the deltaT rejection before hitOffset is lowered (the
 * library's order): a candidate that fails it (typically the triangle a
 * spawned ray just left) then leaves hitOffset alone, so the result is the
 * closest candidate that passes, whatever the visit order.  In the
 * reference's order such a candidate lowers hitOffset first, and every
 * triangle tested after it is rejected.
 */
static int tri_test_core(const orc_ray *r, const float *wp0, const float *wp1, const float *wp2,
                         float *hitOffset, orc_tri_info *info, int guarded)
{
    if (info) { info->reached = 0; info->masked = 0; info->fp64 = 0; info->t = INFINITY; }

    float l0[3], l1[3], l2[3];
    for (int k = 0; k < 3; ++k) {
        l0[k] = wp0[k] - r->o[k];
        l1[k] = wp1[k] - r->o[k];
        l2[k] = wp2[k] - r->o[k];
    }
    float ad[3] = {fabsf(r->d[0]), fabsf(r->d[1]), fabsf(r->d[2])};
    int kz = (ad[0] > ad[1]) ? 0 : 1;
    if (ad[2] > ad[kz]) kz = 2;
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;

    float dx = r->d[kx], dy = r->d[ky], dz = r->d[kz];
    float p0x = l0[kx], p0y = l0[ky], p0z = l0[kz];
    float p1x = l1[kx], p1y = l1[ky], p1z = l1[kz];
    float p2x = l2[kx], p2y = l2[ky], p2z = l2[kz];

    float Sx = -dx / dz;
    float Sy = -dy / dz;
    float Sz = 1.f / dz;
    p0x += Sx * p0z; p0y += Sy * p0z;
    p1x += Sx * p1z; p1y += Sy * p1z;
    p2x += Sx * p2z; p2y += Sy * p2z;

    float e0 = p1x * p2y - p1y * p2x;
    float e1 = p2x * p0y - p2y * p0x;
    float e2 = p0x * p1y - p0y * p1x;
    int used64 = 0;
    if (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f) {
        double p2txp1ty = (double)p2x * (double)p1y;
        double p2typ1tx = (double)p2y * (double)p1x;
        e0 = (float)(p2typ1tx - p2txp1ty);
        double p0txp2ty = (double)p0x * (double)p2y;
        double p0typ2tx = (double)p0y * (double)p2x;
        e1 = (float)(p0typ2tx - p0txp2ty);
        double p1txp0ty = (double)p1x * (double)p0y;
        double p1typ0tx = (double)p1y * (double)p0x;
        e2 = (float)(p1typ0tx - p1txp0ty);
        used64 = 1;
    }
    if ((e0 < 0 || e1 < 0 || e2 < 0) && (e0 > 0 || e1 > 0 || e2 > 0)) return 0;
    float det = e0 + e1 + e2;
    if (det == 0) return 0;

    p0z *= Sz; p1z *= Sz; p2z *= Sz;
    float scaled = e0 * p0z + e1 * p1z + e2 * p2z;
    if (det < 0 && (scaled >= 0 || scaled < r->tmax * det)) return 0;
    else if (det > 0 && (scaled <= 0 || scaled > r->tmax * det)) return 0;

    float invDet = 1 / det;
    float t = scaled * invDet;
    if (info) { info->reached = 1; info->t = t; info->fp64 = used64; }
    if (t > *hitOffset) return 0;
    if (!guarded) *hitOffset = t;

    float maxZt = max3f(fabsf(p0z), fabsf(p1z), fabsf(p2z));
    float deltaZ = orc_error_bound(3) * maxZt;
    float maxXt = max3f(fabsf(p0x), fabsf(p1x), fabsf(p2x));
    float maxYt = max3f(fabsf(p0y), fabsf(p1y), fabsf(p2y));
    float deltaX = orc_error_bound(5) * (maxXt + maxZt);
    float deltaY = orc_error_bound(5) * (maxYt + maxZt);
    float deltaE = 2 * (orc_error_bound(2) * maxXt * maxYt + deltaY * maxXt + deltaX * maxYt);
    float maxE = max3f(fabsf(e0), fabsf(e1), fabsf(e2));
    float deltaT = 3 * (orc_error_bound(3) * maxE * maxZt + deltaE * maxZt + deltaZ * maxE) * fabsf(invDet);
    if (t <= deltaT) {
        if (info) info->masked = 1;
        return 0;
    }
    *hitOffset = t;
    return 1;
}

int orc_tri_intersects(const orc_ray *r, const float *wp0, const float *wp1,
                       const float *wp2, float *hitOffset, orc_tri_info *info)
{
    return tri_test_core(r, wp0, wp1, wp2, hitOffset, info, 0);
}

int orc_tri_intersects_guarded(const orc_ray *r, const float *wp0, const float *wp1, const float *wp2,
                               float *hitOffset, orc_tri_info *info)
{
    return tri_test_core(r, wp0, wp1, wp2, hitOffset, info, 1);
}

/* Bound3f::intersects, geometry/BoundingBox.cu:296-329.  Unqualified
 * min/max on floats resolve to fminf/fmaxf under nvcc. */
int orc_box_intersects(const float lo[3], const float hi[3], const orc_ray *r,
                       float *hitOffset)
{
    float tx1 = (lo[0] - r->o[0]) * r->inv[0];
    float tx2 = (hi[0] - r->o[0]) * r->inv[0];
    float tmin = fminf(tx1, tx2);
    float tmax = fmaxf(tx1, tx2);
    float ty1 = (lo[1] - r->o[1]) * r->inv[1];
    float ty2 = (hi[1] - r->o[1]) * r->inv[1];
    tmin = fmaxf(tmin, fminf(ty1, ty2));
    tmax = fminf(tmax, fmaxf(ty1, ty2));
    float tz1 = (lo[2] - r->o[2]) * r->inv[2];
    float tz2 = (hi[2] - r->o[2]) * r->inv[2];
    tmin = fmaxf(tmin, fminf(tz1, tz2));
    tmax = fminf(tmax, fmaxf(tz1, tz2));
    float h = (tmin > 0) ? tmin : tmax;
    if (hitOffset) *hitOffset = h;
    return tmax >= tmin && h >= 0 && h <= r->tmax;
}

/* ------------------------------------------------------------------ */
/* Sphere                                                               */
/* ------------------------------------------------------------------ */

/* Matrix4x4 operator* (cuda_libraries/Matrix4x4.cu:69-86): accumulate from
 * zero in k order. */
static void mat4_mul(const float *a, const float *b, float *out)
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

/* MESA-derived cofactor inverse, Matrix4x4.cu:303-440 (operand order kept). */
void orc_mat4_inverse(const float *m, float *out)
{
    float inv[16], det;
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    if (det == 0) { memset(out, 0, 16 * sizeof(float)); return; }
    det = (float)(1.0 / det);
    for (int i = 0; i < 16; ++i) out[i] = inv[i] * det;
}

/* Sphere(center, radius), shapes/Sphere.cu:233-245:
 * objectToWorld = translate(c) * rotateZ(90); worldToObject = inverse().
 * rotateZ (Transform.cu:536-547) uses sin/cos of radians(90) =
 * 90 * PI / 180.0 with PI the float constant (math.cu:66-68). */
void orc_sphere_init(orc_sphere *s, const float c[3], float radius)
{
    const float PIf = 3.14159265358979323846f;
    float ang = (float)((90.0f * PIf) / 180.0);
    float st = sinf(ang), ct = cosf(ang);
    float T[16] = {1, 0, 0, c[0], 0, 1, 0, c[1], 0, 0, 1, c[2], 0, 0, 0, 1};
    float R[16] = {ct, -st, 0, 0, st, ct, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    mat4_mul(T, R, s->o2w);
    orc_mat4_inverse(s->o2w, s->w2o);
    s->radius = radius;
    s->center[0] = c[0]; s->center[1] = c[1]; s->center[2] = c[2];
}

typedef struct { float v, e; } efloat; /* ErrorFloat, types.h:92-211 */

static inline efloat ef(float v, float e) { efloat r = {v, e}; return r; }
static inline efloat ef_add(efloat a, efloat b) { return ef(a.v + b.v, a.e + b.e); }
static inline efloat ef_sub(efloat a, efloat b) { return ef(a.v - b.v, a.e + b.e); }
static inline efloat ef_mul(efloat a, efloat b)
{
    efloat r; r.v = a.v * b.v; r.e = fabsf(r.v) * b.e + fabsf(b.v) * a.e; return r;
}
static inline efloat ef_muls(efloat a, float s) /* ErrorFloat * Float */
{
    efloat r; r.v = a.v * s; r.e = fabsf(r.v) * a.e; return r;
}
static inline efloat ef_smul(float s, efloat a) /* Float * ErrorFloat */
{
    efloat r; r.v = s * a.v; r.e = fabsf(r.v) * a.e; return r;
}
static inline efloat ef_div(efloat a, efloat b)
{
    efloat r; r.v = a.v / b.v; r.e = (fabsf(r.v) * b.e + fabsf(b.v) * a.e) / (b.v * b.v); return r;
}
static inline efloat ef_sqrt(efloat a)
{
    efloat r; r.v = sqrtf(a.v); r.e = (float)(0.5 * a.e / r.v); return r;
}

/* quadratic(), cuda_libraries/math.cu:91-115. */
static int ef_quadratic(efloat a, efloat b, efloat c, efloat *s0, efloat *s1)
{
    efloat disc = ef_sub(ef_mul(b, b), ef_mul(ef_smul(4.0f, a), c));
    if (disc.v < 0.0f) return 0;
    efloat rd = ef_sqrt(disc);
    efloat q;
    if (b.v < 0.0f) q = ef_smul(-0.5f, ef_sub(b, rd));
    else q = ef_smul(-0.5f, ef_add(b, rd));
    *s0 = ef_div(q, a);
    *s1 = ef_div(c, q);
    if (s0->v > s1->v) { efloat t = *s0; *s0 = *s1; *s1 = t; }
    return !(s0->v != s0->v) && !(s1->v != s1->v);
}

/* Transform::operator()(Point, absError) const (Transform.cu:335-350). */
static void xform_point_err(const float *m, const float p[3], float out[3], float err[3])
{
    float eb3 = orc_error_bound(3);
    for (int i = 0; i < 3; ++i) {
        float s = fabsf(m[i * 4 + 0] * p[0]) + fabsf(m[i * 4 + 1] * p[1]) + fabsf(m[i * 4 + 2] * p[2]) + fabsf(m[i * 4 + 3]);
        err[i] = s * eb3;
    }
    float xp = m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3];
    float yp = m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7];
    float zp = m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11];
    float wp = m[12] * p[0] + m[13] * p[1] + m[14] * p[2] + m[15];
    float inv = 1.0f / wp; /* Point3::operator/ (Point.cu:147-151) */
    out[0] = inv * xp; out[1] = inv * yp; out[2] = inv * zp;
}

/* Transform::operator()(Vector, absError) const (Transform.cu:363-378). */
static void xform_vec_err(const float *m, const float v[3], float out[3], float err[3])
{
    float eb3 = orc_error_bound(3);
    for (int i = 0; i < 3; ++i) {
        float s = fabsf(m[i * 4 + 0] * v[0]) + fabsf(m[i * 4 + 1] * v[1]) + fabsf(m[i * 4 + 2] * v[2]);
        err[i] = s * eb3;
    }
    out[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
    out[1] = m[4] * v[0] + m[5] * v[1] + m[6] * v[2];
    out[2] = m[8] * v[0] + m[9] * v[1] + m[10] * v[2];
}

/* worldToObject(ray, oErr, dErr) (Transform.cu:386-402) followed by the
 * direction renormalisation of Sphere.cu:287/371 and the quadratic of
 * Sphere.cu:18-39, then checkRayBounds (Sphere.cu:41-81).  The clipping
 * test (Sphere.cu:83-124) never fires for a full sphere built by
 * Sphere(center, radius): zMin = -r, zMax = r and phiMax = 2*PI, while
 * phi is in [0, 2*PI]. */
static int sphere_core(const orc_sphere *s, const orc_ray *r, float *t_hit,
                       float ro[3], float rd[3])
{
    float oErr[3], dErr[3], o[3], d[3];
    xform_point_err(s->w2o, r->o, o, oErr);
    xform_vec_err(s->w2o, r->d, d, dErr);
    float tMax = r->tmax;
    float ls = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (ls > 0) {
        float dt = (fabsf(d[0]) * oErr[0] + fabsf(d[1]) * oErr[1] + fabsf(d[2]) * oErr[2]) / ls;
        o[0] += d[0] * dt; o[1] += d[1] * dt; o[2] += d[2] * dt;
        tMax -= dt;
    }
    /* normalize(): v / length() with Vector3::operator/ multiplying by 1/len */
    float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    float il = (float)1 / len;
    d[0] = d[0] * il; d[1] = d[1] * il; d[2] = d[2] * il;

    efloat ox = ef(o[0], oErr[0]), oy = ef(o[1], oErr[1]), oz = ef(o[2], oErr[2]);
    efloat dx = ef(d[0], dErr[0]), dy = ef(d[1], dErr[1]), dz = ef(d[2], dErr[2]);
    efloat a = ef_add(ef_add(ef_mul(dx, dx), ef_mul(dy, dy)), ef_mul(dz, dz));
    efloat b = ef_smul(2.0f, ef_add(ef_add(ef_mul(dx, ox), ef_mul(dy, oy)), ef_mul(dz, oz)));
    efloat rr = ef(s->radius, 0);
    efloat c = ef_sub(ef_add(ef_add(ef_mul(ox, ox), ef_mul(oy, oy)), ef_mul(oz, oz)), ef_mul(rr, rr));
    efloat t0, t1;
    if (!ef_quadratic(a, b, c, &t0, &t1)) return 0;
    if (t0.v + t0.e > tMax || t1.v - t1.e <= 0) return 0;
    efloat th = t0;
    if (t0.v - t0.e <= 0) {
        th = t1;
        if (t1.v + t1.e > tMax) return 0;
    }
    *t_hit = th.v;
    if (ro) { ro[0] = o[0]; ro[1] = o[1]; ro[2] = o[2]; }
    if (rd) { rd[0] = d[0]; rd[1] = d[1]; rd[2] = d[2]; }
    return 1;
}

int orc_sphere_intersects(const orc_sphere *s, const orc_ray *r, float *hitOffset)
{
    float th;
    if (!sphere_core(s, r, &th, NULL, NULL)) return 0;
    if (th < *hitOffset) { *hitOffset = th; return 1; } /* strict, Sphere.cu:390 */
    return 0;
}

/* Sphere::intersect (Sphere.cu:277-355): the object hit point is
 * reprojected onto the sphere (:66-69); the world point and its error come
 * from transform(objectToWorld, SurfaceInteraction)
 * (SurfaceInteraction.cu:339-357), whose point overload OVERWRITES pError
 * with errorBound(3)-scaled absolute sums (Transform.cu:335-350).  The
 * normal is the outward normal (Sphere.cu:160), carried to world space by
 * the inverse transpose (Transform.cu:201-212). */
/* Sphere::localToSurface (Sphere.cu:269-275) of an object-space point:
 * (theta, phi) in degrees (to_spherical, geometricMath.cu:49-57). */
static void sphere_surface_point(const float hp[3], float uv[2])
{
    float th = atan2f(sqrtf(hp[0] * hp[0] + hp[1] * hp[1]), hp[2]);
    float ph = atan2f(hp[1], hp[0]);
    uv[0] = (float)((double)th * 180.0 / 3.14159265358979323846);
    uv[1] = (float)((double)ph * 180.0 / 3.14159265358979323846);
}

static int sphere_intersect_impl(const orc_sphere *s, const orc_ray *r, float *hitOffset, float p[3], float n[3],
                                 float perr[3], float uv[2]);

int orc_sphere_intersect(const orc_sphere *s, const orc_ray *r, float *hitOffset,
                         float p[3], float n[3], float perr[3])
{
    return sphere_intersect_impl(s, r, hitOffset, p, n, perr, NULL);
}

int orc_sphere_intersect_uv(const orc_sphere *s, const orc_ray *r, float *hitOffset, float p[3], float n[3],
                            float perr[3], float uv[2])
{
    return sphere_intersect_impl(s, r, hitOffset, p, n, perr, uv);
}

static int sphere_intersect_impl(const orc_sphere *s, const orc_ray *r, float *hitOffset, float p[3], float n[3],
                                 float perr[3], float uv[2])
{
    float th, o[3], d[3];
    if (!sphere_core(s, r, &th, o, d)) return 0;
    if (!(th < *hitOffset)) return 0;
    *hitOffset = th;
    float hp[3] = {o[0] + d[0] * th, o[1] + d[1] * th, o[2] + d[2] * th};
    float dist = sqrtf(hp[0] * hp[0] + hp[1] * hp[1] + hp[2] * hp[2]);
    float sc = s->radius / dist;
    hp[0] *= sc; hp[1] *= sc; hp[2] *= sc;
    if (hp[0] == 0 && hp[1] == 0) hp[0] = 1e-5f * s->radius;
    if (uv) sphere_surface_point(hp, uv);
    xform_point_err(s->o2w, hp, p, perr);
    /* outward object normal = hp / r; world normal via inverse transpose of
     * o2w, i.e. w2o^T */
    float on[3] = {hp[0], hp[1], hp[2]};
    float l = sqrtf(on[0] * on[0] + on[1] * on[1] + on[2] * on[2]);
    float il = 1.0f / l;
    on[0] *= il; on[1] *= il; on[2] *= il;
    const float *mi = s->w2o;
    n[0] = mi[0] * on[0] + mi[4] * on[1] + mi[8] * on[2];
    n[1] = mi[1] * on[0] + mi[5] * on[1] + mi[9] * on[2];
    n[2] = mi[2] * on[0] + mi[6] * on[1] + mi[10] * on[2];
    return 1;
}

/* Plane(bottomLeftCorner, axisX, axisY, normal) (shapes/Plane.cu:27-48):
 * the axis lengths are truncated to integers (`int size_X = axisX.length()`),
 * the normal is cross(axisX, axisY) turned towards the given one and
 * normalised, objectToWorld = baseChange(corner, axisX, axisY, normal)
 * (Transform.cu:551-562) and worldToObject its cofactor inverse. */
void orc_plane_init(orc_plane *pl, const float corner[3], const float ax[3], const float ay[3], const float n[3])
{
    float lx = sqrtf(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
    float ly = sqrtf(ay[0] * ay[0] + ay[1] * ay[1] + ay[2] * ay[2]);
    int sx = (int)lx, sy = (int)ly;
    float c[3] = {(ax[1] * ay[2]) - (ax[2] * ay[1]), (ax[2] * ay[0]) - (ax[0] * ay[2]),
                  (ax[0] * ay[1]) - (ax[1] * ay[0])};
    float dn = (c[0] * n[0]) + (c[1] * n[1]) + (c[2] * n[2]);
    float nn[3];
    for (int k = 0; k < 3; ++k) nn[k] = dn > 0 ? c[k] : -c[k];
    /* normalize(v) = v * (1 / length) (Vector.cu:454-457) */
    float ix = 1.0f / lx, iy = 1.0f / ly;
    float X[3] = {ax[0] * ix * (float)sx, ax[1] * ix * (float)sx, ax[2] * ix * (float)sx};
    float Y[3] = {ay[0] * iy * (float)sy, ay[1] * iy * (float)sy, ay[2] * iy * (float)sy};
    float ln = sqrtf(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
    float in = 1.0f / ln;
    float N[3] = {nn[0] * in, nn[1] * in, nn[2] * in};
    float m[16] = {X[0], Y[0], N[0], corner[0], X[1], Y[1], N[1], corner[1], X[2], Y[2], N[2], corner[2], 0, 0, 0, 1};
    memcpy(pl->o2w, m, sizeof m);
    orc_mat4_inverse(pl->o2w, pl->w2o);
}

/* Transform::operator()(Point, pError) (Transform.cu:239-263): the error
 * ADDS errorBound(4) * |row| . |p| * |w| to pError. */
static void xform_point_add_err4(const float *m, const float p[3], float out[3], float err[3])
{
    float eb4 = orc_error_bound(4);
    float xp = m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3];
    float yp = m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7];
    float zp = m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11];
    float wp = m[12] * p[0] + m[13] * p[1] + m[14] * p[2] + m[15];
    for (int i = 0; i < 3; ++i) {
        float sabs = fabsf(m[i * 4 + 0] * p[0]) + fabsf(m[i * 4 + 1] * p[1]) + fabsf(m[i * 4 + 2] * p[2]) +
                     fabsf(m[i * 4 + 3]);
        err[i] += sabs * eb4 * fabsf(wp);
    }
    float inv = 1.0f / wp;
    out[0] = inv * xp; out[1] = inv * yp; out[2] = inv * zp;
}

/* Plane::intersect / intersects (Plane.cu:58-164).  The ray goes to object
 * space with Transform::operator()(Ray) (Transform.cu:216-237: the origin
 * moves to the edge of its error bounds along the direction, which is not
 * normalised, and tMax shrinks by the same amount); a ray whose |d.z| is not
 * above FLOAT_ERROR_MARGIN (1e-7, types.h:39) misses; the distance is
 * Ray::getOffset, (hit.x - origin.x) * invDirection.x (Ray.cu:69-72), so a
 * ray with local d.x == 0 gets NaN and never hits; the hit must lie in
 * [0,1]^2 with 0 < offset <= tMax and offset < hitOffset (strict).  The
 * surface: world point with pError = 1e-4 per axis plus the transform's
 * error, normal (0,0,1) turned against the local direction, taken to world
 * space through the inverse transpose (the rows of worldToObject). */
static int plane_intersect_impl(const orc_plane *pl, const orc_ray *r, float *hitOffset, float p[3], float n[3],
                                float perr[3], float uv[2]);

int orc_plane_intersect(const orc_plane *pl, const orc_ray *r, float *hitOffset, float p[3], float n[3],
                        float perr[3])
{
    return plane_intersect_impl(pl, r, hitOffset, p, n, perr, NULL);
}

int orc_plane_intersect_uv(const orc_plane *pl, const orc_ray *r, float *hitOffset, float p[3], float n[3],
                           float perr[3], float uv[2])
{
    return plane_intersect_impl(pl, r, hitOffset, p, n, perr, uv);
}

static int plane_intersect_impl(const orc_plane *pl, const orc_ray *r, float *hitOffset, float p[3], float n[3],
                                float perr[3], float uv[2])
{
    float oerr[3] = {0, 0, 0}, o[3], d[3];
    xform_point_add_err4(pl->w2o, r->o, o, oerr);
    const float *m = pl->w2o;
    d[0] = m[0] * r->d[0] + m[1] * r->d[1] + m[2] * r->d[2];
    d[1] = m[4] * r->d[0] + m[5] * r->d[1] + m[6] * r->d[2];
    d[2] = m[8] * r->d[0] + m[9] * r->d[1] + m[10] * r->d[2];
    float tmax = r->tmax;
    float ls = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (ls > 0) {
        float dt = (fabsf(d[0]) * oerr[0] + fabsf(d[1]) * oerr[1] + fabsf(d[2]) * oerr[2]) / ls;
        o[0] += d[0] * dt; o[1] += d[1] * dt; o[2] += d[2] * dt;
        tmax -= dt;
    }
    float invx = 1.0f / d[0];
    float cosine = d[2];
    if (!(fabsf(cosine) > 0.0000001)) return 0;
    float sc = o[2] / d[2];
    float hp[3] = {o[0] - sc * d[0], o[1] - sc * d[1], o[2] - sc * d[2]};
    float off = (hp[0] - o[0]) * invx;
    float u = hp[0], v = hp[1];
    if (off <= 0 || off > tmax || u < 0 || u > 1 || v < 0 || v > 1) return 0;
    if (!(off < *hitOffset)) return 0;
    *hitOffset = off;
    if (uv) { uv[0] = u; uv[1] = v; }
    if (p) {
        float on = d[2] > 0.0 ? -1.0f : 1.0f; /* sameHemisphere((0,0,1), d): d.z > 0 */
        float loc[3] = {u, v, 0};
        perr[0] = perr[1] = perr[2] = 0.0001f;
        xform_point_add_err4(pl->o2w, loc, p, perr);
        n[0] = m[8] * on; n[1] = m[9] * on; n[2] = m[10] * on;
    }
    return 1;
}

/* Barycentrics of the watertight test, b_i = e_i * (1/det)
 * (shapes/Triangle.cu:161-290, the same steps as orc_tri_intersects). */
void orc_tri_barycentrics(const orc_ray *r, const float *wp0, const float *wp1, const float *wp2, float b[3])
{
    float l0[3], l1[3], l2[3];
    for (int k = 0; k < 3; ++k) {
        l0[k] = wp0[k] - r->o[k];
        l1[k] = wp1[k] - r->o[k];
        l2[k] = wp2[k] - r->o[k];
    }
    float ad[3] = {fabsf(r->d[0]), fabsf(r->d[1]), fabsf(r->d[2])};
    int kz = (ad[0] > ad[1]) ? 0 : 1;
    if (ad[2] > ad[kz]) kz = 2;
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    float dx = r->d[kx], dy = r->d[ky], dz = r->d[kz];
    float p0x = l0[kx], p0y = l0[ky], p0z = l0[kz];
    float p1x = l1[kx], p1y = l1[ky], p1z = l1[kz];
    float p2x = l2[kx], p2y = l2[ky], p2z = l2[kz];
    float Sx = -dx / dz, Sy = -dy / dz;
    p0x += Sx * p0z; p0y += Sy * p0z;
    p1x += Sx * p1z; p1y += Sy * p1z;
    p2x += Sx * p2z; p2y += Sy * p2z;
    float e0 = p1x * p2y - p1y * p2x;
    float e1 = p2x * p0y - p2y * p0x;
    float e2 = p0x * p1y - p0y * p1x;
    if (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f) {
        e0 = (float)((double)p2y * (double)p1x - (double)p2x * (double)p1y);
        e1 = (float)((double)p0y * (double)p2x - (double)p0x * (double)p2y);
        e2 = (float)((double)p1y * (double)p0x - (double)p1x * (double)p0y);
    }
    float det = e0 + e1 + e2;
    float invDet = 1 / det;
    b[0] = e0 * invDet;
    b[1] = e1 * invDet;
    b[2] = e2 * invDet;
}
