/*
 * bvh.c -- oracle (test infrastructure only): the reference's Morton-pairing
 * BVH and its stack traversal, the brute-force closest hit with the
 * order-independence analysis, and a counting traversal of the library's
 * exported two-wide node layout.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Reference BVH                                                        */
/* ------------------------------------------------------------------ */
typedef struct {
    uint32_t left, right;
    float lo[3], hi[3];
    int32_t object; /* -1 = internal (BVH.cu:95-110) */
} ref_node;

struct orc_refbvh {
    uint32_t nt, root;
    ref_node *nodes;
    float *tri;          /* 9 floats per triangle, input order */
    uint32_t *morton;    /* per input triangle */
    uint32_t *sorted;    /* leaf order */
};

/* expandBits / morton3D, cuda_libraries/geometricMath.cu:16-44.  The clamps
 * are double expressions assigned back to Float. */
static uint32_t expand_bits(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

static uint32_t morton3d(float x, float y, float z)
{
    x = (float)fmin(fmax(x * 1024.0, 0.0), 1023.0);
    y = (float)fmin(fmax(y * 1024.0, 0.0), 1023.0);
    z = (float)fmin(fmax(z * 1024.0, 0.0), 1023.0);
    return expand_bits((uint32_t)x) * 4 + expand_bits((uint32_t)y) * 2 + expand_bits((uint32_t)z);
}

typedef struct { uint32_t code, id; } kv;
static int kv_cmp(const void *a, const void *b)
{
    const kv *x = a, *y = b;
    if (x->code != y->code) return x->code < y->code ? -1 : 1;
    return x->id < y->id ? -1 : (x->id > y->id);
}

/* BVH ctor (BVH.cu:314-341) with sort_by_morton (Triangle.cu:737-776):
 *  - world bound = reduce of triangle bounds with the ZERO box as the
 *    initial value (BoundingBox.cu:16-21), so the origin is always inside;
 *  - centroid = (p0+p1+p2)/3 with Point3::operator/ multiplying by 1/3f;
 *  - unit-cube coordinates (BoundingBox.cu:176-183), Morton code;
 *  - lexicographic sort of (code, id) tuples;
 *  - leaves 0..n-1 in sorted order, parent of node i = n + i/2
 *    (generateHierarchy, BVH.cu:153-201), root = 2n-2;
 *  - bounds bottom-up in id order (BVH.cu:203-212). */
orc_refbvh *orc_refbvh_build(const float *xyz, const uint32_t *idx, uint32_t nt)
{
    orc_refbvh *b = calloc(1, sizeof *b);
    b->nt = nt;
    b->tri = malloc(sizeof(float) * 9 * (nt ? nt : 1));
    b->morton = malloc(sizeof(uint32_t) * (nt ? nt : 1));
    b->sorted = malloc(sizeof(uint32_t) * (nt ? nt : 1));
    for (uint32_t i = 0; i < nt; ++i)
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k) b->tri[i * 9 + v * 3 + k] = xyz[3 * idx[3 * i + v] + k];
    if (nt < 2) {
        /* The reference's generateHierarchy reads an empty queue for n == 1;
         * a single leaf root is the only sensible reading. */
        b->nodes = calloc(1, sizeof(ref_node));
        b->root = 0;
        if (nt == 1) {
            ref_node *n = &b->nodes[0];
            n->object = 0;
            for (int k = 0; k < 3; ++k) {
                float a = b->tri[k], c = b->tri[3 + k], d = b->tri[6 + k];
                n->lo[k] = fminf(fminf(a, c), d);
                n->hi[k] = fmaxf(fmaxf(a, c), d);
            }
            b->sorted[0] = 0;
        } else {
            b->nodes[0].object = -2; /* empty */
        }
        return b;
    }
    float wlo[3] = {0, 0, 0}, whi[3] = {0, 0, 0};
    for (uint32_t i = 0; i < nt; ++i)
        for (int k = 0; k < 3; ++k) {
            const float *t = b->tri + 9 * i;
            float lo = fminf(fminf(t[k], t[3 + k]), t[6 + k]);
            float hi = fmaxf(fmaxf(t[k], t[3 + k]), t[6 + k]);
            wlo[k] = fminf(wlo[k], lo);
            whi[k] = fmaxf(whi[k], hi);
        }
    float diag[3] = {whi[0] - wlo[0], whi[1] - wlo[1], whi[2] - wlo[2]};
    kv *keys = malloc(sizeof(kv) * nt);
    const float third = (float)1 / 3.0f;
    for (uint32_t i = 0; i < nt; ++i) {
        const float *t = b->tri + 9 * i;
        float c[3], u[3];
        for (int k = 0; k < 3; ++k) {
            float s = t[k] + t[3 + k];
            s = s + t[6 + k];
            c[k] = third * s;
            u[k] = (c[k] - wlo[k]) / diag[k];
        }
        b->morton[i] = morton3d(u[0], u[1], u[2]);
        keys[i].code = b->morton[i];
        keys[i].id = i;
    }
    qsort(keys, nt, sizeof(kv), kv_cmp);
    uint32_t nn = 2 * nt - 1;
    b->nodes = calloc(nn, sizeof(ref_node));
    for (uint32_t i = 0; i < nt; ++i) {
        uint32_t id = keys[i].id;
        b->sorted[i] = id;
        ref_node *n = &b->nodes[i];
        n->object = (int32_t)id;
        const float *t = b->tri + 9 * id;
        /* Bound3f(p0, p1) + p2 (Triangle.cu:124-131) */
        for (int k = 0; k < 3; ++k) {
            n->lo[k] = fminf(fminf(t[k], t[3 + k]), t[6 + k]);
            n->hi[k] = fmaxf(fmaxf(t[k], t[3 + k]), t[6 + k]);
        }
    }
    for (uint32_t i = nt; i < nn; ++i) b->nodes[i].object = -1;
    for (uint32_t i = 0; i + 1 < nn; ++i) {
        uint32_t parent = nt + i / 2;
        if (i % 2 == 0) b->nodes[parent].left = i;
        else b->nodes[parent].right = i;
    }
    for (uint32_t i = nt; i < nn; ++i) {
        ref_node *n = &b->nodes[i], *l = &b->nodes[n->left], *r = &b->nodes[n->right];
        for (int k = 0; k < 3; ++k) {
            n->lo[k] = fminf(l->lo[k], r->lo[k]);
            n->hi[k] = fmaxf(l->hi[k], r->hi[k]);
        }
    }
    b->root = nn - 1;
    free(keys);
    return b;
}

void orc_refbvh_free(orc_refbvh *b)
{
    if (!b) return;
    free(b->nodes); free(b->tri); free(b->morton); free(b->sorted); free(b);
}

uint32_t orc_refbvh_root(const orc_refbvh *b) { return b->root; }
void orc_refbvh_sorted_ids(const orc_refbvh *b, uint32_t *out) { memcpy(out, b->sorted, sizeof(uint32_t) * b->nt); }
uint32_t orc_refbvh_morton(const orc_refbvh *b, uint32_t i) { return b->morton[i]; }

static int tri_test(const orc_refbvh *b, uint32_t id, const orc_ray *r, const uint8_t *transparent,
                    int check_shadow, float *h, int guarded)
{
    if (check_shadow && transparent && transparent[id]) return 0; /* Triangle.cu:439-440 */
    const float *t = b->tri + 9 * id;
    return guarded ? orc_tri_intersects_guarded(r, t, t + 3, t + 6, h, NULL)
                   : orc_tri_intersects(r, t, t + 3, t + 6, h, NULL);
}

/* BVH::intersectRay, BVH.cu:355-448: LIFO stack; for each popped node the
 * left child then the right child; leaf children are triangle-tested at
 * once, internal children are box-tested (no hit-distance culling) and
 * pushed, so the right subtree is popped first. */
static int32_t refbvh_walk(const orc_refbvh *b, const orc_ray *r, const uint8_t *transparent, int check_shadow,
                           float *t_out, uint64_t *n_box, uint64_t *n_tri, int guarded)
{
    int32_t hit = -1;
    float h = INFINITY, thit = INFINITY;
    uint64_t nb = 0, ntr = 0;
    if (b->nt == 0) { if (t_out) *t_out = h; return -1; }
    const ref_node *root = &b->nodes[b->root];
    if (root->object >= 0) {
        ++ntr;
        if (tri_test(b, (uint32_t)root->object, r, transparent, check_shadow, &h, guarded)) { hit = root->object; thit = h; }
        if (t_out) *t_out = thit;
        if (n_tri) *n_tri += ntr;
        return hit;
    }
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = b->root;
    while (sp > 0) {
        const ref_node *n = &b->nodes[stack[--sp]];
        uint32_t ch[2] = {n->left, n->right};
        for (int c = 0; c < 2; ++c) {
            const ref_node *cn = &b->nodes[ch[c]];
            if (cn->object >= 0) {
                ++ntr;
                if (tri_test(b, (uint32_t)cn->object, r, transparent, check_shadow, &h, guarded)) { hit = cn->object; thit = h; }
            } else {
                float tmp = h;
                ++nb;
                if (orc_box_intersects(cn->lo, cn->hi, r, &tmp)) {
                    if (sp >= 128) abort(); /* cannot happen: depth <= 32 + 1 */
                    stack[sp++] = ch[c];
                }
            }
        }
    }
    /* the reference re-tests the returned triangle from hitOffset = inf
     * (Scene.cu:209-216), so the reported distance is the winner's own t */
    if (t_out) *t_out = thit;
    if (n_box) *n_box += nb;
    if (n_tri) *n_tri += ntr;
    return hit;
}

int32_t orc_refbvh_intersect(const orc_refbvh *b, const orc_ray *r, const uint8_t *transparent,
                             int check_shadow, float *t_out, uint64_t *n_box, uint64_t *n_tri)
{
    return refbvh_walk(b, r, transparent, check_shadow, t_out, n_box, n_tri, 0);
}

int32_t orc_refbvh_intersect_guarded(const orc_refbvh *b, const orc_ray *r, const uint8_t *transparent,
                                     int check_shadow, float *t_out)
{
    return refbvh_walk(b, r, transparent, check_shadow, t_out, NULL, NULL, 1);
}

/* ------------------------------------------------------------------ */
/* Thread pool helper                                                   */
/* ------------------------------------------------------------------ */
typedef void (*range_fn)(void *ctx, uint32_t begin, uint32_t end, int tid);
typedef struct { range_fn fn; void *ctx; uint32_t b, e; int tid; } job;

static void *job_run(void *p)
{
    job *j = p;
    j->fn(j->ctx, j->b, j->e, j->tid);
    return NULL;
}

static void parallel_for(uint32_t n, int nthreads, range_fn fn, void *ctx)
{
    if (nthreads < 1) nthreads = 1;
    if ((uint32_t)nthreads > n) nthreads = n ? (int)n : 1;
    if (nthreads == 1) { fn(ctx, 0, n, 0); return; }
    pthread_t *th = malloc(sizeof(pthread_t) * nthreads);
    job *jobs = malloc(sizeof(job) * nthreads);
    uint32_t chunk = (n + nthreads - 1) / nthreads;
    for (int i = 0; i < nthreads; ++i) {
        uint32_t b = i * chunk, e = b + chunk;
        if (b > n) b = n;
        if (e > n) e = n;
        jobs[i] = (job){fn, ctx, b, e, i};
        pthread_create(&th[i], NULL, job_run, &jobs[i]);
    }
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th); free(jobs);
}

typedef struct {
    const orc_refbvh *b;
    const float *ox, *oy, *oz, *dx, *dy, *dz, *tmax;
    int32_t *prim; float *t;
} ref_ctx;

static void ref_range(void *p, uint32_t b0, uint32_t e0, int tid)
{
    (void)tid;
    ref_ctx *c = p;
    for (uint32_t i = b0; i < e0; ++i) {
        orc_ray r;
        float o[3] = {c->ox[i], c->oy[i], c->oz[i]}, d[3] = {c->dx[i], c->dy[i], c->dz[i]};
        orc_ray_make(&r, o, d, c->tmax ? c->tmax[i] : INFINITY);
        float th;
        c->prim[i] = orc_refbvh_intersect(c->b, &r, NULL, 0, &th, NULL, NULL);
        if (c->t) c->t[i] = th;
    }
}

void orc_refbvh_trace(const orc_refbvh *b, const float *ox, const float *oy, const float *oz,
                      const float *dx, const float *dy, const float *dz, const float *tmax,
                      uint32_t n, int32_t *prim, float *t, int nthreads)
{
    ref_ctx c = {b, ox, oy, oz, dx, dy, dz, tmax, prim, t};
    parallel_for(n, nthreads, ref_range, &c);
}

/* ------------------------------------------------------------------ */
/* Brute force with order-independence analysis                         */
/* ------------------------------------------------------------------ */
typedef struct {
    const float *xyz; const uint32_t *idx; uint32_t nt;
    const uint8_t *transparent; int check_shadow;
    const float *ox, *oy, *oz, *dx, *dy, *dz, *tmax;
    int32_t *prim; float *t; uint8_t *flags;
} brute_ctx;

static void brute_range(void *p, uint32_t b0, uint32_t e0, int tid)
{
    (void)tid;
    brute_ctx *c = p;
    for (uint32_t i = b0; i < e0; ++i) {
        orc_ray r;
        float o[3] = {c->ox[i], c->oy[i], c->oz[i]}, d[3] = {c->dx[i], c->dy[i], c->dz[i]};
        orc_ray_make(&r, o, d, c->tmax ? c->tmax[i] : INFINITY);
        /* Pass 1: every candidate is evaluated with hitOffset = +inf so its
         * own t and deltaT verdict are known independently of order. */
        float best = INFINITY;
        int32_t best_id = -1;
        int nbest = 0, best64 = 0;
        float min_masked = INFINITY;
        for (uint32_t k = 0; k < c->nt; ++k) {
            if (c->check_shadow && c->transparent && c->transparent[k]) continue;
            const float *p0 = c->xyz + 3 * c->idx[3 * k];
            const float *p1 = c->xyz + 3 * c->idx[3 * k + 1];
            const float *p2 = c->xyz + 3 * c->idx[3 * k + 2];
            float h = INFINITY;
            orc_tri_info info;
            orc_tri_intersects(&r, p0, p1, p2, &h, &info);
            if (!info.reached) continue;
            if (info.masked) {
                if (info.t < min_masked) min_masked = info.t;
                continue;
            }
            if (info.t < best) { best = info.t; best_id = (int32_t)k; nbest = 1; best64 = info.fp64; }
            else if (info.t == best) ++nbest;
        }
        uint8_t f = 0;
        if (nbest > 1) f |= 1;
        if (best_id >= 0 && min_masked <= best) f |= 2;
        if (best64) f |= 4;
        c->prim[i] = best_id;
        if (c->t) c->t[i] = best;
        if (c->flags) c->flags[i] = f;
    }
}

void orc_brute_trace(const float *xyz, const uint32_t *idx, uint32_t nt, const uint8_t *transparent,
                     int check_shadow, const float *ox, const float *oy, const float *oz,
                     const float *dx, const float *dy, const float *dz, const float *tmax,
                     uint32_t n, int32_t *prim, float *t, uint8_t *flags, int nthreads)
{
    brute_ctx c = {xyz, idx, nt, transparent, check_shadow, ox, oy, oz, dx, dy, dz, tmax, prim, t, flags};
    parallel_for(n, nthreads, brute_range, &c);
}

/* ------------------------------------------------------------------ */
/* Counting traversal of the library's exported BVH2 layout              */
/* ------------------------------------------------------------------ */
/* Node (16 floats): [c0.lo.x c0.hi.x c0.lo.y c0.hi.y]
 *                   [c1.lo.x c1.hi.x c1.lo.y c1.hi.y]
 *                   [c0.lo.z c0.hi.z c1.lo.z c1.hi.z]
 *                   [child0 child1 - -]  (int32 bit patterns)
 * child >= 0: internal node index; child < 0: leaf, ~child = first<<3 |
 * (count-1).  Triangle (12 floats): v0.xyz id | v1.xyz 0 | v2.xyz 0.
 * The slab test below is the kernel's, operation for operation. */
static inline int32_t as_i32(float f) { int32_t i; memcpy(&i, &f, 4); return i; }

#define RT_TFAR_SCALE 1.00000012f /* 1 + 2^-23 */
#define RT_HIT_SCALE  1.0000005f  /* slack when culling against the current hit */

static inline void kernel_inv_dir(const float d[3], float inv[3], float oinv[3], const float o[3])
{
    for (int k = 0; k < 3; ++k) {
        float dd = fabsf(d[k]) > 1e-30f ? d[k] : copysignf(1e-30f, d[k]);
        inv[k] = 1.0f / dd;
        oinv[k] = o[k] * inv[k];
    }
}

static inline int kernel_slab(const float *lx, const float *ly, const float *lz, const float inv[3],
                              const float oinv[3], float tlim, float *tnear_out)
{
    float tx0 = fmaf(lx[0], inv[0], -oinv[0]), tx1 = fmaf(lx[1], inv[0], -oinv[0]);
    float ty0 = fmaf(ly[0], inv[1], -oinv[1]), ty1 = fmaf(ly[1], inv[1], -oinv[1]);
    float tz0 = fmaf(lz[0], inv[2], -oinv[2]), tz1 = fmaf(lz[1], inv[2], -oinv[2]);
    float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
    float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
    tf = fminf(tf * RT_TFAR_SCALE, tlim);
    *tnear_out = tn;
    return tn <= tf;
}

typedef struct {
    const float *nodes, *tris;
    const float *ox, *oy, *oz, *dx, *dy, *dz, *tmax;
    int32_t *prim; float *t;
    uint64_t *cnt; /* 2 per thread: nodes, tris */
} b2_ctx;

static void b2_range(void *p, uint32_t b0, uint32_t e0, int tid)
{
    b2_ctx *c = p;
    uint64_t nn = 0, ntr = 0;
    for (uint32_t i = b0; i < e0; ++i) {
        orc_ray r;
        float o[3] = {c->ox[i], c->oy[i], c->oz[i]}, d[3] = {c->dx[i], c->dy[i], c->dz[i]};
        orc_ray_make(&r, o, d, c->tmax ? c->tmax[i] : INFINITY);
        float inv[3], oinv[3];
        kernel_inv_dir(d, inv, oinv, o);
        float h = INFINITY, thit = INFINITY;
        int32_t hit = -1;
        int32_t stack[256];
        int sp = 0, done = 0;
        int32_t cur = 0;
        while (!done) {
            if (cur >= 0) {
                const float *n = c->nodes + 16 * (size_t)cur;
                ++nn;
                float tlim = fminf(h, r.tmax) * RT_HIT_SCALE;
                float t0, t1;
                int h0 = kernel_slab(n + 0, n + 2, n + 8, inv, oinv, tlim, &t0);
                int h1 = kernel_slab(n + 4, n + 6, n + 10, inv, oinv, tlim, &t1);
                int32_t c0 = as_i32(n[12]), c1 = as_i32(n[13]);
                if (h0 && h1) {
                    int32_t nearc = c0, farc = c1;
                    if (t1 < t0) { nearc = c1; farc = c0; }
                    stack[sp++] = farc;
                    cur = nearc;
                } else if (h0) {
                    cur = c0;
                } else if (h1) {
                    cur = c1;
                } else if (sp) {
                    cur = stack[--sp];
                } else {
                    done = 1;
                }
            } else {
                uint32_t enc = ~(uint32_t)cur;
                uint32_t first = enc >> 3, cnt = (enc & 7) + 1;
                for (uint32_t k = 0; k < cnt; ++k) {
                    const float *tr = c->tris + 12 * (size_t)(first + k);
                    ++ntr;
                    if (orc_tri_intersects_guarded(&r, tr, tr + 4, tr + 8, &h, NULL)) {
                        hit = as_i32(tr[3]);
                        thit = h;
                    }
                }
                if (sp) cur = stack[--sp];
                else done = 1;
            }
        }
        c->prim[i] = hit;
        if (c->t) c->t[i] = thit;
    }
    c->cnt[2 * tid] = nn;
    c->cnt[2 * tid + 1] = ntr;
}

void orc_bvh2_trace(const float *nodes, const float *tris, const float *ox, const float *oy,
                    const float *oz, const float *dx, const float *dy, const float *dz,
                    const float *tmax, uint32_t n, int32_t *prim, float *t, uint64_t *n_node,
                    uint64_t *n_tri, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    uint64_t *cnt = calloc(2 * (size_t)nthreads, sizeof(uint64_t));
    b2_ctx c = {nodes, tris, ox, oy, oz, dx, dy, dz, tmax, prim, t, cnt};
    parallel_for(n, nthreads, b2_range, &c);
    uint64_t a = 0, b = 0;
    for (int i = 0; i < nthreads; ++i) { a += cnt[2 * i]; b += cnt[2 * i + 1]; }
    if (n_node) *n_node = a;
    if (n_tri) *n_tri = b;
    free(cnt);
}

/* ------------------------------------------------------------------ */
/* Counting traversal of the library's four-wide quantized layout        */
/* ------------------------------------------------------------------ */
/* Node (16 words): origin.xyz (float), meta = ex | ey<<8 | ez<<16 |
 * valid<<24, child[4] (int32), then lo_x hi_x lo_y hi_y lo_z hi_z words
 * (byte i = child i) and two zero words.  Child i's box along x is
 * [origin.x + lo_x[i]*2^(ex-127), origin.x + hi_x[i]*2^(ex-127)].  The
 * operations below are the kernel's (trav_step4): the slab distances
 * fma(q, s*inv, fma(origin, inv, -o*inv)) (the kernel evaluates only the
 * near and far plane per axis, chosen by the sign of inv; since s*inv has
 * that sign and fma is monotone in q, they are the min and max taken here,
 * so the entry and exit distances are the same numbers), entry distances sorted
 * as bit patterns with misses as 0xffffffff by the same 5-comparator
 * network, the nearest child followed and the others pushed farthest
 * first.  max_sp reports the deepest stack the rays needed, which the
 * library's stack bound must cover. */
static inline uint32_t as_u32(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u32_as_f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

static inline void cswap_key(uint32_t *ka, int32_t *ra, uint32_t *kb, int32_t *rb)
{
    int sw = *kb < *ka;
    uint32_t k0 = sw ? *kb : *ka, k1 = sw ? *ka : *kb;
    int32_t r0 = sw ? *rb : *ra, r1 = sw ? *ra : *rb;
    *ka = k0; *kb = k1; *ra = r0; *rb = r1;
}

typedef struct {
    const float *nodes, *tris;
    const float *ox, *oy, *oz, *dx, *dy, *dz, *tmax;
    int32_t *prim; float *t;
    uint64_t *cnt; /* 3 per thread: nodes, tris, max sp */
} b4_ctx;

static void b4_range(void *p, uint32_t b0, uint32_t e0, int tid)
{
    b4_ctx *c = p;
    uint64_t nn = 0, ntr = 0, msp = 0;
    for (uint32_t i = b0; i < e0; ++i) {
        orc_ray r;
        float o[3] = {c->ox[i], c->oy[i], c->oz[i]}, d[3] = {c->dx[i], c->dy[i], c->dz[i]};
        orc_ray_make(&r, o, d, c->tmax ? c->tmax[i] : INFINITY);
        float inv[3], oinv[3];
        kernel_inv_dir(d, inv, oinv, o);
        float h = INFINITY, thit = INFINITY;
        int32_t hit = -1;
        int32_t stack[1024];
        int sp = 0, done = 0;
        int32_t cur = 0;
        while (!done) {
            if (cur >= 0) {
                const float *n = c->nodes + 16 * (size_t)cur;
                ++nn;
                float tlim = fminf(h, r.tmax) * RT_HIT_SCALE;
                uint32_t meta = as_u32(n[3]);
                float A[3], B[3];
                for (int a = 0; a < 3; ++a) {
                    float sc = u32_as_f(((meta >> (8 * a)) & 0xffu) << 23);
                    A[a] = sc * inv[a];
                    B[a] = fmaf(n[a], inv[a], -oinv[a]);
                }
                uint32_t q[6];
                for (int k = 0; k < 6; ++k) q[k] = as_u32(n[8 + k]);
                uint32_t key[4];
                int32_t ref[4];
                for (int k = 0; k < 4; ++k) ref[k] = as_i32(n[4 + k]);
                for (int k = 0; k < 4; ++k) {
                    int sh = 8 * k;
                    float tx0 = fmaf((float)((q[0] >> sh) & 0xffu), A[0], B[0]);
                    float tx1 = fmaf((float)((q[1] >> sh) & 0xffu), A[0], B[0]);
                    float ty0 = fmaf((float)((q[2] >> sh) & 0xffu), A[1], B[1]);
                    float ty1 = fmaf((float)((q[3] >> sh) & 0xffu), A[1], B[1]);
                    float tz0 = fmaf((float)((q[4] >> sh) & 0xffu), A[2], B[2]);
                    float tz1 = fmaf((float)((q[5] >> sh) & 0xffu), A[2], B[2]);
                    float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
                    float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
                    tf = fminf(tf * RT_TFAR_SCALE, tlim);
                    int hit4 = ((meta >> (24 + k)) & 1u) && tn <= tf;
                    key[k] = hit4 ? as_u32(tn) : 0xffffffffu;
                }
                cswap_key(&key[0], &ref[0], &key[1], &ref[1]);
                cswap_key(&key[2], &ref[2], &key[3], &ref[3]);
                cswap_key(&key[0], &ref[0], &key[2], &ref[2]);
                cswap_key(&key[1], &ref[1], &key[3], &ref[3]);
                cswap_key(&key[1], &ref[1], &key[2], &ref[2]);
                if (key[0] != 0xffffffffu) {
                    for (int k = 3; k >= 1; --k)
                        if (key[k] != 0xffffffffu) {
                            if (sp >= 1024) abort(); /* a 1024-deep stack means a broken tree */
                            stack[sp++] = ref[k];
                        }
                    if ((uint64_t)sp > msp) msp = (uint64_t)sp;
                    cur = ref[0];
                    continue;
                }
                if (sp) cur = stack[--sp];
                else done = 1;
            } else {
                uint32_t enc = ~(uint32_t)cur;
                uint32_t first = enc >> 3, cnt = (enc & 7) + 1;
                for (uint32_t k = 0; k < cnt; ++k) {
                    const float *tr = c->tris + 12 * (size_t)(first + k);
                    ++ntr;
                    if (orc_tri_intersects_guarded(&r, tr, tr + 4, tr + 8, &h, NULL)) {
                        hit = as_i32(tr[3]);
                        thit = h;
                    }
                }
                if (sp) cur = stack[--sp];
                else done = 1;
            }
        }
        c->prim[i] = hit;
        if (c->t) c->t[i] = thit;
    }
    c->cnt[3 * tid] = nn;
    c->cnt[3 * tid + 1] = ntr;
    c->cnt[3 * tid + 2] = msp;
}

void orc_bvh4_trace(const float *nodes4, const float *tris, const float *ox, const float *oy,
                    const float *oz, const float *dx, const float *dy, const float *dz,
                    const float *tmax, uint32_t n, int32_t *prim, float *t, uint64_t *n_node,
                    uint64_t *n_tri, uint32_t *max_sp, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    uint64_t *cnt = calloc(3 * (size_t)nthreads, sizeof(uint64_t));
    b4_ctx c = {nodes4, tris, ox, oy, oz, dx, dy, dz, tmax, prim, t, cnt};
    parallel_for(n, nthreads, b4_range, &c);
    uint64_t a = 0, b = 0, m = 0;
    for (int i = 0; i < nthreads; ++i) {
        a += cnt[3 * i];
        b += cnt[3 * i + 1];
        if (cnt[3 * i + 2] > m) m = cnt[3 * i + 2];
    }
    if (n_node) *n_node = a;
    if (n_tri) *n_tri = b;
    if (max_sp) *max_sp = (uint32_t)m;
    free(cnt);
}
