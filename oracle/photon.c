/*
 * photon.c -- oracle restatement of the photon map search (TEST
 * INFRASTRUCTURE ONLY; see oracle_photon.h).
 */
#include "oracle_photon.h"
#include "oracle_shade.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* kdtree.h: norm(difference(p, e)) with the default Euclidean norm in float */
static float norm3(float a, float b, float c)
{
    float s = 0.0f;
    s += a * a;
    s += b * b;
    s += c * c;
    return sqrtf(s);
}

static float dist_to(const float *xyz, uint32_t e, const float q[3])
{
    return norm3(q[0] - xyz[3 * (size_t)e], q[1] - xyz[3 * (size_t)e + 1], q[2] - xyz[3 * (size_t)e + 2]);
}

/* ---------------- build (kdtree.h build_tree) ---------------- */

static const float *g_xyz;
static int g_axis;

static int cmp_axis(const void *a, const void *b)
{
    const uint32_t i = *(const uint32_t *)a, j = *(const uint32_t *)b;
    const float x = g_xyz[3 * (size_t)i + g_axis], y = g_xyz[3 * (size_t)j + g_axis];
    if (x < y) return -1;
    if (x > y) return 1;
    return i < j ? -1 : (i > j);
}

static uint32_t build(const float *xyz, uint32_t *order, uint8_t *axis, size_t left, size_t right, uint32_t level)
{
    if (right - left <= 1) return level + 1;
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) lo[a] = hi[a] = xyz[3 * (size_t)order[left] + a];
    for (size_t i = left + 1; i < right; ++i)
        for (int a = 0; a < 3; ++a) {
            const float v = xyz[3 * (size_t)order[i] + a];
            if (lo[a] > v) lo[a] = v;
            if (hi[a] < v) hi[a] = v;
        }
    const size_t median = (right + left) / 2;
    int ax = 0;
    float ext = hi[0] - lo[0];
    for (int a = 1; a < 3; ++a)
        if (hi[a] - lo[a] > ext) { ax = a; ext = hi[a] - lo[a]; }
    /* nth_element's partition property, by a full sort of the range */
    g_xyz = xyz;
    g_axis = ax;
    qsort(order + left, right - left, sizeof(uint32_t), cmp_axis);
    axis[median] = (uint8_t)ax;
    const uint32_t d0 = build(xyz, order, axis, left, median, level + 1);
    const uint32_t d1 = build(xyz, order, axis, median + 1, right, level + 1);
    return d0 > d1 ? d0 : d1;
}

uint32_t orc_kd_build(const float *xyz, uint32_t n, uint32_t *order, uint8_t *axis)
{
    for (uint32_t i = 0; i < n; ++i) { order[i] = i; axis[i] = 0; }
    return n ? build(xyz, order, axis, 0, n, 0) : 0;
}

/* ---------------- search (kdtree.h nearest_neighbors_impl) ---------------- */

typedef struct {
    const float *xyz;
    const int32_t *shape;
    const uint32_t *order;
    const uint8_t *axis;
    const float *q;
    int32_t qshape;
    uint32_t k, cnt;
    float maxd;
    uint32_t *idx;
    float *dist;
} Search;

/* The max-heap steps of libstdc++ that kdtree.h's std::make_heap, push_heap
 * and pop_heap run (bits/stl_heap.h), with comp(a, b) = dist(a) < dist(b):
 * the result order of a search is this heap's layout, so they are restated
 * step for step rather than by an equivalent sift-down. */
static void heap_push_up(float *d, uint32_t *v, size_t hole, size_t top, float dv, uint32_t vv)
{
    /* __push_heap: move the hole up while the parent compares less */
    while (hole > top) {
        const size_t parent = (hole - 1) / 2;
        if (!(d[parent] < dv)) break;
        d[hole] = d[parent];
        v[hole] = v[parent];
        hole = parent;
    }
    d[hole] = dv;
    v[hole] = vv;
}

static void heap_adjust(float *d, uint32_t *v, size_t hole, size_t len, float dv, uint32_t vv)
{
    /* __adjust_heap: the hole goes down to a leaf along the larger children
     * (the right one unless the left compares greater), then the value is
     * pushed up from there */
    const size_t top = hole;
    size_t child = hole;
    while (len > 0 && child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (d[child] < d[child - 1]) --child;
        d[hole] = d[child];
        v[hole] = v[child];
        hole = child;
    }
    if ((len & 1) == 0 && len >= 2 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        d[hole] = d[child - 1];
        v[hole] = v[child - 1];
        hole = child - 1;
    }
    heap_push_up(d, v, hole, top, dv, vv);
}

static void heap_make(float *d, uint32_t *v, size_t len)
{
    /* __make_heap */
    if (len < 2) return;
    for (size_t parent = (len - 2) / 2;; --parent) {
        heap_adjust(d, v, parent, len, d[parent], v[parent]);
        if (parent == 0) return;
    }
}

/* push_back(x) onto a full heap of len entries, push_heap, pop_heap,
 * pop_back: d and v hold len + 1 slots. */
static void heap_push_pop(float *d, uint32_t *v, size_t len, float dv, uint32_t vv)
{
    d[len] = dv;
    v[len] = vv;
    heap_push_up(d, v, len, 0, dv, vv);        /* push_heap over len + 1 */
    const float ld = d[len];                    /* pop_heap: value = last, last = front */
    const uint32_t lv = v[len];
    d[len] = d[0];
    v[len] = v[0];
    heap_adjust(d, v, 0, len, ld, lv);          /* pop_back drops slot len */
}

void orc_heap_sequence(const float *dist, uint32_t n, uint32_t k, uint32_t *ids, float *out_dist, uint32_t *count)
{
    /* nearest_neighbors_impl's bookkeeping for a sequence of accepted
     * candidates 0..n-1 with distances dist[] (test hook) */
    float *d = malloc(sizeof(float) * (k + 1));
    uint32_t *v = malloc(sizeof(uint32_t) * (k + 1));
    uint32_t c = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (c < k) {
            d[c] = dist[i];
            v[c] = i;
            if (++c == k) heap_make(d, v, k);
        } else {
            heap_push_pop(d, v, k, dist[i], i);
        }
    }
    for (uint32_t i = 0; i < c; ++i) { ids[i] = v[i]; out_dist[i] = d[i]; }
    *count = c;
    free(d);
    free(v);
}

static void search(Search *s, size_t left, size_t right)
{
    if (right <= left) return;
    const size_t median = (right + left) / 2;
    const uint32_t e = s->order[median];
    const float d = dist_to(s->xyz, e, s->q);
    if (d < s->maxd && (s->qshape == -1 || (s->shape ? s->shape[e] : -1) == s->qshape)) {
        /* push_back; make_heap at k; beyond k push_heap + pop_heap + pop_back,
         * and max_distance = the front's distance (kdtree.h:123-130) */
        if (s->cnt < s->k) {
            s->idx[s->cnt] = e;
            s->dist[s->cnt] = d;
            if (++s->cnt == s->k) heap_make(s->dist, s->idx, s->k);
        } else {
            heap_push_pop(s->dist, s->idx, s->k, d, e);
            s->maxd = s->dist[0];
        }
    }
    if (right - left > 1) {
        const int ax = s->axis[median];
        const float split = s->xyz[3 * (size_t)e + ax];
        float pl[3] = {0.0f, 0.0f, 0.0f};
        pl[ax] = s->q[ax] - split;
        if (s->q[ax] < split) {
            search(s, left, median);
            if (norm3(pl[0], pl[1], pl[2]) < s->maxd) search(s, median + 1, right);
        } else {
            search(s, median + 1, right);
            if (norm3(pl[0], pl[1], pl[2]) < s->maxd) search(s, left, median);
        }
    }
}

uint32_t orc_kd_knn(const float *xyz, const int32_t *shape, const uint32_t *order, const uint8_t *axis, uint32_t n,
                    const float q[3], int32_t qshape, uint32_t k, float radius, uint32_t *idx, float *dist)
{
    if (k == 0) return 0;
    /* the heap needs one slot beyond k for push_heap + pop_heap */
    uint32_t vbuf[k + 1];
    float dbuf[k + 1];
    Search s = {xyz, shape, order, axis, q, qshape, k, 0, radius, vbuf, dbuf};
    search(&s, 0, n);
    memcpy(idx, vbuf, sizeof(uint32_t) * s.cnt);
    memcpy(dist, dbuf, sizeof(float) * s.cnt);
    return s.cnt;
}

/* ---------------- brute force ---------------- */

uint32_t orc_brute_knn(const float *xyz, const int32_t *shape, uint32_t n, const float q[3], int32_t qshape,
                       uint32_t k, float radius, uint32_t *idx, float *dist)
{
    uint32_t cnt = 0;
    for (uint32_t e = 0; e < n; ++e) {
        if (qshape != -1 && (shape ? shape[e] : -1) != qshape) continue;
        const float d = dist_to(xyz, e, q);
        if (!(d < radius)) continue;
        /* insertion into the sorted prefix by (distance, index) */
        if (cnt == k && !(d < dist[k - 1])) continue;
        uint32_t i = cnt < k ? cnt++ : k - 1;
        while (i > 0 && dist[i - 1] > d) { dist[i] = dist[i - 1]; idx[i] = idx[i - 1]; --i; }
        dist[i] = d;
        idx[i] = e;
    }
    return cnt;
}

typedef struct {
    const float *xyz; const int32_t *shape; const uint32_t *order; const uint8_t *axis; uint32_t n;
    const float *q; const int32_t *qshape; uint32_t m, k; float radius; int tree;
    uint32_t *count, *idx; float *dist;
    int t, nt;
} Batch;

static void *batch_worker(void *p)
{
    const Batch *b = (const Batch *)p;
    for (uint32_t i = (uint32_t)b->t; i < b->m; i += (uint32_t)b->nt) {
        const int32_t qs = b->qshape ? b->qshape[i] : -1;
        uint32_t *ix = b->idx + (size_t)i * b->k;
        float *ds = b->dist + (size_t)i * b->k;
        b->count[i] = b->tree ? orc_kd_knn(b->xyz, b->shape, b->order, b->axis, b->n, b->q + 3 * (size_t)i, qs, b->k,
                                           b->radius, ix, ds)
                              : orc_brute_knn(b->xyz, b->shape, b->n, b->q + 3 * (size_t)i, qs, b->k, b->radius, ix,
                                              ds);
    }
    return NULL;
}

void orc_knn_batch(const float *xyz, const int32_t *shape, const uint32_t *order, const uint8_t *axis, uint32_t n,
                   const float *q, const int32_t *qshape, uint32_t m, uint32_t k, float radius, int tree,
                   uint32_t *count, uint32_t *idx, float *dist, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    pthread_t th[64];
    Batch b[64];
    for (int t = 0; t < nthreads; ++t) {
        b[t] = (Batch){xyz, shape, order, axis, n, q, qshape, m, k, radius, tree, count, idx, dist, t, nthreads};
        pthread_create(&th[t], NULL, batch_worker, &b[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}

/* ---------------- density estimate (Scene.cu:607-634, Lambertian) ---------------- */

void orc_photon_gather_lambert(const float *xyz, const int32_t *shape, uint32_t n, const double *rgb, const float *dir,
                               const float *q, const float *nrm, const float *kd, const int32_t *qshape, uint32_t m,
                               uint32_t k, float radius, double *out)
{
    uint32_t *idx = (uint32_t *)malloc(sizeof(uint32_t) * k);
    float *dist = (float *)malloc(sizeof(float) * k);
    const float inv_pi = 1.0f / 3.14159265358979323846f; /* INV_PI = 1 / PI in Float */
    for (uint32_t i = 0; i < m; ++i) {
        const uint32_t c = orc_brute_knn(xyz, shape, n, q + 3 * (size_t)i, qshape ? qshape[i] : -1, k, radius, idx, dist);
        double acc[3] = {0.0, 0.0, 0.0};
        if (c > 0) {
            float maxd = 0.0f;
            for (uint32_t j = 0; j < c; ++j) {
                const uint32_t e = idx[j];
                if (dist[j] > maxd) maxd = dist[j];
                /* lightScattered(-photon direction): |cos| to the shading normal, times color * INV_PI */
                const float w[3] = {-dir[3 * (size_t)e], -dir[3 * (size_t)e + 1], -dir[3 * (size_t)e + 2]};
                const float cosv = fabsf(nrm[3 * (size_t)i] * w[0] + nrm[3 * (size_t)i + 1] * w[1] +
                                         nrm[3 * (size_t)i + 2] * w[2]);
                for (int ch = 0; ch < 3; ++ch) {
                    const double diffuse = (double)kd[3 * (size_t)i + ch] * (double)inv_pi;
                    acc[ch] += rgb[3 * (size_t)e + ch] * (diffuse * (double)cosv);
                }
            }
            const float kernel = 1.0f / (3.14159265358979323846f * maxd * maxd);
            for (int ch = 0; ch < 3; ++ch) acc[ch] = acc[ch] * (double)kernel;
        }
        out[3 * (size_t)i] = acc[0]; out[3 * (size_t)i + 1] = acc[1]; out[3 * (size_t)i + 2] = acc[2];
    }
    free(idx);
    free(dist);
}

/* ---------------- subsurface term (Scene.cu:667-702, geometricMath.cu:451-522) ---------------- */

/* impPaper_dipoleDiffusionAproximation for one channel, float scalars
 * widened to double where the reference's Spectrum meets them. */
static double dipole_rd(double ss, double sa, float g, float fdr, float A, float r)
{
    const double ssp = (double)(1.0f - g) * ss;
    const double stp = ssp + sa;
    const double alpha = ssp / stp;
    const double str = sqrt(3.0 * sa * stp);
    const double zr = sqrt(3.0 * (1.0 - alpha)) / str;
    const double zv = zr * (double)A;
    const double r2 = (double)(r * r);
    const double dr = sqrt(r2 + zr * zr), dv = sqrt(r2 + zv * zv);
    const double sdr = str * dr, sdv = str * dv;
    const double rd = (sdr + 1.0) * exp(-sdr) * zr / (dr * dr * dr) + (sdv + 1.0) * exp(-sdv) * zv / (dv * dv * dv);
    const double c1 = zr * (str + 1.0 / dr);
    return c1 * rd * (double)(1.0f - fdr) * alpha;
}

double orc_dipole_rd(double sigma_s, double sigma_a, float g, float eta, float r)
{
    const float fdr = (float)((-1.440 / (double)(eta * eta)) + (0.710 / (double)eta) + 0.668 + 0.0636 * (double)eta);
    const float A = (1 + fdr) / (1 - fdr);
    return dipole_rd(sigma_s, sigma_a, g, fdr, A, r);
}

/* cosine(v1, v2) (geometricMath.cu:139-142) */
static float cosine3(float ax, float ay, float az, float bx, float by, float bz)
{
    const float d = ax * bx + ay * by + az * bz;
    return d / (sqrtf(ax * ax + ay * ay + az * az) * sqrtf(bx * bx + by * by + bz * bz));
}

typedef struct {
    const float *xyz;
    const int32_t *shape;
    const uint32_t *order;
    const uint8_t *axis;
    const float *q;
    int32_t qshape;
    float radius;
    /* per-photon term inputs */
    const double *rgb;
    const float *pdir, *pnrm;
    const float *params;  /* kd[3], sigma_s[3], sigma_a[3], g, eta */
    float fdr, A;
    double ft_o;
    double acc[3];
    float maxd;
} Range;

/* search_nearest(p, radius, unlimited, shape): the walk of
 * nearest_neighbors_impl with a k never reached, so photons arrive in
 * insertion order and the radius never shrinks. */
static void range_walk(Range *s, size_t left, size_t right)
{
    if (right <= left) return;
    const size_t median = (right + left) / 2;
    const uint32_t e = s->order[median];
    const float d = dist_to(s->xyz, e, s->q);
    if (d < s->radius && (s->qshape == -1 || (s->shape ? s->shape[e] : -1) == s->qshape)) {
        const float *pd = s->pdir + 3 * (size_t)e, *pn = s->pnrm + 3 * (size_t)e;
        const float cos_i = cosine3(pn[0], pn[1], pn[2], -pd[0], -pd[1], -pd[2]);
        const double ft_i = (double)(1 - orc_fresnel_dielectric(cos_i, 1.0f, s->params[10]));
        const float inv_pi = 1.0f / 3.14159265358979323846f;
        for (int c = 0; c < 3; ++c) {
            const double rd = dipole_rd((double)s->params[3 + c], (double)s->params[6 + c], s->params[9], s->fdr, s->A, d);
            const double sd = (double)inv_pi * ft_i * rd * s->ft_o; /* diffusionTermBSSRDF */
            s->acc[c] += s->rgb[3 * (size_t)e + c] * sd * (double)s->params[c];
        }
        if (d > s->maxd) s->maxd = d;
    }
    if (right - left > 1) {
        const int ax = s->axis[median];
        const float split = s->xyz[3 * (size_t)e + ax];
        float pl[3] = {0.0f, 0.0f, 0.0f};
        pl[ax] = s->q[ax] - split;
        if (s->q[ax] < split) {
            range_walk(s, left, median);
            if (norm3(pl[0], pl[1], pl[2]) < s->radius) range_walk(s, median + 1, right);
        } else {
            range_walk(s, median + 1, right);
            if (norm3(pl[0], pl[1], pl[2]) < s->radius) range_walk(s, left, median);
        }
    }
}

void orc_photon_subsurface(const float *xyz, const int32_t *shape, const uint32_t *order, const uint8_t *axis,
                           uint32_t n, const double *rgb, const float *pdir, const float *pnrm, const float *q,
                           const float *qnrm, const float *qview, const int32_t *qshape, uint32_t m, float radius,
                           const float params[11], double *out)
{
    const float eta = params[10];
    const float fdr = (float)((-1.440 / (double)(eta * eta)) + (0.710 / (double)eta) + 0.668 + 0.0636 * (double)eta);
    const float A = (1 + fdr) / (1 - fdr);
    for (uint32_t i = 0; i < m; ++i) {
        const float *qn = qnrm + 3 * (size_t)i, *qv = qview + 3 * (size_t)i;
        Range s;
        memset(&s, 0, sizeof s);
        s.xyz = xyz; s.shape = shape; s.order = order; s.axis = axis; s.q = q + 3 * (size_t)i;
        s.qshape = qshape ? qshape[i] : -1;
        s.radius = radius; s.rgb = rgb; s.pdir = pdir; s.pnrm = pnrm; s.params = params;
        s.fdr = fdr; s.A = A;
        s.ft_o = (double)(1 - orc_fresnel_dielectric(cosine3(qn[0], qn[1], qn[2], qv[0], qv[1], qv[2]), 1.0f, eta));
        range_walk(&s, 0, n);
        /* contribution /= 2 PI maxD^2 (no photon: 0 / 0) */
        const float den = 2.0f * 3.14159265358979323846f * s.maxd * s.maxd;
        for (int c = 0; c < 3; ++c) out[3 * (size_t)i + c] = s.acc[c] / (double)den;
    }
}
