/* lbvh.c -- TEST INFRASTRUCTURE ONLY: CPU restatement of the BVH the
 * library builds on the GPU (rt_bvh_build_device), record for record.
 *
 * The reference builds its tree on the GPU from Morton codes of the
 * triangles (cuda_libraries/BVH.cu:136-151, 314-341, codes from
 * cuda_libraries/geometricMath.cu:16-44).  The device build sorts the
 * triangles by 30-bit Morton code of their centres and builds the binary
 * radix tree over the sorted codes.  This restatement reaches the same tree
 * from its definition, top down and sequentially: a range of sorted
 * triangles splits after the last position whose 62-bit key (code, then
 * index) shares more leading bits with the range's first key than the
 * range's last key does; an internal node is numbered by the end of its
 * range that touches its sibling (left child: its last position, right
 * child: its first), the root is 0.  Subtrees of at most max_leaf triangles
 * become leaves where the SAH says a leaf is no dearer; kept nodes are
 * numbered in ascending radix-tree index.
 * The four-wide collapse is a FIFO breadth-first walk that opens the
 * largest-area internal kid until four, with the quantization searched
 * upward from the smallest exponent (not from a logarithm).
 *
 * Layouts: the library's export formats (16 floats per node, 12 per
 * triangle), see rt_bvh_export / rt_bvh_export4 in include/rt.h. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
    uint32_t code, index;
} LKey;

static int lkey_cmp(const void *a, const void *b)
{
    const LKey *x = a, *y = b;
    if (x->code != y->code) return x->code < y->code ? -1 : 1;
    return x->index < y->index ? -1 : x->index > y->index;
}

static uint32_t spread_bits(uint32_t v)
{
    uint32_t r = 0;
    for (int b = 0; b < 10; ++b) r |= ((v >> b) & 1u) << (3 * b);
    return r;
}

/* leading bits shared by the 62-bit keys of sorted positions i and j */
static int shared_bits(const LKey *k, uint32_t i, uint32_t j)
{
    const uint64_t a = ((uint64_t)k[i].code << 32) | i, b = ((uint64_t)k[j].code << 32) | j;
    const uint64_t x = a ^ b;
    return x ? __builtin_clzll(x) : 64;
}

typedef struct {
    uint32_t lo, hi; /* sorted positions, inclusive */
    int32_t child[2]; /* radix node index, or ~position for a single triangle */
    float box[6];
    double cost;      /* SAH cost of the subtree as built */
    int leaf;         /* the subtree became one leaf */
} RNode;

typedef struct {
    const LKey *key;
    const float *tbox; /* per input triangle: lo xyz, hi xyz */
    RNode *nodes;
    uint32_t max_leaf;
    double trav_cost;
} RTree;

static double area_of(const float *b);

static void grow_box(float *b, const float *c)
{
    for (int a = 0; a < 3; ++a) {
        if (c[a] < b[a]) b[a] = c[a];
        if (c[3 + a] > b[3 + a]) b[3 + a] = c[3 + a];
    }
}

/* builds node `self` over [lo, hi] (hi > lo); boxes are the exact unions of
 * the triangle boxes.  Then the SAH choice: a node over at most max_leaf
 * triangles (not the root) becomes a leaf when triangles x area <=
 * trav_cost x area + the children's costs (a single triangle costs its
 * box's area). */
static void build_range(RTree *t, uint32_t self, uint32_t lo, uint32_t hi)
{
    RNode *n = &t->nodes[self];
    n->lo = lo;
    n->hi = hi;
    const int whole = shared_bits(t->key, lo, hi);
    uint32_t split = lo;
    for (uint32_t s = lo; s < hi; ++s)
        if (shared_bits(t->key, lo, s) > whole) split = s;
    /* left [lo, split], right [split + 1, hi] */
    for (int a = 0; a < 3; ++a) { n->box[a] = INFINITY; n->box[3 + a] = -INFINITY; }
    if (split == lo) {
        n->child[0] = ~(int32_t)lo;
        grow_box(n->box, t->tbox + 6 * (size_t)t->key[lo].index);
    } else {
        n->child[0] = (int32_t)split;
        build_range(t, split, lo, split);
        grow_box(n->box, t->nodes[split].box);
    }
    if (split + 1 == hi) {
        n->child[1] = ~(int32_t)hi;
        grow_box(n->box, t->tbox + 6 * (size_t)t->key[hi].index);
    } else {
        n->child[1] = (int32_t)(split + 1);
        build_range(t, split + 1, split + 1, hi);
        grow_box(n->box, t->nodes[split + 1].box);
    }
    double cc[2];
    for (int c = 0; c < 2; ++c) {
        const int32_t ch = n->child[c];
        cc[c] = ch < 0 ? area_of(t->tbox + 6 * (size_t)t->key[~ch].index) : t->nodes[ch].cost;
    }
    const double area = area_of(n->box);
    const double split_cost = t->trav_cost * area + cc[0] + cc[1];
    const uint32_t size = hi - lo + 1;
    const double leaf_cost = (double)size * area;
    n->leaf = self != 0 && size <= t->max_leaf && leaf_cost <= split_cost;
    n->cost = n->leaf ? leaf_cost : split_cost;
}

static double area_of(const float *b)
{
    double d[3];
    for (int a = 0; a < 3; ++a) {
        d[a] = (double)b[3 + a] - (double)b[a];
        if (d[a] < 0) d[a] = 0;
    }
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
}

static int32_t leaf_link(uint32_t first, uint32_t count) { return (int32_t)~((first << 3) | (count - 1u)); }

static float pad_of(const float *lo, const float *hi)
{
    double diag = 0.0, amax = 0.0;
    for (int k = 0; k < 3; ++k) {
        diag += ((double)hi[k] - lo[k]) * ((double)hi[k] - lo[k]);
        double m = fabs((double)lo[k]);
        if (fabs((double)hi[k]) > m) m = fabs((double)hi[k]);
        if (m > amax) amax = m;
    }
    return (float)((sqrt(diag) + amax) * ldexp(1.0, -19));
}

/* node record in the export layout */
static void put_node2(float *dst, const float b0[6], const float b1[6], int32_t c0, int32_t c1)
{
    memset(dst, 0, 64);
    dst[0] = b0[0]; dst[1] = b0[3]; dst[2] = b0[1]; dst[3] = b0[4];
    dst[4] = b1[0]; dst[5] = b1[3]; dst[6] = b1[1]; dst[7] = b1[4];
    dst[8] = b0[2]; dst[9] = b0[5]; dst[10] = b1[2]; dst[11] = b1[5];
    memcpy(dst + 12, &c0, 4);
    memcpy(dst + 13, &c1, 4);
}

typedef struct {
    float lo[3], hi[3];
    int32_t ref;
} WKid;

static void kids_of(const float *n, WKid *k0, WKid *k1)
{
    k0->lo[0] = n[0]; k0->hi[0] = n[1]; k0->lo[1] = n[2]; k0->hi[1] = n[3]; k0->lo[2] = n[8]; k0->hi[2] = n[9];
    k1->lo[0] = n[4]; k1->hi[0] = n[5]; k1->lo[1] = n[6]; k1->hi[1] = n[7]; k1->lo[2] = n[10]; k1->hi[2] = n[11];
    memcpy(&k0->ref, n + 12, 4);
    memcpy(&k1->ref, n + 13, 4);
}

static double wkid_area(const WKid *k)
{
    float b[6] = {k->lo[0], k->lo[1], k->lo[2], k->hi[0], k->hi[1], k->hi[2]};
    return area_of(b);
}

/* smallest e >= -126 with 255 * 2^e >= extent, searched upward */
static int smallest_exponent(double extent)
{
    int e = -126;
    while (255.0 * ldexp(1.0, e) < extent) ++e;
    return e;
}

int orc_lbvh_build(const float *xyz, uint32_t nv, const uint32_t *idx, uint32_t nt, uint32_t max_leaf,
                   float trav_cost, int wide, float *nodes2, float *tris, float *nodes4, uint32_t *info)
{
    /* info: num_nodes2, num_leaves, max_depth, num_nodes4, depth4, max_stack4 */
    memset(info, 0, 6 * sizeof(uint32_t));
    if (max_leaf == 0) max_leaf = 4;
    float *tbox = malloc(sizeof(float) * 6 * (nt ? nt : 1));
    LKey *key = malloc(sizeof(LKey) * (nt ? nt : 1));
    float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    float smin[3] = {INFINITY, INFINITY, INFINITY}, smax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t t = 0; t < nt; ++t) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int k = 0; k < 3; ++k) {
            const uint32_t v = idx[3 * (size_t)t + k];
            if (v >= nv) { free(tbox); free(key); return -1; }
            for (int a = 0; a < 3; ++a) {
                const float c = xyz[3 * (size_t)v + a];
                if (!isfinite(c)) { free(tbox); free(key); return -1; }
                if (c < lo[a]) lo[a] = c;
                if (c > hi[a]) hi[a] = c;
            }
        }
        for (int a = 0; a < 3; ++a) {
            tbox[6 * (size_t)t + a] = lo[a];
            tbox[6 * (size_t)t + 3 + a] = hi[a];
            const float c = 0.5f * lo[a] + 0.5f * hi[a];
            if (c < cmin[a]) cmin[a] = c;
            if (c > cmax[a]) cmax[a] = c;
            if (lo[a] < smin[a]) smin[a] = lo[a];
            if (hi[a] > smax[a]) smax[a] = hi[a];
        }
    }
    for (uint32_t t = 0; t < nt; ++t) {
        uint32_t q[3];
        for (int a = 0; a < 3; ++a) {
            const float c = 0.5f * tbox[6 * (size_t)t + a] + 0.5f * tbox[6 * (size_t)t + 3 + a];
            const float ext = cmax[a] - cmin[a];
            float u = ext > 0.0f ? (c - cmin[a]) / ext * 1024.0f : 0.0f;
            if (!(u > 0.0f)) u = 0.0f;
            if (u > 1023.0f) u = 1023.0f;
            q[a] = (uint32_t)u;
        }
        key[t].code = spread_bits(q[0]) << 2 | spread_bits(q[1]) << 1 | spread_bits(q[2]);
        key[t].index = t;
    }
    qsort(key, nt, sizeof(LKey), lkey_cmp);

    /* triangles in sorted order, then the degenerate sentinel */
    memset(tris, 0, sizeof(float) * 12 * ((size_t)nt + 1));
    for (uint32_t k = 0; k < nt; ++k) {
        const uint32_t t = key[k].index;
        float *d = tris + 12 * (size_t)k;
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 3; ++a) d[4 * c + a] = xyz[3 * (size_t)idx[3 * (size_t)t + c] + a];
        const int32_t id = (int32_t)t;
        memcpy(d + 3, &id, 4);
    }
    const int32_t none = -1;
    memcpy(tris + 12 * (size_t)nt + 3, &none, 4);

    uint32_t n2 = 0, depth = 0;
    if (nt <= 1) {
        float b0[6] = {0, 0, 0, 0, 0, 0}, b1[6] = {0, 0, 0, 0, 0, 0};
        if (nt == 1) {
            const float pad = pad_of(tbox, tbox + 3);
            for (int a = 0; a < 3; ++a) { b0[a] = tbox[a] - pad; b0[3 + a] = tbox[3 + a] + pad; }
        }
        put_node2(nodes2, b0, b1, leaf_link(0, 1), leaf_link(nt, 1));
        n2 = 1;
        depth = 1;
        info[1] = nt;
    } else {
        RTree tr = {key, tbox, calloc(nt - 1, sizeof(RNode)), max_leaf, (double)(trav_cost > 0 ? trav_cost : 1.2f)};
        build_range(&tr, 0, 0, nt - 1);
        /* kept: the root, and below a kept node every internal child that did
           not become a leaf; numbered in ascending radix-tree index */
        uint32_t *newidx = malloc(sizeof(uint32_t) * (nt - 1));
        uint32_t *lvl = calloc(nt - 1, sizeof(uint32_t));
        uint32_t *stack = malloc(sizeof(uint32_t) * (nt + 1));
        uint32_t sp = 0;
        for (uint32_t i = 0; i < nt - 1; ++i) newidx[i] = UINT32_MAX;
        newidx[0] = 0;
        stack[sp++] = 0;
        while (sp) {
            const RNode *n = &tr.nodes[stack[--sp]];
            for (int c = 0; c < 2; ++c)
                if (n->child[c] >= 0 && !tr.nodes[n->child[c]].leaf) {
                    newidx[n->child[c]] = 0;
                    stack[sp++] = (uint32_t)n->child[c];
                }
        }
        for (uint32_t i = 0; i < nt - 1; ++i)
            if (newidx[i] != UINT32_MAX) newidx[i] = n2++;
        const float pad = pad_of(smin, smax);
        lvl[0] = 1;
        /* parents have smaller depth than children but not smaller indices:
           walk the kept tree from the root with an explicit stack */
        sp = 0;
        stack[sp++] = 0;
        while (sp) {
            const uint32_t i = stack[--sp];
            const RNode *n = &tr.nodes[i];
            if (lvl[i] > depth) depth = lvl[i];
            float b[2][6];
            int32_t link[2];
            for (int c = 0; c < 2; ++c) {
                const int32_t ch = n->child[c];
                if (ch < 0) {
                    const uint32_t pos = (uint32_t)~ch;
                    memcpy(b[c], tbox + 6 * (size_t)key[pos].index, sizeof b[c]);
                    link[c] = leaf_link(pos, 1);
                } else {
                    const RNode *m = &tr.nodes[ch];
                    memcpy(b[c], m->box, sizeof b[c]);
                    if (newidx[ch] != UINT32_MAX) {
                        link[c] = (int32_t)newidx[ch];
                        lvl[ch] = lvl[i] + 1;
                        stack[sp++] = (uint32_t)ch;
                    } else {
                        link[c] = leaf_link(m->lo, m->hi - m->lo + 1);
                    }
                }
                for (int a = 0; a < 3; ++a) { b[c][a] -= pad; b[c][3 + a] += pad; }
            }
            put_node2(nodes2 + 16 * (size_t)newidx[i], b[0], b[1], link[0], link[1]);
        }
        info[1] = n2 + 1;
        free(stack); free(lvl); free(newidx); free(tr.nodes);
    }
    info[0] = n2;
    info[2] = depth;

    if (wide) {
        /* FIFO breadth-first collapse: queue entries (two-wide node, pending
           entries above); a wide node's index is its position in the queue */
        uint32_t *qsrc = malloc(sizeof(uint32_t) * n2), *qabove = malloc(sizeof(uint32_t) * n2);
        uint32_t *qlevel = malloc(sizeof(uint32_t) * n2);
        uint32_t head = 0, tail = 0, need = 0, levels = 0;
        qsrc[tail] = 0; qabove[tail] = 0; qlevel[tail] = 1; ++tail;
        while (head < tail) {
            const uint32_t me = head++;
            WKid kids[4];
            kids_of(nodes2 + 16 * (size_t)qsrc[me], &kids[0], &kids[1]);
            int nk = 2;
            while (nk < 4) {
                int best = -1;
                double best_area = -1.0;
                for (int i = 0; i < nk; ++i)
                    if (kids[i].ref >= 0 && wkid_area(&kids[i]) > best_area) {
                        best_area = wkid_area(&kids[i]);
                        best = i;
                    }
                if (best < 0) break;
                WKid a, b;
                kids_of(nodes2 + 16 * (size_t)kids[best].ref, &a, &b);
                kids[best] = a;
                kids[nk++] = b;
            }
            const uint32_t pending = qabove[me] + (uint32_t)nk - 1u;
            if (pending > need) need = pending;
            if (qlevel[me] > levels) levels = qlevel[me];
            int32_t refs[4];
            for (int i = 0; i < nk; ++i) {
                if (kids[i].ref >= 0) {
                    refs[i] = (int32_t)tail;
                    qsrc[tail] = (uint32_t)kids[i].ref;
                    qabove[tail] = pending;
                    qlevel[tail] = qlevel[me] + 1;
                    ++tail;
                } else {
                    refs[i] = kids[i].ref;
                }
            }
            for (int i = nk; i < 4; ++i) refs[i] = leaf_link(nt, 1);
            uint32_t w[16];
            memset(w, 0, sizeof w);
            uint32_t meta = 0;
            for (int a = 0; a < 3; ++a) {
                float p = kids[0].lo[a];
                for (int i = 1; i < nk; ++i)
                    if (kids[i].lo[a] < p) p = kids[i].lo[a];
                double ext = 0.0;
                for (int i = 0; i < nk; ++i)
                    if ((double)kids[i].hi[a] - (double)p > ext) ext = (double)kids[i].hi[a] - (double)p;
                const int e = smallest_exponent(ext);
                const double sc = ldexp(1.0, e);
                memcpy(&w[a], &p, 4);
                meta |= (uint32_t)(e + 127) << (8 * a);
                uint32_t qlo = 0, qhi = 0;
                for (int i = 0; i < nk; ++i) {
                    /* largest l with p + l*sc <= lo, smallest h with p + h*sc >= hi, in 0..255 */
                    uint32_t l = 0, h = 255;
                    while (l < 255 && (double)p + (double)(l + 1) * sc <= (double)kids[i].lo[a]) ++l;
                    while (h > 0 && (double)p + (double)(h - 1) * sc >= (double)kids[i].hi[a]) --h;
                    qlo |= l << (8 * i);
                    qhi |= h << (8 * i);
                }
                for (int i = nk; i < 4; ++i) qlo |= 255u << (8 * i);
                w[8 + 2 * a] = qlo;
                w[9 + 2 * a] = qhi;
            }
            meta |= (uint32_t)((1u << nk) - 1u) << 24;
            w[3] = meta;
            for (int i = 0; i < 4; ++i) memcpy(&w[4 + i], &refs[i], 4);
            memcpy(nodes4 + 16 * (size_t)me, w, 64);
        }
        info[3] = tail;
        info[4] = levels;
        info[5] = need;
        free(qsrc); free(qabove); free(qlevel);
    }
    free(tbox);
    free(key);
    return 0;
}
