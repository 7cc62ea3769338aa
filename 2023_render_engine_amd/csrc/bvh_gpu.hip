// bvh_gpu.hip -- the BVH built on the GPU from device-resident triangles, in
// the layouts the traversal kernels read (rt_internal.h: Node2, Node4, Tri3).
//
// The reference builds its tree on the GPU from Morton codes of the
// triangles (cuda_libraries/BVH.cu:136-151, 314-341: codes, a sort, leaves
// initialised in parallel, then pairing).  This builder follows the same
// idea with a linear-time radix tree:
//   1. triangle boxes, and the bounds of their centres (k_tri_boxes per
//      block, k_bounds over the blocks: no contended atomics);
//   2. 30-bit Morton codes of the centres in those bounds, sorted with the
//      triangle index as the value (rocPRIM radix sort);
//   3. Karras' binary radix tree over the sorted codes (k_radix_tree: each
//      internal node finds its range and split from the common prefixes of
//      its neighbours; equal codes are split by index);
//   4. boxes and SAH costs bottom-up in passes (k_node_pass: a node is
//      computed once both children were computed by an earlier launch, so
//      nothing is read across work-groups within a launch and no
//      device-scope fence is needed -- on this GPU each such fence writes
//      back the XCD's L2); a subtree of at most max_leaf triangles becomes
//      a leaf when that is no dearer by the SAH;
//   5. the kept internal nodes are numbered by a prefix sum and written as
//      Node2 records with both child boxes (k_emit_node2), and the
//      triangles as Tri3 records in code order plus the degenerate
//      sentinel (k_emit_tris);
//   6. for the four-wide layout, the Node2 tree is collapsed level by level
//      (k_collapse4_count, a prefix sum, k_collapse4): each wide node opens
//      its largest-area internal child until it has four, as the host
//      collapse does, and quantizes the boxes outward; nodes are numbered
//      breadth-first.
// Every step is deterministic: the same input gives the same records.
// The tree is worse than the host's binned SAH (its SAH cost is reported),
// and it is built in milliseconds.
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_scan.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "rt_bvh_gpu.h"
#include "rt_internal.h"

namespace rt {

namespace {

constexpr int kBB = 256;  // block size of the build kernels

__device__ inline uint32_t ord(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float unord(uint32_t u) { return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u); }

__device__ inline uint32_t spread10(uint32_t v)
{
    v &= 1023u;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Per triangle: its box (6 floats: lo xyz, hi xyz).  Each block reduces
// its triangles' bounds into part[block * 12 ..] as ordered ints: [0..2] /
// [3..5] the min / max of the centres, [6..8] / [9..11] the min / max of the
// boxes (k_bounds finishes the reduction).  Indices out of range or
// non-finite coordinates set *bad.
__global__ __launch_bounds__(kBB) void k_tri_boxes(const float *xyz, uint32_t nv, const uint32_t *idx, uint32_t nt,
                                                   float *tbox, uint32_t *part, uint32_t *bad)
{
    __shared__ uint32_t smin[6][kBB], smax[6][kBB];
    const uint32_t t = blockIdx.x * kBB + threadIdx.x;
    uint32_t mn[6], mx[6];
    for (int a = 0; a < 6; ++a) { mn[a] = 0xffffffffu; mx[a] = 0u; }
    if (t < nt) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        bool ok = true;
        for (int k = 0; k < 3; ++k) {
            uint32_t v = idx[3 * (size_t)t + k];
            if (v >= nv) { ok = false; v = 0; }
            for (int a = 0; a < 3; ++a) {
                const float c = xyz[3 * (size_t)v + a];
                if (!isfinite(c)) ok = false;
                lo[a] = fminf(lo[a], c);
                hi[a] = fmaxf(hi[a], c);
            }
        }
        if (!ok) atomicOr(bad, 1u);
        for (int a = 0; a < 3; ++a) {
            tbox[6 * (size_t)t + a] = lo[a];
            tbox[6 * (size_t)t + 3 + a] = hi[a];
            if (ok) {
                mn[a] = mx[a] = ord(0.5f * lo[a] + 0.5f * hi[a]);
                mn[3 + a] = ord(lo[a]);
                mx[3 + a] = ord(hi[a]);
            }
        }
    }
    for (int a = 0; a < 6; ++a) { smin[a][threadIdx.x] = mn[a]; smax[a][threadIdx.x] = mx[a]; }
    __syncthreads();
    for (int s = kBB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int a = 0; a < 6; ++a) {
                smin[a][threadIdx.x] = min(smin[a][threadIdx.x], smin[a][threadIdx.x + s]);
                smax[a][threadIdx.x] = max(smax[a][threadIdx.x], smax[a][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 12) {
        const uint32_t j = threadIdx.x, a = j % 3, q = j / 3;  // q: centre min, centre max, box min, box max
        const uint32_t v = (q & 1u) ? smax[(q >> 1) * 3 + a][0] : smin[(q >> 1) * 3 + a][0];
        part[(size_t)blockIdx.x * 12 + j] = v;
    }
}

// One block: the per-block bounds of k_tri_boxes into cb[0..11].
__global__ __launch_bounds__(kBB) void k_bounds(const uint32_t *part, uint32_t nblocks, uint32_t *cb)
{
    __shared__ uint32_t red[12][kBB];
    uint32_t v[12];
    for (int j = 0; j < 12; ++j) v[j] = ((j / 3) & 1) ? 0u : 0xffffffffu;
    for (uint32_t b = threadIdx.x; b < nblocks; b += kBB)
        for (int j = 0; j < 12; ++j) {
            const uint32_t x = part[(size_t)b * 12 + j];
            v[j] = ((j / 3) & 1) ? max(v[j], x) : min(v[j], x);
        }
    for (int j = 0; j < 12; ++j) red[j][threadIdx.x] = v[j];
    __syncthreads();
    for (int st = kBB / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int j = 0; j < 12; ++j) {
                const uint32_t x = red[j][threadIdx.x + st];
                red[j][threadIdx.x] = ((j / 3) & 1) ? max(red[j][threadIdx.x], x) : min(red[j][threadIdx.x], x);
            }
        __syncthreads();
    }
    if (threadIdx.x < 12) cb[threadIdx.x] = red[threadIdx.x][0];
}

// Outward padding of every child box, the host build's rule: 2^-19 of
// (scene diagonal + largest |coordinate|) per side, so the FP32 slab test is
// conservative with respect to the watertight triangle test.
__host__ __device__ inline float box_pad(const float lo[3], const float hi[3])
{
    double diag = 0.0, amax = 0.0;
    for (int k = 0; k < 3; ++k) {
        diag += ((double)hi[k] - lo[k]) * ((double)hi[k] - lo[k]);
        const double m = fabs((double)lo[k]) > fabs((double)hi[k]) ? fabs((double)lo[k]) : fabs((double)hi[k]);
        amax = amax > m ? amax : m;
    }
    return (float)((sqrt(diag) + amax) * ldexp(1.0, -19));
}

__device__ inline float scene_pad(const uint32_t *cb)
{
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) { lo[a] = unord(cb[6 + a]); hi[a] = unord(cb[9 + a]); }
    return box_pad(lo, hi);
}

__global__ __launch_bounds__(kBB) void k_morton(const float *tbox, uint32_t nt, const uint32_t *cb, uint32_t *key,
                                                uint32_t *val)
{
    const uint32_t t = blockIdx.x * kBB + threadIdx.x;
    if (t >= nt) return;
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
        const float lo = unord(cb[a]), hi = unord(cb[3 + a]);
        const float c = 0.5f * tbox[6 * (size_t)t + a] + 0.5f * tbox[6 * (size_t)t + 3 + a];
        const float ext = hi - lo;
        float u = ext > 0.0f ? (c - lo) / ext * 1024.0f : 0.0f;
        u = u > 0.0f ? (u < 1023.0f ? u : 1023.0f) : 0.0f;  // NaN -> 0
        q[a] = (uint32_t)u;
    }
    key[t] = spread10(q[0]) << 2 | spread10(q[1]) << 1 | spread10(q[2]);
    val[t] = t;
}

// Common-prefix length of sorted codes i and j (equal codes: 32 + the
// common prefix of the indices); -1 outside [0, n).
__device__ inline int prefix(const uint32_t *key, int64_t n, int64_t i, int64_t j)
{
    if (j < 0 || j >= n) return -1;
    const uint32_t a = key[i], b = key[j];
    if (a != b) return __clz(a ^ b);
    return 32 + __clz((uint32_t)i ^ (uint32_t)j);
}

// encode_leaf of rt_internal.h for device code: ~(first << 3 | count - 1).
__device__ inline int32_t leaf_ref(uint32_t first, uint32_t count)
{
    return (int32_t)~((first << kLeafCountBits) | (count - 1u));
}

// Child reference: internal node c, or leaf (sorted position) c | kLeafBit.
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kNotReady = 0xffffffffu;  // pass index of a node not computed yet

__global__ __launch_bounds__(kBB) void k_radix_tree(const uint32_t *key, uint32_t n, uint32_t *child, uint32_t *lo,
                                                    uint32_t *hi, uint32_t *parent)
{
    const int64_t i = (int64_t)blockIdx.x * kBB + threadIdx.x;
    const int64_t N = n;
    if (i >= N - 1) return;
    const int d = prefix(key, N, i, i + 1) - prefix(key, N, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = prefix(key, N, i, i - d);
    int64_t lmax = 2;
    while (prefix(key, N, i, i + lmax * d) > dmin) lmax <<= 1;
    int64_t l = 0;
    for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
        if (prefix(key, N, i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = prefix(key, N, i, j);
    int64_t s = 0;
    for (int64_t div = 2;; div <<= 1) {
        const int64_t t = (l + div - 1) / div;
        if (prefix(key, N, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    const int64_t g = i + s * d + (d < 0 ? -1 : 0);
    const uint32_t a = (uint32_t)(i < j ? i : j), b = (uint32_t)(i < j ? j : i);
    const uint32_t c0 = (a == (uint32_t)g) ? ((uint32_t)g | kLeafBit) : (uint32_t)g;
    const uint32_t c1 = (b == (uint32_t)g + 1) ? ((uint32_t)(g + 1) | kLeafBit) : (uint32_t)(g + 1);
    child[2 * i] = c0;
    child[2 * i + 1] = c1;
    lo[i] = a;
    hi[i] = b;
    // parents: internal nodes at [0, n-1), leaves at [n-1, 2n-1)
    parent[(c0 & kLeafBit) ? (N - 1 + (c0 & ~kLeafBit)) : c0] = (uint32_t)i;
    parent[(c1 & kLeafBit) ? (N - 1 + (c1 & ~kLeafBit)) : c1] = (uint32_t)i;
}

__device__ inline double area6(const float b[6])
{
    double d[3];
    for (int a = 0; a < 3; ++a) d[a] = fmax(0.0, (double)b[3 + a] - (double)b[a]);
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
}

// Bottom-up boxes and SAH costs in passes: pass p computes every internal
// node whose children are leaves or were computed in an earlier pass
// (ready[c] < p).  All it reads was written by earlier launches, so no
// fence or coherent load is needed; the root is done after as many passes
// as the tree is high.  Per node: its box (the union of its children's),
// its SAH cost (cost of a triangle test 1, of a node visit trav_cost, per
// unit of area) and its choice: a node over at most max_leaf triangles
// becomes a leaf when (triangles x area) does not exceed (trav_cost x area
// + its children's costs); and the height of the kept tree below it (0
// for a node that became a leaf).  The root never becomes a leaf.
__global__ __launch_bounds__(kBB) void k_node_pass(const uint32_t *child, const uint32_t *lo, const uint32_t *hi,
                                                   const float *tbox, const uint32_t *val, uint32_t n,
                                                   uint32_t pass, uint32_t max_leaf, double trav_cost, float *nbox,
                                                   double *cost, uint32_t *height, uint32_t *leafc, uint32_t *ready)
{
    const uint32_t i = blockIdx.x * kBB + threadIdx.x;
    if (i >= n - 1 || ready[i] != kNotReady) return;
    const uint32_t c[2] = {child[2 * i], child[2 * i + 1]};
    for (int k = 0; k < 2; ++k)
        if (!(c[k] & kLeafBit) && !(ready[c[k]] < pass)) return;
    float b[2][6];
    double cc[2];
    uint32_t hh[2];
    for (int k = 0; k < 2; ++k) {
        if (c[k] & kLeafBit) {
            const float *src = tbox + 6 * (size_t)val[c[k] & ~kLeafBit];
            for (int j = 0; j < 6; ++j) b[k][j] = src[j];
            cc[k] = area6(b[k]);
            hh[k] = 0;
        } else {
            for (int j = 0; j < 6; ++j) b[k][j] = nbox[6 * (size_t)c[k] + j];
            cc[k] = cost[c[k]];
            hh[k] = height[c[k]];
        }
    }
    float box[6];
    for (int a = 0; a < 3; ++a) {
        box[a] = fminf(b[0][a], b[1][a]);
        box[3 + a] = fmaxf(b[0][3 + a], b[1][3 + a]);
    }
    const double area = area6(box);
    const double split = trav_cost * area + cc[0] + cc[1];
    const uint32_t size = hi[i] - lo[i] + 1;
    const double as_leaf = (double)size * area;
    const bool leaf = i != 0 && size <= max_leaf && as_leaf <= split;
    for (int j = 0; j < 6; ++j) nbox[6 * (size_t)i + j] = box[j];
    cost[i] = leaf ? as_leaf : split;
    height[i] = leaf ? 0u : 1u + (hh[0] > hh[1] ? hh[0] : hh[1]);
    leafc[i] = leaf ? 1u : 0u;
    ready[i] = pass;
}

// Kept internal nodes: the root, and every node neither of whose own
// choice nor any ancestor's choice was a leaf (only nodes over at most
// max_leaf triangles choose; the walk up stops at the first larger one).
__global__ __launch_bounds__(kBB) void k_keep(const uint32_t *lo, const uint32_t *hi, const uint32_t *parent,
                                              const uint32_t *leafc, uint32_t n, uint32_t max_leaf, uint32_t *keep)
{
    const uint32_t i = blockIdx.x * kBB + threadIdx.x;
    if (i >= n - 1) return;
    uint32_t k = 1;
    if (i != 0) {
        if (leafc[i]) {
            k = 0;
        } else {
            for (uint32_t p = parent[i]; p != 0 && hi[p] - lo[p] + 1 <= max_leaf; p = parent[p])
                if (leafc[p]) { k = 0; break; }
        }
    }
    keep[i] = k;
}

// Node2 record of each kept node, its children as node indices or leaf
// ranges and their boxes padded outward; SAH terms (1.2 * area of an
// internal node, count * area of a leaf, unpadded, as the host build sums
// them) added per block into *sah.
__global__ __launch_bounds__(kBB) void k_emit_node2(const uint32_t *child, const uint32_t *lo, const uint32_t *hi,
                                                    const uint32_t *keep, const uint32_t *newidx, const float *tbox,
                                                    const uint32_t *val, const float *nbox, uint32_t n,
                                                    const uint32_t *cb, Node2 *nodes, double *sah)
{
    __shared__ double part[kBB];
    const uint32_t i = blockIdx.x * kBB + threadIdx.x;
    double acc = 0.0;
    if (i < n - 1 && keep[i]) {
        const float pad = scene_pad(cb);
        Node2 nd{};
        float bx[2][6];
        for (int c = 0; c < 2; ++c) {
            const uint32_t ch = child[2 * i + c];
            int32_t ref;
            if (ch & kLeafBit) {
                const uint32_t pos = ch & ~kLeafBit;
                const float *src = tbox + 6 * (size_t)val[pos];
                for (int k = 0; k < 6; ++k) bx[c][k] = src[k];
                ref = leaf_ref(pos, 1);
                acc += area6(bx[c]);
            } else {
                for (int k = 0; k < 6; ++k) bx[c][k] = nbox[6 * (size_t)ch + k];
                if (keep[ch]) {
                    ref = (int32_t)newidx[ch];
                } else {
                    ref = leaf_ref(lo[ch], hi[ch] - lo[ch] + 1);
                    acc += (double)(hi[ch] - lo[ch] + 1) * area6(bx[c]);
                }
            }
            nd.child[c] = ref;
            for (int a = 0; a < 3; ++a) { bx[c][a] -= pad; bx[c][3 + a] += pad; }
        }
        nd.f[0] = bx[0][0]; nd.f[1] = bx[0][3]; nd.f[2] = bx[0][1]; nd.f[3] = bx[0][4];
        nd.f[4] = bx[1][0]; nd.f[5] = bx[1][3]; nd.f[6] = bx[1][1]; nd.f[7] = bx[1][4];
        nd.f[8] = bx[0][2]; nd.f[9] = bx[0][5]; nd.f[10] = bx[1][2]; nd.f[11] = bx[1][5];
        nodes[newidx[i]] = nd;
        float own[6];
        for (int k = 0; k < 6; ++k) own[k] = nbox[6 * (size_t)i + k];
        acc += 1.2 * area6(own);
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kBB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && part[0] != 0.0) atomicAdd(sah, part[0]);
}

// Triangles in code order (leaf order) and the degenerate sentinel after
// them.  An index past the vertices (the build then fails) reads vertex 0.
__global__ __launch_bounds__(kBB) void k_emit_tris(const float *xyz, uint32_t nv, const uint32_t *idx,
                                                   const uint32_t *val, uint32_t n, Tri3 *tris)
{
    const uint32_t k = blockIdx.x * kBB + threadIdx.x;
    if (k > n) return;
    Tri3 t{};
    if (k == n) {
        t.id = -1;
    } else {
        const uint32_t tri = val[k];
        const float *v[3];
        for (int c = 0; c < 3; ++c) {
            const uint32_t vi = idx[3 * (size_t)tri + c];
            v[c] = xyz + 3 * (size_t)(vi < nv ? vi : 0u);
        }
        for (int a = 0; a < 3; ++a) { t.v0[a] = v[0][a]; t.v1[a] = v[1][a]; t.v2[a] = v[2][a]; }
        t.id = (int32_t)tri;
    }
    tris[k] = t;
}

// ---- four-wide collapse (bvh_build.cpp's Collapser, level by level) ----
struct Kid {
    float lo[3], hi[3];
    int32_t ref;
};

__device__ inline double kid_area(const Kid &k)
{
    double d[3];
    for (int a = 0; a < 3; ++a) d[a] = fmax(0.0, (double)k.hi[a] - (double)k.lo[a]);
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
}

__device__ inline void node2_kids(const Node2 &n, Kid &k0, Kid &k1)
{
    k0.lo[0] = n.f[0]; k0.hi[0] = n.f[1]; k0.lo[1] = n.f[2]; k0.hi[1] = n.f[3]; k0.lo[2] = n.f[8]; k0.hi[2] = n.f[9];
    k1.lo[0] = n.f[4]; k1.hi[0] = n.f[5]; k1.lo[1] = n.f[6]; k1.hi[1] = n.f[7]; k1.lo[2] = n.f[10]; k1.hi[2] = n.f[11];
    k0.ref = n.child[0];
    k1.ref = n.child[1];
}

// Smallest power-of-two scale 2^e (e >= -126) with 255 * 2^e >= extent.
__device__ inline int scale_exponent(double extent)
{
    if (!(extent > 0.0)) return -126;
    int e = (int)ceil(log2(extent / 255.0));
    if (e < -126) e = -126;
    while (255.0 * ldexp(1.0, e) < extent) ++e;
    while (e > -126 && 255.0 * ldexp(1.0, e - 1) >= extent) --e;
    return e;
}

// The kids of a wide node: its two-wide node's two children, then the
// internal kid of largest surface area is opened (replaced by its two
// children) until there are four or none is internal.
__device__ inline int open_kids(const Node2 *n2, uint32_t src, Kid kids[4])
{
    node2_kids(n2[src], kids[0], kids[1]);
    int nk = 2;
    while (nk < 4) {
        int best = -1;
        double best_area = -1.0;
        for (int i = 0; i < nk; ++i)
            if (kids[i].ref >= 0 && kid_area(kids[i]) > best_area) {
                best_area = kid_area(kids[i]);
                best = i;
            }
        if (best < 0) break;
        Kid a, b;
        node2_kids(n2[kids[best].ref], a, b);
        kids[best] = a;
        kids[nk++] = b;
    }
    return nk;
}

// One level, first pass: internal kids per frontier entry.
__global__ __launch_bounds__(kBB) void k_collapse4_count(const Node2 *n2, const uint2 *front, uint32_t nfront,
                                                         uint32_t *cnt)
{
    const uint32_t f = blockIdx.x * kBB + threadIdx.x;
    if (f >= nfront) return;
    Kid kids[4];
    const int nk = open_kids(n2, front[f].x, kids);
    uint32_t ni = 0;
    for (int i = 0; i < nk; ++i) ni += kids[i].ref >= 0 ? 1u : 0u;
    cnt[f] = ni;
}

// One level, second pass: frontier entry f (two-wide node, pending entries
// above it) becomes wide node base + f; its internal kids, in slot order,
// become next-level entries off[f], off[f] + 1, ... and wide nodes
// base + nfront + off[f] + j.  The numbering is breadth-first and does not
// depend on scheduling.  The four-wide traversal holds at most
// (pending above) + kids - 1 entries at a node, so the deepest such sum is
// its stack bound (*need), the host collapse's figure.
__global__ __launch_bounds__(kBB) void k_collapse4(const Node2 *n2, const uint2 *front, uint32_t nfront,
                                                   const uint32_t *off, uint32_t base, uint32_t sentinel, Node4 *n4,
                                                   uint2 *next, uint32_t *need)
{
    const uint32_t f = blockIdx.x * kBB + threadIdx.x;
    if (f >= nfront) return;
    const uint2 fe = front[f];
    Kid kids[4];
    const int nk = open_kids(n2, fe.x, kids);
    const uint32_t pending = fe.y + (uint32_t)nk - 1u;
    atomicMax(need, pending);
    int32_t refs[4];
    uint32_t q = off[f];
    for (int i = 0; i < nk; ++i) {
        if (kids[i].ref >= 0) {
            next[q] = make_uint2((uint32_t)kids[i].ref, pending);
            refs[i] = (int32_t)(base + nfront + q);
            ++q;
        } else {
            refs[i] = kids[i].ref;
        }
    }
    for (int i = nk; i < 4; ++i) refs[i] = leaf_ref(sentinel, 1);
    Node4 n{};
    uint32_t meta = 0;
    for (int a = 0; a < 3; ++a) {
        float p = kids[0].lo[a];
        for (int i = 1; i < nk; ++i) p = fminf(p, kids[i].lo[a]);
        double ext = 0.0;
        for (int i = 0; i < nk; ++i) ext = fmax(ext, (double)kids[i].hi[a] - (double)p);
        const int e = scale_exponent(ext);
        const double sc = ldexp(1.0, e);
        n.origin[a] = p;
        meta |= (uint32_t)(e + 127) << (8 * a);
        uint32_t qlo = 0, qhi = 0;
        for (int i = 0; i < nk; ++i) {
            double l = floor(((double)kids[i].lo[a] - (double)p) / sc);
            double h = ceil(((double)kids[i].hi[a] - (double)p) / sc);
            l = fmin(fmax(l, 0.0), 255.0);
            h = fmin(fmax(h, 0.0), 255.0);
            while (l > 0.0 && (double)p + l * sc > (double)kids[i].lo[a]) l -= 1.0;
            while (h < 255.0 && (double)p + h * sc < (double)kids[i].hi[a]) h += 1.0;
            qlo |= (uint32_t)l << (8 * i);
            qhi |= (uint32_t)h << (8 * i);
        }
        for (int i = nk; i < 4; ++i) qlo |= 255u << (8 * i);
        n.q[2 * a] = qlo;
        n.q[2 * a + 1] = qhi;
    }
    meta |= (uint32_t)((1u << nk) - 1u) << 24;
    n.meta = meta;
    for (int i = 0; i < 4; ++i) n.child[i] = refs[i];
    n4[base + f] = n;
}

unsigned blocks(uint64_t n) { return (unsigned)std::max<uint64_t>(1, (n + kBB - 1) / kBB); }

struct Arena {
    char *base = nullptr;
    ~Arena() { if (base) (void)hipFree(base); }
};

}  // namespace

hipError_t build_bvh_device(const float *xyz, uint32_t nv, const uint32_t *idx, uint32_t nt, const rt_bvh_opts &o,
                            hipStream_t s, DeviceBvhBuild &out, std::string &err)
{
    const auto t0 = std::chrono::steady_clock::now();
    out = DeviceBvhBuild{};
    const uint32_t max_leaf = o.max_leaf ? o.max_leaf : 4;  // 1..kMaxLeafCount, checked by the caller
    const bool wide = o.width != 2;
    hipError_t e = hipSuccess;
    auto fail = [&](hipError_t x, const char *what) {
        err = std::string("rt_bvh_build_device: ") + what + ": " + hipGetErrorString(x);
        return x;
    };
    // outputs: two-wide nodes (at least one), triangles + sentinel
    const uint32_t n2_cap = nt > 1 ? nt - 1 : 1;
    if ((e = hipMalloc(&out.nodes2, (size_t)n2_cap * sizeof(Node2))) != hipSuccess) return fail(e, "nodes");
    if ((e = hipMalloc(&out.tris, ((size_t)nt + 1) * sizeof(Tri3))) != hipSuccess) return fail(e, "triangles");
    out.num_tris = nt;
    if (nt <= 1) {
        // one node on the host: the triangle (if any) and the sentinel
        std::vector<float> v(9, 0.0f);
        uint32_t ii[3] = {0, 0, 0};
        if (nt == 1) {
            if ((e = hipMemcpyAsync(ii, idx, 12, hipMemcpyDeviceToHost, s)) != hipSuccess ||
                (e = hipStreamSynchronize(s)) != hipSuccess)
                return fail(e, "input");
            for (int c = 0; c < 3; ++c) {
                if (ii[c] >= nv) { err = "rt_bvh_build_device: triangle index out of range"; return hipErrorInvalidValue; }
                if ((e = hipMemcpyAsync(&v[3 * c], xyz + 3 * (size_t)ii[c], 12, hipMemcpyDeviceToHost, s)) !=
                    hipSuccess)
                    return fail(e, "input");
            }
            if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail(e, "input");
            for (float x : v)
                if (!std::isfinite(x)) { err = "rt_bvh_build_device: non-finite coordinate"; return hipErrorInvalidValue; }
        }
        Node2 root{};
        Tri3 tr[2]{};
        tr[0].id = nt ? 0 : -1;
        for (int a = 0; a < 3; ++a) { tr[0].v0[a] = v[a]; tr[0].v1[a] = v[3 + a]; tr[0].v2[a] = v[6 + a]; }
        tr[1].id = -1;
        float lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min({v[a], v[3 + a], v[6 + a]});
            hi[a] = std::max({v[a], v[3 + a], v[6 + a]});
        }
        double area = 0.0;
        if (nt == 1) {
            const double d[3] = {(double)hi[0] - lo[0], (double)hi[1] - lo[1], (double)hi[2] - lo[2]};
            area = d[0] * d[1] + d[1] * d[2] + d[2] * d[0];
            const float pad = box_pad(lo, hi);
            for (int a = 0; a < 3; ++a) { lo[a] -= pad; hi[a] += pad; }
        } else {
            for (int a = 0; a < 3; ++a) lo[a] = hi[a] = 0.0f;
        }
        root.f[0] = lo[0]; root.f[1] = hi[0]; root.f[2] = lo[1]; root.f[3] = hi[1]; root.f[8] = lo[2]; root.f[9] = hi[2];
        root.child[0] = encode_leaf(nt ? 0u : nt, 1);
        root.child[1] = encode_leaf(nt, 1);
        if ((e = hipMemcpyAsync(out.nodes2, &root, sizeof root, hipMemcpyHostToDevice, s)) != hipSuccess ||
            (e = hipMemcpyAsync(out.tris, tr + (nt ? 0 : 1), (nt + 1) * sizeof(Tri3), hipMemcpyHostToDevice, s)) !=
                hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess)
            return fail(e, "upload");
        out.num_nodes2 = 1;
        out.num_leaves = nt;
        out.max_depth = 1;
        out.sah_cost = area > 0.0 ? 1.0 : 0.0;
    } else {
        const size_t N = nt;
        const size_t nb = blocks(N);
        size_t off = 0;
        auto take = [&](size_t bytes) { size_t r = off; off += (bytes + 255) & ~(size_t)255; return r; };
        const size_t o_tbox = take(N * 24), o_cb = take(16 * 4), o_part = take(nb * 48), o_key = take(N * 4),
                     o_key2 = take(N * 4), o_val = take(N * 4), o_val2 = take(N * 4), o_child = take(2 * N * 4),
                     o_lo = take(N * 4), o_hi = take(N * 4), o_parent = take(2 * N * 4), o_nbox = take(N * 24),
                     o_cost = take(N * 8), o_height = take(N * 4), o_leafc = take(N * 4), o_ready = take(N * 4),
                     o_keep = take(N * 4), o_new = take(N * 4);
        size_t sort_bytes = 0, scan_bytes = 0;
        if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                                    (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)N, 0, 30, s)) !=
                hipSuccess ||
            (e = hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                                  (int)(N - 1), s)) != hipSuccess)
            return fail(e, "sizing");
        const size_t o_tmp = take(std::max(sort_bytes, scan_bytes));
        Arena ar;
        if ((e = hipMalloc(&ar.base, off)) != hipSuccess) return fail(e, "work space");
        char *b = ar.base;
        float *tbox = (float *)(b + o_tbox), *nbox = (float *)(b + o_nbox);
        // cb: centre and box bounds (12 words), the bad-input flag, then the
        // SAH sum (a double at word 14)
        uint32_t *cb = (uint32_t *)(b + o_cb), *bad = cb + 12;
        double *sah = (double *)(cb + 14), *cost = (double *)(b + o_cost);
        uint32_t *part = (uint32_t *)(b + o_part), *key = (uint32_t *)(b + o_key), *key2 = (uint32_t *)(b + o_key2),
                 *val = (uint32_t *)(b + o_val), *val2 = (uint32_t *)(b + o_val2), *child = (uint32_t *)(b + o_child),
                 *lo = (uint32_t *)(b + o_lo), *hi = (uint32_t *)(b + o_hi), *parent = (uint32_t *)(b + o_parent),
                 *height = (uint32_t *)(b + o_height), *leafc = (uint32_t *)(b + o_leafc),
                 *ready = (uint32_t *)(b + o_ready), *keep = (uint32_t *)(b + o_keep), *newidx = (uint32_t *)(b + o_new);
        if ((e = hipMemsetAsync(cb, 0, 16 * 4, s)) != hipSuccess ||
            (e = hipMemsetAsync(ready, 0xff, N * 4, s)) != hipSuccess)
            return fail(e, "clear");
        const double trav_cost = o.trav_cost > 0 ? o.trav_cost : 1.2f;
        hipLaunchKernelGGL(k_tri_boxes, dim3(nb), dim3(kBB), 0, s, xyz, nv, idx, nt, tbox, part, bad);
        hipLaunchKernelGGL(k_bounds, dim3(1), dim3(kBB), 0, s, part, (uint32_t)nb, cb);
        hipLaunchKernelGGL(k_morton, dim3(nb), dim3(kBB), 0, s, tbox, nt, cb, key, val);
        if ((e = hipGetLastError()) != hipSuccess) return fail(e, "boxes");
        if ((e = hipcub::DeviceRadixSort::SortPairs(b + o_tmp, sort_bytes, key, key2, val, val2, (int)N, 0, 30, s)) !=
            hipSuccess)
            return fail(e, "sort");
        hipLaunchKernelGGL(k_radix_tree, dim3(blocks(N - 1)), dim3(kBB), 0, s, key2, nt, child, lo, hi, parent);
        if ((e = hipGetLastError()) != hipSuccess) return fail(e, "radix tree");
        // passes in batches of eight until the root is done; a radix tree of
        // 30-bit codes split further by 32-bit positions is at most 62 high
        uint32_t pass = 0, root_ready = kNotReady;
        while (root_ready == kNotReady) {
            if (pass >= 64) {
                err = "rt_bvh_build_device: radix tree higher than 64 levels";
                return hipErrorUnknown;
            }
            for (int k = 0; k < 8; ++k) {
                ++pass;
                hipLaunchKernelGGL(k_node_pass, dim3(blocks(N - 1)), dim3(kBB), 0, s, child, lo, hi, tbox, val2, nt,
                                   pass, max_leaf, trav_cost, nbox, cost, height, leafc, ready);
            }
            if ((e = hipGetLastError()) != hipSuccess ||
                (e = hipMemcpyAsync(&root_ready, ready, 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
                (e = hipStreamSynchronize(s)) != hipSuccess)
                return fail(e, "boxes");
        }
        hipLaunchKernelGGL(k_keep, dim3(blocks(N - 1)), dim3(kBB), 0, s, lo, hi, parent, leafc, nt, max_leaf, keep);
        if ((e = hipGetLastError()) != hipSuccess) return fail(e, "collapse");
        if ((e = hipcub::DeviceScan::ExclusiveSum(b + o_tmp, scan_bytes, keep, newidx, (int)(N - 1), s)) != hipSuccess)
            return fail(e, "scan");
        hipLaunchKernelGGL(k_emit_node2, dim3(blocks(N - 1)), dim3(kBB), 0, s, child, lo, hi, keep, newidx, tbox, val2,
                           nbox, nt, cb, out.nodes2, sah);
        hipLaunchKernelGGL(k_emit_tris, dim3(blocks(N + 1)), dim3(kBB), 0, s, xyz, nv, idx, val2, nt, out.tris);
        if ((e = hipGetLastError()) != hipSuccess) return fail(e, "emit");
        uint32_t last[2] = {0, 0}, bad_h = 0, height_h = 0;
        double sah_h = 0.0;
        float root_box[6];
        if ((e = hipMemcpyAsync(last, newidx + (N - 2), 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipMemcpyAsync(last + 1, keep + (N - 2), 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipMemcpyAsync(&bad_h, bad, 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipMemcpyAsync(&height_h, height, 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipMemcpyAsync(&sah_h, sah, 8, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipMemcpyAsync(root_box, nbox, 24, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess)
            return fail(e, "read-back");
        if (bad_h) { err = "rt_bvh_build_device: triangle index out of range or non-finite coordinate"; return hipErrorInvalidValue; }
        out.num_nodes2 = last[0] + last[1];
        out.max_depth = height_h;
        double ra = 0.0;
        {
            double d[3];
            for (int a = 0; a < 3; ++a) d[a] = std::max(0.0, (double)root_box[3 + a] - (double)root_box[a]);
            ra = 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
        }
        out.sah_cost = ra > 0 ? sah_h / ra : 0.0;
        // a binary tree of m kept nodes has m + 1 leaves
        out.num_leaves = out.num_nodes2 + 1;
    }
    if (wide) {
        // level by level from the root, breadth-first numbering.  Each wide
        // node takes at least one two-wide node of its own, so num_nodes2
        // bounds both the wide nodes and any level's frontier.
        const uint32_t cap4 = out.num_nodes2;
        if ((e = hipMalloc(&out.nodes4, (size_t)cap4 * sizeof(Node4))) != hipSuccess) return fail(e, "wide nodes");
        size_t scan_bytes = 0;
        if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                                  (int)cap4, s)) != hipSuccess)
            return fail(e, "sizing");
        size_t off = 0;
        auto take = [&](size_t bytes) { size_t r = off; off += (bytes + 255) & ~(size_t)255; return r; };
        const size_t o_f0 = take((size_t)cap4 * 8), o_f1 = take((size_t)cap4 * 8), o_cnt = take((size_t)cap4 * 4),
                     o_off = take((size_t)cap4 * 4), o_need = take(4), o_tmp = take(scan_bytes);
        Arena fr;
        if ((e = hipMalloc(&fr.base, off)) != hipSuccess) return fail(e, "frontier");
        uint2 *front[2] = {(uint2 *)(fr.base + o_f0), (uint2 *)(fr.base + o_f1)};
        uint32_t *cnt = (uint32_t *)(fr.base + o_cnt), *offs = (uint32_t *)(fr.base + o_off),
                 *need = (uint32_t *)(fr.base + o_need);
        const uint2 root = make_uint2(0u, 0u);
        if ((e = hipMemcpyAsync(front[0], &root, sizeof root, hipMemcpyHostToDevice, s)) != hipSuccess ||
            (e = hipMemsetAsync(need, 0, 4, s)) != hipSuccess)
            return fail(e, "frontier");
        uint32_t nfront = 1, level = 0, base = 0;
        while (nfront) {
            const uint32_t cur = level & 1u;
            hipLaunchKernelGGL(k_collapse4_count, dim3(blocks(nfront)), dim3(kBB), 0, s, out.nodes2, front[cur], nfront,
                               cnt);
            if ((e = hipGetLastError()) != hipSuccess ||
                (e = hipcub::DeviceScan::ExclusiveSum(fr.base + o_tmp, scan_bytes, cnt, offs, (int)nfront, s)) !=
                    hipSuccess)
                return fail(e, "collapse");
            hipLaunchKernelGGL(k_collapse4, dim3(blocks(nfront)), dim3(kBB), 0, s, out.nodes2, front[cur], nfront,
                               offs, base, out.num_tris, out.nodes4, front[cur ^ 1u], need);
            uint32_t tail[2] = {0, 0};
            if ((e = hipGetLastError()) != hipSuccess ||
                (e = hipMemcpyAsync(tail, offs + nfront - 1, 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
                (e = hipMemcpyAsync(tail + 1, cnt + nfront - 1, 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
                (e = hipStreamSynchronize(s)) != hipSuccess)
                return fail(e, "collapse");
            base += nfront;
            nfront = tail[0] + tail[1];
            ++level;
        }
        uint32_t need_h = 0;
        if ((e = hipMemcpyAsync(&need_h, need, 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess)
            return fail(e, "collapse");
        out.num_nodes4 = base;
        out.depth4 = level;
        out.max_stack4 = need_h;
    }
    out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return hipSuccess;
}

namespace {
__global__ __launch_bounds__(kBB) void k_set_flags(Tri3 *tris, uint32_t n, const uint8_t *flags)
{
    const uint32_t k = blockIdx.x * kBB + threadIdx.x;
    if (k >= n) return;
    tris[k].flags = flags[tris[k].id] ? 1u : 0u;
}
}  // namespace

hipError_t set_transparent_device(Tri3 *tris, uint32_t n, const uint8_t *flags, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_set_flags, dim3(blocks(n)), dim3(kBB), 0, s, tris, n, flags);
    return hipGetLastError();
}

void free_bvh_device(DeviceBvhBuild &b)
{
    if (b.nodes2) (void)hipFree(b.nodes2);
    if (b.nodes4) (void)hipFree(b.nodes4);
    if (b.tris) (void)hipFree(b.tris);
    b.nodes2 = nullptr;
    b.nodes4 = nullptr;
    b.tris = nullptr;
}

}  // namespace rt
