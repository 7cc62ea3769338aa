// photon.hip -- the photon map on the GPU: the kd-tree built by
// photon_build.cpp in HBM, and a batched k-nearest search that keeps the
// semantics of KDTree::nearest_neighbors (cuda_libraries/kdtree.h, as
// Scene::search_nearest calls it, integration/Scene.cu:704-722):
//   * a photon is taken if its distance is strictly below the current search
//     radius and, for a query with a shape id, if it lies on that shape;
//   * the first k taken fill the result, which then becomes a max-heap on
//     distance; each later one replaces the farthest and the radius shrinks
//     to the heap's new maximum;
//   * the near child is searched first; the far child only if the distance
//     to the splitting plane, as the norm of (0, .., p[axis] - split, .., 0),
//     is still below the radius after the near child is done.
// Distances are float: the square root of ((0 + dx*dx) + dy*dy) + dz*dz
// (the library is built with -ffp-contract=off), as the reference's norm.
//
// Queries are taken in the Morton order of their position (a radix sort in
// the map's work space), one lane per query; the recursion becomes an
// explicit stack of index ranges in private memory (at most the tree's depth
// + 1 entries).  For k <= 128 the heap's distances and one-byte slot numbers
// live in LDS and the photons of the slots in the query's output slice;
// beyond, the heap lives in the output slice.
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>

#include "rt_internal.h"
#include "rt_photon.h"
#include "rt_photon_render.h"
#include "rt_shade.h"

struct rt_ctx;
namespace rt {
void ctx_retain(rt_ctx *c);
void ctx_release(rt_ctx *c);
int ctx_device(rt_ctx *c);
void *ctx_stream_ptr(rt_ctx *c);
std::recursive_mutex &ctx_lock(rt_ctx *c);
}  // namespace rt

using rt::set_error;

#define RT_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            set_error(std::string(#call) + " failed: " + hipGetErrorString(e_));        \
            return e_ == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP;                 \
        }                                                                               \
    } while (0)

struct rt_photon_map {
    rt_ctx *ctx = nullptr;
    rt::PhotonNode *nodes = nullptr;  // tree order
    uint32_t *order = nullptr;        // tree position -> input index
    uint32_t *sig = nullptr;          // tree position -> shape signature of its subtree
    uint32_t n = 0;
    uint32_t depth = 0;
    float lo[3] = {0, 0, 0}, scale[3] = {0, 0, 0};  // Morton grid of the photons' bounding box
    // query ordering work space, grown on demand: Morton keys and query ids
    // (two of each for the radix sort) and the sort's temporary storage
    void *work = nullptr;
    size_t work_bytes = 0;
    void *gather = nullptr;  // k-nearest results of rt_photon_gather_lambert, grown on demand
    size_t gather_bytes = 0;
    // the photons' radiance (3 doubles) and direction (3 floats) in tree
    // order, copied at the start of each render that uses the map, so a
    // query's photons (a few subtrees) are read from a few cache lines
    double *trgb = nullptr;
    float *tdir = nullptr;
};

namespace {

constexpr int kThreads = 64;      // one wavefront per block: its LDS heaps take k * 512 bytes
constexpr uint32_t kStack = 32;    // >= tree depth + 1 (checked on the host)
constexpr uint32_t kMaxK = 1024;
constexpr uint32_t kLdsK = 128;    // k up to this keeps its heap in LDS

__device__ inline float norm3(float a, float b, float c)
{
    float s = 0.0f;
    s += a * a;
    s += b * b;
    s += c * c;
    return sqrtf(s);
}

__device__ inline uint32_t spread10(uint32_t v)
{
    v &= 1023u;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// 30-bit Morton code of each query in the photons' bounding box (outside
// points clamp to its faces), and the identity permutation to sort with it.
// With m_dev, only the first *m_dev of the m rows are queries: the rest get
// the largest key and sort last.
__global__ __launch_bounds__(256) void k_query_morton(const float *q, uint32_t m, float lx, float ly, float lz,
                                                      float sx, float sy, float sz, uint32_t *key, uint32_t *id,
                                                      const uint32_t *m_dev)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    if (m_dev && i >= *m_dev) {
        key[i] = 0xffffffffu;
        id[i] = i;
        return;
    }
    const float c[3] = {(q[3 * (size_t)i] - lx) * sx, (q[3 * (size_t)i + 1] - ly) * sy, (q[3 * (size_t)i + 2] - lz) * sz};
    uint32_t u[3];
    for (int a = 0; a < 3; ++a) {
        const float v = c[a] > 0.0f ? (c[a] < 1023.0f ? c[a] : 1023.0f) : 0.0f;  // NaN -> 0
        u[a] = (uint32_t)v;
    }
    key[i] = spread10(u[0]) << 2 | spread10(u[1]) << 1 | spread10(u[2]);
    id[i] = i;
}

struct KnnArgs {
    const uint32_t *perm;     // lane -> query (Morton order), or null
    const rt::PhotonNode *nodes;
    const uint32_t *order;    // tree position -> input index, or null: results are tree positions
    const uint32_t *sig;      // subtree shape signatures
    uint32_t n;
    const float *queries;     // m * 3
    const int32_t *qshape;    // m or null
    uint32_t m, k;
    float radius;
    uint32_t *count;          // m
    uint32_t *index;          // m * k: entry i of query q at index[q * qstride + i * istride]
    float *dist;              // m * k, row q at q * k
    const uint32_t *m_dev;    // if set: only the first *m_dev of the m lanes are queries
    uint32_t qstride = 0, istride = 1;  // qstride 0: k (rows of k entries)
    float *qmax = nullptr;    // if set: each query's farthest distance here instead of its row of distances
};

// Strided view of a heap: distance i at d[i * sd], id i at v[i * sv] (a
// photon's tree position, or a slot number that names where it is kept).
template <typename V>
struct HeapRef {
    float *d;
    V *v;
    uint32_t sd, sv;
    __device__ float &dist(uint32_t i) const { return d[i * sd]; }
    __device__ V &id(uint32_t i) const { return v[i * sv]; }
};

// The max-heap steps of libstdc++'s std::make_heap / push_heap / pop_heap
// (bits/stl_heap.h) that kdtree.h's search runs, with comp(a, b) =
// dist(a) < dist(b).  The search returns its vector in this heap's layout, so
// the steps are kept as they are rather than replaced by an equivalent
// sift-down (which would order ties and equal-set layouts differently).
// __push_heap: the hole moves up while its parent compares less.
template <typename H>
__device__ inline void heap_push_up(const H &h, uint32_t hole, uint32_t top, float d, uint32_t v)
{
    while (hole > top) {
        const uint32_t parent = (hole - 1) / 2;
        if (!(h.dist(parent) < d)) break;
        h.dist(hole) = h.dist(parent);
        h.id(hole) = h.id(parent);
        hole = parent;
    }
    h.dist(hole) = d;
    h.id(hole) = v;
}

// __adjust_heap over [0, len): the hole goes down to a leaf along the larger
// children, then the value is pushed up from there.
template <typename H>
__device__ inline void heap_adjust(const H &h, uint32_t hole, uint32_t len, float d, uint32_t v)
{
    const uint32_t top = hole;
    uint32_t child = hole;
    while (len > 0 && child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (h.dist(child) < h.dist(child - 1)) --child;
        h.dist(hole) = h.dist(child);
        h.id(hole) = h.id(child);
        hole = child;
    }
    if ((len & 1u) == 0 && len >= 2 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        h.dist(hole) = h.dist(child - 1);
        h.id(hole) = h.id(child - 1);
        hole = child - 1;
    }
    heap_push_up(h, hole, top, d, v);
}

// __make_heap over the k entries found so far.
template <typename H>
__device__ inline void heap_make(const H &h, uint32_t len)
{
    if (len < 2) return;
    for (uint32_t parent = (len - 2) / 2;; --parent) {
        heap_adjust(h, parent, len, h.dist(parent), h.id(parent));
        if (parent == 0) return;
    }
}

// push_back + push_heap + pop_heap + pop_back on a full heap of len entries.
// The transient entry len lives in (xd, xv), so the heap needs no slot
// beyond len.
template <typename H>
__device__ inline void heap_push_pop(const H &h, uint32_t len, float d, uint32_t v)
{
    float xd = d;
    uint32_t xv = v;
    if (len > 0) {
        const uint32_t parent = (len - 1) / 2;
        if (h.dist(parent) < d) {  // the first push_heap step moves the parent into slot len
            xd = h.dist(parent);
            xv = h.id(parent);
            heap_push_up(h, parent, 0, d, v);
        }
    }
    // pop_heap: the front goes to slot len and is dropped; slot len's entry
    // is adjusted down from the root
    heap_adjust(h, 0, len, xd, xv);
}

// Where the heap lives.  kHeapLdsSlot: distances in LDS, slot-major (entry
// i of lane t at i * 64 + t, so the lanes of the wavefront hit distinct
// banks), each with a one-byte slot number beside it; the photon of slot s
// is kept in word s of the query's distance slice, written once when it
// enters.  The first k photons take slots 0 .. k-1; later each takes the
// slot of the entry dropped (the heap's root) when it is not dropped itself.
// So the heap's sift
// steps move distances and slot numbers in LDS only, and the photons' ids
// are looked up once at the end.  (Ids moved along with the distances in
// the output slice cost a dependent global load and store per sift level:
// 24.6 M queries/s at k = 100.)  kHeapGlobal: both in the output slice.
enum : int { kHeapGlobal = 0, kHeapLdsSlot = 1 };
template <int MODE, uint32_t STACK>
__global__ __launch_bounds__(kThreads) void k_photon_knn(KnnArgs a)
{
    constexpr bool SLOT = MODE == kHeapLdsSlot;
    extern __shared__ uint32_t lds_heap[];  // k * kThreads distances, then k * kThreads slot bytes
    float *lds_d = reinterpret_cast<float *>(lds_heap);
    uint8_t *lds_s = reinterpret_cast<uint8_t *>(lds_heap + a.k * kThreads);
    const uint32_t lane = blockIdx.x * kThreads + threadIdx.x;
    if (lane >= (a.m_dev ? min(a.m, *a.m_dev) : a.m)) return;
    const uint32_t q = a.perm ? a.perm[lane] : lane;
    const float p[3] = {a.queries[3 * (size_t)q], a.queries[3 * (size_t)q + 1], a.queries[3 * (size_t)q + 2]};
    const uint32_t want = a.qshape ? (uint32_t)(a.qshape[q] + 1) : 0u;  // 0: any shape
    const uint32_t want_bit = rt::photon_shape_bit(want);
    uint32_t *out_v = a.index + (size_t)q * (a.qstride ? a.qstride : a.k);
    float *out_d = a.dist + (size_t)q * a.k;
    // kHeapLdsSlot: out_d[s] holds the tree position of slot s (as float bits) until the end
    using H = HeapRef<typename std::conditional<SLOT, uint8_t, uint32_t>::type>;
    H h;
    if constexpr (SLOT) h = H{lds_d + threadIdx.x, lds_s + threadIdx.x, (uint32_t)kThreads, (uint32_t)kThreads};
    else h = H{out_d, out_v, 1u, 1u};
    float maxd = a.radius;
    uint32_t cnt = 0;

    // kdtree.h's recursion as a loop: the near child is searched next; the
    // far child is stacked with the distance to its splitting plane and
    // searched once the near side is done, if that distance is still below
    // the radius then.  Only far children are stacked (at most depth - 1
    // entries).  The stack lives in scratch (an array of structs, which the
    // compiler does not turn into registers indexed through selects), with
    // its top entry in registers: a pop loads the next top while the popped
    // subtree is walked, so the walk does not wait on scratch memory before
    // its next node.  (k_photon_subsurface walks the same way; the two loops
    // are kept apart because a shared walker taking lambdas compiled to
    // slower code.)
    struct Far {
        uint32_t l, r;
        float pd;
    } st[STACK], top{};
    uint32_t sp = 0;  // entries, the top included
    uint32_t l = 0, r = a.n;
    for (;;) {
        if (r <= l) {
            // this subtree is done: the last stacked far child whose plane
            // is still within the radius
            while (sp > 0 && !(top.pd < maxd)) {
                --sp;
                if (sp > 0) top = st[sp - 1];
            }
            if (sp == 0) break;
            l = top.l;
            r = top.r;
            --sp;
            if (sp > 0) top = st[sp - 1];
        }
        const uint32_t med = l + (r - l) / 2;
        // a subtree with no photon of the wanted shape takes none, so the
        // heap and radius leave it as they entered: passing over it keeps
        // every later step of the walk, and so the result, the same
        if (want && !(a.sig[med] & want_bit)) { r = l; continue; }
        const rt::PhotonNode nd = a.nodes[med];
        const float d = norm3(p[0] - nd.x, p[1] - nd.y, p[2] - nd.z);
        const bool take = d < maxd && (want == 0u || (nd.meta >> 2) == want);
        if (r - l > 1) {
            const uint32_t ax = nd.meta & 3u;
            const float split = ax == 0 ? nd.x : (ax == 1 ? nd.y : nd.z);
            const float diff = p[ax] - split;
            const float pd = norm3(ax == 0 ? diff : 0.0f, ax == 1 ? diff : 0.0f, ax == 2 ? diff : 0.0f);
            uint32_t fl, fr;
            if (p[ax] < split) { fl = med + 1; fr = r; r = med; }
            else { fl = l; fr = med; l = med + 1; }
            if (fr > fl) {
                if (sp >= STACK) break;  // unreachable: the host checks depth - 1 <= STACK
                if (sp > 0) st[sp - 1] = top;
                top = Far{fl, fr, pd};
                ++sp;
            }
        } else {
            r = l;
        }
        // the stacked plane distance is compared with the radius only when
        // popped, so the heap step may follow the push
        if (take) {
            // push_back; make_heap at k; beyond k push_heap + pop_heap +
            // pop_back and max_distance = the front's (kdtree.h:123-130)
            if (cnt < a.k) {
                h.dist(cnt) = d;
                if constexpr (SLOT) {
                    out_d[cnt] = __uint_as_float(med);
                    h.id(cnt) = (uint8_t)cnt;
                } else {
                    h.id(cnt) = med;
                }
                ++cnt;
                if (cnt == a.k) heap_make(h, a.k);
            } else {
                if constexpr (SLOT) {
                    // The farthest of the k + 1 is dropped.  That is the root,
                    // unless the new photon is farther: the radius becomes
                    // the root's distance only at the first replacement, so
                    // the first photon after make_heap can be, and then it
                    // rises to the root and is dropped itself.  It moves
                    // with the spare slot number k, which names no photon.
                    const uint32_t slot = h.dist(0) < d ? a.k : h.id(0);
                    if (slot < a.k) out_d[slot] = __uint_as_float(med);
                    heap_push_pop(h, a.k, d, slot);
                } else {
                    heap_push_pop(h, a.k, d, med);
                }
                maxd = h.dist(0);
            }
        }
    }
    if constexpr (SLOT) {
        for (uint32_t i = 0; i < cnt; ++i) {
            const uint32_t pos = __float_as_uint(out_d[h.id(i)]);
            out_v[i * a.istride] = a.order ? a.order[pos] : pos;
        }
        if (a.qmax) {
            float mx = 0.0f;
            for (uint32_t i = 0; i < cnt; ++i) mx = fmaxf(mx, h.dist(i));
            a.qmax[q] = mx;
        } else {
            for (uint32_t i = 0; i < cnt; ++i) out_d[i] = h.dist(i);  // over the slots, all read above
        }
    } else {
        if (a.order)
            for (uint32_t i = 0; i < cnt; ++i) out_v[i] = a.order[out_v[i]];
    }
    a.count[q] = cnt;
}

// combinePhotonContributions (integration/Scene.cu:607-634) on a Lambertian
// surface: sum over the query's photons of radiance * kd * INV_PI *
// |cos(n, -photon direction)| (LambertianBSDF::lightReflected), times
// 1 / (PI * maxDistance^2) in float, maxDistance the farthest photon found.
__global__ __launch_bounds__(256) void k_photon_gather_lambert(const double *rgb, const float *dir, const float *nrm,
                                                               const float *kd, uint32_t m, uint32_t k,
                                                               const uint32_t *count, const uint32_t *index,
                                                               const float *dist, double *out)
{
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    if (q >= m) return;
    const uint32_t c = count[q];
    double acc[3] = {0.0, 0.0, 0.0};
    if (c > 0) {
        const float inv_pi = 1.0f / 3.14159265358979323846f;
        const double f0 = (double)kd[3 * (size_t)q] * (double)inv_pi, f1 = (double)kd[3 * (size_t)q + 1] * (double)inv_pi,
                     f2 = (double)kd[3 * (size_t)q + 2] * (double)inv_pi;
        const float nx = nrm[3 * (size_t)q], ny = nrm[3 * (size_t)q + 1], nz = nrm[3 * (size_t)q + 2];
        float maxd = 0.0f;
        for (uint32_t i = 0; i < c; ++i) {
            const uint32_t e = index[(size_t)q * k + i];
            maxd = fmaxf(maxd, dist[(size_t)q * k + i]);
            const float cosv = fabsf(-(nx * dir[3 * (size_t)e] + ny * dir[3 * (size_t)e + 1] + nz * dir[3 * (size_t)e + 2]));
            acc[0] += rgb[3 * (size_t)e] * (f0 * (double)cosv);
            acc[1] += rgb[3 * (size_t)e + 1] * (f1 * (double)cosv);
            acc[2] += rgb[3 * (size_t)e + 2] * (f2 * (double)cosv);
        }
        const float kernel = 1.0f / (3.14159265358979323846f * maxd * maxd);
        acc[0] *= (double)kernel; acc[1] *= (double)kernel; acc[2] *= (double)kernel;
    }
    out[3 * (size_t)q] = acc[0];
    out[3 * (size_t)q + 1] = acc[1];
    out[3 * (size_t)q + 2] = acc[2];
}

// combinePhotonContributions for any material: the photon's radiance times
// SurfaceInteraction::lightScattered(-photon direction), i.e. the material's
// lightReflected (f * |cos wi|, the renderer's bsdf_f) with wo the view
// direction, both in the shading frame of the query's normal; Lambertian and
// Trowbridge materials only (plain: lightReflected = 0; the rest are not
// scattering).  Sum in double in the search's order, times 1 / (PI * maxd^2)
// in float.
// index: entry i of query q at index[q * qstride + i * istride]; qmax (if
// set): each query's farthest distance, read instead of its row of dist.
__global__ __launch_bounds__(256) void k_photon_gather_mat(const double *rgb, const float *dir, const float *nrm,
                                                           const float *view, const rt_material *mats, uint32_t m,
                                                           uint32_t k, const uint32_t *count, const uint32_t *index,
                                                           uint32_t qstride, uint32_t istride, const float *dist,
                                                           const float *qmax, double *out, const uint32_t *m_dev)
{
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    if (q >= (m_dev ? min(m, *m_dev) : m)) return;
    const uint32_t c = count[q];
    const rt_material mt = mats[q];
    double acc[3] = {0.0, 0.0, 0.0};
    if (c > 0 && rt::is_scattering(mt.type)) {
        const rt::V3 n = rt::mk(nrm[3 * (size_t)q], nrm[3 * (size_t)q + 1], nrm[3 * (size_t)q + 2]);
        rt::V3 t, b;
        rt::onb(n, t, b);
        const rt::V3 vw = rt::mk(view[3 * (size_t)q], view[3 * (size_t)q + 1], view[3 * (size_t)q + 2]);
        const rt::V3 wo = rt::mk(rt::dot(vw, t), rt::dot(vw, b), rt::dot(vw, n));
        float maxd = qmax ? qmax[q] : 0.0f;
        const uint32_t *ix = index + (size_t)q * qstride;
        for (uint32_t i = 0; i < c; ++i) {
            const uint32_t e = ix[(size_t)i * istride];
            if (!qmax) maxd = fmaxf(maxd, dist[(size_t)q * k + i]);
            const rt::V3 w = rt::mk(-dir[3 * (size_t)e], -dir[3 * (size_t)e + 1], -dir[3 * (size_t)e + 2]);
            const rt::V3 wi = rt::mk(rt::dot(w, t), rt::dot(w, b), rt::dot(w, n));
            const rt::V3 f = rt::bsdf_f(mt, wo, wi);
            acc[0] += rgb[3 * (size_t)e] * (double)f.x;
            acc[1] += rgb[3 * (size_t)e + 1] * (double)f.y;
            acc[2] += rgb[3 * (size_t)e + 2] * (double)f.z;
        }
        const float kernel = 1.0f / (3.14159265358979323846f * maxd * maxd);
        acc[0] *= (double)kernel; acc[1] *= (double)kernel; acc[2] *= (double)kernel;
    }
    out[3 * (size_t)q] = acc[0];
    out[3 * (size_t)q + 1] = acc[1];
    out[3 * (size_t)q + 2] = acc[2];
}

// impPaper_dipoleDiffusionAproximation (geometricMath.cu:451-483) for one
// channel: float scalars widened where the reference's Spectrum (double)
// meets them.
__device__ inline double dipole_rd(double ss, double sa, float g, float fdr, float A, float r)
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

// cosine(v1, v2) (geometricMath.cu:139-142)
__device__ inline float cosine3(float ax, float ay, float az, float bx, float by, float bz)
{
    const float d = ax * bx + ay * by + az * bz;
    return d / (sqrtf(ax * ax + ay * ay + az * az) * sqrtf(bx * bx + by * by + bz * bz));
}

struct SssArgs {
    const uint32_t *perm;
    const rt::PhotonNode *nodes;
    const uint32_t *order;
    const uint32_t *sig;
    uint32_t n;
    const double *rgb;
    const float *pdir, *pnrm;
    const float *queries, *qnrm, *qview;
    const int32_t *qshape;
    uint32_t m;
    float radius;
    rt_sss_params prm;
    double *out;
};

// Scene::combinePhotonsSubsurface over every photon strictly within the
// radius on the query's shape (the unlimited search never forms a heap, so
// photons are summed in the walk's insertion order).
__global__ __launch_bounds__(kThreads) void k_photon_subsurface(SssArgs a)
{
    const uint32_t lane = blockIdx.x * kThreads + threadIdx.x;
    if (lane >= a.m) return;
    const uint32_t q = a.perm ? a.perm[lane] : lane;
    const float p[3] = {a.queries[3 * (size_t)q], a.queries[3 * (size_t)q + 1], a.queries[3 * (size_t)q + 2]};
    const uint32_t want = a.qshape ? (uint32_t)(a.qshape[q] + 1) : 0u;
    const uint32_t want_bit = rt::photon_shape_bit(want);
    const float eta = a.prm.eta, g = a.prm.g;
    const float fdr = (float)((-1.440 / (double)(eta * eta)) + (0.710 / (double)eta) + 0.668 + 0.0636 * (double)eta);
    const float A = (1 + fdr) / (1 - fdr);
    // Ft of the camera side: 1 - fresnelDielectric(cos(normal, view), 1, eta)
    const float cos_o = cosine3(a.qnrm[3 * (size_t)q], a.qnrm[3 * (size_t)q + 1], a.qnrm[3 * (size_t)q + 2],
                                a.qview[3 * (size_t)q], a.qview[3 * (size_t)q + 1], a.qview[3 * (size_t)q + 2]);
    const double ft_o = (double)(1 - rt::fresnel_dielectric(cos_o, 1.0f, eta));
    const float inv_pi = 1.0f / 3.14159265358979323846f;
    double acc[3] = {0.0, 0.0, 0.0};
    float maxd = 0.0f;
    // the k-nearest search's walk (near child next, far children stacked in
    // scratch) with a fixed radius
    struct Far {
        uint32_t l, r;
        float pd;
    } st[kStack], top{};
    uint32_t sp = 0;  // entries, the top (in registers) included
    uint32_t l = 0, r = a.n;
    for (;;) {
        if (r <= l) {
            while (sp > 0 && !(top.pd < a.radius)) {
                --sp;
                if (sp > 0) top = st[sp - 1];
            }
            if (sp == 0) break;
            l = top.l;
            r = top.r;
            --sp;
            if (sp > 0) top = st[sp - 1];
        }
        const uint32_t med = l + (r - l) / 2;
        if (want && !(a.sig[med] & want_bit)) { r = l; continue; }  // no photon of the shape below: nothing to sum
        const rt::PhotonNode nd = a.nodes[med];
        const float d = norm3(p[0] - nd.x, p[1] - nd.y, p[2] - nd.z);
        if (d < a.radius && (want == 0u || (nd.meta >> 2) == want)) {
            const uint32_t e = a.order[med];
            const float *pd = a.pdir + 3 * (size_t)e, *pn = a.pnrm + 3 * (size_t)e;
            const float cos_i = cosine3(pn[0], pn[1], pn[2], -pd[0], -pd[1], -pd[2]);
            const double ft_i = (double)(1 - rt::fresnel_dielectric(cos_i, 1.0f, eta));
            for (int c = 0; c < 3; ++c) {
                const double rd = dipole_rd((double)a.prm.sigma_s[c], (double)a.prm.sigma_a[c], g, fdr, A, d);
                const double sd = (double)inv_pi * ft_i * rd * ft_o;  // diffusionTermBSSRDF (geometricMath.cu:500-522)
                acc[c] += a.rgb[3 * (size_t)e + c] * sd * (double)a.prm.kd[c];
            }
            maxd = fmaxf(maxd, d);
        }
        if (r - l > 1) {
            const uint32_t ax = nd.meta & 3u;
            const float split = ax == 0 ? nd.x : (ax == 1 ? nd.y : nd.z);
            const float diff = p[ax] - split;
            const float pdd = norm3(ax == 0 ? diff : 0.0f, ax == 1 ? diff : 0.0f, ax == 2 ? diff : 0.0f);
            uint32_t fl, fr;
            if (p[ax] < split) { fl = med + 1; fr = r; r = med; }
            else { fl = l; fr = med; l = med + 1; }
            if (fr > fl) {
                if (sp >= kStack) break;  // unreachable: the host checks depth + 1 <= kStack
                if (sp > 0) st[sp - 1] = top;
                top = Far{fl, fr, pdd};
                ++sp;
            }
        } else {
            r = l;
        }
    }
    // contribution /= 2 PI maxD^2 (no photon: 0 / 0)
    const float den = 2.0f * 3.14159265358979323846f * maxd * maxd;
    for (int c = 0; c < 3; ++c) a.out[3 * (size_t)q + c] = acc[c] / (double)den;
}

// Radiance and direction of tree position i = those of photon order[i].
__global__ __launch_bounds__(256) void k_photon_tree_copy(const uint32_t *order, uint32_t n, const double *rgb,
                                                          const float *dir, double *trgb, float *tdir)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const size_t e = order[i];
    for (int c = 0; c < 3; ++c) {
        trgb[3 * (size_t)i + c] = rgb[3 * e + c];
        tdir[3 * (size_t)i + c] = dir[3 * e + c];
    }
}

// The search kernel for a k and a tree depth: the LDS heap up to kLdsK,
// and the smaller register stack when the tree is shallow enough for it
// (every map of up to 2^25 photons).
constexpr uint32_t kStackSmall = 24;
void launch_knn(const KnnArgs &a, uint32_t depth, uint32_t blocks, hipStream_t s)
{
    const bool small = depth <= kStackSmall + 1;
    if (a.k <= kLdsK) {
        const size_t lds = (size_t)a.k * kThreads * 5;  // float distance + slot byte per entry
        if (small) k_photon_knn<kHeapLdsSlot, kStackSmall><<<blocks, kThreads, lds, s>>>(a);
        else k_photon_knn<kHeapLdsSlot, kStack><<<blocks, kThreads, lds, s>>>(a);
    } else {
        if (small) k_photon_knn<kHeapGlobal, kStackSmall><<<blocks, kThreads, 0, s>>>(a);
        else k_photon_knn<kHeapGlobal, kStack><<<blocks, kThreads, 0, s>>>(a);
    }
}

}  // namespace

// Morton order of nq query points in the map's work space (perm: lane ->
// query); the work space is shared by every search on the map.
static rt_status sort_queries(rt_photon_map *m, const float *queries, uint32_t nq, hipStream_t stream,
                              const uint32_t **perm, const char *who)
{
    size_t sort_bytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                                      (uint32_t *)nullptr, (uint32_t *)nullptr, (int)nq, 0, 30,
                                                      stream);
    const size_t ids = ((size_t)nq * sizeof(uint32_t) + 255) & ~(size_t)255;
    const size_t need = 4 * ids + sort_bytes;
    if (e == hipSuccess && need > m->work_bytes) {
        if (m->work) {
            e = hipStreamSynchronize(stream);
            if (e == hipSuccess) e = hipFree(m->work);
            m->work = nullptr;
            m->work_bytes = 0;
        }
        if (e == hipSuccess) e = hipMalloc(&m->work, need);
        if (e == hipSuccess) m->work_bytes = need;
    }
    if (e != hipSuccess) {
        set_error(std::string(who) + ": query sort: " + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP;
    }
    uint8_t *w = (uint8_t *)m->work;
    uint32_t *key0 = (uint32_t *)w, *key1 = (uint32_t *)(w + ids), *id0 = (uint32_t *)(w + 2 * ids),
             *id1 = (uint32_t *)(w + 3 * ids);
    k_query_morton<<<(nq + 255) / 256, 256, 0, stream>>>(queries, nq, m->lo[0], m->lo[1], m->lo[2], m->scale[0],
                                                         m->scale[1], m->scale[2], key0, id0, nullptr);
    e = hipGetLastError();
    if (e == hipSuccess)
        e = hipcub::DeviceRadixSort::SortPairs(w + 4 * ids, sort_bytes, key0, key1, id0, id1, (int)nq, 0, 30, stream);
    if (e != hipSuccess) {
        set_error(std::string(who) + ": query sort: " + hipGetErrorString(e));
        return RT_ERR_HIP;
    }
    *perm = id1;
    return RT_OK;
}

namespace rt {

rt_ctx *photon_map_ctx(const rt_photon_map *m) { return m->ctx; }

// Work space of a render's searches: the query sort (keys and ids twice,
// the sort's temporary storage) and the k-nearest results, for cap queries.
static void render_layout(uint32_t cap, uint32_t k, size_t sort_bytes, size_t &ids, size_t &work, size_t &cb,
                          size_t &ib, size_t &gather)
{
    ids = ((size_t)cap * sizeof(uint32_t) + 255) & ~(size_t)255;
    work = 4 * ids + sort_bytes;
    cb = ((size_t)cap * 4 + 255) & ~(size_t)255;
    ib = ((size_t)cap * k * 4 + 255) & ~(size_t)255;
    gather = 2 * cb + 2 * ib;  // counts, indices, distances (the search's slot pool), farthest distances
}

hipError_t photon_render_reserve(rt_photon_map *m, uint32_t cap, uint32_t k, const double *photon_rgb,
                                 const float *photon_dir, hipStream_t s, std::string &err)
{
    size_t sort_bytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                                      (uint32_t *)nullptr, (uint32_t *)nullptr, (int)cap, 0, 32, s);
    if (e != hipSuccess) { err = std::string("photon search sizing: ") + hipGetErrorString(e); return e; }
    size_t ids, work, cb, ib, gather;
    render_layout(cap, k, sort_bytes, ids, work, cb, ib, gather);
    if (work > m->work_bytes || gather > m->gather_bytes) {
        if ((e = hipStreamSynchronize(s)) != hipSuccess) { err = hipGetErrorString(e); return e; }
    }
    if (work > m->work_bytes) {
        if (m->work) (void)hipFree(m->work);
        m->work = nullptr;
        m->work_bytes = 0;
        if ((e = hipMalloc(&m->work, work)) != hipSuccess) { err = "photon search work space"; return e; }
        m->work_bytes = work;
    }
    if (gather > m->gather_bytes) {
        if (m->gather) (void)hipFree(m->gather);
        m->gather = nullptr;
        m->gather_bytes = 0;
        if ((e = hipMalloc(&m->gather, gather)) != hipSuccess) { err = "photon search results"; return e; }
        m->gather_bytes = gather;
    }
    if (m->n && !m->trgb) {
        if ((e = hipMalloc(&m->trgb, (size_t)m->n * 24)) != hipSuccess) { err = "photon radiance in tree order"; return e; }
        if ((e = hipMalloc(&m->tdir, (size_t)m->n * 12)) != hipSuccess) { err = "photon directions in tree order"; return e; }
    }
    if (m->n) {
        k_photon_tree_copy<<<(m->n + 255) / 256, 256, 0, s>>>(m->order, m->n, photon_rgb, photon_dir, m->trgb,
                                                              m->tdir);
        if ((e = hipGetLastError()) != hipSuccess) { err = "photon tree-order copy"; return e; }
    }
    return hipSuccess;
}

hipError_t photon_gather_render(rt_photon_map *m, const float *points, const float *normals, const float *views,
                                const rt_material *mats, const int32_t *qshape, uint32_t cap,
                                const uint32_t *n_dev, uint32_t k, float radius, double *out, hipStream_t s)
{
    if (cap == 0) return hipSuccess;
    size_t sort_bytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                                      (uint32_t *)nullptr, (uint32_t *)nullptr, (int)cap, 0, 32, s);
    if (e != hipSuccess) return e;
    size_t ids, work, cb, ib, gather;
    render_layout(cap, k, sort_bytes, ids, work, cb, ib, gather);
    if (work > m->work_bytes || gather > m->gather_bytes) return hipErrorInvalidValue;  // reserve first
    uint8_t *w = (uint8_t *)m->work;
    uint32_t *key0 = (uint32_t *)w, *key1 = (uint32_t *)(w + ids), *id0 = (uint32_t *)(w + 2 * ids),
             *id1 = (uint32_t *)(w + 3 * ids);
    k_query_morton<<<(cap + 255) / 256, 256, 0, s>>>(points, cap, m->lo[0], m->lo[1], m->lo[2], m->scale[0],
                                                      m->scale[1], m->scale[2], key0, id0, n_dev);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipcub::DeviceRadixSort::SortPairs(w + 4 * ids, sort_bytes, key0, key1, id0, id1, (int)cap, 0, 32, s)) !=
        hipSuccess)
        return e;
    uint8_t *gw = (uint8_t *)m->gather;
    uint32_t *cnt = (uint32_t *)gw, *idx = (uint32_t *)(gw + cb);
    float *dist = (float *)(gw + cb + ib), *qmax = (float *)(gw + cb + 2 * ib);
    // Results as tree positions, read against the tree-ordered copies.  With
    // the heap in LDS the indices are stored entry-major (entry i of every
    // query together), so the density estimate's lanes read them coalesced,
    // and only each query's farthest distance is written.
    KnnArgs a{id1, m->nodes, nullptr, m->sig, m->n, points, qshape, cap, k, radius, cnt, idx, dist, n_dev};
    const bool lds_heap = k <= kLdsK;
    if (lds_heap) {
        a.qstride = 1;
        a.istride = cap;
        a.qmax = qmax;
    }
    const unsigned blocks = (cap + kThreads - 1) / kThreads;
    launch_knn(a, m->depth, blocks, s);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_photon_gather_mat<<<(cap + 255) / 256, 256, 0, s>>>(m->trgb, m->tdir, normals, views, mats, cap, k, cnt, idx,
                                                          lds_heap ? 1u : k, lds_heap ? cap : 1u, dist,
                                                          lds_heap ? qmax : nullptr, out, n_dev);
    return hipGetLastError();
}

}  // namespace rt

extern "C" {

rt_status rt_photon_map_build(rt_ctx *ctx, const float *xyz, const int32_t *shape, uint32_t n, rt_photon_map **out)
{
    if (!out) { set_error("rt_photon_map_build: out is null"); return RT_ERR_INVALID; }
    *out = nullptr;
    if (!ctx || (n && !xyz)) { set_error("rt_photon_map_build: null argument"); return RT_ERR_INVALID; }
    if (n > (1u << 30)) { set_error("rt_photon_map_build: at most 2^30 photons"); return RT_ERR_INVALID; }
    for (size_t i = 0; i < 3 * (size_t)n; ++i)
        if (!std::isfinite(xyz[i])) { set_error("rt_photon_map_build: non-finite photon location"); return RT_ERR_INVALID; }
    if (shape)
        for (uint32_t i = 0; i < n; ++i)
            if (shape[i] < -1 || shape[i] > (1 << 30) - 2) {
                set_error("rt_photon_map_build: shape ids must be -1 .. 2^30 - 2");
                return RT_ERR_INVALID;
            }
    rt::HostPhotonTree t;
    rt::build_photon_tree(xyz, shape, n, t);
    if (t.depth + 1 > kStack) { set_error("rt_photon_map_build: tree too deep"); return RT_ERR_INVALID; }
    rt_photon_map *m = new (std::nothrow) rt_photon_map();
    if (!m) { set_error("rt_photon_map_build: host allocation failed"); return RT_ERR_OOM; }
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(ctx));
    hipError_t e = hipSetDevice(rt::ctx_device(ctx));
    if (e == hipSuccess && n) e = hipMalloc(&m->nodes, (size_t)n * sizeof(rt::PhotonNode));
    if (e == hipSuccess && n) e = hipMalloc(&m->order, (size_t)n * sizeof(uint32_t));
    if (e == hipSuccess && n) e = hipMalloc(&m->sig, (size_t)n * sizeof(uint32_t));
    if (e == hipSuccess && n) e = hipMemcpy(m->nodes, t.nodes.data(), (size_t)n * sizeof(rt::PhotonNode), hipMemcpyHostToDevice);
    if (e == hipSuccess && n) e = hipMemcpy(m->order, t.order.data(), (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && n) e = hipMemcpy(m->sig, t.sig.data(), (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (m->nodes) (void)hipFree(m->nodes);
        if (m->order) (void)hipFree(m->order);
        if (m->sig) (void)hipFree(m->sig);
        delete m;
        set_error(std::string("rt_photon_map_build: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP;
    }
    m->ctx = ctx;
    m->n = n;
    m->depth = t.depth;
    if (n) {
        float hi[3];
        for (int a = 0; a < 3; ++a) m->lo[a] = hi[a] = xyz[a];
        for (size_t i = 1; i < n; ++i)
            for (int a = 0; a < 3; ++a) {
                m->lo[a] = std::min(m->lo[a], xyz[3 * i + a]);
                hi[a] = std::max(hi[a], xyz[3 * i + a]);
            }
        for (int a = 0; a < 3; ++a) m->scale[a] = hi[a] > m->lo[a] ? 1024.0f / (hi[a] - m->lo[a]) : 0.0f;
    }
    rt::ctx_retain(ctx);
    *out = m;
    return RT_OK;
}

rt_status rt_photon_map_destroy(rt_photon_map *m)
{
    if (!m) return RT_OK;
    {
        std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(m->ctx));
        (void)hipSetDevice(rt::ctx_device(m->ctx));
        (void)hipStreamSynchronize((hipStream_t)rt::ctx_stream_ptr(m->ctx));
        if (m->nodes) (void)hipFree(m->nodes);
        if (m->order) (void)hipFree(m->order);
        if (m->sig) (void)hipFree(m->sig);
        if (m->trgb) (void)hipFree(m->trgb);
        if (m->tdir) (void)hipFree(m->tdir);
        if (m->work) (void)hipFree(m->work);
        if (m->gather) (void)hipFree(m->gather);
    }
    rt::ctx_release(m->ctx);
    delete m;
    return RT_OK;
}

rt_status rt_photon_map_get_info(const rt_photon_map *m, uint32_t *n, uint32_t *depth)
{
    if (!m) { set_error("rt_photon_map_get_info: null map"); return RT_ERR_INVALID; }
    if (n) *n = m->n;
    if (depth) *depth = m->depth;
    return RT_OK;
}

rt_status rt_photon_map_knn(rt_photon_map *m, const float *queries, const int32_t *query_shape, uint32_t nq,
                            uint32_t k, float max_distance, uint32_t *count, uint32_t *index, float *distance,
                            void *stream_)
{
    if (!m || (nq && (!queries || !count || !index || !distance))) {
        set_error("rt_photon_map_knn: null argument");
        return RT_ERR_INVALID;
    }
    if (k == 0 || k > kMaxK) { set_error("rt_photon_map_knn: k must be 1..1024"); return RT_ERR_INVALID; }
    if (!(max_distance > 0.0f)) { set_error("rt_photon_map_knn: max_distance must be positive"); return RT_ERR_INVALID; }
    if ((size_t)nq * k > (1ull << 31) || nq > (1u << 30)) {
        set_error("rt_photon_map_knn: at most 2^30 queries and queries * k <= 2^31");
        return RT_ERR_INVALID;
    }
    if (nq == 0) return RT_OK;
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(m->ctx));
    RT_HIP(hipSetDevice(rt::ctx_device(m->ctx)));
    hipStream_t stream = stream_ ? (hipStream_t)stream_ : (hipStream_t)rt::ctx_stream_ptr(m->ctx);
    // queries in Morton order of their position, so the lanes of a wavefront
    // walk neighbouring parts of the tree (results go to the callers' slots)
    const uint32_t *id1 = nullptr;
    rt_status sst = sort_queries(m, queries, nq, stream, &id1, "rt_photon_map_knn");
    if (sst != RT_OK) return sst;
    KnnArgs a{id1, m->nodes, m->order, m->sig, m->n, queries, query_shape, nq, k, max_distance, count, index, distance,
              nullptr};
    const unsigned blocks = (nq + kThreads - 1) / kThreads;
    launch_knn(a, m->depth, blocks, stream);
    RT_HIP(hipGetLastError());
    return RT_OK;
}

rt_status rt_photon_gather_lambert(rt_photon_map *m, const double *photon_rgb, const float *photon_dir,
                                   const float *points, const float *normals, const float *kd,
                                   const int32_t *query_shape, uint32_t nq, uint32_t k, float max_distance,
                                   double *out, void *stream_)
{
    if (!m || (nq && (!points || !normals || !kd || !out)) || (m->n && nq && (!photon_rgb || !photon_dir))) {
        set_error("rt_photon_gather_lambert: null argument");
        return RT_ERR_INVALID;
    }
    if (nq == 0) return RT_OK;
    if (k == 0 || k > kMaxK || (size_t)nq * k > (1ull << 31) || nq > (1u << 30)) {
        set_error("rt_photon_gather_lambert: k must be 1..1024, queries <= 2^30, queries * k <= 2^31");
        return RT_ERR_INVALID;
    }
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(m->ctx));
    RT_HIP(hipSetDevice(rt::ctx_device(m->ctx)));
    hipStream_t stream = stream_ ? (hipStream_t)stream_ : (hipStream_t)rt::ctx_stream_ptr(m->ctx);
    const size_t cb = ((size_t)nq * 4 + 255) & ~(size_t)255, ib = ((size_t)nq * k * 4 + 255) & ~(size_t)255;
    const size_t need = cb + 2 * ib;
    if (need > m->gather_bytes) {
        if (m->gather) {
            RT_HIP(hipStreamSynchronize(stream));
            RT_HIP(hipFree(m->gather));
            m->gather = nullptr;
            m->gather_bytes = 0;
        }
        RT_HIP(hipMalloc(&m->gather, need));
        m->gather_bytes = need;
    }
    uint8_t *w = (uint8_t *)m->gather;
    uint32_t *cnt = (uint32_t *)w, *idx = (uint32_t *)(w + cb);
    float *dist = (float *)(w + cb + ib);
    rt_status st = rt_photon_map_knn(m, points, query_shape, nq, k, max_distance, cnt, idx, dist, stream);
    if (st != RT_OK) return st;
    k_photon_gather_lambert<<<(nq + 255) / 256, 256, 0, stream>>>(photon_rgb, photon_dir, normals, kd, nq, k, cnt,
                                                                  idx, dist, out);
    RT_HIP(hipGetLastError());
    return RT_OK;
}

// k nearest into the map's gather buffers (shared by the density estimates)
static rt_status gather_knn(rt_photon_map *m, const float *points, const int32_t *query_shape, uint32_t nq,
                            uint32_t k, float max_distance, hipStream_t stream, uint32_t **cnt, uint32_t **idx,
                            float **dist)
{
    const size_t cb = ((size_t)nq * 4 + 255) & ~(size_t)255, ib = ((size_t)nq * k * 4 + 255) & ~(size_t)255;
    const size_t need = cb + 2 * ib;
    if (need > m->gather_bytes) {
        if (m->gather) {
            RT_HIP(hipStreamSynchronize(stream));
            RT_HIP(hipFree(m->gather));
            m->gather = nullptr;
            m->gather_bytes = 0;
        }
        RT_HIP(hipMalloc(&m->gather, need));
        m->gather_bytes = need;
    }
    uint8_t *w = (uint8_t *)m->gather;
    *cnt = (uint32_t *)w;
    *idx = (uint32_t *)(w + cb);
    *dist = (float *)(w + cb + ib);
    return rt_photon_map_knn(m, points, query_shape, nq, k, max_distance, *cnt, *idx, *dist, stream);
}

rt_status rt_photon_gather(rt_photon_map *m, const double *photon_rgb, const float *photon_dir, const float *points,
                           const float *normals, const float *views, const rt_material *materials,
                           const int32_t *query_shape, uint32_t nq, uint32_t k, float max_distance, double *out,
                           void *stream_)
{
    if (!m || (nq && (!points || !normals || !views || !materials || !out)) ||
        (m->n && nq && (!photon_rgb || !photon_dir))) {
        set_error("rt_photon_gather: null argument");
        return RT_ERR_INVALID;
    }
    if (nq == 0) return RT_OK;
    if (k == 0 || k > kMaxK || (size_t)nq * k > (1ull << 31) || nq > (1u << 30)) {
        set_error("rt_photon_gather: k must be 1..1024, queries <= 2^30, queries * k <= 2^31");
        return RT_ERR_INVALID;
    }
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(m->ctx));
    RT_HIP(hipSetDevice(rt::ctx_device(m->ctx)));
    hipStream_t stream = stream_ ? (hipStream_t)stream_ : (hipStream_t)rt::ctx_stream_ptr(m->ctx);
    uint32_t *cnt, *idx;
    float *dist;
    rt_status st = gather_knn(m, points, query_shape, nq, k, max_distance, stream, &cnt, &idx, &dist);
    if (st != RT_OK) return st;
    k_photon_gather_mat<<<(nq + 255) / 256, 256, 0, stream>>>(photon_rgb, photon_dir, normals, views, materials, nq,
                                                              k, cnt, idx, k, 1u, dist, nullptr, out, nullptr);
    RT_HIP(hipGetLastError());
    return RT_OK;
}

rt_status rt_photon_subsurface(rt_photon_map *m, const double *photon_rgb, const float *photon_dir,
                               const float *photon_nrm, const float *points, const float *normals,
                               const float *views, const int32_t *query_shape, uint32_t nq, float radius,
                               const rt_sss_params *params, double *out, void *stream_)
{
    if (!m || !params || (nq && (!points || !normals || !views || !out)) ||
        (m->n && nq && (!photon_rgb || !photon_dir || !photon_nrm))) {
        set_error("rt_photon_subsurface: null argument");
        return RT_ERR_INVALID;
    }
    if (!(radius > 0.0f)) { set_error("rt_photon_subsurface: radius must be positive"); return RT_ERR_INVALID; }
    if (nq > (1u << 30)) { set_error("rt_photon_subsurface: at most 2^30 queries"); return RT_ERR_INVALID; }
    if (nq == 0) return RT_OK;
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(m->ctx));
    RT_HIP(hipSetDevice(rt::ctx_device(m->ctx)));
    hipStream_t stream = stream_ ? (hipStream_t)stream_ : (hipStream_t)rt::ctx_stream_ptr(m->ctx);
    const uint32_t *perm = nullptr;
    rt_status st = sort_queries(m, points, nq, stream, &perm, "rt_photon_subsurface");
    if (st != RT_OK) return st;
    SssArgs a{perm, m->nodes, m->order, m->sig, m->n, photon_rgb, photon_dir, photon_nrm, points, normals, views,
              query_shape, nq, radius, *params, out};
    k_photon_subsurface<<<(nq + kThreads - 1) / kThreads, kThreads, 0, stream>>>(a);
    RT_HIP(hipGetLastError());
    return RT_OK;
}

}  // extern "C"
