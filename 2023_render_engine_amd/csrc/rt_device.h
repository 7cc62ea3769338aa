// rt_device.h -- device-side geometry shared by the HIP kernels: the
// reference's watertight ray/triangle test in exact FP32 operation order, the
// fused slab test of the two-wide BVH and the two-wide traversal itself.
//
// Every translation unit that includes this file is compiled with
// -ffp-contract=off: the only fused multiply-adds are the explicit fmaf()
// calls of the slab test, whose rounding the box padding absorbs.
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstdint>

#include "rt_trace.h"

namespace rt {

// errorBound(n) = (n * eps/2) / (1 - n * eps/2) in double, rounded to float
// (cuda_libraries/math.cu:117-120, types.h:23).
__host__ __device__ constexpr float error_bound(int n)
{
    return (float)((n * ((double)FLT_EPSILON * 0.5)) / (1 - n * ((double)FLT_EPSILON * 0.5)));
}

constexpr float kTfarScale = 1.00000012f;  // 1 + 2^-23: slab rounding slack
constexpr float kHitScale = 1.0000005f;    // cull against the current hit with slack

// Per-ray constants of the watertight test (shapes/Triangle.cu:452-495):
// the permutation by the largest |d| component and the shear.  They depend
// on the ray only, so they are hoisted out of the triangle loop; the values
// are the same FP32 numbers the reference computes per call.
struct TriRay {
    float ox, oy, oz;  // origin, permuted (x <- kx, y <- ky, z <- kz)
    float Sx, Sy, Sz;
    float tmax;        // ray.maximumOffset
    int kx, ky, kz;
};

__device__ __forceinline__ float sel3(int k, float a, float b, float c)
{
    return k == 0 ? a : (k == 1 ? b : c);
}

__device__ __forceinline__ TriRay make_tri_ray(float ox, float oy, float oz, float dx, float dy,
                                               float dz, float tmax)
{
    TriRay r;
    float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    int kz = (ax > ay) ? 0 : 1;
    if (az > (kz == 0 ? ax : ay)) kz = 2;
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    r.kx = kx; r.ky = ky; r.kz = kz;
    float pdx = sel3(kx, dx, dy, dz), pdy = sel3(ky, dx, dy, dz), pdz = sel3(kz, dx, dy, dz);
    r.Sx = -pdx / pdz;
    r.Sy = -pdy / pdz;
    r.Sz = 1.f / pdz;
    r.ox = sel3(kx, ox, oy, oz);
    r.oy = sel3(ky, ox, oy, oz);
    r.oz = sel3(kz, ox, oy, oz);
    r.tmax = tmax;
    return r;
}

__device__ __forceinline__ float fmax3(float a, float b, float c) { return fmaxf(a, fmaxf(b, c)); }

// Triangle::intersects (shapes/Triangle.cu:433-643), FP32 with the FP64
// edge-function fallback.  One step is reordered: the conservative deltaT
// rejection (:575-593) comes before hitOffset is lowered (:569-573), so a
// rejected candidate -- typically the triangle a spawned ray just left, met
// again at t ~ 0 -- leaves hitOffset alone and the result is the closest
// candidate that passes, in any visit order.  In the reference's order that
// candidate lowers hitOffset first and every triangle tested after it is
// rejected, so the outcome depends on the tree's visit order; a nearest-first
// walk would meet it first almost always and lose the bounce.  Rays with no
// such candidate get the reference's result unchanged.  Returns true on
// acceptance.
__device__ __forceinline__ bool tri_intersects(const TriRay &r, const float4 a, const float4 b,
                                               const float4 c, float &hitOffset)
{
    // translate (Triangle.cu:452-454), then permute (:482-484)
    float p0x = sel3(r.kx, a.x, a.y, a.z) - r.ox;
    float p0y = sel3(r.ky, a.x, a.y, a.z) - r.oy;
    float p0z = sel3(r.kz, a.x, a.y, a.z) - r.oz;
    float p1x = sel3(r.kx, b.x, b.y, b.z) - r.ox;
    float p1y = sel3(r.ky, b.x, b.y, b.z) - r.oy;
    float p1z = sel3(r.kz, b.x, b.y, b.z) - r.oz;
    float p2x = sel3(r.kx, c.x, c.y, c.z) - r.ox;
    float p2y = sel3(r.ky, c.x, c.y, c.z) - r.oy;
    float p2z = sel3(r.kz, c.x, c.y, c.z) - r.oz;
    // shear (:490-495)
    p0x += r.Sx * p0z; p0y += r.Sy * p0z;
    p1x += r.Sx * p1z; p1y += r.Sy * p1z;
    p2x += r.Sx * p2z; p2y += r.Sy * p2z;
    // edge functions (:500-522)
    float e0 = p1x * p2y - p1y * p2x;
    float e1 = p2x * p0y - p2y * p0x;
    float e2 = p0x * p1y - p0y * p1x;
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
    }
    if ((e0 < 0 || e1 < 0 || e2 < 0) && (e0 > 0 || e1 > 0 || e2 > 0)) return false;
    float det = e0 + e1 + e2;
    if (det == 0) return false;
    p0z *= r.Sz; p1z *= r.Sz; p2z *= r.Sz;
    float scaled = e0 * p0z + e1 * p1z + e2 * p2z;
    if (det < 0 && (scaled >= 0 || scaled < r.tmax * det)) return false;
    else if (det > 0 && (scaled <= 0 || scaled > r.tmax * det)) return false;
    float invDet = 1 / det;
    float t = scaled * invDet;
    if (t > hitOffset) return false;
    // conservative deltaT rejection (:575-593), before hitOffset is lowered
    float maxZt = fmax3(fabsf(p0z), fabsf(p1z), fabsf(p2z));
    float deltaZ = error_bound(3) * maxZt;
    float maxXt = fmax3(fabsf(p0x), fabsf(p1x), fabsf(p2x));
    float maxYt = fmax3(fabsf(p0y), fabsf(p1y), fabsf(p2y));
    float deltaX = error_bound(5) * (maxXt + maxZt);
    float deltaY = error_bound(5) * (maxYt + maxZt);
    float deltaE = 2 * (error_bound(2) * maxXt * maxYt + deltaY * maxXt + deltaX * maxYt);
    float maxE = fmax3(fabsf(e0), fabsf(e1), fabsf(e2));
    float deltaT = 3 * (error_bound(3) * maxE * maxZt + deltaE * maxZt + deltaZ * maxE) * fabsf(invDet);
    if (t <= deltaT) return false;
    hitOffset = t;
    return true;
}

// Slab-test constants.  A zero direction component is replaced by +-1e-30
// so no 0 * inf NaN can arise; the padded boxes keep the test conservative.
struct BoxRay {
    float ix, iy, iz;     // 1 / d
    float oix, oiy, oiz;  // o / d
};

__device__ __forceinline__ BoxRay make_box_ray(float ox, float oy, float oz, float dx, float dy, float dz)
{
    BoxRay b;
    float ddx = fabsf(dx) > 1e-30f ? dx : copysignf(1e-30f, dx);
    float ddy = fabsf(dy) > 1e-30f ? dy : copysignf(1e-30f, dy);
    float ddz = fabsf(dz) > 1e-30f ? dz : copysignf(1e-30f, dz);
    b.ix = 1.0f / ddx; b.iy = 1.0f / ddy; b.iz = 1.0f / ddz;
    b.oix = ox * b.ix; b.oiy = oy * b.iy; b.oiz = oz * b.iz;
    return b;
}

// Returns tnear; the box is entered iff tnear <= tfar (written to *hit).
__device__ __forceinline__ float slab(const BoxRay &r, float lx, float hx, float ly, float hy,
                                      float lz, float hz, float tlim, bool *hit)
{
    float tx0 = fmaf(lx, r.ix, -r.oix), tx1 = fmaf(hx, r.ix, -r.oix);
    float ty0 = fmaf(ly, r.iy, -r.oiy), ty1 = fmaf(hy, r.iy, -r.oiy);
    float tz0 = fmaf(lz, r.iz, -r.oiz), tz1 = fmaf(hz, r.iz, -r.oiz);
    float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
    float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
    tf = fminf(tf * kTfarScale, tlim);
    *hit = tn <= tf;
    return tn;
}

// Leaf reference decoding: child < 0 is ~(first << 3 | count - 1).
__device__ __forceinline__ void decode_leaf(int32_t c, uint32_t &first, uint32_t &count)
{
    uint32_t enc = ~(uint32_t)c;
    first = enc >> 3;
    count = (enc & 7u) + 1u;
}

// ---------------------------------------------------------------------
// Persistent traversal state machine, shared by the ray-query kernels and
// the renderer's extend / connect kernels.  Each lane walks its ray one node
// visit (or one leaf) per step; the stack's first entries live in LDS laid
// out [depth][lane] with a stride of kTravBlock entries, so a wavefront's
// pushes hit distinct banks.  Its depth is the tree's internal depth + 1,
// which bounds the pending entries, so it cannot overflow.
// ---------------------------------------------------------------------
constexpr int kTravBlock = 128;
// Tuning constants of the persistent traversal (build-time, -D overridable).
#ifndef RT_REFILL_LANES
#define RT_REFILL_LANES 16
#endif
#ifndef RT_CLAIM_CHUNK
#define RT_CLAIM_CHUNK 128
#endif
#ifndef RT_CLAIM_MIN
#define RT_CLAIM_MIN 64
#endif
#ifndef RT_LEAF_BATCH
#define RT_LEAF_BATCH 16
#endif
#ifndef RT_NODE_STEPS
#define RT_NODE_STEPS 4
#endif
#ifndef RT_LEAF_BATCH_ANY
#define RT_LEAF_BATCH_ANY 16
#endif
constexpr uint32_t kRefillLanes = RT_REFILL_LANES;  // refill a wavefront when at least this many lanes are idle

// The traversal stack of a lane: its first L entries in LDS ([depth][lane],
// stride kTravBlock), deeper ones (up to kMaxStack, the most pending entries
// a tree the host accepts can need) in the lane's private memory.  The ray
// tracing kernel keeps all of it in LDS (L = kMaxStack: no private part), as
// its one-block-size launch is not LDS-limited at the depths real trees need;
// the path tracer's extend and connect kernels keep the first 24 entries in
// LDS (a 1M-triangle four-wide tree needs 36 but its rays rarely use more than
// ~20), which leaves room for more workgroups per CU; 24 measured best
// against 16, 20, 28 and 36 on the 1080p procedural scene.
#ifndef RT_TRACE_LDS_STACK
#define RT_TRACE_LDS_STACK 128
#endif
#ifndef RT_PATH_LDS_STACK
#define RT_PATH_LDS_STACK 24
#endif
constexpr int kMaxStack = (int)kMaxTraversalStack;  // the host refuses trees that need more

template <int L>
struct TravStack {
    static_assert(L > 0 && L <= kMaxStack, "the LDS part holds 1..kMaxStack entries");
    static constexpr int kLds = L;
    static constexpr int kTail = kMaxStack - L;  // private entries
    int32_t *lds;   // this lane's column of the workgroup's LDS stack
    int32_t *tail;  // this lane's private overflow entries (kTail of them)
};
using TraceStack = TravStack<RT_TRACE_LDS_STACK>;
using PathStack = TravStack<RT_PATH_LDS_STACK>;

// The private part's storage in a kernel (one entry when there is none, as
// an array cannot be empty; it is then never touched).
template <class S>
using StackTail = int32_t[S::kTail > 0 ? S::kTail : 1];

template <class S>
__device__ __forceinline__ void stack_put(const S &st, int sp, int32_t v)
{
    if constexpr (S::kTail == 0) st.lds[sp * kTravBlock] = v;
    else if (sp < S::kLds) st.lds[sp * kTravBlock] = v;
    else st.tail[sp - S::kLds] = v;
}

template <class S>
__device__ __forceinline__ int32_t stack_get(const S &st, int sp)
{
    if constexpr (S::kTail == 0) return st.lds[sp * kTravBlock];
    else {
        if (sp < S::kLds) return st.lds[sp * kTravBlock];
        return st.tail[sp - S::kLds];
    }
}

// LDS bytes of a workgroup of `block` lanes whose tree needs `stack_cap`
// entries per lane.
template <class S>
inline size_t stack_lds_bytes(uint32_t stack_cap, uint32_t block)
{
    const uint32_t lds = stack_cap < (uint32_t)S::kLds ? stack_cap : (uint32_t)S::kLds;
    return (size_t)lds * block * sizeof(int32_t);
}

// No node or leaf: a reference that no tree produces (a leaf reference is
// ~(first << 3 | count - 1) with first + count <= 2^28).
constexpr int32_t kNoRef = INT32_MIN;

struct TravState {
    TriRay tr;
    BoxRay br;
    float tmax, h, thit;
    int32_t hit, cur;
    int32_t leaf;  // four-wide traversal: a reached leaf waiting for the wavefront's leaf phase
    int sp;
    bool occ;
};

// Closest hit: the triangle test sees ray.maximumOffset = tmax and a running
// hitOffset starting at +inf (BVH.cu:362).  Any hit: shadow rays have
// maximumOffset = +inf and each triangle is judged on its own
// (order-independent), occluding iff accepted at t <= tmax.
__device__ __forceinline__ void trav_begin(TravState &s, float ox, float oy, float oz, float dx, float dy, float dz,
                                           float tmax, bool any, int32_t root = 0)
{
    s.tmax = tmax;
    s.tr = make_tri_ray(ox, oy, oz, dx, dy, dz, any ? INFINITY : tmax);
    s.br = make_box_ray(ox, oy, oz, dx, dy, dz);
    s.h = INFINITY;
    s.thit = INFINITY;
    s.hit = -1;
    s.cur = root;
    s.leaf = kNoRef;
    s.sp = 0;
    s.occ = false;
}

// One traversal step.  Returns true when the ray is finished.  OPAQUE: the
// closest hit skips transparent triangles, as Triangle::intersects does for
// shadow checks (shapes/Triangle.cu:439-440) under
// BVH::intersectRays(..., checkShadows = true) (BVH.cu:541-560).
template <bool ANY, bool STATS, class Stack, bool OPAQUE = false>
__device__ __forceinline__ bool trav_step(TravState &s, const float4 *__restrict__ nodes,
                                          const float4 *__restrict__ tris, const Stack &stack, int cap,
                                          uint32_t &n_node, uint32_t &n_tri)
{
    if (s.cur >= 0) {
        if (STATS) ++n_node;
        const float4 *np = nodes + 4 * (size_t)s.cur;
        const float4 a = np[0], b = np[1], c = np[2], d = np[3];
        const float tlim = ANY ? s.tmax * kHitScale : fminf(s.h, s.tmax) * kHitScale;
        bool h0, h1;
        const float t0 = slab(s.br, a.x, a.y, a.z, a.w, c.x, c.y, tlim, &h0);
        const float t1 = slab(s.br, b.x, b.y, b.z, b.w, c.z, c.w, tlim, &h1);
        const int32_t c0 = __float_as_int(d.x), c1 = __float_as_int(d.y);
        if (h0 && h1) {
            const bool swap = t1 < t0;
            if (s.sp < cap) {
                stack_put(stack, s.sp, swap ? c0 : c1);
                ++s.sp;
            }
            s.cur = swap ? c1 : c0;
            return false;
        }
        if (h0) { s.cur = c0; return false; }
        if (h1) { s.cur = c1; return false; }
    } else {
        uint32_t first, count;
        decode_leaf(s.cur, first, count);
        for (uint32_t k = 0; k < count; ++k) {
            const float4 *tp = tris + 3 * (size_t)(first + k);
            const float4 va = tp[0], vb = tp[1], vc = tp[2];
            if (STATS) ++n_tri;
            if (ANY) {
                if (__float_as_uint(vb.w) & 1u) continue;  // transparent (Triangle.cu:439-440)
                float hh = INFINITY;
                if (tri_intersects(s.tr, va, vb, vc, hh) && !(hh > s.tmax)) {
                    s.occ = true;
                    return true;
                }
            } else if (!(OPAQUE && (__float_as_uint(vb.w) & 1u)) && tri_intersects(s.tr, va, vb, vc, s.h)) {
                s.hit = __float_as_int(va.w);
                s.thit = s.h;
            }
        }
    }
    if (s.sp > 0) {
        --s.sp;
        s.cur = stack_get(stack, s.sp);
        return false;
    }
    return true;
}

// One step over the four-wide quantized tree (Node4, rt_internal.h): the
// four child boxes are dequantized with one FMA per bound against the
// node's origin and power-of-two scale, hit children are sorted by entry
// distance (a 5-comparator network on the distance bits; misses sort last),
// the nearest is followed and the others are pushed farthest first.  Any
// hit (shadow rays, most of which are not occluded and so visit every node
// they enter whatever the order) only finds the nearest (3 comparators) and
// pushes the others as they come.
__device__ __forceinline__ void cswap_key(uint32_t &ka, int32_t &ra, uint32_t &kb, int32_t &rb)
{
    const bool sw = kb < ka;
    const uint32_t k0 = sw ? kb : ka, k1 = sw ? ka : kb;
    const int32_t r0 = sw ? rb : ra, r1 = sw ? ra : rb;
    ka = k0; kb = k1; ra = r0; rb = r1;
}

// Leaves are not intersected the moment a lane reaches one: the lane parks
// the leaf and keeps walking interior nodes, and the wavefront intersects
// parked leaves together once kLeafBatch lanes hold one, or as soon as some
// lane cannot go on without it (a second leaf reached, or no nodes left).
// Leaf work (up to four triangle tests) then runs with most lanes active
// instead of one or two lanes holding the other 62 at every step.  Closest
// hits are unchanged (the nearest accepted triangle does not depend on the
// order of the tests, only ties would); culling against the running hit
// distance starts a little later, so a ray may visit a few more nodes.
constexpr uint32_t kLeafBatch = RT_LEAF_BATCH;        // closest hit
constexpr uint32_t kLeafBatchAny = RT_LEAF_BATCH_ANY;  // any hit (shadow rays end at their first occluder)
constexpr uint32_t kNodeSteps = RT_NODE_STEPS;         // interior-node visits per traversal step

template <bool ANY, bool STATS, class Stack, bool OPAQUE = false>
__device__ __forceinline__ bool trav_step4(TravState &s, const float4 *__restrict__ nodes,
                                           const float4 *__restrict__ tris, const Stack &stack, int cap,
                                           uint32_t &n_node, uint32_t &n_tri)
{
    // kNodeSteps interior-node visits (each followed by parking a reached
    // leaf) per step, so the per-step refill and leaf-phase checks are paid
    // once for several visits
#pragma unroll
    for (uint32_t it = 0; it < kNodeSteps; ++it) {
        if (s.cur >= 0) {
            if (STATS) ++n_node;
            const float4 *np = nodes + 4 * (size_t)s.cur;
            const float4 a = np[0], b = np[1], c = np[2], d = np[3];
            const float tlim = ANY ? s.tmax * kHitScale : fminf(s.h, s.tmax) * kHitScale;
            const uint32_t meta = __float_as_uint(a.w);
            const float Ax = __uint_as_float((meta & 0xffu) << 23) * s.br.ix;
            const float Ay = __uint_as_float(((meta >> 8) & 0xffu) << 23) * s.br.iy;
            const float Az = __uint_as_float(((meta >> 16) & 0xffu) << 23) * s.br.iz;
            const float Bx = fmaf(a.x, s.br.ix, -s.br.oix);
            const float By = fmaf(a.y, s.br.iy, -s.br.oiy);
            const float Bz = fmaf(a.z, s.br.iz, -s.br.oiz);
            // Near and far planes per axis by the sign of 1/d: A has that sign
            // and fma(q, A, B) is monotone in q, so this selects exactly the
            // min and max of the two plane distances without computing both.
            const uint32_t qlx = __float_as_uint(c.x), qhx = __float_as_uint(c.y);
            const uint32_t qly = __float_as_uint(c.z), qhy = __float_as_uint(c.w);
            const uint32_t qlz = __float_as_uint(d.x), qhz = __float_as_uint(d.y);
            const bool nx = s.br.ix < 0, ny = s.br.iy < 0, nz = s.br.iz < 0;
            const uint32_t qnx = nx ? qhx : qlx, qfx = nx ? qlx : qhx;
            const uint32_t qny = ny ? qhy : qly, qfy = ny ? qly : qhy;
            const uint32_t qnz = nz ? qhz : qlz, qfz = nz ? qlz : qhz;
            uint32_t key[4];
            int32_t ref[4] = {__float_as_int(b.x), __float_as_int(b.y), __float_as_int(b.z), __float_as_int(b.w)};
    #pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int sh = 8 * i;
                const float txn = fmaf((float)((qnx >> sh) & 0xffu), Ax, Bx);
                const float txf = fmaf((float)((qfx >> sh) & 0xffu), Ax, Bx);
                const float tyn = fmaf((float)((qny >> sh) & 0xffu), Ay, By);
                const float tyf = fmaf((float)((qfy >> sh) & 0xffu), Ay, By);
                const float tzn = fmaf((float)((qnz >> sh) & 0xffu), Az, Bz);
                const float tzf = fmaf((float)((qfz >> sh) & 0xffu), Az, Bz);
                const float tn = fmaxf(fmaxf(txn, tyn), fmaxf(tzn, 0.0f));
                float tf = fminf(fminf(txf, tyf), tzf);
                tf = fminf(tf * kTfarScale, tlim);
                const bool hit = tn <= tf;  // unused slots hold empty boxes
                key[i] = hit ? __float_as_uint(tn) : 0xffffffffu;
            }
            cswap_key(key[0], ref[0], key[1], ref[1]);
            cswap_key(key[2], ref[2], key[3], ref[3]);
            cswap_key(key[0], ref[0], key[2], ref[2]);
            if constexpr (!ANY) {
                cswap_key(key[1], ref[1], key[3], ref[3]);
                cswap_key(key[1], ref[1], key[2], ref[2]);
            }
            if (key[0] != 0xffffffffu) {
    #pragma unroll
                for (int i = 3; i >= 1; --i)
                    if (key[i] != 0xffffffffu && s.sp < cap) {
                        stack_put(stack, s.sp, ref[i]);
                        ++s.sp;
                    }
                s.cur = ref[0];
            } else if (s.sp > 0) {
                --s.sp;
                s.cur = stack_get(stack, s.sp);
            } else {
                s.cur = kNoRef;
            }
        }
        // park a reached leaf while the slot is free and walk on
        if (s.cur < 0 && s.cur != kNoRef && s.leaf == kNoRef) {
            s.leaf = s.cur;
            if (s.sp > 0) {
                --s.sp;
                s.cur = stack_get(stack, s.sp);
            } else {
                s.cur = kNoRef;
            }
        }
    }
    // leaf phase of the wavefront (the lanes stepping here)
    const bool want = s.leaf != kNoRef;
    const bool stuck = want && s.cur < 0;  // a second leaf, or nothing left to walk
    if (__ballot(stuck) != 0 || (uint32_t)__popcll(__ballot(want)) >= (ANY ? kLeafBatchAny : kLeafBatch)) {
        if (want) {
            uint32_t first, count;
            decode_leaf(s.leaf, first, count);
            s.leaf = kNoRef;
            for (uint32_t k = 0; k < count; ++k) {
                const float4 *tp = tris + 3 * (size_t)(first + k);
                const float4 va = tp[0], vb = tp[1], vc = tp[2];
                if (STATS) ++n_tri;
                if (ANY) {
                    if (__float_as_uint(vb.w) & 1u) continue;  // transparent (Triangle.cu:439-440)
                    float hh = INFINITY;
                    if (tri_intersects(s.tr, va, vb, vc, hh) && !(hh > s.tmax)) {
                        s.occ = true;
                        return true;
                    }
                } else if (!(OPAQUE && (__float_as_uint(vb.w) & 1u)) && tri_intersects(s.tr, va, vb, vc, s.h)) {
                    s.hit = __float_as_int(va.w);
                    s.thit = s.h;
                }
            }
        }
    }
    return s.cur == kNoRef && s.leaf == kNoRef;
}

// Dispatch on the tree layout.
template <bool WIDE, bool ANY, bool STATS, class Stack, bool OPAQUE = false>
__device__ __forceinline__ bool trav_step_any(TravState &s, const float4 *__restrict__ nodes,
                                              const float4 *__restrict__ tris, const Stack &stack, int cap,
                                              uint32_t &n_node, uint32_t &n_tri)
{
    if constexpr (WIDE) return trav_step4<ANY, STATS, Stack, OPAQUE>(s, nodes, tris, stack, cap, n_node, n_tri);
    else return trav_step<ANY, STATS, Stack, OPAQUE>(s, nodes, tris, stack, cap, n_node, n_tri);
}

// ---------------------------------------------------------------------
// Work striping.  Device-scope atomics on one address serialise at about
// 11 ns each on MI355X (0.09 G/s from the whole chip, against 15 G/s spread
// over many addresses), so a counter that every wavefront touches once per
// refill or append caps a kernel near 90 M operations/s.  Work lists and ray
// ranges are therefore cut into kStripes stripes, each with its own counter
// on its own 128-byte line.  A workgroup works on stripe blockIdx % kStripes
// (workgroups are dispatched round-robin over the 8 XCDs, so a stripe stays
// on one XCD and its L2), and a persistent wavefront that finds its stripe
// drained claims work from the others.
// ---------------------------------------------------------------------
// kStripes and kCtrStride are in rt_trace.h (the host sizes the counters)
static_assert(kStripes == 64, "stripe scans use one wavefront lane per stripe");

// A persistent wavefront's position in the stripes (wavefront-uniform).
// A claim takes, by guided self-scheduling, the items left in the stripe
// (as last seen) over the launch's wavefronts per stripe, rounded down to a
// power-of-two share and held between kClaimMin and kClaimChunk: a large
// launch claims kClaimChunk at a time, while a small one, and the last items
// of every stripe, spread over as many wavefronts as the items allow instead
// of queueing two or more rays on each lane of a few wavefronts.  The
// wavefront hands claimed items to its idle lanes from the local range
// [lb, le) of stripe ls, so partial refills cost no atomic.
constexpr uint32_t kClaimChunk = RT_CLAIM_CHUNK;
constexpr uint32_t kClaimMin = RT_CLAIM_MIN;
static_assert(kClaimMin >= 1 && kClaimMin <= kClaimChunk, "claim sizes run from kClaimMin to kClaimChunk");

struct StripeCursor {
    uint32_t s;       // stripe being claimed from
    uint32_t n;       // items in that stripe
    uint32_t salt;    // per-wavefront pseudo-random state for picking a stripe to help
    uint32_t left;    // items of stripe s not yet claimed, as last seen (an upper bound)
    uint32_t wlog;    // log2 of the launch's wavefronts per stripe (rounded down)
    uint32_t ls, lb, le;  // claimed, not yet handed out: items [lb, le) of stripe ls
    bool empty;       // all of stripe s is claimed: scan before the next claim
    bool done;        // every stripe is claimed
};

// No lane will get more work: every stripe is claimed and the local range
// is handed out.
__device__ __forceinline__ bool cursor_finished(const StripeCursor &c) { return c.done && c.lb == c.le; }

template <class Count>
__device__ __forceinline__ void cursor_init(StripeCursor &c, uint32_t home, Count count)
{
    c.s = home;
    c.n = count(home);
    c.left = c.n;
    const uint32_t per_stripe = (gridDim.x * (blockDim.x >> 6)) / kStripes;
    c.wlog = per_stripe ? 31u - (uint32_t)__clz(per_stripe) : 0u;
    c.salt = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 0x9E3779B1u + 0x7F4A7C15u;
    c.ls = home;
    c.lb = 0;
    c.le = 0;
    c.empty = false;
    c.done = false;
}

// One pass over every stripe's claim counter (lane j reads stripe j): move
// to a stripe with unclaimed items, drawn at random among them so that the
// wavefronts whose own stripe ran dry spread over the rest instead of all
// queueing on one counter; or mark the wavefront done.  Counters only grow
// and item counts are fixed during a launch, so a stale read can only
// under-count claims: "done" is never concluded early.
template <class Count>
__device__ __forceinline__ bool cursor_scan(StripeCursor &c, uint32_t *fetch, Count count)
{
    const uint32_t lane = __lane_id();
    const uint32_t claimed =
        __hip_atomic_load(fetch + lane * kCtrStride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t n = count(lane);
    const bool open = claimed < n;
    const uint64_t left = __ballot(open);
    if (left == 0) {
        c.done = true;
        return false;
    }
    // pick = floor(u * open / 2^24) for a 24-bit draw u: always < open, and no
    // integer division (the compiler lowers a 24-bit `%` through a float
    // reciprocal that can overshoot the quotient by one for dividends above
    // 2^23, which would leave no lane selected)
    c.salt = c.salt * 1664525u + 1013904223u;
    const uint32_t pick = ((c.salt >> 8) * (uint32_t)__popcll(left)) >> 24;
    const uint32_t rank = (uint32_t)__popcll(left & ((1ull << lane) - 1ull));
    const uint64_t sel = __ballot(open && rank == pick);
    // `sel` has exactly one bit; the first open stripe stands in should it
    // ever be empty, so the stripe index is always one of the 64
    const uint32_t t = (uint32_t)(__ffsll((unsigned long long)(sel ? sel : left)) - 1);
    c.s = t;
    c.n = (uint32_t)__shfl((int)n, (int)t);
    c.left = c.n - (uint32_t)__shfl((int)claimed, (int)t);
    c.empty = false;
    return true;
}

// Items of the next claim (wavefront-uniform).
__device__ __forceinline__ uint32_t claim_size(const StripeCursor &c)
{
    const uint32_t k = c.left >> c.wlog;
    return k < kClaimMin ? kClaimMin : (k > kClaimChunk ? kClaimChunk : k);
}

// Wavefront-uniform refill of idle lanes: when at least kRefillLanes lanes
// are idle (or all are), they take consecutive items of the local range;
// an empty local range is first refilled with one atomic claim of
// claim_size items from the cursor's stripe.  Returns the lane's item
// offset within `stripe`, or UINT32_MAX.  The claim loop is
// wavefront-uniform and ends: every turn either claims items or finds a
// stripe drained, and claims are finite.
template <class Count>
__device__ __forceinline__ uint32_t wave_claim(StripeCursor &c, uint32_t *fetch, Count count, bool has,
                                               uint32_t &stripe)
{
    const uint64_t idle = __ballot(!has);
    const uint32_t nidle = (uint32_t)__popcll(idle);
    if (!(nidle >= kRefillLanes || nidle == 64u)) return 0xffffffffu;
    const uint32_t lane = __lane_id();
    if (c.lb == c.le) {
        if (c.done) return 0xffffffffu;
        const uint32_t leader = (uint32_t)(__ffsll((unsigned long long)idle) - 1);
        for (;;) {
            if (c.empty && !cursor_scan(c, fetch, count)) return 0xffffffffu;
            const uint32_t want = claim_size(c);
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(fetch + c.s * kCtrStride, want);
            base = (uint32_t)__shfl((int)base, (int)leader);
            if (base >= c.n) {
                c.empty = true;
                continue;
            }
            c.ls = c.s;
            c.lb = base;
            c.le = c.n - base > want ? base + want : c.n;
            c.left = c.n - c.le;
            if (c.le == c.n) c.empty = true;
            break;
        }
    }
    const uint32_t take = c.le - c.lb < nidle ? c.le - c.lb : nidle;
    const uint32_t b0 = c.lb;
    c.lb += take;
    stripe = c.ls;
    if (has) return 0xffffffffu;
    const uint32_t r = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
    return r < take ? b0 + r : 0xffffffffu;
}

}  // namespace rt
