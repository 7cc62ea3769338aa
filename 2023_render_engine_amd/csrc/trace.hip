// trace.hip -- closest-hit and any-hit traversal of the two-wide BVH.
//
// Rays arrive as structure-of-arrays FP32 streams.  Each visit reads one
// 64-byte node holding both child boxes; both slabs are tested, the far
// child is pushed on a per-lane stack kept in LDS (laid out [depth][lane] so
// a wavefront's pushes hit 64 distinct banks) and the near child is
// followed.  The stack depth is the tree's internal depth, which bounds the
// number of pending entries, so it cannot overflow.
//
// Scheduling: persistent wavefronts with dynamic ray fetch.  A launch is
// sized to fill the chip; each 64-lane wavefront keeps its lanes busy by
// refilling lanes whose ray finished from a global ray counter (one atomic
// per refill per wavefront), so a few long rays no longer hold 63 idle lanes
// hostage (incoherent rays differ in visit count by an order of magnitude).
//
// Replaces BVH::intersectRay (cuda_libraries/BVH.cu:355-448), which visits
// both children in fixed order with no hit-distance culling and a 64-entry
// per-thread stack of Node pointers, one thread per ray.
#include "rt_device.h"
#include "rt_trace.h"

namespace rt {

constexpr int kTraceBlock = 128;
constexpr uint32_t kRefillLanes = 16;  // refill when at least this many lanes are idle

// Per-lane traversal state.
struct TravState {
    TriRay tr;
    BoxRay br;
    float tmax, h, thit;
    int32_t hit, cur;
    int sp;
    bool occ;
};

__device__ __forceinline__ void trav_init(TravState &s, const RaysDev &rays, uint32_t i, bool any)
{
    const float ox = rays.ox[i], oy = rays.oy[i], oz = rays.oz[i];
    const float dx = rays.dx[i], dy = rays.dy[i], dz = rays.dz[i];
    s.tmax = rays.tmax ? rays.tmax[i] : INFINITY;
    // Closest hit: the triangle test sees ray.maximumOffset = tmax and a
    // running hitOffset starting at +inf (BVH.cu:362).  Any hit: shadow rays
    // have maximumOffset = +inf and each triangle is judged on its own
    // (order-independent), occluding iff accepted at t <= tmax.
    s.tr = make_tri_ray(ox, oy, oz, dx, dy, dz, any ? INFINITY : s.tmax);
    s.br = make_box_ray(ox, oy, oz, dx, dy, dz);
    s.h = INFINITY;
    s.thit = INFINITY;
    s.hit = -1;
    s.cur = 0;
    s.sp = 0;
    s.occ = false;
}

// One traversal step (one node visit or one leaf).  Returns true when the
// ray is finished.
template <bool ANY, bool STATS>
__device__ __forceinline__ bool trav_step(TravState &s, const float4 *__restrict__ nodes,
                                          const float4 *__restrict__ tris, int32_t *stack, int cap,
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
                stack[s.sp * kTraceBlock] = swap ? c0 : c1;
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
            } else if (tri_intersects(s.tr, va, vb, vc, s.h)) {
                s.hit = __float_as_int(va.w);
                s.thit = s.h;
            }
        }
    }
    if (s.sp > 0) {
        --s.sp;
        s.cur = stack[s.sp * kTraceBlock];
        return false;
    }
    return true;
}

// STATS = true counts node visits and triangle tests (for pricing the
// algorithmic bytes of a launch); it is never the timed kernel.
template <bool ANY, bool STATS>
__global__ __launch_bounds__(kTraceBlock) void k_trace(const float4 *__restrict__ nodes,
                                                       const float4 *__restrict__ tris,
                                                       RaysDev rays, uint32_t n,
                                                       int32_t *__restrict__ prim,
                                                       float *__restrict__ tout,
                                                       uint8_t *__restrict__ occluded, int cap,
                                                       uint32_t *__restrict__ ray_counter,
                                                       unsigned long long *__restrict__ counters)
{
    extern __shared__ int32_t s_stack[];
    int32_t *stack = s_stack + threadIdx.x;
    const uint32_t lane = __lane_id();
    TravState s;
    bool has = false;
    bool exhausted = false;  // wavefront-uniform: the ray counter passed n
    uint32_t ray = 0;
    uint32_t n_node = 0, n_tri = 0;
    for (;;) {
        // refill idle lanes (wavefront-uniform decision)
        const uint64_t idle = __ballot(!has);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (!exhausted && (nidle >= kRefillLanes || nidle == 64u)) {
            const uint32_t leader = (uint32_t)(__ffsll((unsigned long long)idle) - 1);
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(ray_counter, nidle);
            base = __shfl(base, (int)leader);
            if (base + nidle >= n) exhausted = true;
            if (!has) {
                const uint32_t r = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                if (r < n) {
                    ray = r;
                    trav_init(s, rays, r, ANY);
                    has = true;
                }
            }
        }
        if (__ballot(has) == 0) {
            if (exhausted) break;
            continue;
        }
        if (has && trav_step<ANY, STATS>(s, nodes, tris, stack, cap, n_node, n_tri)) {
            if (ANY) {
                occluded[ray] = s.occ ? 1 : 0;
            } else {
                prim[ray] = s.hit;
                if (tout) tout[ray] = s.thit;
            }
            has = false;
        }
    }
    if (STATS) {
        atomicAdd(&counters[0], (unsigned long long)n_node);
        atomicAdd(&counters[1], (unsigned long long)n_tri);
    }
}

static int g_cus = 0;

hipError_t launch_trace(const DevBvh &bvh, const RaysDev &rays, uint32_t n, int32_t *prim, float *t,
                        uint8_t *occluded, bool any, hipStream_t stream, unsigned long long *counters,
                        uint32_t *ray_counter)
{
    if (n == 0) return hipSuccess;
    if (g_cus == 0) {
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&g_cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess || g_cus <= 0) g_cus = 256;
    }
    const int cap = (int)bvh.max_depth + 1;
    const size_t lds = (size_t)cap * kTraceBlock * sizeof(int32_t);
    // persistent launch: as many workgroups as fit (LDS-limited, at most 16
    // per CU), never more than the rays need
    size_t per_cu = lds ? (160u * 1024u) / lds : 16;
    if (per_cu > 16) per_cu = 16;
    if (per_cu < 1) per_cu = 1;
    size_t blocks = (size_t)g_cus * per_cu;
    const size_t need = (n + kTraceBlock - 1) / kTraceBlock;
    if (blocks > need) blocks = need;
    const dim3 grid((unsigned)blocks);
    const float4 *nd = (const float4 *)bvh.nodes, *tr = (const float4 *)bvh.tris;
    if (counters) {
        if (any)
            hipLaunchKernelGGL((k_trace<true, true>), grid, dim3(kTraceBlock), lds, stream, nd, tr, rays, n, prim, t, occluded, cap, ray_counter, counters);
        else
            hipLaunchKernelGGL((k_trace<false, true>), grid, dim3(kTraceBlock), lds, stream, nd, tr, rays, n, prim, t, occluded, cap, ray_counter, counters);
    } else {
        if (any)
            hipLaunchKernelGGL((k_trace<true, false>), grid, dim3(kTraceBlock), lds, stream, nd, tr, rays, n, prim, t, occluded, cap, ray_counter, counters);
        else
            hipLaunchKernelGGL((k_trace<false, false>), grid, dim3(kTraceBlock), lds, stream, nd, tr, rays, n, prim, t, occluded, cap, ray_counter, counters);
    }
    return hipGetLastError();
}

size_t trace_lds_bytes(uint32_t max_depth)
{
    return (size_t)(max_depth + 1) * kTraceBlock * sizeof(int32_t);
}

}  // namespace rt
