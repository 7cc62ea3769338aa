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
// refilling lanes whose ray finished (one atomic per refill per wavefront),
// so a few long rays no longer hold 63 idle lanes hostage (incoherent rays
// differ in visit count by an order of magnitude).  The rays are cut into
// kStripes contiguous ranges with one claim counter each (rt_device.h), so
// the refills of the whole chip do not queue on one address.
//
// Replaces BVH::intersectRay (cuda_libraries/BVH.cu:355-448), which visits
// both children in fixed order with no hit-distance culling and a 64-entry
// per-thread stack of Node pointers, one thread per ray.
#include "rt_device.h"
#include "rt_trace.h"

namespace rt {

constexpr int kTraceBlock = kTravBlock;

__device__ __forceinline__ void trav_init(TravState &s, const RaysDev &rays, uint32_t i, bool any)
{
    trav_begin(s, rays.ox[i], rays.oy[i], rays.oz[i], rays.dx[i], rays.dy[i], rays.dz[i],
               rays.tmax ? rays.tmax[i] : INFINITY, any);
}

// STATS = true counts node visits and triangle tests (for pricing the
// algorithmic bytes of a launch); it is never the timed kernel.  OPAQUE
// (closest hit only) skips transparent triangles: the closest opaque hit of
// BVH::intersectRays(..., checkShadows = true).
template <bool WIDE, bool ANY, bool STATS, bool OPAQUE = false>
__global__ __launch_bounds__(kTraceBlock) void k_trace(const float4 *__restrict__ nodes,
                                                       const float4 *__restrict__ tris,
                                                       RaysDev rays, uint32_t n,
                                                       int32_t *__restrict__ prim,
                                                       float *__restrict__ tout,
                                                       uint8_t *__restrict__ occluded, int cap,
                                                       uint32_t lc, uint32_t *__restrict__ ray_counter,
                                                       unsigned long long *__restrict__ counters)
{
    extern __shared__ int32_t s_stack[];
    StackTail<TraceStack> s_tail;
    const TraceStack stack{s_stack + threadIdx.x, s_tail};
    // the rays are dealt to the stripes in chunks of 2^lc consecutive rays
    // (chunk c to stripe c % kStripes), so every stripe holds a sample of
    // the whole batch and the stripes take about equally long
    const uint32_t nfull = n >> lc, rem = n & ((1u << lc) - 1u);
    const uint32_t last = nfull & (kStripes - 1u);
    auto count = [&](uint32_t j) {
        uint32_t k = ((nfull >> 6) + (j < last ? 1u : 0u)) << lc;
        if (rem && j == last) k += rem;
        return k;
    };
    auto ray_of = [&](uint32_t j, uint32_t off) {
        return ((((off >> lc) << 6) + j) << lc) + (off & ((1u << lc) - 1u));
    };
    StripeCursor cur;
    cursor_init(cur, blockIdx.x % kStripes, count);
    TravState s;
    bool has = false;
    uint32_t ray = 0;
    uint32_t n_node = 0, n_tri = 0;
    for (;;) {
        uint32_t stripe = 0;
        const uint32_t r = wave_claim(cur, ray_counter, count, has, stripe);
        if (r != 0xffffffffu) {
            ray = ray_of(stripe, r);
            trav_init(s, rays, ray, ANY);
            has = true;
        }
        if (__ballot(has) == 0) {
            if (cursor_finished(cur)) break;
            continue;
        }
        if (has && trav_step_any<WIDE, ANY, STATS, TraceStack, OPAQUE>(s, nodes, tris, stack, cap, n_node, n_tri)) {
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
                        uint32_t *ray_counter, bool opaque)
{
    if (n == 0) return hipSuccess;
    if (g_cus == 0) {
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&g_cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess || g_cus <= 0) g_cus = 256;
    }
    const int cap = (int)bvh.stack_cap;
    // stripe chunk 2^lc: a power of two in [64, 4096], about n / 2048 rays
    uint32_t lc = 6;
    while (lc < 12 && ((uint64_t)2048 << lc) <= n) ++lc;
    const size_t lds = trace_lds_bytes(bvh.stack_cap);
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
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, grid, dim3(kTraceBlock), lds, stream, nd, tr, rays, n, prim, t, occluded, cap,
                           lc, ray_counter, counters);
    };
    const bool wide = bvh.wide != 0, stats = counters != nullptr;
    if (opaque && !any && !stats) {
        wide ? go(k_trace<true, false, false, true>) : go(k_trace<false, false, false, true>);
        return hipGetLastError();
    }
    if (wide) {
        if (stats) any ? go(k_trace<true, true, true>) : go(k_trace<true, false, true>);
        else any ? go(k_trace<true, true, false>) : go(k_trace<true, false, false>);
    } else {
        if (stats) any ? go(k_trace<false, true, true>) : go(k_trace<false, false, true>);
        else any ? go(k_trace<false, true, false>) : go(k_trace<false, false, false>);
    }
    return hipGetLastError();
}

__global__ void k_clear_claims(uint32_t *ctr)
{
    if (blockIdx.x == 0 && threadIdx.x < kStripes) ctr[threadIdx.x * kCtrStride] = 0;
}

hipError_t clear_claims(uint32_t *ctr, hipStream_t stream)
{
    hipLaunchKernelGGL(k_clear_claims, dim3(1), dim3(kStripes), 0, stream, ctr);
    return hipGetLastError();
}

size_t trace_lds_bytes(uint32_t stack_cap)
{
    return stack_lds_bytes<TraceStack>(stack_cap, kTraceBlock);
}

}  // namespace rt
