// trace.hip -- closest-hit and any-hit traversal of the two-wide BVH.
//
// One ray per lane, 64-lane wavefronts.  Rays arrive as structure-of-arrays
// FP32 streams (coalesced 256-byte loads per wavefront per component).  Each
// visit reads one 64-byte node holding both child boxes; both slabs are
// tested, the far child is pushed on a per-lane stack kept in LDS (laid out
// [depth][lane] so a wavefront's pushes hit 64 distinct banks) and the near
// child is followed.  The stack depth is the tree's internal depth, which
// bounds the number of pending entries, so it cannot overflow.
//
// Replaces BVH::intersectRay (cuda_libraries/BVH.cu:355-448), which visits
// both children in fixed order with no hit-distance culling and a 64-entry
// per-thread stack of Node pointers.
#include "rt_device.h"
#include "rt_trace.h"

namespace rt {

constexpr int kTraceBlock = 128;

// STATS = true counts node visits and triangle tests (for pricing the
// algorithmic bytes of a launch); it is never the timed kernel.
template <bool ANY, bool STATS>
__global__ __launch_bounds__(kTraceBlock) void k_trace(const float4 *__restrict__ nodes,
                                                       const float4 *__restrict__ tris,
                                                       RaysDev rays, uint32_t n,
                                                       int32_t *__restrict__ prim,
                                                       float *__restrict__ tout,
                                                       uint8_t *__restrict__ occluded, int cap,
                                                       unsigned long long *__restrict__ counters)
{
    extern __shared__ int32_t s_stack[];
    const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= n) return;
    int32_t *stack = s_stack + threadIdx.x;

    const float ox = rays.ox[i], oy = rays.oy[i], oz = rays.oz[i];
    const float dx = rays.dx[i], dy = rays.dy[i], dz = rays.dz[i];
    const float tmax = rays.tmax ? rays.tmax[i] : INFINITY;

    // Closest hit: the triangle test sees ray.maximumOffset = tmax and a
    // running hitOffset starting at +inf (BVH.cu:362).  Any hit: shadow
    // rays have maximumOffset = +inf and each triangle is judged on its own
    // (order-independent), occluding iff accepted at t <= tmax.
    const TriRay tr = make_tri_ray(ox, oy, oz, dx, dy, dz, ANY ? INFINITY : tmax);
    const BoxRay br = make_box_ray(ox, oy, oz, dx, dy, dz);

    float h = INFINITY, thit = INFINITY;
    int32_t hit = -1;
    bool occ = false;
    int sp = 0;
    int32_t cur = 0;
    uint32_t n_node = 0, n_tri = 0;
    for (;;) {
        if (cur >= 0) {
            if (STATS) ++n_node;
            const float4 *np = nodes + 4 * (size_t)cur;
            const float4 a = np[0], b = np[1], c = np[2], d = np[3];
            const float tlim = ANY ? tmax * kHitScale : fminf(h, tmax) * kHitScale;
            bool h0, h1;
            const float t0 = slab(br, a.x, a.y, a.z, a.w, c.x, c.y, tlim, &h0);
            const float t1 = slab(br, b.x, b.y, b.z, b.w, c.z, c.w, tlim, &h1);
            const int32_t c0 = __float_as_int(d.x), c1 = __float_as_int(d.y);
            if (h0 && h1) {
                const bool swap = t1 < t0;
                const int32_t nearc = swap ? c1 : c0, farc = swap ? c0 : c1;
                if (sp < cap) {
                    stack[sp * kTraceBlock] = farc;
                    ++sp;
                }
                cur = nearc;
            } else if (h0) {
                cur = c0;
            } else if (h1) {
                cur = c1;
            } else if (sp > 0) {
                --sp;
                cur = stack[sp * kTraceBlock];
            } else {
                break;
            }
        } else {
            uint32_t first, count;
            decode_leaf(cur, first, count);
            for (uint32_t k = 0; k < count; ++k) {
                const float4 *tp = tris + 3 * (size_t)(first + k);
                const float4 va = tp[0], vb = tp[1], vc = tp[2];
                if (STATS) ++n_tri;
                if (ANY) {
                    if (__float_as_uint(vb.w) & 1u) continue;  // transparent (Triangle.cu:439-440)
                    float hh = INFINITY;
                    if (tri_intersects(tr, va, vb, vc, hh) && !(hh > tmax)) {
                        occ = true;
                        break;
                    }
                } else {
                    if (tri_intersects(tr, va, vb, vc, h)) {
                        hit = __float_as_int(va.w);
                        thit = h;
                    }
                }
            }
            if (ANY && occ) break;
            if (sp > 0) {
                --sp;
                cur = stack[sp * kTraceBlock];
            } else {
                break;
            }
        }
    }
    if (ANY) {
        occluded[i] = occ ? 1 : 0;
    } else {
        prim[i] = hit;
        if (tout) tout[i] = thit;
    }
    if (STATS) {
        atomicAdd(&counters[0], (unsigned long long)n_node);
        atomicAdd(&counters[1], (unsigned long long)n_tri);
    }
}

hipError_t launch_trace(const DevBvh &bvh, const RaysDev &rays, uint32_t n, int32_t *prim, float *t,
                        uint8_t *occluded, bool any, hipStream_t stream, unsigned long long *counters)
{
    if (n == 0) return hipSuccess;
    const int cap = (int)bvh.max_depth + 1;
    const size_t lds = (size_t)cap * kTraceBlock * sizeof(int32_t);
    const dim3 grid((n + kTraceBlock - 1) / kTraceBlock);
    const float4 *nd = (const float4 *)bvh.nodes, *tr = (const float4 *)bvh.tris;
    if (counters) {
        if (any)
            hipLaunchKernelGGL((k_trace<true, true>), grid, dim3(kTraceBlock), lds, stream, nd, tr, rays, n, prim, t, occluded, cap, counters);
        else
            hipLaunchKernelGGL((k_trace<false, true>), grid, dim3(kTraceBlock), lds, stream, nd, tr, rays, n, prim, t, occluded, cap, counters);
    } else {
        if (any)
            hipLaunchKernelGGL((k_trace<true, false>), grid, dim3(kTraceBlock), lds, stream, nd, tr, rays, n, prim, t, occluded, cap, counters);
        else
            hipLaunchKernelGGL((k_trace<false, false>), grid, dim3(kTraceBlock), lds, stream, nd, tr, rays, n, prim, t, occluded, cap, counters);
    }
    return hipGetLastError();
}

size_t trace_lds_bytes(uint32_t max_depth)
{
    return (size_t)(max_depth + 1) * kTraceBlock * sizeof(int32_t);
}

}  // namespace rt
