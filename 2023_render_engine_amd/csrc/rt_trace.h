// rt_trace.h -- host-visible launch interface of the traversal kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace rt {

// Work striping of the persistent kernels and the wavefront queues: work is
// cut into kStripes stripes whose counters sit kCtrStride u32 words (one
// 128-byte line) apart, so device atomics spread over 64 addresses.
constexpr uint32_t kStripes = 64;
constexpr uint32_t kCtrStride = 32;
constexpr uint32_t kClaimWords = kStripes * kCtrStride;  // one striped counter set
// Largest traversal-stack bound (pending entries per lane) the kernels hold:
// the first entries in LDS, the rest in the lane's private memory.
constexpr uint32_t kMaxTraversalStack = 128;

struct RaysDev {
    const float *ox, *oy, *oz, *dx, *dy, *dz, *tmax;
};

// Device copy of a built BVH (see rt_internal.h for the record layouts).
struct DevBvh {
    void *nodes = nullptr;  // Node4[num_nodes] when wide, else Node2[num_nodes]
    void *tris = nullptr;   // Tri3[num_tris + 1]
    uint32_t num_nodes = 0, num_tris = 0;
    uint32_t stack_cap = 0; // traversal stack entries per lane
    uint32_t wide = 0;      // 1: four-wide quantized nodes
};

// counters != nullptr selects the counting variant (2 x u64: node visits,
// triangle tests, accumulated).  ray_counter: kClaimWords device words
// (kStripes claim counters, kCtrStride words apart), cleared by clear_claims
// before the launch, that the persistent kernel claims rays from; two
// launches in flight at once need different counters.  opaque (closest hit,
// no counters): transparent triangles are skipped.
hipError_t launch_trace(const DevBvh &bvh, const RaysDev &rays, uint32_t n, int32_t *prim, float *t,
                        uint8_t *occluded, bool any, hipStream_t stream, unsigned long long *counters,
                        uint32_t *ray_counter, bool opaque = false);
size_t trace_lds_bytes(uint32_t stack_cap);
// zeroes the kStripes claim counters of ray_counter (a kernel, so it can be
// captured into a graph)
hipError_t clear_claims(uint32_t *ray_counter, hipStream_t stream);

}  // namespace rt
