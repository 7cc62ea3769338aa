// rt_trace.h -- host-visible launch interface of the traversal kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace rt {

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
// triangle tests, accumulated).  ray_counter: one device u32, zero at launch,
// that the persistent kernel fetches rays from; two launches in flight at
// once need different counters.
hipError_t launch_trace(const DevBvh &bvh, const RaysDev &rays, uint32_t n, int32_t *prim, float *t,
                        uint8_t *occluded, bool any, hipStream_t stream, unsigned long long *counters,
                        uint32_t *ray_counter);
size_t trace_lds_bytes(uint32_t stack_cap);

}  // namespace rt
