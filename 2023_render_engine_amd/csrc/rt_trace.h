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
    void *nodes = nullptr;  // Node2[num_nodes]
    void *tris = nullptr;   // Tri3[num_tris + 1]
    uint32_t num_nodes = 0, num_tris = 0, max_depth = 0;
};

// counters != nullptr selects the counting variant (2 x u64: node visits,
// triangle tests, accumulated).
hipError_t launch_trace(const DevBvh &bvh, const RaysDev &rays, uint32_t n, int32_t *prim, float *t,
                        uint8_t *occluded, bool any, hipStream_t stream,
                        unsigned long long *counters = nullptr);
size_t trace_lds_bytes(uint32_t max_depth);

}  // namespace rt
