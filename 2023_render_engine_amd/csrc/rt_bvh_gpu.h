// rt_bvh_gpu.h -- the BVH builder that runs on the GPU (bvh_gpu.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

#include "rt_internal.h"

namespace rt {

// A tree built on the device: the two-wide nodes always (they are what
// rt_bvh_export copies), the four-wide nodes when the layout is four-wide,
// and the triangles in leaf order plus the degenerate sentinel.
struct DeviceBvhBuild {
    Node2 *nodes2 = nullptr;
    Node4 *nodes4 = nullptr;
    Tri3 *tris = nullptr;
    uint32_t num_tris = 0, num_nodes2 = 0, num_nodes4 = 0, num_leaves = 0;
    uint32_t max_depth = 0;   // kept two-wide levels on the longest root-leaf path
    uint32_t depth4 = 0;      // levels of the four-wide tree
    uint32_t max_stack4 = 0;  // pending-entry bound of the four-wide traversal
    double sah_cost = 0.0, build_ms = 0.0;
};

// Builds from device-resident vertices (nv x 3 floats) and triangle indices
// (nt x 3) on stream s; synchronizes s.  On failure returns the HIP error
// (hipErrorInvalidValue for bad input) with err filled; what was allocated
// stays in out for free_bvh_device.
hipError_t build_bvh_device(const float *xyz, uint32_t nv, const uint32_t *idx, uint32_t nt, const rt_bvh_opts &o,
                            hipStream_t s, DeviceBvhBuild &out, std::string &err);
void free_bvh_device(DeviceBvhBuild &b);

// tris[k].flags = flags[tris[k].id] for the n real triangles (flags is a
// device array indexed by triangle id).
hipError_t set_transparent_device(Tri3 *tris, uint32_t n, const uint8_t *flags, hipStream_t s);

}  // namespace rt
