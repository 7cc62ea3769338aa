// rt_photon_trace.h -- the photon pre-pass (photon_trace.hip), called by the
// scene entry point in scene.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#include "../../include/rt.h"
#include "rt_render.h"

namespace rt {

// Selection probability p[l] of every point light (power over the total)
// and their running sum cdf[l], in double.
void photon_light_table(const rt_point_light *lights, uint32_t nl, std::vector<double> &p, std::vector<double> &cdf);

// Traces o.photons photon paths through the scene; writes the stored photons
// (at most capacity, sorted by path then bounce) to the device arrays (key,
// if not null: path << 5 | bounce) and their number to *count.  More than capacity: RT_ERR_INVALID with *count
// the number the pass produced.  Synchronises the stream.
rt_status trace_photons(const SceneDev &sc, const rt_point_light *lights, uint32_t nl, const rt_photon_opts &o,
                        uint32_t capacity, float *location, float *direction, float *normal, double *radiance,
                        int32_t *shape, uint32_t *key, uint32_t *count, hipStream_t s, const char *who);

}  // namespace rt
