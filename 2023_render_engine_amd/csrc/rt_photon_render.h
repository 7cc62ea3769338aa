// rt_photon_render.h -- the photon map's search and density estimate driven
// from inside the renderer's bounce loop (photon.hip), with the query count
// on the device so an iteration needs no host round trip.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

#include "../../include/rt.h"

namespace rt {

// Sizes the map's work space for up to `cap` queries of k photons and
// copies the photons' radiance (3 doubles) and direction (3 floats, both in
// the map's input order) into tree order on stream s (may allocate and
// synchronize the stream; call before the loop).
hipError_t photon_render_reserve(rt_photon_map *m, uint32_t cap, uint32_t k, const double *photon_rgb,
                                 const float *photon_dir, hipStream_t s, std::string &err);

// The k nearest photons within radius (on the query's shape) and
// combinePhotonContributions for the first *n_dev of `cap` queries:
// points, normals (shading) and views (towards the viewer) 3 floats each,
// the query's material (textured kd) and shape id; out: 3 doubles per query.
// Launches only (no allocation, no synchronisation).
// The photons' radiance and direction are the copies the reserve made.
hipError_t photon_gather_render(rt_photon_map *m, const float *points, const float *normals, const float *views,
                                const rt_material *mats, const int32_t *qshape, uint32_t cap,
                                const uint32_t *n_dev, uint32_t k, float radius, double *out, hipStream_t s);

// The map's context (renders check it against the scene's).
rt_ctx *photon_map_ctx(const rt_photon_map *m);

}  // namespace rt
