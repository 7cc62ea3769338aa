// rt_waves.h -- host-side owner of a wavefront renderer's device buffers
// (path state, shadow queue, queues and counters), shared by single-scene
// rendering (scene.hip) and batches (batch.hip).
#pragma once

#include <cstdint>

#include "rt_render.h"

namespace rt {

struct WaveBuffers {
    uint32_t capacity = 0;    // paths per wave
    uint32_t cap_lights = 0;  // light slots per path
    void *arena = nullptr;
    PathsDev paths{};
    ShadowDev shadow{};
    QueuesDev queues{};
};

// Grows `w` to at least P paths with NL light slots each (one allocation,
// 256-byte aligned pieces).  `who` names the caller in the error text.
rt_status ensure_wave_buffers(WaveBuffers &w, uint32_t P, uint32_t NL, const char *who);
void free_wave_buffers(WaveBuffers &w);

}  // namespace rt
