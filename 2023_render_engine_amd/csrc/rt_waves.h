// rt_waves.h -- host-side owner of a wavefront renderer's device buffers
// (path state, shadow queue, queues and counters), shared by single-scene
// rendering (scene.hip) and batches (batch.hip).
#pragma once

#include <cstdint>

#include "rt_render.h"

namespace rt {

struct WaveBuffers {
    uint32_t capacity = 0;    // path slots
    uint32_t cap_lights = 0;  // light slots per path
    void *arena = nullptr;
    PathsDev paths{};
    ShadowDev shadow{};
    QueuesDev queues{};
    float *res = nullptr;     // per-path radiance of a chunk (3 floats each)
    uint64_t res_cap = 0;     // paths res holds
    uint32_t *host_counters = nullptr;  // pinned, kNumCounters words
};

// Grows `w` to at least P path slots with NL light slots each (one
// allocation, 256-byte aligned pieces).  `who` names the caller in the
// error text.
rt_status ensure_wave_buffers(WaveBuffers &w, uint32_t P, uint32_t NL, const char *who);
// Grows the chunk result buffer to R paths.
rt_status ensure_result_buffer(WaveBuffers &w, uint64_t R, const char *who);
void free_wave_buffers(WaveBuffers &w);
// rt_render_stats of a finished frame from its stats words (kNumStats) and,
// under RT_FLAG_PROFILE, the elapsed times of prof's event pairs; resets
// prof for the next frame.
rt_status fill_stats(rt_render_stats &st, const unsigned long long *counts, KernelProfile &prof, uint32_t flags,
                     const char *who);

// Paths per chunk: the result buffer holds 12 bytes per path (3 GiB at
// most).  Each chunk ends with a drain whose last iterations hold few live
// paths; on the 32-scene batch (383M paths per frame) 2^26-path chunks
// measured 618 ms per frame, 2^27 607, 2^28 604 and one 2^29 chunk 599.
constexpr uint64_t kChunkPaths = 1ull << 28;

}  // namespace rt
