// rt_cache.h -- device layout of the neural radiance cache (cache.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace rt {

constexpr uint32_t kCacheWidth = 64;     // MLP width and padded input width
constexpr uint32_t kCacheLayers = 5;     // weight matrices: 4 hidden 64x64 + output 64x16 (3 used)
constexpr uint32_t kCacheOut = 16;       // padded output width (one MFMA tile)
constexpr uint32_t kCacheInStride = 12;  // query record: pos 3, normal 3, view dir 3, roughness 1, pad 2
constexpr uint32_t kCacheMaxLevels = 16;
constexpr uint32_t kCacheParams = 4 * kCacheWidth * kCacheWidth + kCacheWidth * kCacheOut;  // 17408

__host__ __device__ constexpr uint32_t kCacheLayerOffset(uint32_t l) { return l * kCacheWidth * kCacheWidth; }

// Parameters and their optimiser state (device pointers) plus the encoding's
// level table.
struct CacheDev {
    uint32_t levels, table_size;
    uint32_t res[kCacheMaxLevels];
    uint32_t dense[kCacheMaxLevels];
    float lo[3], inv_ext[3];
    // MLP: W (in x out, row-major) BF16 for the backward pass, W^T (out x in)
    // BF16 for the forward pass, FP32 master, gradient and Adam moments
    uint16_t *w, *w_t;
    float *w32, *grad, *m, *v;
    // hash grid: [levels][table][2] features
    uint32_t *grid_bf16;  // two BF16 per entry
    float *grid32, *grid_grad, *grid_m, *grid_v;
};

struct CacheOptsDev {
    float lr, beta1, beta2, eps;
};

// Scratch for one batch of up to `rows` queries.
struct CacheWork {
    uint32_t rows;
    uint16_t *act[kCacheLayers];  // act[0] = encoding, act[l] = ReLU output of layer l-1
    float *y;                     // [rows][16] output layer
    uint16_t *dy16;               // [rows][16] loss gradient
    uint16_t *dact[2];            // [rows][64] ping-pong data gradients
    float *dx32;                  // [rows][64] encoding gradient
    // the grid gradient visits the batch in Morton order of position:
    // keys and sample indices (two buffers each for the sort) and the sort's
    // temporary storage (cache_sort_bytes(rows) bytes)
    uint32_t *skey[2], *sidx[2];
    void *sort_tmp;
    size_t sort_bytes;
};

// Temporary bytes the batch's Morton sort needs for `rows` samples.
size_t cache_sort_bytes(uint32_t rows);

// n is the row count (and grid size); with n_dev the kernels read the
// actual count from device memory (capped at n), so a launch sequence can be
// fixed before the count is known.
hipError_t cache_forward(const CacheDev &c, const CacheWork &w, const float *in, uint32_t n, hipStream_t s,
                         const uint32_t *n_dev = nullptr);
hipError_t cache_infer(const CacheDev &c, const CacheWork &w, const float *in, uint32_t n, float *rgb,
                       hipStream_t s, const uint32_t *n_dev = nullptr);
// Forward + loss + backward; accumulates gradients (c.grad, c.grid_grad) and
// adds the summed loss to *loss (device, may be null).  With n_dev the row
// count is min(n, *n_dev), read on the device.
hipError_t cache_backward(const CacheDev &c, const CacheWork &w, const float *in, const float *target,
                          uint32_t n, float *loss, hipStream_t s, const uint32_t *n_dev = nullptr);
// One Adam step over every parameter; clears the gradients.
hipError_t cache_adam(const CacheDev &c, const CacheOptsDev &o, uint32_t step, hipStream_t s);
hipError_t cache_fill(float *p, uint32_t count, float value, hipStream_t s);

}  // namespace rt
