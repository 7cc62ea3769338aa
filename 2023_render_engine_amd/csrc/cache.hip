// cache.hip -- online-trained neural radiance cache for gfx950.
//
// Model: a multiresolution hash-grid encoding of position (Mueller et al.,
// "Instant Neural Graphics Primitives with a Multiresolution Hash Encoding",
// SIGGRAPH 2022: L levels, 2^log2T entries of 2 BF16 features per level,
// trilinear interpolation, spatial hash x*1 ^ y*2654435761 ^ z*805459861)
// concatenated with the surface normal, view direction and roughness, then a
// bias-free MLP 64 -> 64 -> 64 -> 64 -> 64 -> 3 with ReLU on the hidden
// layers, trained online with Adam on a relative L2 loss (the loss of
// Mueller et al., "Real-time Neural Radiance Caching for Path Tracing",
// SIGGRAPH 2021).  The reference renderer has no counterpart (its learned
// component is a per-shape polynomial regression,
// sampling/ShapeLightDistribution.cu:20-345).
//
// Matmuls run on the matrix cores: v_mfma_f32_16x16x16_bf16 (BF16 inputs,
// FP32 accumulation).  One wavefront owns a 16-row strip of the output;
// activations live in HBM between layers as BF16 (64 columns = 128 B per
// row), weights are read through the L1/L2 caches.  Master weights, Adam
// moments and gradients are FP32.
#include <hip/hip_runtime.h>

#include <cstdint>

#include <hipcub/device/device_radix_sort.hpp>

#include "rt_cache.h"

namespace rt {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    // round to nearest even (no NaN inputs here)
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Wavefront MFMA: D[16x16] += A[16xK] B[Kx16].  Lane l supplies
// A[l%16][k0 + 4(l/16) + 0..3] and B[k0 + 4(l/16) + 0..3][l%16]; D comes back
// as D[4(l/16) + r][l%16], r = 0..3.
__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------
// hash-grid encoding
// ---------------------------------------------------------------------
__device__ __forceinline__ uint32_t grid_index(const CacheDev &c, uint32_t l, uint32_t x, uint32_t y, uint32_t z)
{
    const uint32_t res1 = c.res[l] + 1;
    if (c.dense[l]) return x + res1 * (y + res1 * z);
    return (x * 1u ^ y * 2654435761u ^ z * 805459861u) & (c.table_size - 1u);
}

// position -> level coordinates; writes the 8 corner indices and weights
__device__ __forceinline__ void grid_corners(const CacheDev &c, uint32_t l, float px, float py, float pz,
                                             uint32_t idx[8], float w[8])
{
    const float r = (float)c.res[l];
    const float x = px * r, y = py * r, z = pz * r;
    const float fx0 = floorf(x), fy0 = floorf(y), fz0 = floorf(z);
    const uint32_t x0 = (uint32_t)fx0, y0 = (uint32_t)fy0, z0 = (uint32_t)fz0;
    const float fx = x - fx0, fy = y - fy0, fz = z - fz0;
    const uint32_t rmax = c.res[l];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
        const uint32_t xi = min(x0 + dx, rmax), yi = min(y0 + dy, rmax), zi = min(z0 + dz, rmax);
        idx[k] = grid_index(c, l, xi, yi, zi);
        w[k] = (dx ? fx : 1 - fx) * (dy ? fy : 1 - fy) * (dz ? fz : 1 - fz);
    }
}

__device__ __forceinline__ void normalise_pos(const CacheDev &c, const float *q, float &px, float &py, float &pz)
{
    px = fminf(fmaxf((q[0] - c.lo[0]) * c.inv_ext[0], 0.0f), 1.0f);
    py = fminf(fmaxf((q[1] - c.lo[1]) * c.inv_ext[1], 0.0f), 1.0f);
    pz = fminf(fmaxf((q[2] - c.lo[2]) * c.inv_ext[2], 0.0f), 1.0f);
}

// X[n][64] (BF16): 2 features per level, normal, view direction, roughness,
// a constant 1 (the bias-free MLP's offset input) and zero padding.
__global__ __launch_bounds__(256) void k_encode(CacheDev c, const float *__restrict__ in, uint32_t n,
                                                const uint32_t *__restrict__ n_dev, uint16_t *__restrict__ X)
{
    if (n_dev) n = min(n, *n_dev);
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float *q = in + (size_t)i * kCacheInStride;
        float px, py, pz;
        normalise_pos(c, q, px, py, pz);
        alignas(16) uint16_t row[kCacheWidth];
        for (uint32_t l = 0; l < c.levels; ++l) {
            uint32_t idx[8];
            float w[8];
            grid_corners(c, l, px, py, pz, idx, w);
            const uint32_t *tab = (const uint32_t *)c.grid_bf16 + (size_t)l * c.table_size;
            float f0 = 0, f1 = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t v = tab[idx[k]];
                f0 += w[k] * bf2f((uint16_t)(v & 0xffffu));
                f1 += w[k] * bf2f((uint16_t)(v >> 16));
            }
            row[2 * l] = f2bf(f0);
            row[2 * l + 1] = f2bf(f1);
        }
        uint32_t o = 2 * c.levels;
        for (int k = 3; k < 10; ++k) row[o++] = f2bf(q[k]);  // normal, view dir, roughness
        row[o++] = f2bf(1.0f);
        for (; o < kCacheWidth; ++o) row[o] = 0;
        uint4 *dst = (uint4 *)(X + (size_t)i * kCacheWidth);
        const uint4 *src = (const uint4 *)row;
#pragma unroll
        for (int k = 0; k < kCacheWidth / 8; ++k) dst[k] = src[k];
    }
}

// ---------------------------------------------------------------------
// forward layer: Y[n][NOUT] = act(X[n][64] * W), W^T stored [NOUT][64] BF16
// MODE 0: ReLU, BF16 out (row stride 64); MODE 1: linear, FP32 out (NOUT = 16)
// ---------------------------------------------------------------------
template <int NOUT, int MODE>
__global__ __launch_bounds__(256) void k_fwd(const uint16_t *__restrict__ X, const uint16_t *__restrict__ Wt,
                                             uint32_t n, const uint32_t *__restrict__ n_dev, void *__restrict__ Y)
{
    if (n_dev) n = min(n, *n_dev);
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x / 64;
    const uint32_t li = lane & 15, lq = lane >> 4;
    const uint32_t strips = (n + 15) / 16;
    for (uint32_t s = blockIdx.x * 4 + wave; s < strips; s += gridDim.x * 4) {
        const uint32_t row = s * 16 + li;
        f32x4 acc[NOUT / 16];
#pragma unroll
        for (int t = 0; t < NOUT / 16; ++t) acc[t] = f32x4{0, 0, 0, 0};
#pragma unroll
        for (int k0 = 0; k0 < kCacheWidth; k0 += 16) {
            bf16x4 a = {0, 0, 0, 0};
            if (row < n) a = *(const bf16x4 *)(X + (size_t)row * kCacheWidth + k0 + 4 * lq);
#pragma unroll
            for (int t = 0; t < NOUT / 16; ++t) {
                const bf16x4 b = *(const bf16x4 *)(Wt + (size_t)(t * 16 + li) * kCacheWidth + k0 + 4 * lq);
                acc[t] = mfma16(a, b, acc[t]);
            }
        }
#pragma unroll
        for (int t = 0; t < NOUT / 16; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t orow = s * 16 + 4 * lq + r;
                const uint32_t ocol = t * 16 + li;
                if (orow < n) {
                    if (MODE == 0)
                        ((uint16_t *)Y)[(size_t)orow * kCacheWidth + ocol] = f2bf(fmaxf(acc[t][r], 0.0f));
                    else
                        ((float *)Y)[(size_t)orow * NOUT + ocol] = acc[t][r];
                }
            }
    }
}

// ---------------------------------------------------------------------
// backward through a layer (data gradient):
//   dX[n][64] = (dY[n][NIN] * W^T) (.) [A > 0]   (A = the layer's input
//   activation; MASK = 0 for the encoding input, which has no ReLU)
// W stored [64 in][NIN out] row-major BF16 is the B operand (B[k][j] = W[j][k]).
// OUTF32 writes FP32 (the encoding gradient), otherwise BF16.
// ---------------------------------------------------------------------
template <int NIN, bool MASK, bool OUTF32>
__global__ __launch_bounds__(256) void k_bwd_data(const uint16_t *__restrict__ dY, const uint16_t *__restrict__ W,
                                                  const uint16_t *__restrict__ A, uint32_t n,
                                                  const uint32_t *__restrict__ n_dev, void *__restrict__ dX)
{
    if (n_dev) n = min(n, *n_dev);
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x / 64;
    const uint32_t li = lane & 15, lq = lane >> 4;
    const uint32_t strips = (n + 15) / 16;
    for (uint32_t s = blockIdx.x * 4 + wave; s < strips; s += gridDim.x * 4) {
        const uint32_t row = s * 16 + li;
        f32x4 acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = f32x4{0, 0, 0, 0};
#pragma unroll
        for (int k0 = 0; k0 < NIN; k0 += 16) {
            bf16x4 a = {0, 0, 0, 0};
            if (row < n) a = *(const bf16x4 *)(dY + (size_t)row * NIN + k0 + 4 * lq);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const bf16x4 b = *(const bf16x4 *)(W + (size_t)(t * 16 + li) * NIN + k0 + 4 * lq);
                acc[t] = mfma16(a, b, acc[t]);
            }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t orow = s * 16 + 4 * lq + r;
                const uint32_t ocol = t * 16 + li;
                if (orow >= n) continue;
                float v = acc[t][r];
                if (MASK && !(bf2f(A[(size_t)orow * kCacheWidth + ocol]) > 0.0f)) v = 0.0f;
                if (OUTF32)
                    ((float *)dX)[(size_t)orow * kCacheWidth + ocol] = v;
                else
                    ((uint16_t *)dX)[(size_t)orow * kCacheWidth + ocol] = f2bf(v);
            }
    }
}

// ---------------------------------------------------------------------
// weight gradient: dW[64][NOUT] += X^T dY over all rows.  A workgroup stages
// 64-row chunks of X and dY in LDS, wave w accumulates dW rows 16w..16w+15
// in registers and adds them to the FP32 gradient once at the end.
// ---------------------------------------------------------------------
template <int NOUT>
__global__ __launch_bounds__(256) void k_wgrad(const uint16_t *__restrict__ X, const uint16_t *__restrict__ dY,
                                               uint32_t n, const uint32_t *__restrict__ n_dev, float *__restrict__ dW)
{
    if (n_dev) n = min(n, *n_dev);
    __shared__ uint16_t sX[64 * kCacheWidth];
    __shared__ uint16_t sD[64 * NOUT];
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x / 64;
    const uint32_t li = lane & 15, lq = lane >> 4;
    f32x4 acc[NOUT / 16];
#pragma unroll
    for (int t = 0; t < NOUT / 16; ++t) acc[t] = f32x4{0, 0, 0, 0};
    const uint32_t chunks = (n + 63) / 64;
    for (uint32_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
        const uint32_t r0 = ch * 64;
        // stage (zero rows past n)
        for (uint32_t e = threadIdx.x; e < 64 * kCacheWidth / 8; e += 256) {
            const uint32_t r = e / (kCacheWidth / 8), c8 = e % (kCacheWidth / 8);
            uint4 v = {0, 0, 0, 0};
            if (r0 + r < n) v = *(const uint4 *)(X + (size_t)(r0 + r) * kCacheWidth + c8 * 8);
            *(uint4 *)(sX + r * kCacheWidth + c8 * 8) = v;
        }
        for (uint32_t e = threadIdx.x; e < 64 * NOUT / 8; e += 256) {
            const uint32_t r = e / (NOUT / 8), c8 = e % (NOUT / 8);
            uint4 v = {0, 0, 0, 0};
            if (r0 + r < n) v = *(const uint4 *)(dY + (size_t)(r0 + r) * NOUT + c8 * 8);
            *(uint4 *)(sD + r * NOUT + c8 * 8) = v;
        }
        __syncthreads();
        // dW[16w + i][j] += sum_m X[m][16w + i] dY[m][j]; A[i][m] = X[m][16w+i], B[m][j] = dY[m][j]
#pragma unroll
        for (int m0 = 0; m0 < 64; m0 += 16) {
            bf16x4 a;
#pragma unroll
            for (int q = 0; q < 4; ++q) a[q] = (short)sX[(m0 + 4 * lq + q) * kCacheWidth + 16 * wave + li];
#pragma unroll
            for (int t = 0; t < NOUT / 16; ++t) {
                bf16x4 b;
#pragma unroll
                for (int q = 0; q < 4; ++q) b[q] = (short)sD[(m0 + 4 * lq + q) * NOUT + t * 16 + li];
                acc[t] = mfma16(a, b, acc[t]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < NOUT / 16; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t wr = 16 * wave + 4 * lq + r, wc = t * 16 + li;
            atomicAdd(&dW[wr * NOUT + wc], acc[t][r]);
        }
}

// ---------------------------------------------------------------------
// relative L2 loss (Mueller et al. 2021): l = (y - t)^2 / (sg(y)^2 + 0.01),
// mean over rows and the 3 channels; writes dY (BF16, [n][16]) and adds the
// loss sum to *loss.  With n_dev the row count (and the mean's scale) comes
// from device memory.
// ---------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ Y, const float *__restrict__ target,
                                              uint32_t n, const uint32_t *__restrict__ n_dev, float scale,
                                              uint16_t *__restrict__ dY, float *__restrict__ loss)
{
    if (n_dev) {
        n = min(n, *n_dev);
        scale = n ? 1.0f / (3.0f * (float)n) : 0.0f;
    }
    float lsum = 0;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        alignas(16) uint16_t g[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) g[c] = 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float y = Y[(size_t)i * 16 + c], t = target[(size_t)i * 3 + c];
            const float den = y * y + 0.01f;
            const float d = y - t;
            lsum += d * d / den;
            g[c] = f2bf(2.0f * d / den * scale);
        }
        uint4 *dst = (uint4 *)(dY + (size_t)i * 16);
        dst[0] = ((uint4 *)g)[0];
        dst[1] = ((uint4 *)g)[1];
    }
    // wavefront reduction, one atomic per wavefront
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_down(lsum, off);
    if (__lane_id() == 0 && loss) atomicAdd(loss, lsum);
}

// Morton key of each sample's normalised position (10 bits per axis; rows
// past the count get the largest key and sort last), and its index.
__global__ __launch_bounds__(256) void k_sample_morton(CacheDev c, const float *__restrict__ in, uint32_t n,
                                                       const uint32_t *__restrict__ n_dev, uint32_t *__restrict__ key,
                                                       uint32_t *__restrict__ idx)
{
    const uint32_t count = n_dev ? min(n, *n_dev) : n;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        uint32_t k = 0xffffffffu;
        if (i < count) {
            float p[3];
            normalise_pos(c, in + (size_t)i * kCacheInStride, p[0], p[1], p[2]);
            k = 0;
            for (int a = 0; a < 3; ++a) {
                uint32_t q = (uint32_t)fminf(p[a] * 1024.0f, 1023.0f);
                q &= 1023u;
                q = (q | (q << 16)) & 0x030000FFu;
                q = (q | (q << 8)) & 0x0300F00Fu;
                q = (q | (q << 4)) & 0x030C30C3u;
                q = (q | (q << 2)) & 0x09249249u;
                k |= q << (2 - a);
            }
        }
        key[i] = k;
        idx[i] = i;
    }
}

// hash-grid gradient: scatter dX[:, 0..2L) through the trilinear weights.
// Block (level l, run jb) takes the samples at positions jb*256 .. +255 of
// the batch's Morton order, so its lanes sit close together in space and
// share grid cells on all but the finest levels.  Their 8 corner updates
// are first merged in an LDS table keyed by cell (open addressing, 4096
// slots for at most 2048 distinct cells, so a probe always ends), and each
// distinct cell then takes one device-scope atomic per feature: on a
// surface-sampled batch that is 2.4x fewer atomics overall and 50-130x fewer
// on the coarse levels, where one atomic per corner serialised on a few
// thousand addresses.
constexpr uint32_t kGridSlots = 4096;
constexpr uint32_t kGridEmpty = 0xffffffffu;

__global__ __launch_bounds__(256) void k_grid_bwd(CacheDev c, const float *__restrict__ in, uint32_t n,
                                                  const uint32_t *__restrict__ n_dev, uint32_t runs,
                                                  const uint32_t *__restrict__ order, const float *__restrict__ dX,
                                                  float *__restrict__ dgrid)
{
    __shared__ uint32_t skey[kGridSlots];
    __shared__ float sv0[kGridSlots], sv1[kGridSlots];
    const uint32_t count = n_dev ? min(n, *n_dev) : n;
    const uint32_t l = blockIdx.x / runs, jb = blockIdx.x - l * runs;
    if (l >= c.levels || jb * 256 >= count) return;  // whole block: before any barrier
    for (uint32_t t = threadIdx.x; t < kGridSlots; t += 256) {
        skey[t] = kGridEmpty;
        sv0[t] = 0.0f;
        sv1[t] = 0.0f;
    }
    __syncthreads();
    const uint32_t j = jb * 256 + threadIdx.x;
    if (j < count) {
        const uint32_t i = order[j];
        const float g0 = dX[(size_t)i * kCacheWidth + 2 * l], g1 = dX[(size_t)i * kCacheWidth + 2 * l + 1];
        if (g0 != 0.0f || g1 != 0.0f) {
            float px, py, pz;
            normalise_pos(c, in + (size_t)i * kCacheInStride, px, py, pz);
            uint32_t idx[8];
            float w[8];
            grid_corners(c, l, px, py, pz, idx, w);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                uint32_t h = (idx[k] * 2654435761u) >> 20;  // 12 bits
                for (;;) {
                    const uint32_t prev = atomicCAS(&skey[h], kGridEmpty, idx[k]);
                    if (prev == kGridEmpty || prev == idx[k]) {
                        atomicAdd(&sv0[h], w[k] * g0);
                        atomicAdd(&sv1[h], w[k] * g1);
                        break;
                    }
                    h = (h + 1) & (kGridSlots - 1);
                }
            }
        }
    }
    __syncthreads();
    float *tab = dgrid + (size_t)l * c.table_size * 2;
    for (uint32_t t = threadIdx.x; t < kGridSlots; t += 256) {
        const uint32_t key = skey[t];
        if (key != kGridEmpty) {
            atomicAdd(&tab[2 * (size_t)key], sv0[t]);
            atomicAdd(&tab[2 * (size_t)key + 1], sv1[t]);
        }
    }
}

// Adam (Kingma & Ba 2015) on FP32 master weights; refreshes the BF16 copies.
// For the MLP it also writes the transposed forward copy.
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ w, float *__restrict__ g, float *__restrict__ m,
                                              float *__restrict__ v, uint32_t count, float lr, float b1, float b2,
                                              float eps, float bc1, float bc2, uint16_t *__restrict__ wbf,
                                              uint16_t *__restrict__ wbf_t, uint32_t mlp)
{
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < count; i += gridDim.x * 256) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1 - b1) * gi;
        const float vi = b2 * v[i] + (1 - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
        w[i] = wi;
        g[i] = 0.0f;
        const uint16_t h = f2bf(wi);
        wbf[i] = h;
        if (mlp) {
            // layer l: W stored [64][nout] at offset kCacheLayerOffset(l); the
            // transposed copy [nout][64] at the same offset in wbf_t
            uint32_t l = 0, off = 0, nout = kCacheWidth;
            while (l + 1 < kCacheLayers && i >= off + kCacheWidth * nout) {
                off += kCacheWidth * nout;
                ++l;
                nout = (l + 1 == kCacheLayers) ? kCacheOut : kCacheWidth;
            }
            const uint32_t k = (i - off) / nout, j = (i - off) % nout;
            wbf_t[off + j * kCacheWidth + k] = h;
        }
    }
}

__global__ __launch_bounds__(256) void k_fill(float *__restrict__ p, uint32_t count, float value)
{
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < count; i += gridDim.x * 256) p[i] = value;
}

// inference output: rgb = max(Y[:, 0..2], 0)
__global__ __launch_bounds__(256) void k_out_rgb(const float *__restrict__ Y, uint32_t n,
                                                 const uint32_t *__restrict__ n_dev, float *__restrict__ rgb)
{
    if (n_dev) n = min(n, *n_dev);
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
#pragma unroll
        for (int c = 0; c < 3; ++c) rgb[(size_t)i * 3 + c] = fmaxf(Y[(size_t)i * 16 + c], 0.0f);
}

// ---------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------
static dim3 grid_for(uint32_t work, uint32_t per_block, uint32_t cap = 4096)
{
    uint32_t b = (work + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return dim3(b);
}

hipError_t cache_forward(const CacheDev &c, const CacheWork &w, const float *in, uint32_t n, hipStream_t s,
                         const uint32_t *n_dev)
{
    if (n == 0) return hipSuccess;
    const dim3 blk(256);
    hipLaunchKernelGGL(k_encode, grid_for(n, 256), blk, 0, s, c, in, n, n_dev, w.act[0]);
    const dim3 g = grid_for((n + 15) / 16, 4);
    for (uint32_t l = 0; l + 1 < kCacheLayers; ++l)
        hipLaunchKernelGGL((k_fwd<64, 0>), g, blk, 0, s, w.act[l], c.w_t + kCacheLayerOffset(l), n, n_dev,
                           (void *)w.act[l + 1]);
    hipLaunchKernelGGL((k_fwd<16, 1>), g, blk, 0, s, w.act[kCacheLayers - 1],
                       c.w_t + kCacheLayerOffset(kCacheLayers - 1), n, n_dev, (void *)w.y);
    return hipGetLastError();
}

hipError_t cache_infer(const CacheDev &c, const CacheWork &w, const float *in, uint32_t n, float *rgb, hipStream_t s,
                       const uint32_t *n_dev)
{
    hipError_t e = cache_forward(c, w, in, n, s, n_dev);
    if (e != hipSuccess || n == 0) return e;
    hipLaunchKernelGGL(k_out_rgb, grid_for(n, 256), dim3(256), 0, s, w.y, n, n_dev, rgb);
    return hipGetLastError();
}

hipError_t cache_backward(const CacheDev &c, const CacheWork &w, const float *in, const float *target, uint32_t n,
                          float *loss, hipStream_t s, const uint32_t *n_dev)
{
    if (n == 0) return hipSuccess;
    const dim3 blk(256);
    hipError_t e = cache_forward(c, w, in, n, s, n_dev);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_loss, grid_for(n, 256), blk, 0, s, w.y, target, n, n_dev, 1.0f / (3.0f * (float)n), w.dy16,
                       loss);
    const dim3 g = grid_for((n + 15) / 16, 4);
    const dim3 gw = grid_for((n + 63) / 64, 1, 512);
    const uint32_t L = kCacheLayers;
    // output layer
    hipLaunchKernelGGL((k_wgrad<16>), gw, blk, 0, s, w.act[L - 1], w.dy16, n, n_dev, c.grad + kCacheLayerOffset(L - 1));
    hipLaunchKernelGGL((k_bwd_data<16, true, false>), g, blk, 0, s, w.dy16, c.w + kCacheLayerOffset(L - 1),
                       w.act[L - 1], n, n_dev, (void *)w.dact[0]);
    // hidden layers l = L-2 .. 1: input act[l], output grad in dact
    uint32_t cur = 0;
    for (int l = (int)L - 2; l >= 1; --l) {
        hipLaunchKernelGGL((k_wgrad<64>), gw, blk, 0, s, w.act[l], w.dact[cur], n, n_dev,
                           c.grad + kCacheLayerOffset(l));
        hipLaunchKernelGGL((k_bwd_data<64, true, false>), g, blk, 0, s, w.dact[cur], c.w + kCacheLayerOffset(l),
                           w.act[l], n, n_dev, (void *)w.dact[cur ^ 1]);
        cur ^= 1;
    }
    // first layer: weight gradient and the encoding gradient (FP32)
    hipLaunchKernelGGL((k_wgrad<64>), gw, blk, 0, s, w.act[0], w.dact[cur], n, n_dev, c.grad + kCacheLayerOffset(0));
    hipLaunchKernelGGL((k_bwd_data<64, false, true>), g, blk, 0, s, w.dact[cur], c.w + kCacheLayerOffset(0),
                       w.act[0], n, n_dev, (void *)w.dx32);
    // the grid gradient in Morton order of the samples' positions
    hipLaunchKernelGGL(k_sample_morton, grid_for(n, 256), blk, 0, s, c, in, n, n_dev, w.skey[0], w.sidx[0]);
    size_t tb = w.sort_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(w.sort_tmp, tb, w.skey[0], w.skey[1], w.sidx[0], w.sidx[1], (int)n, 0, 32,
                                           s);
    if (e != hipSuccess) return e;
    const uint32_t runs = (n + 255) / 256;
    hipLaunchKernelGGL(k_grid_bwd, dim3(runs * c.levels), blk, 0, s, c, in, n, n_dev, runs, w.sidx[1], w.dx32,
                       c.grid_grad);
    return hipGetLastError();
}

size_t cache_sort_bytes(uint32_t rows)
{
    size_t b = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, b, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                           (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)rows, 0, 32) !=
        hipSuccess)
        return 0;
    return b;
}

hipError_t cache_adam(const CacheDev &c, const CacheOptsDev &o, uint32_t step, hipStream_t s)
{
    const float bc1 = 1.0f - powf(o.beta1, (float)step);
    const float bc2 = 1.0f - powf(o.beta2, (float)step);
    const dim3 blk(256);
    hipLaunchKernelGGL(k_adam, grid_for(kCacheParams, 256), blk, 0, s, c.w32, c.grad, c.m, c.v, kCacheParams, o.lr,
                       o.beta1, o.beta2, o.eps, bc1, bc2, c.w, c.w_t, 1u);
    const uint32_t gp = c.levels * c.table_size * 2;
    hipLaunchKernelGGL(k_adam, grid_for(gp, 256), blk, 0, s, c.grid32, c.grid_grad, c.grid_m, c.grid_v, gp, o.lr,
                       o.beta1, o.beta2, o.eps, bc1, bc2, (uint16_t *)c.grid_bf16, (uint16_t *)nullptr, 0u);
    return hipGetLastError();
}

hipError_t cache_fill(float *p, uint32_t count, float value, hipStream_t s)
{
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, grid_for(count, 256), dim3(256), 0, s, p, count, value);
    return hipGetLastError();
}

}  // namespace rt
