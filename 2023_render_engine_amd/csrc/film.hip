// film.hip -- film output on the GPU: the tone-mapping operators of
// camera/Film.cu:156-283 applied to every pixel sum, and the pixel arrays
// that Film::toneMapBMP / toneMapPPM (Film.cu:285-337) hand to the BMP writer
// (camera/BMP_Format.cu) and the PPM stream.
//
// Two passes over the film, which stays in HBM:
//   k_film_max   grid-stride max of every channel and of the luminance of
//                every pixel sum (getMaxChannel / getMaxLuminance,
//                Film.cu:120-154); one partial pair per block, no atomics,
//                no initialisation; k_film_fold (one block) folds the pairs;
//   k_film_map   one lane per pixel of an output row (2-D grid: x across the
//                row, y = row), three byte stores, the row's BMP padding
//                written by its last lane.
// Arithmetic is double precision in the reference's operation order (the
// library is built with -ffp-contract=off), so the bytes match a host
// evaluation wherever the value before truncation is not within an ulp or
// so of a multiple of 1/255 (the gamma operator's power is the only non
// correctly rounded operation).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>

#include "rt_internal.h"

struct rt_ctx;
namespace rt {
int ctx_device(rt_ctx *c);
void *ctx_stream_ptr(rt_ctx *c);
std::recursive_mutex &ctx_lock(rt_ctx *c);
void *ctx_film_scratch(rt_ctx *c, size_t bytes);
}  // namespace rt

using rt::set_error;

#define RT_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            set_error(std::string(#call) + " failed: " + hipGetErrorString(e_));        \
            return e_ == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP;                 \
        }                                                                               \
    } while (0)

namespace {

constexpr int kThreads = 256;
constexpr uint32_t kPartials = 1024;      // blocks of k_film_max (the scratch holds kPartials + 1 pairs)
constexpr uint32_t kMaxSide = 1u << 15;   // width and height limit (BMP stores int32; keeps sizes in 32 bits)

struct FilmIn {
    const double *sums;      // width * height * 3
    const uint32_t *nvalid;  // null: sums are the pixel sums
    double samples;          // Film samplesLoadedFromFile + samplesComputedPixel0
};

// Pixel sum i: the sums, or the renderer's estimate sums * (samples / nvalid)
// (integration/pathTracing.cu:474-486; 0 where no sample was valid).
__device__ inline void pixel_sum(const FilmIn &f, size_t i, double p[3])
{
    double s = 1.0;
    if (f.nvalid) {
        const uint32_t n = f.nvalid[i];
        s = n ? f.samples / (double)n : 0.0;
    }
    p[0] = f.sums[3 * i];
    p[1] = f.sums[3 * i + 1];
    p[2] = f.sums[3 * i + 2];
    if (f.nvalid) { p[0] = p[0] * s; p[1] = p[1] * s; p[2] = p[2] * s; }
}

// Spectrum::brightness (radiometry/Spectrum.cu:68-75): float weights held in doubles.
__device__ inline double brightness(const double p[3])
{
    const double wr = 0.212671f, wg = 0.715160f, wb = 0.072169f;
    return wr * p[0] + wg * p[1] + wb * p[2];
}

__device__ inline double dmax(double a, double b) { return b > a ? b : a; }

// math.cu:44-47 clamp as max(a, min(b, x)): a NaN yields the upper bound.
__device__ inline double clamp01(double x)
{
    const double m = x < 1.0 ? x : 1.0;
    return m > 0.0 ? m : 0.0;
}

// Film.cu:202-211; the float constants meet in float before they widen.
__device__ inline double uncharted2_partial(double x)
{
    const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
    const float cb = C * B, de = D * E, df = D * F, ef = E / F;
    return ((x * (x * (double)A + (double)cb) + (double)de) / (x * (x * (double)A + (double)B) + (double)df)) -
           (double)ef;
}

// Film.cu:224-229; the Float overload of operator* narrows 0.983729 to float.
__device__ inline double rtt_and_odt_fit(double v)
{
    const float k = 0.983729f;
    const double a = v * (v + 0.0245786) - 0.000090537;
    const double b = v * (v * (double)k + 0.4329510) + 0.238081;
    return a / b;
}

struct Maxima {
    double max, max_lum;  // already divided by the sample count
};

// toneMaxPixel (Film.cu:266-283).
__device__ inline void tonemap(double p[3], const Maxima &m, double n, int mode)
{
    p[0] = p[0] / n; p[1] = p[1] / n; p[2] = p[2] / n;
    switch (mode) {
    case RT_TONE_EQUALIZE:
        p[0] = p[0] / m.max; p[1] = p[1] / m.max; p[2] = p[2] / m.max;
        break;
    case RT_TONE_REINHARD: {
        const double l_old = brightness(p);
        if (l_old == 0) break;
        const double numerator = l_old * (1.0 + (l_old / (m.max_lum * m.max_lum)));
        const double l_new = numerator / (1.0 + l_old);
        const double s = l_new / l_old;
        p[0] = clamp01(p[0] * s); p[1] = clamp01(p[1] * s); p[2] = clamp01(p[2] * s);
        break;
    }
    case RT_TONE_REINHARD_JODIE: {
        // lerp(x, a, b) = (1 - x) * a + x * b with a = b = pixel / (1 + luminance)
        const double lum = brightness(p);
        for (int c = 0; c < 3; ++c) {
            const double rp = p[c] / (1 + lum);
            const double x = p[c] / (p[c] + 1.0);
            p[c] = clamp01((1.0 - x) * rp + x * rp);
        }
        break;
    }
    case RT_TONE_ACES_FILMIC: {
        const double r = p[0], g = p[1], b = p[2];
        double c0 = 0.59719 * r + 0.35458 * g + 0.04823 * b;
        double c1 = 0.07600 * r + 0.90834 * g + 0.01566 * b;
        double c2 = 0.02840 * r + 0.13383 * g + 0.83777 * b;
        c0 = rtt_and_odt_fit(c0); c1 = rtt_and_odt_fit(c1); c2 = rtt_and_odt_fit(c2);
        p[0] = clamp01(1.60475 * c0 + -0.53108 * c1 + -0.07367 * c2);
        p[1] = clamp01(-0.10208 * c0 + 1.10813 * c1 + -0.00605 * c2);
        p[2] = clamp01(-0.00327 * c0 + -0.07276 * c1 + 1.07602 * c2);
        break;
    }
    case RT_TONE_UNCHARTED2: {
        const float exposure_bias = 2.0f;
        const double white_scale = 1.0 / uncharted2_partial(11.2);
        for (int c = 0; c < 3; ++c) p[c] = uncharted2_partial(p[c] * (double)exposure_bias) * white_scale;
        break;
    }
    default:  // RT_TONE_GAMMA_CLAMP
        for (int c = 0; c < 3; ++c) p[c] = pow(p[c] / m.max, 1.0 / 2.2);
        break;
    }
}

// (unsigned char)(v * 255) as the reference's x86-64 host code computes it:
// a 32-bit truncating conversion (out of range or NaN: 0x80000000), low byte.
__device__ inline uint32_t to_byte(double v)
{
    const double x = v * 255;
    const int32_t i = (x > -2147483649.0 && x < 2147483648.0) ? (int32_t)x : INT32_MIN;
    return (uint32_t)i & 0xffu;
}

__device__ inline void fold(double &a, double &b)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a = dmax(a, __shfl_xor(a, off));
        b = dmax(b, __shfl_xor(b, off));
    }
}

__global__ __launch_bounds__(kThreads) void k_film_max(FilmIn f, size_t npix, double2 *partial)
{
    double m = 0.0, ml = 0.0;
    const size_t stride = (size_t)gridDim.x * kThreads;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < npix; i += stride) {
        double p[3];
        pixel_sum(f, i, p);
        if (p[0] > m) m = p[0];
        if (p[1] > m) m = p[1];
        if (p[2] > m) m = p[2];
        const double l = brightness(p);
        if (l > ml) ml = l;
    }
    fold(m, ml);
    __shared__ double2 s[kThreads / 64];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = make_double2(m, ml);
    __syncthreads();
    if (threadIdx.x == 0) {
        double2 r = s[0];
        for (int w = 1; w < kThreads / 64; ++w) { r.x = dmax(r.x, s[w].x); r.y = dmax(r.y, s[w].y); }
        partial[blockIdx.x] = r;
    }
}

// Folds the partial pairs of k_film_max into partial[npart] (one block).
__global__ __launch_bounds__(kThreads) void k_film_fold(double2 *partial, uint32_t npart)
{
    __shared__ double2 s[kThreads / 64];
    double m = 0.0, ml = 0.0;
    for (uint32_t k = threadIdx.x; k < npart; k += kThreads) {
        const double2 v = partial[k];
        m = dmax(m, v.x);
        ml = dmax(ml, v.y);
    }
    fold(m, ml);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = make_double2(m, ml);
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kThreads / 64; ++w) { m = dmax(m, s[w].x); ml = dmax(ml, s[w].y); }
        partial[npart] = make_double2(m, ml);
    }
}

// Byte of pow(x, 1 / 2.2) (the gamma operator) without a double pow in the
// common case; x = c / samples / max, and xa = c * (1 / (samples * max)),
// which is within a few ulp of it.  A float estimate of xa^(1/2.2) (relative
// error below 1e-5 for every value the float range holds) refined by one
// Newton step on y^11 = xa^5 (1 / 2.2 = 5 / 11) is within 1e-9 of the true
// power; where 255 * y lies more than 1e-4 from an integer, its truncation
// is the one the reference's pow-multiply-truncate chain gives.  Near an
// integer, and for xa outside (0, 1] or an estimate that under- or
// overflows, the double pow of the exact x decides.
__device__ inline uint32_t gamma_byte(double c, double samples, double max, double inv)
{
    const double x = c * inv;
    if (x > 0.0 && x <= 1.0) {
        const double y0 = (double)exp2f(log2f((float)x) * (float)(1.0 / 2.2));
        const double x2 = x * x, x5 = x2 * x2 * x;
        const double y2 = y0 * y0, y4 = y2 * y2, y10 = y4 * y4 * y2;
        const double y = y0 - (y10 * y0 - x5) / (11.0 * y10);
        const double t = y * 255.0;
        const double k = rint(t);
        if (fabs(t - k) > 1e-4 || k == 0.0) return (uint32_t)(int32_t)t & 0xffu;
    }
    return to_byte(pow(c / samples / max, 1.0 / 2.2));
}

struct MapOut {
    uint8_t *pixels;
    uint32_t width, height;
    uint32_t row_bytes;  // BMP: 3 * width rounded up to 4; PPM: 3 * width
    int bmp;
};

// One lane per pixel of output row blockIdx.y; BMP rows run bottom-up, so
// output row j holds film row height-1-j, and the row's last lane writes its
// zero padding.
__global__ __launch_bounds__(kThreads) void k_film_map(FilmIn f, const double2 *total, MapOut o, int mode)
{
    const uint32_t x = blockIdx.x * kThreads + threadIdx.x;
    const uint32_t j = blockIdx.y;
    if (x >= o.width) return;
    const double2 t = *total;
    const Maxima mx{t.x / f.samples, t.y / f.samples};
    const uint32_t y = o.bmp ? o.height - 1u - j : j;
    double p[3];
    pixel_sum(f, (size_t)y * o.width + x, p);
    uint32_t r, g, b;
    if (mode == RT_TONE_GAMMA_CLAMP) {
        const double inv = 1.0 / (f.samples * mx.max);
        r = gamma_byte(p[0], f.samples, mx.max, inv);
        g = gamma_byte(p[1], f.samples, mx.max, inv);
        b = gamma_byte(p[2], f.samples, mx.max, inv);
    } else {
        tonemap(p, mx, f.samples, mode);
        r = to_byte(p[0]);
        g = to_byte(p[1]);
        b = to_byte(p[2]);
    }
    uint8_t *row = o.pixels + (size_t)j * o.row_bytes;
    uint8_t *dst = row + 3u * x;
    dst[0] = (uint8_t)(o.bmp ? b : r);
    dst[1] = (uint8_t)g;
    dst[2] = (uint8_t)(o.bmp ? r : b);
    if (x == o.width - 1u)
        for (uint32_t k = 3u * o.width; k < o.row_bytes; ++k) row[k] = 0;
}

uint32_t bmp_row_bytes(uint32_t w) { return (3u * w + 3u) & ~3u; }

// ptr must be device (or managed) memory of `device`; where the runtime
// knows the allocation's range (plain hipMalloc blocks), it must also reach
// ptr + bytes.  Memory whose range the runtime does not report (virtual
// memory mappings) is accepted on its attributes alone.
bool device_buffer_ok(const void *ptr, size_t bytes, int device)
{
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, ptr) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if ((at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged) || at.device != device) return false;
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)ptr) != hipSuccess) {
        (void)hipGetLastError();
        return true;
    }
    const uintptr_t p = (uintptr_t)ptr, b0 = (uintptr_t)base;
    return p >= b0 && p - b0 <= size && bytes <= size - (p - b0);
}

}  // namespace

extern "C" {

size_t rt_image_pixel_bytes(rt_image_format format, uint32_t width, uint32_t height)
{
    if (width == 0 || height == 0 || width > kMaxSide || height > kMaxSide) return 0;
    if (format == RT_IMAGE_BMP) return (size_t)bmp_row_bytes(width) * height;
    if (format == RT_IMAGE_PPM) return (size_t)3 * width * height;
    return 0;
}

size_t rt_image_header(rt_image_format format, uint32_t width, uint32_t height, uint8_t *out, size_t cap)
{
    if (width == 0 || height == 0 || width > kMaxSide || height > kMaxSide) return 0;
    uint8_t h[64] = {0};
    size_t len = 0;
    auto u16 = [&](size_t at, uint32_t v) { h[at] = (uint8_t)v; h[at + 1] = (uint8_t)(v >> 8); };
    auto u32 = [&](size_t at, uint32_t v) { for (int k = 0; k < 4; ++k) h[at + k] = (uint8_t)(v >> (8 * k)); };
    if (format == RT_IMAGE_BMP) {
        // BitmapImage (BMP_Format.cu): 14-byte file header + 40-byte info
        // header; Image_Size keeps the reference's (width + 4 - width % 4) * height
        const uint32_t image_size = (width + 4u - (width % 4u)) * height;
        h[0] = 'B'; h[1] = 'M';
        u32(2, 14u + 40u + image_size);
        u32(10, 54u);
        u32(14, 40u);
        u32(18, width);
        u32(22, height);
        u16(26, 1u);
        u16(28, 24u);
        u32(30, 0u);
        u32(34, image_size);
        u32(38, 3780u);
        u32(42, 3780u);
        len = 54;
    } else if (format == RT_IMAGE_PPM) {
        const std::string s = "P6\n" + std::to_string(width) + " " + std::to_string(height) + "\n255\n";
        len = s.size();
        for (size_t k = 0; k < len; ++k) h[k] = (uint8_t)s[k];
    } else {
        return 0;
    }
    if (out)
        for (size_t k = 0; k < len && k < cap; ++k) out[k] = h[k];
    return len;
}

}  // extern "C"

// The partial maxima of rt_film_tonemap_noalloc: one array per GPU, kept for
// the life of the process and shared by every such call on that GPU.  Host
// code never passes the array's name to a kernel (on the host it is a
// stand-in, not a device address): each device's address is resolved once
// with hipGetSymbolAddress and used from then on.
constexpr uint32_t kScratchPairs = kPartials + 1;  // 1025 pairs of doubles, 16.4 KB
static __device__ __align__(256) double2 scratch[kScratchPairs];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_lock;
double2 *g_scratch[kMaxScratchDevices] = {};

// The device's copy of `scratch` (that device current), or null with the
// error text set.
double2 *scratch_on(int device)
{
    if (device < 0 || device >= kMaxScratchDevices) {
        set_error("rt_film_tonemap_noalloc: device index beyond the scratch table");
        return nullptr;
    }
    std::lock_guard<std::mutex> g(g_scratch_lock);
    if (!g_scratch[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error(std::string("rt_film_tonemap_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return nullptr;
        }
        g_scratch[device] = (double2 *)ptr;
    }
    return g_scratch[device];
}

// Shared body of rt_film_tonemap and rt_film_tonemap_noalloc: `partial` is
// null for the context's scratch (allocated on first use), else a device
// buffer of kPartials + 1 pairs.
rt_status tonemap(const char *who, bool noalloc, rt_ctx *ctx, const double *sums, const uint32_t *nvalid,
                  uint32_t width, uint32_t height, uint64_t samples, rt_tone_mapping mode, rt_image_format format,
                  uint8_t *pixels, size_t pixel_bytes, void *stream_)
{
    const std::string name(who);
    if (!ctx || !sums || !pixels) { set_error(name + ": null argument"); return RT_ERR_INVALID; }
    if ((int)mode < RT_TONE_EQUALIZE || (int)mode > RT_TONE_GAMMA_CLAMP) {
        set_error(name + ": unknown tone mapping");
        return RT_ERR_INVALID;
    }
    const size_t need = rt_image_pixel_bytes(format, width, height);
    if (need == 0) {
        set_error(name + ": image format unknown or size outside 1..32768 per side");
        return RT_ERR_INVALID;
    }
    if (pixel_bytes < need) { set_error(name + ": pixel buffer too small"); return RT_ERR_INVALID; }
    if ((uintptr_t)pixels % 4u || (uintptr_t)sums % 8u || (uintptr_t)nvalid % 4u) {
        set_error(name + ": pixels must be 4-byte aligned, sums 8-byte, nvalid 4-byte");
        return RT_ERR_INVALID;
    }
    if (samples == 0 || samples > (1ull << 53)) {
        set_error(name + ": samples must be 1..2^53 (the film divides by it)");
        return RT_ERR_INVALID;
    }
    const size_t npix = (size_t)width * height;
    // partial pairs of the max pass (one per block) plus the folded pair
    const uint32_t nmax = (uint32_t)std::min<size_t>(kPartials, (npix + kThreads - 1) / kThreads);
    if (noalloc && nmax + 1 > kScratchPairs) {
        set_error(name + ": the film needs more partial maxima than the scratch array holds");
        return RT_ERR_INVALID;
    }
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(ctx));
    RT_HIP(hipSetDevice(rt::ctx_device(ctx)));
    const int dev = rt::ctx_device(ctx);
    if (!device_buffer_ok(sums, npix * 3 * sizeof(double), dev) ||
        (nvalid && !device_buffer_ok(nvalid, npix * sizeof(uint32_t), dev)) || !device_buffer_ok(pixels, need, dev)) {
        set_error(name + ": a buffer is not device memory of the film's size on this context's device");
        return RT_ERR_INVALID;
    }
    double2 *partial = nullptr;
    if (noalloc) {
        partial = scratch_on(dev);
        if (!partial) return RT_ERR_HIP;
    } else {
        partial = (double2 *)rt::ctx_film_scratch(ctx, (kPartials + 1) * sizeof(double2));
        if (!partial) { set_error(name + ": device allocation failed"); return RT_ERR_OOM; }
    }
    hipStream_t stream = stream_ ? (hipStream_t)stream_ : (hipStream_t)rt::ctx_stream_ptr(ctx);

    const FilmIn f{sums, nvalid, (double)samples};
    k_film_max<<<nmax, kThreads, 0, stream>>>(f, npix, partial);
    RT_HIP(hipGetLastError());
    k_film_fold<<<1, kThreads, 0, stream>>>(partial, nmax);
    RT_HIP(hipGetLastError());

    MapOut o{};
    o.pixels = pixels;
    o.width = width;
    o.height = height;
    o.bmp = format == RT_IMAGE_BMP;
    o.row_bytes = o.bmp ? bmp_row_bytes(width) : 3u * width;
    const dim3 grid((width + kThreads - 1) / kThreads, height);
    k_film_map<<<grid, kThreads, 0, stream>>>(f, partial + nmax, o, (int)mode);
    RT_HIP(hipGetLastError());
    return RT_OK;
}

}  // namespace

extern "C" {

rt_status rt_film_tonemap(rt_ctx *ctx, const double *sums, const uint32_t *nvalid, uint32_t width, uint32_t height,
                          uint64_t samples, rt_tone_mapping mode, rt_image_format format, uint8_t *pixels,
                          size_t pixel_bytes, void *stream)
{
    return tonemap("rt_film_tonemap", false, ctx, sums, nvalid, width, height, samples, mode, format, pixels,
                   pixel_bytes, stream);
}

rt_status rt_film_tonemap_noalloc(rt_ctx *ctx, const double *sums, const uint32_t *nvalid, uint32_t width,
                                  uint32_t height, uint64_t samples, rt_tone_mapping mode, rt_image_format format,
                                  uint8_t *pixels, size_t pixel_bytes, void *stream)
{
    return tonemap("rt_film_tonemap_noalloc", true, ctx, sums, nvalid, width, height, samples, mode, format, pixels,
                   pixel_bytes, stream);
}

}  // extern "C"
