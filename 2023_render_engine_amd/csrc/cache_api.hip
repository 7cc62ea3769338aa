// cache_api.hip -- C ABI of the neural radiance cache (rt_cache_*).
#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "rt_cache.h"
#include "rt_internal.h"
#include "rt_pcg.h"

struct rt_ctx;
namespace rt {
void ctx_retain(rt_ctx *c);
void ctx_release(rt_ctx *c);
int ctx_device(rt_ctx *c);
void *ctx_stream_ptr(rt_ctx *c);
std::recursive_mutex &ctx_lock(rt_ctx *c);
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

struct rt_cache {
    rt_ctx *ctx = nullptr;
    rt_cache_opts opts{};
    rt::CacheDev dev{};
    rt::CacheWork work{};
    void *arena = nullptr;
    float *d_loss = nullptr;
    uint32_t step = 0;
    // divergence watch over the steps of rt_render_cache_frame: a moving
    // average of the step losses; a diverged cache is reinitialised and the
    // next frame renders plain
    double loss_ema = 0.0;
    uint32_t watched = 0;
    uint32_t resets = 0;
    bool fallback_pending = false;
};

// accessors for the renderer (scene.hip)
namespace rt {
const CacheDev &cache_device(const rt_cache *c) { return c->dev; }
rt_ctx *cache_ctx(const rt_cache *c) { return c->ctx; }
uint32_t cache_batch(const rt_cache *c) { return c->work.rows; }
}  // namespace rt

namespace {

uint16_t host_f2bf(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

size_t grid_params(const rt::CacheDev &d) { return (size_t)d.levels * d.table_size * 2; }

// Uploads FP32 parameters and derives every BF16 copy on the host.
rt_status upload_params(rt_cache *c, const float *mlp, const float *grid, hipStream_t s)
{
    const rt::CacheDev &d = c->dev;
    std::vector<uint16_t> w(rt::kCacheParams), wt(rt::kCacheParams);
    for (uint32_t l = 0; l < rt::kCacheLayers; ++l) {
        const uint32_t off = rt::kCacheLayerOffset(l);
        const uint32_t nout = (l + 1 == rt::kCacheLayers) ? rt::kCacheOut : rt::kCacheWidth;
        for (uint32_t k = 0; k < rt::kCacheWidth; ++k)
            for (uint32_t j = 0; j < nout; ++j) {
                const uint16_t h = host_f2bf(mlp[off + k * nout + j]);
                w[off + k * nout + j] = h;
                wt[off + j * rt::kCacheWidth + k] = h;
            }
    }
    const size_t gp = grid_params(d);
    std::vector<uint16_t> gb(gp);
    for (size_t i = 0; i < gp; ++i) gb[i] = host_f2bf(grid[i]);
    RT_HIP(hipMemcpyAsync(d.w32, mlp, rt::kCacheParams * 4, hipMemcpyHostToDevice, s));
    RT_HIP(hipMemcpyAsync(d.w, w.data(), rt::kCacheParams * 2, hipMemcpyHostToDevice, s));
    RT_HIP(hipMemcpyAsync(d.w_t, wt.data(), rt::kCacheParams * 2, hipMemcpyHostToDevice, s));
    RT_HIP(hipMemcpyAsync(d.grid32, grid, gp * 4, hipMemcpyHostToDevice, s));
    RT_HIP(hipMemcpyAsync(d.grid_bf16, gb.data(), gp * 2, hipMemcpyHostToDevice, s));
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

// Fresh parameters and optimiser state: Glorot-uniform MLP (zero in the
// unused output columns), grid features U(-1e-4, 1e-4) (Mueller et al. 2022),
// zero Adam moments and gradients, step count 0.
rt_status init_params(rt_cache *c, uint32_t seed, hipStream_t s)
{
    const rt::CacheDev &d = c->dev;
    const size_t gp = grid_params(d);
    rt_status st = RT_OK;
    float *zero_ptrs[] = {d.grad, d.m, d.v};
    for (float *p : zero_ptrs)
        if (rt::cache_fill(p, rt::kCacheParams, 0.0f, s) != hipSuccess) st = RT_ERR_HIP;
    float *gzero[] = {d.grid_grad, d.grid_m, d.grid_v};
    for (float *p : gzero)
        if (rt::cache_fill(p, (uint32_t)gp, 0.0f, s) != hipSuccess) st = RT_ERR_HIP;
    if (st != RT_OK) { set_error("cache parameter initialisation: fill kernel failed"); return st; }
    std::vector<float> mlp(rt::kCacheParams, 0.0f), grid(gp);
    rt::Pcg32 rng = rt::pcg32_seed(seed, 0xcace);
    for (uint32_t l = 0; l < rt::kCacheLayers; ++l) {
        const uint32_t off = rt::kCacheLayerOffset(l);
        const bool last = l + 1 == rt::kCacheLayers;
        const uint32_t nout = last ? rt::kCacheOut : rt::kCacheWidth, used = last ? 3u : nout;
        const float a = std::sqrt(6.0f / (float)(rt::kCacheWidth + used));
        for (uint32_t k = 0; k < rt::kCacheWidth; ++k)
            for (uint32_t j = 0; j < used; ++j) mlp[off + k * nout + j] = (2.0f * rt::pcg32_float(rng) - 1.0f) * a;
    }
    for (size_t i = 0; i < gp; ++i) grid[i] = (2.0f * rt::pcg32_float(rng) - 1.0f) * 1e-4f;
    c->step = 0;
    return upload_params(c, mlp.data(), grid.data(), s);
}

}  // namespace

extern "C" {

rt_status rt_cache_create(rt_ctx *ctx, const rt_cache_opts *opts, rt_cache **out)
{
    if (!ctx || !opts || !out) { set_error("rt_cache_create: null argument"); return RT_ERR_INVALID; }
    *out = nullptr;
    const rt_cache_opts &o = *opts;
    if (o.levels < 1 || o.levels > rt::kCacheMaxLevels || o.log2_table < 8 || o.log2_table > 24 || o.base_res < 2 ||
        !(o.per_level_scale >= 1.0f) || o.batch == 0 || o.batch > (1u << 22)) {
        set_error("rt_cache_create: levels 1..16, log2_table 8..24, base_res >= 2, scale >= 1, batch 1..2^22");
        return RT_ERR_INVALID;
    }
    for (int k = 0; k < 3; ++k)
        if (!(o.aabb_hi[k] > o.aabb_lo[k])) { set_error("rt_cache_create: empty bounding box"); return RT_ERR_INVALID; }
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(ctx));
    RT_HIP(hipSetDevice(rt::ctx_device(ctx)));
    hipStream_t s = (hipStream_t)rt::ctx_stream_ptr(ctx);
    rt_cache *c = new (std::nothrow) rt_cache();
    if (!c) { set_error("rt_cache_create: host allocation failed"); return RT_ERR_OOM; }
    c->ctx = ctx;
    c->opts = o;
    rt::CacheDev &d = c->dev;
    d.levels = o.levels;
    d.table_size = 1u << o.log2_table;
    for (uint32_t l = 0; l < o.levels; ++l) {
        // N_l = floor(N_min * b^l) (Mueller et al. 2022, eq. 2-3)
        const double r = std::floor((double)o.base_res * std::pow((double)o.per_level_scale, (double)l));
        d.res[l] = (uint32_t)r;
        const double cells = std::pow(r + 1.0, 3.0);
        d.dense[l] = cells <= (double)d.table_size ? 1u : 0u;
    }
    for (int k = 0; k < 3; ++k) {
        d.lo[k] = o.aabb_lo[k];
        d.inv_ext[k] = 1.0f / (o.aabb_hi[k] - o.aabb_lo[k]);
    }
    const uint32_t rows = o.batch;
    const size_t gp = grid_params(d);
    size_t total = 0;
    auto take = [&](size_t b) { size_t off = total; total += (b + 255) & ~(size_t)255; return off; };
    const size_t o_w = take(rt::kCacheParams * 2), o_wt = take(rt::kCacheParams * 2);
    const size_t o_w32 = take(rt::kCacheParams * 4), o_g = take(rt::kCacheParams * 4);
    const size_t o_m = take(rt::kCacheParams * 4), o_v = take(rt::kCacheParams * 4);
    const size_t o_gb = take(gp * 2), o_g32 = take(gp * 4), o_gg = take(gp * 4), o_gm = take(gp * 4),
                 o_gv = take(gp * 4);
    size_t o_act[rt::kCacheLayers];
    for (uint32_t l = 0; l < rt::kCacheLayers; ++l) o_act[l] = take((size_t)rows * rt::kCacheWidth * 2);
    const size_t o_y = take((size_t)rows * 16 * 4), o_dy = take((size_t)rows * 16 * 2);
    const size_t o_da0 = take((size_t)rows * 64 * 2), o_da1 = take((size_t)rows * 64 * 2);
    const size_t o_dx = take((size_t)rows * 64 * 4), o_loss = take(4);
    const size_t sort_bytes = std::max<size_t>(rt::cache_sort_bytes(rows), 1);
    const size_t o_sk0 = take((size_t)rows * 4), o_sk1 = take((size_t)rows * 4), o_si0 = take((size_t)rows * 4),
                 o_si1 = take((size_t)rows * 4), o_stmp = take(sort_bytes);
    char *base = nullptr;
    hipError_t e = hipMalloc(&base, total);
    if (e != hipSuccess) {
        delete c;
        set_error(std::string("rt_cache_create: device allocation: ") + hipGetErrorString(e));
        return RT_ERR_OOM;
    }
    c->arena = base;
    d.w = (uint16_t *)(base + o_w); d.w_t = (uint16_t *)(base + o_wt);
    d.w32 = (float *)(base + o_w32); d.grad = (float *)(base + o_g);
    d.m = (float *)(base + o_m); d.v = (float *)(base + o_v);
    d.grid_bf16 = (uint32_t *)(base + o_gb); d.grid32 = (float *)(base + o_g32);
    d.grid_grad = (float *)(base + o_gg); d.grid_m = (float *)(base + o_gm); d.grid_v = (float *)(base + o_gv);
    rt::CacheWork &w = c->work;
    w.rows = rows;
    for (uint32_t l = 0; l < rt::kCacheLayers; ++l) w.act[l] = (uint16_t *)(base + o_act[l]);
    w.y = (float *)(base + o_y);
    w.dy16 = (uint16_t *)(base + o_dy);
    w.dact[0] = (uint16_t *)(base + o_da0);
    w.dact[1] = (uint16_t *)(base + o_da1);
    w.dx32 = (float *)(base + o_dx);
    w.skey[0] = (uint32_t *)(base + o_sk0);
    w.skey[1] = (uint32_t *)(base + o_sk1);
    w.sidx[0] = (uint32_t *)(base + o_si0);
    w.sidx[1] = (uint32_t *)(base + o_si1);
    w.sort_tmp = base + o_stmp;
    w.sort_bytes = sort_bytes;
    c->d_loss = (float *)(base + o_loss);
    rt_status st = init_params(c, o.seed, s);
    if (st != RT_OK) {
        (void)hipFree(c->arena);
        delete c;
        return st;
    }
    rt::ctx_retain(ctx);
    *out = c;
    return RT_OK;
}

rt_status rt_cache_destroy(rt_cache *c)
{
    if (!c) return RT_OK;
    rt_ctx *ctx = c->ctx;
    {
        std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(ctx));
        (void)hipSetDevice(rt::ctx_device(ctx));
        (void)hipStreamSynchronize((hipStream_t)rt::ctx_stream_ptr(ctx));
        if (c->arena) (void)hipFree(c->arena);
        delete c;
    }
    rt::ctx_release(ctx);
    return RT_OK;
}

rt_status rt_cache_get_info(const rt_cache *c, rt_cache_info *info)
{
    if (!c || !info) { set_error("rt_cache_get_info: null argument"); return RT_ERR_INVALID; }
    info->levels = c->dev.levels;
    info->table_size = c->dev.table_size;
    for (uint32_t l = 0; l < 16; ++l) {
        info->res[l] = l < c->dev.levels ? c->dev.res[l] : 0;
        info->dense[l] = l < c->dev.levels ? c->dev.dense[l] : 0;
    }
    info->mlp_params = rt::kCacheParams;
    info->grid_params = (uint64_t)grid_params(c->dev);
    info->steps = c->step;
    return RT_OK;
}

rt_status rt_cache_query(rt_cache *c, const float *in, uint32_t n, float *rgb, void *stream_)
{
    if (!c || (n && (!in || !rgb))) { set_error("rt_cache_query: null argument"); return RT_ERR_INVALID; }
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(c->ctx));
    RT_HIP(hipSetDevice(rt::ctx_device(c->ctx)));
    hipStream_t s = stream_ ? (hipStream_t)stream_ : (hipStream_t)rt::ctx_stream_ptr(c->ctx);
    for (uint32_t i0 = 0; i0 < n; i0 += c->work.rows) {
        const uint32_t m = std::min(c->work.rows, n - i0);
        RT_HIP(rt::cache_infer(c->dev, c->work, in + (size_t)i0 * rt::kCacheInStride, m, rgb + (size_t)i0 * 3, s));
    }
    return RT_OK;
}

static rt_status train_batch(rt_cache *c, const float *in, const float *target, uint32_t n, hipStream_t s)
{
    if (n > c->work.rows) { set_error("rt_cache_train_step: batch larger than the cache's batch size"); return RT_ERR_INVALID; }
    RT_HIP(rt::cache_fill(c->d_loss, 1, 0.0f, s));
    RT_HIP(rt::cache_backward(c->dev, c->work, in, target, n, c->d_loss, s));
    return RT_OK;
}

}  // extern "C"

namespace rt {
// One optimiser step on min(n, batch, *n_dev) rows whose count is known only
// on the device; the loss sum is left in *loss_dev.  No host synchronisation.
rt_status cache_train_rows_dev(rt_cache *c, const float *in, const float *target, uint32_t n, const uint32_t *n_dev,
                               hipStream_t s, float **loss_dev)
{
    n = std::min(n, c->work.rows);
    if (loss_dev) *loss_dev = c->d_loss;
    RT_HIP(rt::cache_fill(c->d_loss, 1, 0.0f, s));
    if (n == 0) return RT_OK;
    RT_HIP(rt::cache_backward(c->dev, c->work, in, target, n, c->d_loss, s, n_dev));
    ++c->step;
    rt::CacheOptsDev od{c->opts.lr, c->opts.beta1, c->opts.beta2, c->opts.eps};
    RT_HIP(rt::cache_adam(c->dev, od, c->step, s));
    return RT_OK;
}

// `steps` Adam steps on `total` harvested (record, target) rows: each step
// takes the next min(batch, total) rows, wrapping around, so a small harvest
// is revisited rather than padded with empty steps.  Synchronises to read the
// last step's loss, which feeds the divergence watch: a non-finite loss, or
// one above 100 times the moving average of the watched steps (after 16 of
// them), reinitialises the cache with fresh parameters and asks for the next
// frame to be rendered without it.
rt_status cache_train_harvest(rt_cache *c, const float *in, const float *target, uint32_t total, uint32_t steps,
                              hipStream_t s, uint32_t *steps_done, float *mean_loss, bool *reset)
{
    *steps_done = 0;
    *mean_loss = std::numeric_limits<float>::quiet_NaN();
    *reset = false;
    if (total == 0 || steps == 0) return RT_OK;
    const uint32_t n = std::min(c->work.rows, total);
    uint32_t last = n;
    for (uint32_t k = 0; k < steps; ++k) {
        const uint32_t off = (uint32_t)(((uint64_t)k * n) % total);
        last = std::min(n, total - off);
        rt_status st = cache_train_rows_dev(c, in + (size_t)off * kCacheInStride, target + (size_t)off * 3, last,
                                            nullptr, s, nullptr);
        if (st != RT_OK) return st;
    }
    float l = 0;
    RT_HIP(hipMemcpyAsync(&l, c->d_loss, 4, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    *steps_done = steps;
    const float mean = l / (3.0f * (float)last);
    *mean_loss = mean;
    bool diverged = !std::isfinite(mean);
    if (!diverged && c->watched >= 16 && (double)mean > 100.0 * c->loss_ema) diverged = true;
    if (!diverged) {
        c->loss_ema = c->watched ? 0.9 * c->loss_ema + 0.1 * (double)mean : (double)mean;
        ++c->watched;
        return RT_OK;
    }
    ++c->resets;
    rt_status st = init_params(c, c->opts.seed + 0x9E3779B1u * c->resets, s);
    if (st != RT_OK) return st;
    c->loss_ema = 0.0;
    c->watched = 0;
    c->fallback_pending = true;
    *reset = true;
    return RT_OK;
}

// The frame about to render falls back to plain path tracing when the cache
// was reset after diverging (the flag is consumed).
bool cache_take_fallback(rt_cache *c)
{
    const bool f = c->fallback_pending;
    c->fallback_pending = false;
    return f;
}

uint32_t cache_resets(const rt_cache *c) { return c->resets; }

}  // namespace rt

extern "C" {

rt_status rt_cache_train_step(rt_cache *c, const float *in, const float *target, uint32_t n, float *loss,
                              void *stream_)
{
    if (!c || (n && (!in || !target))) { set_error("rt_cache_train_step: null argument"); return RT_ERR_INVALID; }
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(c->ctx));
    RT_HIP(hipSetDevice(rt::ctx_device(c->ctx)));
    hipStream_t s = stream_ ? (hipStream_t)stream_ : (hipStream_t)rt::ctx_stream_ptr(c->ctx);
    if (n == 0) return RT_OK;
    rt_status st = train_batch(c, in, target, n, s);
    if (st != RT_OK) return st;
    ++c->step;
    rt::CacheOptsDev od{c->opts.lr, c->opts.beta1, c->opts.beta2, c->opts.eps};
    RT_HIP(rt::cache_adam(c->dev, od, c->step, s));
    if (loss) {
        float l = 0;
        RT_HIP(hipMemcpyAsync(&l, c->d_loss, 4, hipMemcpyDeviceToHost, s));
        RT_HIP(hipStreamSynchronize(s));
        *loss = l / (3.0f * (float)n);
    }
    return RT_OK;
}

rt_status rt_cache_gradients(rt_cache *c, const float *in, const float *target, uint32_t n, float *mlp_grad,
                             float *grid_grad, float *loss)
{
    if (!c || !in || !target || !mlp_grad || !grid_grad) { set_error("rt_cache_gradients: null argument"); return RT_ERR_INVALID; }
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(c->ctx));
    RT_HIP(hipSetDevice(rt::ctx_device(c->ctx)));
    hipStream_t s = (hipStream_t)rt::ctx_stream_ptr(c->ctx);
    rt_status st = train_batch(c, in, target, n, s);
    if (st != RT_OK) return st;
    const size_t gp = grid_params(c->dev);
    float l = 0;
    RT_HIP(hipMemcpyAsync(mlp_grad, c->dev.grad, rt::kCacheParams * 4, hipMemcpyDeviceToHost, s));
    RT_HIP(hipMemcpyAsync(grid_grad, c->dev.grid_grad, gp * 4, hipMemcpyDeviceToHost, s));
    RT_HIP(hipMemcpyAsync(&l, c->d_loss, 4, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    if (loss) *loss = l / (3.0f * (float)n);
    RT_HIP(rt::cache_fill(c->dev.grad, rt::kCacheParams, 0.0f, s));
    RT_HIP(rt::cache_fill(c->dev.grid_grad, (uint32_t)gp, 0.0f, s));
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

rt_status rt_cache_get_params(const rt_cache *c, float *mlp, float *grid)
{
    if (!c || !mlp || !grid) { set_error("rt_cache_get_params: null argument"); return RT_ERR_INVALID; }
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(c->ctx));
    RT_HIP(hipSetDevice(rt::ctx_device(c->ctx)));
    hipStream_t s = (hipStream_t)rt::ctx_stream_ptr(c->ctx);
    RT_HIP(hipMemcpyAsync(mlp, c->dev.w32, rt::kCacheParams * 4, hipMemcpyDeviceToHost, s));
    RT_HIP(hipMemcpyAsync(grid, c->dev.grid32, grid_params(c->dev) * 4, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

rt_status rt_cache_set_params(rt_cache *c, const float *mlp, const float *grid)
{
    if (!c || !mlp || !grid) { set_error("rt_cache_set_params: null argument"); return RT_ERR_INVALID; }
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(c->ctx));
    RT_HIP(hipSetDevice(rt::ctx_device(c->ctx)));
    return upload_params(c, mlp, grid, (hipStream_t)rt::ctx_stream_ptr(c->ctx));
}

}  // extern "C"
