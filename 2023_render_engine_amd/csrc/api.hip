// api.hip -- the C ABI (include/rt.h): contexts, device memory helpers, BVH
// objects and ray queries.  Host code only; kernels live in trace.hip and
// render.hip.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <atomic>
#include <mutex>
#include <new>
#include <string>

#include "rt_bvh_gpu.h"
#include "rt_internal.h"
#include "rt_trace.h"

// A context is reference counted: every BVH and scene built on it holds a
// reference, so rt_ctx_destroy before them only marks it closed and the last
// holder frees it.
struct rt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::recursive_mutex lock;
    std::atomic<int> refs{1};
    void *film_scratch = nullptr;  // partial maxima of the film output kernels, allocated on first use
};

namespace rt {

void ctx_retain(rt_ctx *c) { c->refs.fetch_add(1); }

void ctx_release(rt_ctx *c)
{
    if (c->refs.fetch_sub(1) == 1) {
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        if (c->film_scratch) (void)hipFree(c->film_scratch);
        (void)hipStreamDestroy(c->stream);
        delete c;
    }
}

}  // namespace rt

struct rt_bvh {
    rt_ctx *ctx = nullptr;
    rt::DevBvh dev;
    rt::HostBvh host;  // kept for export and info; tris carry transparency
    // built on the device (rt_bvh_build_device): the arrays live there and
    // host holds only the scalar figures
    bool on_device = false;
    rt::DeviceBvhBuild gpu;
    uint32_t num_nodes2 = 0, num_nodes4 = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    // ray-claim counters of the persistent kernels (rt::kClaimWords words per
    // launch): a ring, so launches that are in flight together on different
    // streams do not share one
    uint32_t *ray_counters = nullptr;
    uint32_t seq = 0;
    uint32_t *next_counter() { return ray_counters + (size_t)(seq++ % kRayCounterRing) * rt::kClaimWords; }
    static constexpr uint32_t kRayCounterRing = 64;
    uint32_t width = 4;
};

// Traversal stack entries a lane needs for this tree's layout.
static uint32_t stack_cap(const rt_bvh *b)
{
    return b->width == 4 ? b->host.max_stack4 + 1 : b->host.max_depth + 1;
}

namespace rt {

int ctx_device(rt_ctx *c) { return c->device; }
void *ctx_stream_ptr(rt_ctx *c) { return (void *)c->stream; }
std::recursive_mutex &ctx_lock(rt_ctx *c) { return c->lock; }
DevBvh bvh_device(const rt_bvh *b) { return b->dev; }

// Device block of `bytes` owned by the context (the caller holds ctx_lock and
// has made the context's device current); null if the allocation fails.
void *ctx_film_scratch(rt_ctx *c, size_t bytes)
{
    if (!c->film_scratch && hipMalloc(&c->film_scratch, bytes) != hipSuccess) c->film_scratch = nullptr;
    return c->film_scratch;
}

}  // namespace rt

namespace rt {

static thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }

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

extern "C" {

const char *rt_last_error(void) { return rt::g_error.c_str(); }

uint32_t rt_version(void) { return (1u << 16) | 0u; }

rt_status rt_ctx_create(int device, rt_ctx **out)
{
    if (!out) { set_error("rt_ctx_create: out is null"); return RT_ERR_INVALID; }
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        set_error("rt_ctx_create: no HIP device visible");
        return RT_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= count) {
        set_error("rt_ctx_create: device index out of range");
        return RT_ERR_INVALID;
    }
    RT_HIP(hipSetDevice(device));
    rt_ctx *c = new (std::nothrow) rt_ctx();
    if (!c) { set_error("rt_ctx_create: host allocation failed"); return RT_ERR_OOM; }
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        set_error(std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
        return RT_ERR_HIP;
    }
    *out = c;
    return RT_OK;
}

rt_status rt_ctx_destroy(rt_ctx *ctx)
{
    if (!ctx) return RT_OK;
    rt::ctx_release(ctx);
    return RT_OK;
}

rt_status rt_ctx_synchronize(rt_ctx *ctx)
{
    if (!ctx) { set_error("rt_ctx_synchronize: null context"); return RT_ERR_INVALID; }
    RT_HIP(hipSetDevice(ctx->device));
    RT_HIP(hipStreamSynchronize(ctx->stream));
    return RT_OK;
}

void *rt_ctx_stream(rt_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

rt_status rt_malloc(rt_ctx *ctx, size_t bytes, void **out)
{
    if (!ctx || !out) { set_error("rt_malloc: null argument"); return RT_ERR_INVALID; }
    RT_HIP(hipSetDevice(ctx->device));
    RT_HIP(hipMalloc(out, bytes ? bytes : 1));
    return RT_OK;
}

rt_status rt_free(rt_ctx *ctx, void *ptr)
{
    if (!ctx) { set_error("rt_free: null context"); return RT_ERR_INVALID; }
    if (!ptr) return RT_OK;
    RT_HIP(hipSetDevice(ctx->device));
    RT_HIP(hipFree(ptr));
    return RT_OK;
}

rt_status rt_memcpy_h2d(rt_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx || (bytes && (!dst || !src))) { set_error("rt_memcpy_h2d: null argument"); return RT_ERR_INVALID; }
    RT_HIP(hipSetDevice(ctx->device));
    RT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    RT_HIP(hipStreamSynchronize(ctx->stream));
    return RT_OK;
}

rt_status rt_memcpy_d2h(rt_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx || (bytes && (!dst || !src))) { set_error("rt_memcpy_d2h: null argument"); return RT_ERR_INVALID; }
    RT_HIP(hipSetDevice(ctx->device));
    RT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    RT_HIP(hipStreamSynchronize(ctx->stream));
    return RT_OK;
}

/* ------------------------------------------------------------------ */
/* BVH                                                                  */
/* ------------------------------------------------------------------ */
static rt_status upload_bvh(rt_bvh *b)
{
    const bool wide = b->width == 4;
    const void *src = wide ? (const void *)b->host.nodes4.data() : (const void *)b->host.nodes.data();
    const size_t nn = wide ? b->host.nodes4.size() : b->host.nodes.size();
    const size_t nb = nn * 64;
    const size_t tb = b->host.tris.size() * sizeof(rt::Tri3);
    b->num_nodes2 = (uint32_t)b->host.nodes.size();
    b->num_nodes4 = (uint32_t)b->host.nodes4.size();
    RT_HIP(hipMalloc(&b->dev.nodes, nb));
    RT_HIP(hipMalloc(&b->dev.tris, tb));
    RT_HIP(hipMalloc(&b->ray_counters, (size_t)rt_bvh::kRayCounterRing * rt::kClaimWords * sizeof(uint32_t)));
    RT_HIP(hipMemcpyAsync(b->dev.nodes, src, nb, hipMemcpyHostToDevice, b->ctx->stream));
    RT_HIP(hipMemcpyAsync(b->dev.tris, b->host.tris.data(), tb, hipMemcpyHostToDevice, b->ctx->stream));
    RT_HIP(hipStreamSynchronize(b->ctx->stream));
    b->dev.num_nodes = (uint32_t)nn;
    b->dev.num_tris = b->host.num_tris;
    b->dev.stack_cap = stack_cap(b);
    b->dev.wide = wide ? 1u : 0u;
    return RT_OK;
}

rt_status rt_bvh_build(rt_ctx *ctx, const float *xyz, uint32_t nv, const uint32_t *idx, uint32_t nt,
                       const rt_bvh_opts *opts, rt_bvh **out)
{
    if (!out) { set_error("rt_bvh_build: out is null"); return RT_ERR_INVALID; }
    *out = nullptr;
    rt_bvh_opts o{4, 32, 1.2f, 4};
    if (opts) o = *opts;
    if (o.width == 0) o.width = 4;
    if (o.width != 2 && o.width != 4) { set_error("rt_bvh_build: width must be 2 or 4"); return RT_ERR_INVALID; }
    rt_bvh *b = new (std::nothrow) rt_bvh();
    if (!b) { set_error("rt_bvh_build: host allocation failed"); return RT_ERR_OOM; }
    b->ctx = ctx;
    b->width = o.width;
    std::string err;
    if (!rt::build_bvh2(xyz, nv, idx, nt, o, b->host, err)) {
        delete b;
        set_error(err);
        return RT_ERR_INVALID;
    }
    rt::build_bvh4(b->host);
    // The kernels hold a traversal stack of kMaxTraversalStack entries per
    // lane: refuse trees whose pending-entry bound exceeds it.
    if (stack_cap(b) > rt::kMaxTraversalStack) {
        delete b;
        set_error("rt_bvh_build: tree depth exceeds the traversal stack");
        return RT_ERR_UNSUPPORTED;
    }
    b->num_nodes2 = (uint32_t)b->host.nodes.size();
    b->num_nodes4 = (uint32_t)b->host.nodes4.size();
    if (!ctx) {  // host-only tree: info and export work, traces are refused
        *out = b;
        return RT_OK;
    }
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    hipError_t he = hipSetDevice(ctx->device);
    rt_status s = he == hipSuccess ? upload_bvh(b) : RT_ERR_HIP;
    if (he != hipSuccess) set_error(std::string("hipSetDevice: ") + hipGetErrorString(he));
    if (s != RT_OK) {
        if (b->dev.nodes) (void)hipFree(b->dev.nodes);
        if (b->dev.tris) (void)hipFree(b->dev.tris);
        delete b;
        return s;
    }
    (void)hipEventCreate(&b->ev0);
    (void)hipEventCreate(&b->ev1);
    rt::ctx_retain(ctx);
    *out = b;
    return RT_OK;
}

rt_status rt_bvh_build_device(rt_ctx *ctx, const float *xyz, uint32_t nv, const uint32_t *idx, uint32_t nt,
                              const rt_bvh_opts *opts, rt_bvh **out)
{
    if (!out) { set_error("rt_bvh_build_device: out is null"); return RT_ERR_INVALID; }
    *out = nullptr;
    if (!ctx) { set_error("rt_bvh_build_device: ctx is null"); return RT_ERR_INVALID; }
    if (nt && (!xyz || !idx || nv == 0)) { set_error("rt_bvh_build_device: null or empty triangle arrays"); return RT_ERR_INVALID; }
    if (nt > (1u << 28)) { set_error("rt_bvh_build_device: more than 2^28 triangles"); return RT_ERR_INVALID; }
    rt_bvh_opts o{4, 32, 1.2f, 4};
    if (opts) o = *opts;
    if (o.width == 0) o.width = 4;
    if (o.max_leaf == 0) o.max_leaf = 4;
    if (o.width != 2 && o.width != 4) { set_error("rt_bvh_build_device: width must be 2 or 4"); return RT_ERR_INVALID; }
    if (o.max_leaf > rt::kMaxLeafCount) { set_error("rt_bvh_build_device: max_leaf must be 1..8"); return RT_ERR_INVALID; }
    rt_bvh *b = new (std::nothrow) rt_bvh();
    if (!b) { set_error("rt_bvh_build_device: host allocation failed"); return RT_ERR_OOM; }
    b->ctx = ctx;
    b->width = o.width;
    b->on_device = true;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    auto drop = [&](rt_status st) {
        rt::free_bvh_device(b->gpu);
        if (b->ray_counters) (void)hipFree(b->ray_counters);
        delete b;
        return st;
    };
    hipError_t he = hipSetDevice(ctx->device);
    if (he != hipSuccess) {
        set_error(std::string("hipSetDevice: ") + hipGetErrorString(he));
        return drop(RT_ERR_HIP);
    }
    std::string err;
    he = rt::build_bvh_device(xyz, nv, idx, nt, o, ctx->stream, b->gpu, err);
    if (he != hipSuccess) {
        set_error(err);
        return drop(he == hipErrorInvalidValue ? RT_ERR_INVALID : he == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP);
    }
    const rt::DeviceBvhBuild &d = b->gpu;
    b->host.num_tris = d.num_tris;
    b->host.num_leaves = d.num_leaves;
    b->host.max_depth = d.max_depth;
    b->host.max_stack4 = d.max_stack4;
    b->host.depth4 = d.depth4;
    b->host.sah_cost = d.sah_cost;
    b->host.build_ms = d.build_ms;
    b->num_nodes2 = d.num_nodes2;
    b->num_nodes4 = d.num_nodes4;
    if (stack_cap(b) > rt::kMaxTraversalStack) {
        set_error("rt_bvh_build_device: tree depth exceeds the traversal stack");
        return drop(RT_ERR_UNSUPPORTED);
    }
    he = hipMalloc(&b->ray_counters, (size_t)rt_bvh::kRayCounterRing * rt::kClaimWords * sizeof(uint32_t));
    if (he != hipSuccess) {
        set_error(std::string("rt_bvh_build_device: ray counters: ") + hipGetErrorString(he));
        return drop(RT_ERR_OOM);
    }
    const bool wide = b->width == 4;
    b->dev.nodes = wide ? (void *)d.nodes4 : (void *)d.nodes2;
    b->dev.tris = d.tris;
    b->dev.num_nodes = wide ? d.num_nodes4 : d.num_nodes2;
    b->dev.num_tris = d.num_tris;
    b->dev.stack_cap = stack_cap(b);
    b->dev.wide = wide ? 1u : 0u;
    (void)hipEventCreate(&b->ev0);
    (void)hipEventCreate(&b->ev1);
    rt::ctx_retain(ctx);
    *out = b;
    return RT_OK;
}

rt_status rt_bvh_destroy(rt_bvh *b)
{
    if (!b) return RT_OK;
    if (!b->ctx) {
        delete b;
        return RT_OK;
    }
    rt_ctx *ctx = b->ctx;
    {
        std::lock_guard<std::recursive_mutex> g(ctx->lock);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        if (b->on_device) {
            rt::free_bvh_device(b->gpu);
        } else {
            if (b->dev.nodes) (void)hipFree(b->dev.nodes);
            if (b->dev.tris) (void)hipFree(b->dev.tris);
        }
        if (b->ray_counters) (void)hipFree(b->ray_counters);
        if (b->ev0) (void)hipEventDestroy(b->ev0);
        if (b->ev1) (void)hipEventDestroy(b->ev1);
        delete b;
    }
    rt::ctx_release(ctx);
    return RT_OK;
}

rt_status rt_bvh_get_info(const rt_bvh *b, rt_bvh_info *info)
{
    if (!b || !info) { set_error("rt_bvh_get_info: null argument"); return RT_ERR_INVALID; }
    info->num_tris = b->host.num_tris;
    info->num_nodes = b->num_nodes2;
    info->num_leaves = b->host.num_leaves;
    info->max_depth = b->host.max_depth;
    info->node_bytes = (uint64_t)b->num_nodes2 * sizeof(rt::Node2);
    info->tri_bytes = (uint64_t)b->host.num_tris * sizeof(rt::Tri3);
    info->build_ms = b->host.build_ms;
    info->sah_cost = b->host.sah_cost;
    info->width = b->width;
    info->num_nodes4 = b->num_nodes4;
    info->depth4 = b->host.depth4;
    info->stack_cap = stack_cap(b);
    return RT_OK;
}

rt_status rt_bvh_set_transparent(rt_bvh *b, const uint8_t *flags)
{
    if (!b || (!flags && b->host.num_tris)) { set_error("rt_bvh_set_transparent: null argument"); return RT_ERR_INVALID; }
    if (!b->ctx) {
        for (auto &t : b->host.tris)
            if (t.id >= 0) t.flags = flags[t.id] ? 1u : 0u;
        return RT_OK;
    }
    std::lock_guard<std::recursive_mutex> g(b->ctx->lock);
    RT_HIP(hipSetDevice(b->ctx->device));
    if (b->on_device) {
        // the triangles are only on the device: flags go up, a kernel applies
        // them by triangle id
        const uint32_t nt = b->host.num_tris;
        if (nt == 0) return RT_OK;
        uint8_t *df = nullptr;
        RT_HIP(hipMalloc(&df, nt));
        hipError_t e = hipMemcpyAsync(df, flags, nt, hipMemcpyHostToDevice, b->ctx->stream);
        if (e == hipSuccess) e = rt::set_transparent_device(b->gpu.tris, nt, df, b->ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->ctx->stream);
        (void)hipFree(df);
        if (e != hipSuccess) {
            set_error(std::string("rt_bvh_set_transparent: ") + hipGetErrorString(e));
            return RT_ERR_HIP;
        }
        return RT_OK;
    }
    for (auto &t : b->host.tris)
        if (t.id >= 0) t.flags = flags[t.id] ? 1u : 0u;
    const size_t tb = b->host.tris.size() * sizeof(rt::Tri3);
    RT_HIP(hipMemcpyAsync(b->dev.tris, b->host.tris.data(), tb, hipMemcpyHostToDevice, b->ctx->stream));
    RT_HIP(hipStreamSynchronize(b->ctx->stream));
    return RT_OK;
}

rt_status rt_bvh_export(const rt_bvh *b, float *nodes, float *tris)
{
    if (!b || !nodes || !tris) { set_error("rt_bvh_export: null argument"); return RT_ERR_INVALID; }
    if (b->on_device) {
        std::lock_guard<std::recursive_mutex> g(b->ctx->lock);
        RT_HIP(hipSetDevice(b->ctx->device));
        RT_HIP(hipMemcpyAsync(nodes, b->gpu.nodes2, (size_t)b->num_nodes2 * sizeof(rt::Node2), hipMemcpyDeviceToHost,
                              b->ctx->stream));
        RT_HIP(hipMemcpyAsync(tris, b->gpu.tris, ((size_t)b->host.num_tris + 1) * sizeof(rt::Tri3),
                              hipMemcpyDeviceToHost, b->ctx->stream));
        RT_HIP(hipStreamSynchronize(b->ctx->stream));
        return RT_OK;
    }
    std::memcpy(nodes, b->host.nodes.data(), b->host.nodes.size() * sizeof(rt::Node2));
    // num_tris records plus the sentinel
    std::memcpy(tris, b->host.tris.data(), b->host.tris.size() * sizeof(rt::Tri3));
    return RT_OK;
}

rt_status rt_bvh_export4(const rt_bvh *b, float *nodes4)
{
    if (!b || !nodes4) { set_error("rt_bvh_export4: null argument"); return RT_ERR_INVALID; }
    if (b->on_device) {
        if (!b->num_nodes4) return RT_OK;  // a two-wide device build has no wide nodes
        std::lock_guard<std::recursive_mutex> g(b->ctx->lock);
        RT_HIP(hipSetDevice(b->ctx->device));
        RT_HIP(hipMemcpyAsync(nodes4, b->gpu.nodes4, (size_t)b->num_nodes4 * sizeof(rt::Node4), hipMemcpyDeviceToHost,
                              b->ctx->stream));
        RT_HIP(hipStreamSynchronize(b->ctx->stream));
        return RT_OK;
    }
    std::memcpy(nodes4, b->host.nodes4.data(), b->host.nodes4.size() * sizeof(rt::Node4));
    return RT_OK;
}

rt_status rt_bvh_last_kernel_ms(const rt_bvh *b, float *ms)
{
    if (!b || !ms) { set_error("rt_bvh_last_kernel_ms: null argument"); return RT_ERR_INVALID; }
    if (!b->timed) { set_error("rt_bvh_last_kernel_ms: no traced call yet"); return RT_ERR_INVALID; }
    RT_HIP(hipEventSynchronize(b->ev1));
    RT_HIP(hipEventElapsedTime(ms, b->ev0, b->ev1));
    return RT_OK;
}

static bool rays_valid(const rt_rays *r)
{
    return r && r->ox && r->oy && r->oz && r->dx && r->dy && r->dz;
}

// Shared body of rt_trace_closest / rt_trace_any.
static rt_status trace_common(rt_bvh *b, const rt_rays *rays, uint32_t n, int32_t *prim, float *t,
                              uint8_t *occ, bool any, uint32_t flags, void *stream_)
{
    if (!b) { set_error("rt_trace: null bvh"); return RT_ERR_INVALID; }
    if (!b->ctx) { set_error("rt_trace: host-only BVH (built without a context)"); return RT_ERR_INVALID; }
    if (n == 0) return RT_OK;
    if (!rays_valid(rays) || (any ? !occ : !prim)) { set_error("rt_trace: null ray or output array"); return RT_ERR_INVALID; }
    if (flags & ~(uint32_t)(RT_FLAG_HOST_PTRS | RT_FLAG_SHADOW)) { set_error("rt_trace: unknown flag"); return RT_ERR_INVALID; }
    const bool opaque = (flags & RT_FLAG_SHADOW) != 0;  // any hit skips transparent triangles always
    std::lock_guard<std::recursive_mutex> g(b->ctx->lock);
    RT_HIP(hipSetDevice(b->ctx->device));
    hipStream_t stream = stream_ ? (hipStream_t)stream_ : b->ctx->stream;

    if (!(flags & RT_FLAG_HOST_PTRS)) {
        rt::RaysDev rd{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmax};
        uint32_t *rc = b->next_counter();
        RT_HIP(rt::clear_claims(rc, stream));
        RT_HIP(hipEventRecord(b->ev0, stream));
        RT_HIP(rt::launch_trace(b->dev, rd, n, prim, t, occ, any, stream, nullptr, rc, opaque));
        RT_HIP(hipEventRecord(b->ev1, stream));
        b->timed = true;
        return RT_OK;
    }

    // Host pointers: stage through device memory and block (the reference's
    // BVH::intersectRaysGpu contract, BVH.cu:562-584).
    const size_t fb = (size_t)n * sizeof(float);
    const int ncomp = rays->tmax ? 7 : 6;
    char *buf = nullptr;
    const size_t out_bytes = any ? (size_t)n : (size_t)n * (sizeof(int32_t) + (t ? sizeof(float) : 0));
    RT_HIP(hipMalloc(&buf, ncomp * fb + out_bytes));
    const float *src[7] = {rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmax};
    float *dst[7];
    for (int k = 0; k < ncomp; ++k) {
        dst[k] = (float *)(buf + k * fb);
        hipError_t e = hipMemcpyAsync(dst[k], src[k], fb, hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) { (void)hipFree(buf); set_error(std::string("hipMemcpyAsync: ") + hipGetErrorString(e)); return RT_ERR_HIP; }
    }
    char *out = buf + ncomp * fb;
    rt::RaysDev rd{dst[0], dst[1], dst[2], dst[3], dst[4], dst[5], ncomp == 7 ? dst[6] : nullptr};
    int32_t *dprim = any ? nullptr : (int32_t *)out;
    float *dt = (!any && t) ? (float *)(out + (size_t)n * sizeof(int32_t)) : nullptr;
    uint8_t *docc = any ? (uint8_t *)out : nullptr;
    uint32_t *rc = b->next_counter();
    hipError_t e = rt::clear_claims(rc, stream);
    (void)hipEventRecord(b->ev0, stream);
    if (e == hipSuccess) e = rt::launch_trace(b->dev, rd, n, dprim, dt, docc, any, stream, nullptr, rc, opaque);
    (void)hipEventRecord(b->ev1, stream);
    b->timed = true;
    if (e == hipSuccess) {
        if (any) e = hipMemcpyAsync(occ, docc, n, hipMemcpyDeviceToHost, stream);
        else {
            e = hipMemcpyAsync(prim, dprim, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess && t) e = hipMemcpyAsync(t, dt, fb, hipMemcpyDeviceToHost, stream);
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    (void)hipFree(buf);
    if (e != hipSuccess) { set_error(std::string("rt_trace (host pointers): ") + hipGetErrorString(e)); return RT_ERR_HIP; }
    return RT_OK;
}

rt_status rt_trace_stats(rt_bvh *b, const rt_rays *rays, uint32_t n, int any, uint64_t *node_visits,
                         uint64_t *tri_tests)
{
    if (!b || !b->ctx || !rays_valid(rays) || !node_visits || !tri_tests) {
        set_error("rt_trace_stats: null argument or host-only BVH");
        return RT_ERR_INVALID;
    }
    std::lock_guard<std::recursive_mutex> g(b->ctx->lock);
    RT_HIP(hipSetDevice(b->ctx->device));
    hipStream_t stream = b->ctx->stream;
    char *buf = nullptr;
    const size_t outb = any ? (size_t)n : (size_t)n * 8;
    const size_t claims = 256, head = claims + rt::kClaimWords * sizeof(uint32_t);
    RT_HIP(hipMalloc(&buf, head + outb));
    hipError_t e = hipMemsetAsync(buf, 0, head, stream);
    rt::RaysDev rd{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmax};
    char *out = buf + head;
    if (e == hipSuccess)
        e = rt::launch_trace(b->dev, rd, n, any ? nullptr : (int32_t *)out, any ? nullptr : (float *)(out + 4 * (size_t)n),
                             any ? (uint8_t *)out : nullptr, any != 0, stream, (unsigned long long *)buf,
                             (uint32_t *)(buf + claims));
    unsigned long long c[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(c, buf, 16, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    (void)hipFree(buf);
    if (e != hipSuccess) { set_error(std::string("rt_trace_stats: ") + hipGetErrorString(e)); return RT_ERR_HIP; }
    *node_visits = c[0];
    *tri_tests = c[1];
    return RT_OK;
}

rt_status rt_trace_closest(rt_bvh *b, const rt_rays *rays, uint32_t n, int32_t *prim, float *t,
                           uint32_t flags, void *stream)
{
    return trace_common(b, rays, n, prim, t, nullptr, false, flags, stream);
}

rt_status rt_trace_any(rt_bvh *b, const rt_rays *rays, uint32_t n, uint8_t *occluded, uint32_t flags,
                       void *stream)
{
    return trace_common(b, rays, n, nullptr, nullptr, occluded, true, flags, stream);
}

}  // extern "C"
