// batch.hip -- many scenes in one launch sequence (rt_batch_*): a device
// table of the scenes, one wave stream over the concatenation of their
// sample index spaces, graph capture of whole frames and the autotuner of
// wave size and grid.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "rt_internal.h"
#include "rt_render.h"
#include "rt_waves.h"

struct rt_ctx;
namespace rt {
void ctx_retain(rt_ctx *c);
void ctx_release(rt_ctx *c);
int ctx_device(rt_ctx *c);
void *ctx_stream_ptr(rt_ctx *c);
std::recursive_mutex &ctx_lock(rt_ctx *c);
bool scene_committed(const rt_scene *s);
rt_ctx *scene_ctx(const rt_scene *s);
const SceneDev &scene_device(const rt_scene *s);
uint64_t scene_stream_base(const rt_scene *s);
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

// What a captured iteration block depends on (the sample offset, the
// chunk's first path and count and the scene table are read from device
// memory at replay; the wave buffers are fixed while the key holds).
struct GraphKey {
    uint32_t max_depth = 0, rr_depth = 0, capacity = 0, grid = 0;
    void *arena = nullptr, *res = nullptr, *hint = nullptr;
    bool operator==(const GraphKey &o) const
    {
        return max_depth == o.max_depth && rr_depth == o.rr_depth && capacity == o.capacity && grid == o.grid &&
               arena == o.arena && res == o.res && hint == o.hint;
    }
};

constexpr uint32_t kBlockIters = 4;  // iterations per captured block (even)

}  // namespace

struct rt_batch {
    rt_ctx *ctx = nullptr;
    std::vector<rt_scene *> scenes;
    std::vector<rt::SceneDev> host_table;   // as uploaded (seeds included)
    std::vector<uint32_t> npix;
    uint32_t stack_cap = 0, max_lights = 1;
    uint64_t paths_per_sample = 0;           // sum of pixels
    rt::SceneDev *d_table = nullptr;
    void *d_nodes = nullptr, *d_tris = nullptr;  // the scenes' trees, pooled
    uint32_t *d_root = nullptr;              // root node of scene s in d_nodes
    uint64_t *d_bases = nullptr;             // per call: n + 1 path-index bases
    uint32_t *d_hint = nullptr;              // scene at each 2^16-path block start
    size_t hint_cap = 0;
    uint32_t *d_spp_offset = nullptr;
    uint64_t *d_chunk = nullptr;             // {first path, count} of the running chunk
    std::vector<uint64_t> bases;             // host copy for the current spp
    uint32_t bases_spp = 0;
    rt::WaveBuffers wb;
    uint32_t tuned_capacity = 0, tuned_grid = 0;
    // graph of the last captured frame
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    GraphKey key;
    rt_render_stats stats{};
    rt::KernelProfile prof;  // event pairs of profiled frames (RT_FLAG_PROFILE)
};

namespace {

// Copies one scene's tree into the batch pool: node references move by
// nbase, leaf references' first triangle by tbase.  Two-wide nodes keep
// their children in words 12-13; four-wide nodes in words 4-7.
__global__ void k_pool_nodes(const int32_t *__restrict__ src, int32_t *__restrict__ dst, uint32_t n, int32_t nbase,
                             uint32_t tbase, uint32_t wide)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t *a = src + 16 * (size_t)i;
    int32_t *d = dst + 16 * (size_t)i;
    for (int k = 0; k < 16; ++k) d[k] = a[k];
    auto rebase = [&](int32_t c) {
        if (c >= 0) return c + nbase;
        const uint32_t enc = ~(uint32_t)c;
        return (int32_t)~((((enc >> rt::kLeafCountBits) + tbase) << rt::kLeafCountBits) |
                          (enc & (rt::kMaxLeafCount - 1u)));
    };
    if (wide) {
        // every slot holds a reference (unused ones point at the sentinel leaf)
        for (int k = 0; k < 4; ++k) d[4 + k] = rebase(a[4 + k]);
    } else {
        d[12] = rebase(a[12]);
        d[13] = rebase(a[13]);
    }
}

// Pools the scenes' trees (nodes and triangle records, each scene's block
// after the previous one's) and records each scene's root.
rt_status pool_trees(rt_batch *b, hipStream_t s)
{
    const size_t n = b->host_table.size();
    std::vector<uint32_t> root(n), tbase(n);
    uint64_t nn = 0, nt = 0;
    for (size_t i = 0; i < n; ++i) {
        const rt::DevBvh &v = b->host_table[i].bvh;
        root[i] = (uint32_t)nn;
        tbase[i] = (uint32_t)nt;
        nn += v.num_nodes;
        nt += (uint64_t)v.num_tris + 1;  // records plus the sentinel
    }
    if (nn >= (1ull << 31) || nt >= (1ull << (31 - rt::kLeafCountBits))) {
        set_error("rt_batch_create: the scenes' trees are too large to pool");
        return RT_ERR_UNSUPPORTED;
    }
    RT_HIP(hipMalloc(&b->d_nodes, std::max<uint64_t>(1, nn) * 64));
    RT_HIP(hipMalloc(&b->d_tris, std::max<uint64_t>(1, nt) * sizeof(rt::Tri3)));
    RT_HIP(hipMalloc(&b->d_root, n * sizeof(uint32_t)));
    for (size_t i = 0; i < n; ++i) {
        const rt::DevBvh &v = b->host_table[i].bvh;
        if (v.num_nodes) {
            hipLaunchKernelGGL(k_pool_nodes, dim3((v.num_nodes + 255) / 256), dim3(256), 0, s, (const int32_t *)v.nodes,
                               (int32_t *)b->d_nodes + 16 * (size_t)root[i], v.num_nodes, (int32_t)root[i], tbase[i],
                               v.wide);
            RT_HIP(hipGetLastError());
        }
        RT_HIP(hipMemcpyAsync((char *)b->d_tris + (size_t)tbase[i] * sizeof(rt::Tri3), v.tris,
                              ((size_t)v.num_tris + 1) * sizeof(rt::Tri3), hipMemcpyDeviceToDevice, s));
    }
    RT_HIP(hipMemcpyAsync(b->d_root, root.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    RT_HIP(hipStreamSynchronize(s));
    return RT_OK;
}

void drop_graph(rt_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->key = GraphKey{};
}

// Uploads the scene table when the seeds change (seeds live in the table).
rt_status upload_table(rt_batch *b, const uint32_t *seeds, hipStream_t s)
{
    bool changed = false;
    for (size_t i = 0; i < b->scenes.size(); ++i) {
        const uint32_t sd = seeds ? seeds[i] : 1u;
        if (b->host_table[i].seed != sd) {
            b->host_table[i].seed = sd;
            changed = true;
        }
    }
    if (changed || !b->d_table) {
        if (!b->d_table) RT_HIP(hipMalloc(&b->d_table, b->host_table.size() * sizeof(rt::SceneDev)));
        RT_HIP(hipMemcpyAsync(b->d_table, b->host_table.data(), b->host_table.size() * sizeof(rt::SceneDev),
                              hipMemcpyHostToDevice, s));
        RT_HIP(hipStreamSynchronize(s));
    }
    return RT_OK;
}

rt_status upload_bases(rt_batch *b, uint32_t spp, hipStream_t s)
{
    if (b->bases_spp == spp && !b->bases.empty()) return RT_OK;
    const size_t n = b->scenes.size();
    b->bases.assign(n + 1, 0);
    for (size_t i = 0; i < n; ++i) b->bases[i + 1] = b->bases[i] + (uint64_t)b->npix[i] * spp;
    // scene of each 2^16-path block start
    const uint64_t total = b->bases[n];
    std::vector<uint32_t> hint((size_t)((total >> rt::kHintShift) + 1), 0);
    uint32_t sid = 0;
    for (size_t k = 0; k < hint.size(); ++k) {
        const uint64_t g = (uint64_t)k << rt::kHintShift;
        while (sid + 1 < n && b->bases[sid + 1] <= g) ++sid;
        hint[k] = sid;
    }
    if (hint.size() > b->hint_cap) {
        if (b->d_hint) RT_HIP(hipFree(b->d_hint));
        b->d_hint = nullptr;
        b->hint_cap = 0;
        RT_HIP(hipMalloc(&b->d_hint, hint.size() * sizeof(uint32_t)));
        b->hint_cap = hint.size();
    }
    RT_HIP(hipMemcpyAsync(b->d_bases, b->bases.data(), (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    RT_HIP(hipMemcpyAsync(b->d_hint, hint.data(), hint.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    RT_HIP(hipStreamSynchronize(s));
    b->bases_spp = spp;
    return RT_OK;
}

rt::BatchLaunch batch_launch(rt_batch *b, const std::vector<rt::FilmDev> &films)
{
    rt::BatchLaunch bl{};
    bl.dev = rt::BatchDev{b->d_table, b->d_bases, (uint32_t)b->scenes.size(), b->d_nodes, b->d_tris, b->d_root,
                          b->d_hint};
    bl.bases = b->bases.data();
    bl.npix = b->npix.data();
    bl.films = films.data();
    bl.stack_cap = b->stack_cap;
    return bl;
}

rt::WaveParams batch_params(rt_batch *b, const rt_batch_opts &o, uint32_t grid)
{
    rt::WaveParams wp{};
    wp.spp_offset = o.spp_offset;
    wp.max_depth = o.max_depth;
    wp.rr_depth = o.rr_depth;
    wp.cache_depth = rt::kNoCache;
    wp.grid = grid;
    wp.spp_offset_dev = b->d_spp_offset;
    wp.chunk_dev = b->d_chunk;
    return wp;
}

// One frame: every chunk of the batch's path index space through the
// streaming loop.  With `graph`, the iteration block is replayed from the
// captured graph (captured on first use for the current key).
hipError_t run_frame(rt_batch *b, const rt_batch_opts &o, uint32_t P, uint32_t grid,
                     const std::vector<rt::FilmDev> &films, bool graph, hipStream_t s,
                     rt::KernelProfile *prof = nullptr)
{
    const rt::BatchLaunch bl = batch_launch(b, films);
    rt::ChunkLaunch cl{};
    cl.scene = &b->host_table[0];
    cl.paths = b->wb.paths;
    cl.shadow = b->wb.shadow;
    cl.queues = b->wb.queues;
    cl.capacity = P;
    cl.res = b->wb.res;
    cl.batch = &bl;
    cl.prof = graph ? nullptr : prof;
    rt::WaveParams wp = batch_params(b, o, grid);
    if (graph) {
        GraphKey k;
        k.max_depth = o.max_depth; k.rr_depth = o.rr_depth; k.capacity = P; k.grid = grid;
        k.arena = b->wb.arena; k.res = b->wb.res; k.hint = b->d_hint;
        if (!b->exec || !(k == b->key)) {
            drop_graph(b);
            hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) return e;
            e = rt::chunk_iterate(cl, wp, kBlockIters, s);
            hipGraph_t gr = nullptr;
            hipError_t e2 = hipStreamEndCapture(s, &gr);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (gr) (void)hipGraphDestroy(gr);
                return e != hipSuccess ? e : e2;
            }
            b->graph = gr;
            e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
            if (e != hipSuccess) return e;
            b->key = k;
        }
    }
    const uint64_t total = b->bases.back();
    const uint64_t chunk = o.chunk_paths ? o.chunk_paths : rt::kChunkPaths;
    const uint64_t R = std::min<uint64_t>(std::min<uint64_t>(total, chunk), b->wb.res_cap);
    uint32_t *hc = b->wb.host_counters;
    for (uint64_t first = 0; first < total; first += R) {
        wp.first_path = first;
        wp.count = (uint32_t)std::min<uint64_t>(R, total - first);
        hipError_t e = rt::set_u64(b->d_chunk, first, s);
        if (e == hipSuccess) e = rt::set_u64(b->d_chunk + 1, wp.count, s);
        if (e == hipSuccess) e = rt::chunk_begin(cl, wp, s);
        if (e != hipSuccess) return e;
        const uint64_t rounds = ((uint64_t)wp.count + P - 1) / P + 1;
        const uint64_t max_iters = rounds * o.max_depth + kBlockIters;
        bool done = false;
        for (uint64_t it = 0; it < max_iters && !done; it += kBlockIters) {
            e = graph ? hipGraphLaunch(b->exec, s) : rt::chunk_iterate(cl, wp, kBlockIters, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(hc, b->wb.queues.counters, rt::kCounterWords * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) return e;
            done = hc[rt::kCNext * rt::kCtrStride] >= wp.count && hc[rt::kCLive * rt::kCtrStride] == 0;
        }
        if (!done) return hipErrorUnknown;  // cannot happen: the bound above holds
        e = rt::chunk_film(cl, films[0], wp, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

rt_status check_opts(const char *fn, const rt_batch *b, const rt_batch_opts *o)
{
    const std::string name(fn);
    if (!b || !o) { set_error(name + ": null argument"); return RT_ERR_INVALID; }
    if (o->spp == 0) { set_error(name + ": spp must be >= 1"); return RT_ERR_INVALID; }
    if (o->max_depth == 0 || o->max_depth > 0xffffu) { set_error(name + ": max_depth must be 1..65535"); return RT_ERR_INVALID; }
    if (o->path_capacity > (1u << 26)) { set_error(name + ": path_capacity above 2^26"); return RT_ERR_INVALID; }
    if (o->grid > 65536) { set_error(name + ": grid above 65536"); return RT_ERR_INVALID; }
    return RT_OK;
}

uint32_t wave_size(const rt_batch *b, const rt_batch_opts &o, uint64_t total)
{
    uint32_t P = o.path_capacity ? o.path_capacity : (b->tuned_capacity ? b->tuned_capacity : (1u << 24));
    if ((uint64_t)P > total) P = (uint32_t)std::max<uint64_t>(1, total);
    return rt::stripe_capacity(P);
}

}  // namespace

extern "C" {

rt_status rt_batch_create(rt_ctx *ctx, rt_scene *const *scenes, uint32_t n, rt_batch **out)
{
    if (!ctx || !scenes || !out || n == 0) { set_error("rt_batch_create: null argument or no scenes"); return RT_ERR_INVALID; }
    if (n > 4096) { set_error("rt_batch_create: more than 4096 scenes"); return RT_ERR_INVALID; }
    for (uint32_t i = 0; i < n; ++i) {
        if (!scenes[i]) { set_error("rt_batch_create: null scene"); return RT_ERR_INVALID; }
        if (!rt::scene_committed(scenes[i])) { set_error("rt_batch_create: scene " + std::to_string(i) + " not committed"); return RT_ERR_INVALID; }
        if (rt::scene_ctx(scenes[i]) != ctx) { set_error("rt_batch_create: scene " + std::to_string(i) + " belongs to another context"); return RT_ERR_INVALID; }
    }
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(ctx));
    RT_HIP(hipSetDevice(rt::ctx_device(ctx)));
    rt_batch *b = new (std::nothrow) rt_batch();
    if (!b) { set_error("rt_batch_create: host allocation failed"); return RT_ERR_OOM; }
    b->ctx = ctx;
    b->scenes.assign(scenes, scenes + n);
    for (uint32_t i = 0; i < n; ++i) {
        rt::SceneDev d = rt::scene_device(scenes[i]);
        d.seed = 0xffffffffu;  // forces the first upload
        d.stream_base = rt::scene_stream_base(scenes[i]);
        b->host_table.push_back(d);
        b->npix.push_back(d.cam.width * d.cam.height);
        b->stack_cap = std::max(b->stack_cap, d.bvh.stack_cap);
        if (d.bvh.wide != b->host_table[0].bvh.wide) {
            delete b;
            set_error("rt_batch_create: scenes mix two-wide and four-wide trees");
            return RT_ERR_INVALID;
        }
        b->max_lights = std::max(b->max_lights, d.nlights);
        b->paths_per_sample += (uint64_t)d.cam.width * d.cam.height;
    }
    if (b->stack_cap > rt::kMaxTraversalStack) {
        delete b;
        set_error("rt_batch_create: BVH too deep");
        return RT_ERR_UNSUPPORTED;
    }
    auto release = [&]() {
        if (b->d_bases) (void)hipFree(b->d_bases);
        if (b->d_spp_offset) (void)hipFree(b->d_spp_offset);
        if (b->d_chunk) (void)hipFree(b->d_chunk);
        if (b->d_nodes) (void)hipFree(b->d_nodes);
        if (b->d_tris) (void)hipFree(b->d_tris);
        if (b->d_root) (void)hipFree(b->d_root);
        delete b;
    };
    hipError_t e = hipMalloc(&b->d_bases, (n + 1) * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(&b->d_spp_offset, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&b->d_chunk, 2 * sizeof(uint64_t));
    if (e != hipSuccess) {
        release();
        set_error(std::string("rt_batch_create: ") + hipGetErrorString(e));
        return RT_ERR_OOM;
    }
    const rt_status pst = pool_trees(b, (hipStream_t)rt::ctx_stream_ptr(ctx));
    if (pst != RT_OK) {
        release();
        return pst;
    }
    rt::ctx_retain(ctx);
    *out = b;
    return RT_OK;
}

rt_status rt_batch_destroy(rt_batch *b)
{
    if (!b) return RT_OK;
    rt_ctx *ctx = b->ctx;
    {
        std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(ctx));
        (void)hipSetDevice(rt::ctx_device(ctx));
        (void)hipStreamSynchronize((hipStream_t)rt::ctx_stream_ptr(ctx));
        drop_graph(b);
        rt::free_wave_buffers(b->wb);
        if (b->d_table) (void)hipFree(b->d_table);
        if (b->d_bases) (void)hipFree(b->d_bases);
        if (b->d_spp_offset) (void)hipFree(b->d_spp_offset);
        if (b->d_chunk) (void)hipFree(b->d_chunk);
        if (b->d_nodes) (void)hipFree(b->d_nodes);
        if (b->d_tris) (void)hipFree(b->d_tris);
        if (b->d_root) (void)hipFree(b->d_root);
        if (b->d_hint) (void)hipFree(b->d_hint);
        delete b;
    }
    rt::ctx_release(ctx);
    return RT_OK;
}

rt_status rt_batch_render(rt_batch *b, const rt_batch_opts *opts, const uint32_t *seeds, double *const *accum,
                          uint32_t *const *nvalid, uint32_t flags, void *stream_)
{
    rt_status st = check_opts("rt_batch_render", b, opts);
    if (st != RT_OK) return st;
    if (!accum || !nvalid) { set_error("rt_batch_render: null film arrays"); return RT_ERR_INVALID; }
    if (flags & RT_FLAG_HOST_PTRS) { set_error("rt_batch_render: films must be device pointers"); return RT_ERR_UNSUPPORTED; }
    if (flags & ~(uint32_t)(RT_FLAG_GRAPH | RT_FLAG_PROFILE | RT_FLAG_COUNT)) { set_error("rt_batch_render: unknown flag"); return RT_ERR_INVALID; }
    const bool measure = (flags & (RT_FLAG_PROFILE | RT_FLAG_COUNT)) != 0;
    if (measure && (flags & RT_FLAG_GRAPH)) {
        set_error("rt_batch_render: RT_FLAG_PROFILE / RT_FLAG_COUNT frames run eagerly (not with RT_FLAG_GRAPH)");
        return RT_ERR_INVALID;
    }
    const size_t n = b->scenes.size();
    std::vector<rt::FilmDev> films(n);
    for (size_t i = 0; i < n; ++i) {
        if (!accum[i] || !nvalid[i]) { set_error("rt_batch_render: null film of scene " + std::to_string(i)); return RT_ERR_INVALID; }
        films[i] = rt::FilmDev{accum[i], nvalid[i]};
    }
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(b->ctx));
    RT_HIP(hipSetDevice(rt::ctx_device(b->ctx)));
    hipStream_t s = stream_ ? (hipStream_t)stream_ : (hipStream_t)rt::ctx_stream_ptr(b->ctx);
    const uint64_t total = b->paths_per_sample * opts->spp;
    const uint32_t P = wave_size(b, *opts, total);
    const uint32_t grid = opts->grid ? opts->grid : (b->tuned_grid ? b->tuned_grid : 2048);
    st = rt::ensure_wave_buffers(b->wb, P, b->max_lights, "rt_batch_render");
    const uint64_t chunk = opts->chunk_paths ? opts->chunk_paths : rt::kChunkPaths;
    if (st == RT_OK) st = rt::ensure_result_buffer(b->wb, std::min<uint64_t>(total, chunk), "rt_batch_render");
    if (st != RT_OK) return st;
    st = upload_table(b, seeds, s);
    if (st != RT_OK) return st;
    st = upload_bases(b, opts->spp, s);
    if (st != RT_OK) return st;
    auto t0 = std::chrono::steady_clock::now();
    RT_HIP(hipMemsetAsync(b->wb.queues.stats, 0, rt::kNumStats * 8, s));
    RT_HIP(rt::set_u32(b->d_spp_offset, opts->spp_offset, s));
    b->prof.events = (flags & RT_FLAG_PROFILE) != 0;
    b->prof.count = (flags & RT_FLAG_COUNT) != 0;
    b->prof.used = 0;
    {
        hipError_t e = run_frame(b, *opts, P, grid, films, (flags & RT_FLAG_GRAPH) != 0, s,
                                 measure ? &b->prof : nullptr);
        if (e != hipSuccess) {
            set_error(std::string("rt_batch_render: ") + hipGetErrorString(e));
            return RT_ERR_HIP;
        }
    }
    unsigned long long counts[rt::kNumStats];
    RT_HIP(hipMemcpyAsync(counts, b->wb.queues.stats, sizeof counts, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    b->stats = rt_render_stats{};
    b->stats.paths = total;
    b->stats.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rt::fill_stats(b->stats, counts, b->prof, flags, "rt_batch_render");
}

rt_status rt_batch_autotune(rt_batch *b, const rt_batch_opts *opts, rt_tune_entry *best, rt_tune_entry *entries,
                            uint32_t max_entries, uint32_t *n_entries)
{
    rt_status st = check_opts("rt_batch_autotune", b, opts);
    if (st != RT_OK) return st;
    if (n_entries) *n_entries = 0;
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(b->ctx));
    RT_HIP(hipSetDevice(rt::ctx_device(b->ctx)));
    hipStream_t s = (hipStream_t)rt::ctx_stream_ptr(b->ctx);
    const size_t n = b->scenes.size();
    // scratch films
    std::vector<rt::FilmDev> films(n);
    size_t bytes = 0;
    for (size_t i = 0; i < n; ++i) bytes += (size_t)b->npix[i] * (24 + 4) + 512;
    char *scratch = nullptr;
    RT_HIP(hipMalloc(&scratch, bytes));
    size_t off = 0;
    for (size_t i = 0; i < n; ++i) {
        films[i].accum = (double *)(scratch + off);
        off += ((size_t)b->npix[i] * 24 + 255) & ~(size_t)255;
        films[i].nvalid = (uint32_t *)(scratch + off);
        off += ((size_t)b->npix[i] * 4 + 255) & ~(size_t)255;
    }
    rt_batch_opts o = *opts;
    // trial frames of opts->spp samples per pixel (1..4): with more than one,
    // waves larger than one sample's paths are among the candidates
    o.spp = std::max(1u, std::min(opts->spp, 4u));
    o.spp_offset = 0x40000000u;  // far from rendering samples
    const uint64_t total = b->paths_per_sample * o.spp;
    std::vector<uint32_t> caps;
    for (uint32_t P = 1u << 19; P <= (1u << 24); P <<= 1) {
        const uint32_t c = rt::stripe_capacity((uint32_t)std::min<uint64_t>(P, total));
        if (caps.empty() || caps.back() != c) caps.push_back(c);
    }
    const uint32_t grids[4] = {1024, 2048, 4096, 8192};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    rt_tune_entry top{0, 0, 1e30f};
    uint32_t count = 0;
    rt_status result = RT_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) result = RT_ERR_HIP;
    if (result == RT_OK) result = upload_table(b, nullptr, s);
    if (result == RT_OK) result = upload_bases(b, o.spp, s);
    for (uint32_t P : caps) {
        if (result != RT_OK) break;
        result = rt::ensure_wave_buffers(b->wb, P, b->max_lights, "rt_batch_autotune");
        if (result == RT_OK)
            result = rt::ensure_result_buffer(b->wb, std::min<uint64_t>(total, rt::kChunkPaths), "rt_batch_autotune");
        for (uint32_t gsz : grids) {
            if (result != RT_OK) break;
            float ms_best = 1e30f;
            for (int rep = 0; rep < 2 && result == RT_OK; ++rep) {
                if (rt::set_u32(b->d_spp_offset, o.spp_offset + (uint32_t)rep, s) != hipSuccess ||
                    hipEventRecord(e0, s) != hipSuccess || run_frame(b, o, P, gsz, films, false, s) != hipSuccess ||
                    hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                    set_error("rt_batch_autotune: trial launch failed");
                    result = RT_ERR_HIP;
                    break;
                }
                float ms = 0;
                (void)hipEventElapsedTime(&ms, e0, e1);
                ms_best = std::min(ms_best, ms);
            }
            if (result != RT_OK) break;
            if (entries && count < max_entries) entries[count] = rt_tune_entry{P, gsz, ms_best};
            ++count;
            if (ms_best < top.ms) top = rt_tune_entry{P, gsz, ms_best};
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    (void)hipFree(scratch);
    b->bases_spp = 0;  // bases were set up for the trial frames
    if (result != RT_OK) return result;
    b->tuned_capacity = top.path_capacity;
    b->tuned_grid = top.grid;
    drop_graph(b);
    if (best) *best = top;
    if (n_entries) *n_entries = std::min(count, max_entries);
    return RT_OK;
}

rt_status rt_batch_get_stats(const rt_batch *b, rt_render_stats *st)
{
    if (!b || !st) { set_error("rt_batch_get_stats: null argument"); return RT_ERR_INVALID; }
    *st = b->stats;
    return RT_OK;
}

}  // extern "C"
