// scene.hip -- host side of the scene API (rt_scene_*) and rt_render:
// gathers meshes, spheres, materials and lights, composes the camera and
// sphere transforms exactly as the reference does, builds the BVH, owns the
// wavefront buffers and launches waves.
#include <hip/hip_runtime.h>

#include <limits>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "rt_cache.h"
#include "rt_internal.h"
#include "rt_photon_render.h"
#include "rt_photon_trace.h"
#include "rt_render.h"
#include "rt_waves.h"

struct rt_ctx;
namespace rt {
void ctx_retain(rt_ctx *c);
void ctx_release(rt_ctx *c);
int ctx_device(rt_ctx *c);
void *ctx_stream_ptr(rt_ctx *c);
std::recursive_mutex &ctx_lock(rt_ctx *c);
DevBvh bvh_device(const rt_bvh *b);
const CacheDev &cache_device(const rt_cache *c);
rt_ctx *cache_ctx(const rt_cache *c);
uint32_t cache_batch(const rt_cache *c);
rt_status cache_train_harvest(rt_cache *c, const float *in, const float *target, uint32_t total, uint32_t steps,
                              hipStream_t s, uint32_t *steps_done, float *mean_loss, bool *reset);
bool cache_take_fallback(rt_cache *c);
rt_status cache_train_rows_dev(rt_cache *c, const float *in, const float *target, uint32_t n, const uint32_t *n_dev,
                               hipStream_t s, float **loss_dev);
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

// ---- reference matrix algebra (transformations/Transform.cu,
//      cuda_libraries/Matrix4x4.cu) in FP32, row-major ----
void mat_mul(const float *a, const float *b, float *out)
{
    float r[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float s = 0;
            for (int k = 0; k < 4; ++k) s += a[i * 4 + k] * b[k * 4 + j];
            r[i * 4 + j] = s;
        }
    std::memcpy(out, r, sizeof r);
}

// Cofactor inverse with the reference's term order (Matrix4x4.cu:303-440).
void mat_inverse(const float *m, float *out)
{
    float v[16];
    v[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    v[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    v[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    v[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    v[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    v[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    v[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    v[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    v[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    v[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    v[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    v[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    v[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    v[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    v[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    v[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    float det = m[0] * v[0] + m[1] * v[4] + m[2] * v[8] + m[3] * v[12];
    if (det == 0) {
        std::memset(out, 0, 16 * sizeof(float));
        return;
    }
    det = (float)(1.0 / det);
    for (int i = 0; i < 16; ++i) out[i] = v[i] * det;
}

struct F3 {
    float x, y, z;
};
F3 f3(float x, float y, float z) { return F3{x, y, z}; }
F3 fsub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
F3 fcross(F3 a, F3 b)
{
    return f3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
F3 fnrm(F3 a)
{
    float l = std::sqrt((a.x * a.x) + (a.y * a.y) + (a.z * a.z));
    float inv = (float)1 / l;
    return f3(a.x * inv, a.y * inv, a.z * inv);
}

// ProjectiveCamera(inverse(lookAt(p, look, up)), translate(cts), window)
// (FileImporting.cu:404-436, ProjectiveCamera.cu:26-57, Transform.cu:611-640).
void compose_camera(const rt_camera &c, rt::CameraDev &out)
{
    F3 pos = f3(c.pos[0], c.pos[1], c.pos[2]);
    F3 look = f3(c.lookat[0], c.lookat[1], c.lookat[2]);
    F3 up = f3(c.up[0], c.up[1], c.up[2]);
    F3 dir = fnrm(fsub(look, pos));
    F3 right = fnrm(fcross(fnrm(up), dir));
    F3 nup = fcross(dir, right);
    float c2w[16] = {right.x, nup.x, dir.x, pos.x, right.y, nup.y, dir.y, pos.y,
                     right.z, nup.z, dir.z, pos.z, 0, 0, 0, 1};
    float la[16];
    mat_inverse(c2w, la);          // the Transform lookAt returns (its mInv is dropped)
    mat_inverse(la, out.c2w);      // the camera's cameraToWorld = inverse(lookAt)
    float vfs = c.frame_y;
    vfs /= 2;
    float hfs = (vfs / (float)c.height) * (float)c.width;
    float pminx = -hfs, pminy = -vfs, pmaxx = hfs, pmaxy = vfs;
    float T[16] = {1, 0, 0, -pminx, 0, 1, 0, -pmaxy, 0, 0, 1, 0, 0, 0, 0, 1};
    float S1[16] = {1 / (pmaxx - pminx), 0, 0, 0, 0, 1 / (pminy - pmaxy), 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    float S2[16] = {(float)c.width, 0, 0, 0, 0, (float)c.height, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    float tmp[16], s2r[16], r2s[16];
    mat_mul(S1, T, tmp);
    mat_mul(S2, tmp, s2r);
    mat_inverse(s2r, r2s);
    float cts[16] = {1, 0, 0, c.cts[0], 0, 1, 0, c.cts[1], 0, 0, 1, c.cts[2], 0, 0, 0, 1};
    float icts[16];
    mat_inverse(cts, icts);
    mat_mul(icts, r2s, out.r2c);
    out.width = c.width;
    out.height = c.height;
}

// Sphere(center, radius): translate(c) * rotateZ(90) (Sphere.cu:233-245,
// Transform.cu:475-488, 536-547; radians() at math.cu:66-68).
void compose_sphere(const float c[3], float r, rt::SphereDev &s)
{
    const float PIf = 3.14159265358979323846f;
    float ang = (float)((90.0f * PIf) / 180.0);
    float st = std::sin(ang), ct = std::cos(ang);
    float T[16] = {1, 0, 0, c[0], 0, 1, 0, c[1], 0, 0, 1, c[2], 0, 0, 0, 1};
    float R[16] = {ct, -st, 0, 0, st, ct, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    mat_mul(T, R, s.o2w);
    mat_inverse(s.o2w, s.w2o);
    s.center[0] = c[0]; s.center[1] = c[1]; s.center[2] = c[2];
    s.radius = r;
}

// Plane(corner, axisX, axisY, normal) (Plane.cu:27-48): the axis lengths
// truncated to integers, the normal cross(axisX, axisY) turned towards the
// given one and normalised (normalize = v * (1/length), Vector.cu:454-457),
// objectToWorld = baseChange(corner, axisX, axisY, normal)
// (Transform.cu:551-562), worldToObject its cofactor inverse.  Returns
// false for an axis shorter than 1 (a zero-size plane).
bool compose_plane(const float c[3], const float ax[3], const float ay[3], const float n[3], rt::PlaneDev &p)
{
    const float lx = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
    const float ly = std::sqrt(ay[0] * ay[0] + ay[1] * ay[1] + ay[2] * ay[2]);
    const int sx = (int)lx, sy = (int)ly;
    if (sx < 1 || sy < 1) return false;
    const float cr[3] = {(ax[1] * ay[2]) - (ax[2] * ay[1]), (ax[2] * ay[0]) - (ax[0] * ay[2]),
                         (ax[0] * ay[1]) - (ax[1] * ay[0])};
    const float dn = (cr[0] * n[0]) + (cr[1] * n[1]) + (cr[2] * n[2]);
    float nn[3];
    for (int k = 0; k < 3; ++k) nn[k] = dn > 0 ? cr[k] : -cr[k];
    const float ix = 1.0f / lx, iy = 1.0f / ly;
    const float X[3] = {ax[0] * ix * (float)sx, ax[1] * ix * (float)sx, ax[2] * ix * (float)sx};
    const float Y[3] = {ay[0] * iy * (float)sy, ay[1] * iy * (float)sy, ay[2] * iy * (float)sy};
    const float ln = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
    if (!(ln > 0)) return false;
    const float in = 1.0f / ln;
    const float N[3] = {nn[0] * in, nn[1] * in, nn[2] * in};
    const float m[16] = {X[0], Y[0], N[0], c[0], X[1], Y[1], N[1], c[1], X[2], Y[2], N[2], c[2], 0, 0, 0, 1};
    for (int k = 0; k < 16; ++k) p.o2w[k] = m[k];
    mat_inverse(p.o2w, p.w2o);
    return true;
}

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
};

}  // namespace

struct rt_scene {
    rt_ctx *ctx = nullptr;
    // host description
    std::vector<float> xyz;         // 9 floats per triangle, id order
    std::vector<uint32_t> tri_mat;
    std::vector<rt_material> mats;
    std::vector<rt::SphereDev> spheres;
    std::vector<rt::PlaneDev> planes;
    std::vector<float> tri_uv;       // 6 floats per triangle once a mesh brings texture coordinates
    bool has_uv = false;
    std::vector<float> tri_nrm;      // 9 floats per triangle (corner normals) once a mesh brings normals;
    bool has_nrm = false;            // a NaN first component: the triangle has none
    std::vector<rt::TexDev> textures;
    std::vector<uint32_t> texels;    // 0x00BBGGRR
    std::vector<rt_point_light> lights;
    rt_camera cam{};
    bool has_cam = false;
    bool committed = false;
    // device state
    rt_bvh *bvh = nullptr;
    float *d_tri_xyz = nullptr;
    uint32_t *d_tri_mat = nullptr;
    rt_material *d_mat = nullptr;
    rt_point_light *d_pl = nullptr;
    rt::PlaneDev *d_plane = nullptr;
    float *d_tri_uv = nullptr;
    float *d_tri_nrm = nullptr;
    rt::TexDev *d_tex = nullptr;
    uint32_t *d_texels = nullptr;
    rt::SceneDev dev{};
    // wavefront buffers (grown on demand)
    rt::WaveBuffers wb;
    rt_render_stats stats{};
    rt::KernelProfile prof;  // event pairs of profiled frames (RT_FLAG_PROFILE)
    uint64_t stream_base = 0;
    // radiance-cache inference scratch (one row per path slot), on demand
    void *cache_arena = nullptr;
    rt::CacheWork cwork{};
    // photon-mapping query rows (one per path slot), on demand
    void *photon_arena = nullptr;
    uint32_t photon_rows = 0;
    // the BVH of the next commit
    rt_bvh_opts bvh_opts{4, 32, 1.2f, 4};
    bool bvh_on_device = false;
};

namespace rt {

void free_wave_buffers(WaveBuffers &w)
{
    if (w.arena) (void)hipFree(w.arena);
    if (w.res) (void)hipFree(w.res);
    if (w.host_counters) (void)hipHostFree(w.host_counters);
    w = WaveBuffers{};
}

rt_status ensure_result_buffer(WaveBuffers &w, uint64_t R, const char *who)
{
    if (!w.host_counters) {
        hipError_t e = hipHostMalloc((void **)&w.host_counters, kCounterWords * sizeof(uint32_t), hipHostMallocDefault);
        if (e != hipSuccess) {
            w.host_counters = nullptr;
            set_error(std::string(who) + ": pinned counters: " + hipGetErrorString(e));
            return RT_ERR_HIP;
        }
    }
    if (w.res_cap >= R && w.res) return RT_OK;
    if (w.res) (void)hipFree(w.res);
    w.res = nullptr;
    w.res_cap = 0;
    hipError_t e = hipMalloc(&w.res, std::max<uint64_t>(1, R) * 12);
    if (e != hipSuccess) {
        w.res = nullptr;
        set_error(std::string(who) + ": result buffer: " + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP;
    }
    w.res_cap = R;
    return RT_OK;
}

static void free_arena(rt_scene *s)
{
    free_wave_buffers(s->wb);
    if (s->cache_arena) (void)hipFree(s->cache_arena);
    s->cache_arena = nullptr;
    s->cwork = rt::CacheWork{};
    if (s->photon_arena) (void)hipFree(s->photon_arena);
    s->photon_arena = nullptr;
    s->photon_rows = 0;
}

// Photon query rows for `rows` path slots: point, normal, view (3 floats
// each), material, shape, slot, the estimate (3 doubles) and the
// throughput on arrival (3 floats, hybrid).
static rt_status ensure_photon_rows(rt_scene *s, uint32_t rows, PhotonLaunch &ph)
{
    auto bytes = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t v3 = bytes((size_t)rows * 12), mat = bytes((size_t)rows * sizeof(rt_material)),
                 u = bytes((size_t)rows * 4), rgb = bytes((size_t)rows * 24);
    // point, normal, view, throughput on arrival: four rows of 3 floats
    if (!s->photon_arena || s->photon_rows < rows) {
        if (s->photon_arena) (void)hipFree(s->photon_arena);
        s->photon_arena = nullptr;
        s->photon_rows = 0;
        hipError_t e = hipMalloc(&s->photon_arena, 4 * v3 + mat + 2 * u + rgb);
        if (e != hipSuccess) {
            set_error(std::string("rt_render_photon: photon query rows: ") + hipGetErrorString(e));
            return RT_ERR_OOM;
        }
        s->photon_rows = rows;
    }
    char *b = (char *)s->photon_arena;
    ph.qp = (float *)b;
    ph.qn = (float *)(b + v3);
    ph.qv = (float *)(b + 2 * v3);
    ph.qm = (rt_material *)(b + 3 * v3);
    ph.qs = (int32_t *)(b + 3 * v3 + mat);
    ph.qslot = (uint32_t *)(b + 3 * v3 + mat + u);
    ph.qrgb = (double *)(b + 3 * v3 + mat + 2 * u);
    ph.qbeta = (float *)(b + 3 * v3 + mat + 2 * u + rgb);
    return RT_OK;
}

// Inference activations for `rows` cache queries (forward only).
static rt_status ensure_cache_work(rt_scene *s, uint32_t rows)
{
    if (s->cache_arena && s->cwork.rows >= rows) return RT_OK;
    if (s->cache_arena) (void)hipFree(s->cache_arena);
    s->cache_arena = nullptr;
    s->cwork = CacheWork{};
    const size_t act = ((size_t)rows * kCacheWidth * 2 + 255) & ~(size_t)255;
    const size_t y = (size_t)rows * kCacheOut * 4;
    char *base = nullptr;
    hipError_t e = hipMalloc(&base, act * kCacheLayers + y);
    if (e != hipSuccess) {
        set_error(std::string("rt_render_cached: cache scratch: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP;
    }
    s->cache_arena = base;
    s->cwork.rows = rows;
    for (uint32_t l = 0; l < kCacheLayers; ++l) s->cwork.act[l] = (uint16_t *)(base + l * act);
    s->cwork.y = (float *)(base + kCacheLayers * act);
    return RT_OK;
}

// One allocation for all wavefront buffers, carved in 256-byte aligned
// pieces: paths (SoA), shadow queue (capacity * lights entries), queues.
rt_status ensure_wave_buffers(WaveBuffers &w, uint32_t P, uint32_t NL, const char *who)
{
    NL = std::max<uint32_t>(1, NL);
    if (w.capacity >= P && w.cap_lights >= NL) return RT_OK;
    if (w.arena) (void)hipFree(w.arena);
    w.arena = nullptr;
    w.capacity = 0;
    const size_t SQ = (size_t)P * NL;
    size_t total = 0;
    auto take = [&](size_t bytes) {
        size_t off = total;
        total += (bytes + 255) & ~(size_t)255;
        return off;
    };
    struct Piece { void **dst; size_t off; };
    std::vector<Piece> pieces;
    auto add = [&](void **dst, size_t bytes) { pieces.push_back({dst, take(bytes)}); };
    PathsDev &p = w.paths;
    add((void **)&p.rec, (size_t)P * kPathQuads * sizeof(float4));
    add((void **)&p.nee_c, (size_t)P * NL * 12);
    add((void **)&p.cin, (size_t)P * 48);
    add((void **)&p.crgb, (size_t)P * 12);
    add((void **)&p.cslot, (size_t)P * 4);
    add((void **)&p.cqin, (size_t)P * 48);
    add((void **)&p.tin, (size_t)P * 48);
    add((void **)&p.tLb, (size_t)P * 12);
    add((void **)&p.tB, (size_t)P * 12);
    add((void **)&p.ttgt, (size_t)P * 12);
    ShadowDev &sh = w.shadow;
    float **sf[] = {&sh.ox, &sh.oy, &sh.oz, &sh.dx, &sh.dy, &sh.dz, &sh.tmax};
    for (float **f : sf) add((void **)f, SQ * 4);
    add((void **)&sh.slot, SQ * 4);
    add((void **)&sh.root, SQ * 4);
    QueuesDev &q = w.queues;
    add((void **)&q.active[0], (size_t)P * 4);
    add((void **)&q.active[1], (size_t)P * 4);
    for (uint32_t k = 0; k < kNumQ; ++k) add((void **)&q.q[k], (size_t)P * 4);
    add((void **)&q.term, (size_t)P * 4);
    add((void **)&q.sctr, kStripedWords * 4);
    add((void **)&q.counters, kCounterWords * 4);
    add((void **)&q.stats, kNumStats * 8);
    char *base = nullptr;
    hipError_t e = hipMalloc(&base, total);
    if (e != hipSuccess) {
        set_error(std::string(who) + ": wavefront buffers (" + std::to_string(total >> 20) +
                  " MiB): " + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP;
    }
    for (auto &pc : pieces) *pc.dst = base + pc.off;
    p.nee_stride = NL;
    w.arena = base;
    w.capacity = P;
    w.cap_lights = NL;
    return RT_OK;
}

// rt_render_stats of a frame from its stats words and, when profiled, the
// elapsed times of its event pairs (the stream has been synchronised).
rt_status fill_stats(rt_render_stats &st, const unsigned long long *counts, KernelProfile &prof, uint32_t flags,
                     const char *who)
{
    st.extension_rays = counts[kStatExt];
    st.shadow_rays = counts[kStatShadow];
    if (flags & RT_FLAG_COUNT) {
        st.ext_node_visits = counts[kStatExtNodes];
        st.ext_tri_tests = counts[kStatExtTris];
        st.shadow_node_visits = counts[kStatShNodes];
        st.shadow_tri_tests = counts[kStatShTris];
        st.shadow_occluded = counts[kStatOccluded];
    }
    if (flags & RT_FLAG_PROFILE) {
        for (size_t i = 0; i < prof.used; ++i) {
            float ms = 0;
            hipError_t e = hipEventElapsedTime(&ms, prof.ev[2 * i], prof.ev[2 * i + 1]);
            if (e != hipSuccess) {
                set_error(std::string(who) + ": kernel timing: " + hipGetErrorString(e));
                return RT_ERR_HIP;
            }
            if (prof.kind[i] == 0) { st.extend_ms += ms; ++st.extend_launches; }
            else { st.connect_ms += ms; ++st.connect_launches; }
        }
    }
    prof.used = 0;
    return RT_OK;
}

static rt_status ensure_buffers(rt_scene *s, uint32_t P, const char *who)
{
    if (s->wb.capacity >= P && s->wb.cap_lights >= std::max<uint32_t>(1, s->dev.nlights)) return RT_OK;
    free_arena(s);  // the cache scratch is sized to the wave capacity too
    return ensure_wave_buffers(s->wb, P, s->dev.nlights, who);
}

}  // namespace rt

extern "C" {

rt_status rt_scene_create(rt_ctx *ctx, rt_scene **out)
{
    if (!ctx || !out) { set_error("rt_scene_create: null argument"); return RT_ERR_INVALID; }
    rt_scene *s = new (std::nothrow) rt_scene();
    if (!s) { set_error("rt_scene_create: host allocation failed"); return RT_ERR_OOM; }
    s->ctx = ctx;
    rt::ctx_retain(ctx);
    *out = s;
    return RT_OK;
}

static void free_device(rt_scene *s)
{
    if (s->bvh) rt_bvh_destroy(s->bvh);
    s->bvh = nullptr;
    if (s->d_tri_xyz) (void)hipFree(s->d_tri_xyz);
    if (s->d_tri_mat) (void)hipFree(s->d_tri_mat);
    if (s->d_mat) (void)hipFree(s->d_mat);
    if (s->d_pl) (void)hipFree(s->d_pl);
    if (s->d_plane) (void)hipFree(s->d_plane);
    if (s->d_tri_uv) (void)hipFree(s->d_tri_uv);
    if (s->d_tri_nrm) (void)hipFree(s->d_tri_nrm);
    s->d_tri_nrm = nullptr;
    if (s->d_tex) (void)hipFree(s->d_tex);
    if (s->d_texels) (void)hipFree(s->d_texels);
    s->d_tri_uv = nullptr; s->d_tex = nullptr; s->d_texels = nullptr;
    s->d_tri_xyz = nullptr; s->d_tri_mat = nullptr; s->d_mat = nullptr; s->d_pl = nullptr; s->d_plane = nullptr;
    rt::free_arena(s);
    s->committed = false;
}

rt_status rt_scene_destroy(rt_scene *s)
{
    if (!s) return RT_OK;
    rt_ctx *ctx = s->ctx;
    {
        std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(ctx));
        (void)hipSetDevice(rt::ctx_device(ctx));
        (void)hipStreamSynchronize((hipStream_t)rt::ctx_stream_ptr(ctx));
        free_device(s);
        delete s;
    }
    rt::ctx_release(ctx);
    return RT_OK;
}

rt_status rt_scene_add_material(rt_scene *s, const rt_material *m, uint32_t *id)
{
    if (!s || !m) { set_error("rt_scene_add_material: null argument"); return RT_ERR_INVALID; }
    if (m->type > RT_MAT_PLAIN) { set_error("rt_scene_add_material: unknown material type"); return RT_ERR_INVALID; }
    if (m->type == RT_MAT_DIELECTRIC && !(m->ior > 0)) { set_error("rt_scene_add_material: ior must be > 0"); return RT_ERR_INVALID; }
    if ((m->type == RT_MAT_TROWBRIDGE || m->type == RT_MAT_TROWBRIDGE_BASIC) && !(m->roughness >= 0)) {
        set_error("rt_scene_add_material: roughness must be >= 0");
        return RT_ERR_INVALID;
    }
    s->mats.push_back(*m);
    if (id) *id = (uint32_t)s->mats.size() - 1;
    s->committed = false;
    return RT_OK;
}

rt_status rt_scene_add_mesh(rt_scene *s, const float *xyz, uint32_t nv, const uint32_t *idx, uint32_t nt,
                            uint32_t material, uint32_t *first_id)
{
    return rt_scene_add_mesh_uv(s, xyz, nv, idx, nt, nullptr, material, first_id);
}

rt_status rt_scene_add_mesh_uv(rt_scene *s, const float *xyz, uint32_t nv, const uint32_t *idx, uint32_t nt,
                               const float *uv, uint32_t material, uint32_t *first_id)
{
    return rt_scene_add_mesh_normals(s, xyz, nv, idx, nt, uv, nullptr, material, first_id);
}

rt_status rt_scene_add_mesh_normals(rt_scene *s, const float *xyz, uint32_t nv, const uint32_t *idx, uint32_t nt,
                                    const float *uv, const float *normals, uint32_t material, uint32_t *first_id)
{
    if (!s || (nt && (!xyz || !idx))) { set_error("rt_scene_add_mesh: null argument"); return RT_ERR_INVALID; }
    if (material >= s->mats.size()) { set_error("rt_scene_add_mesh: unknown material id"); return RT_ERR_INVALID; }
    const size_t base = s->tri_mat.size();
    if (base + nt >= (1u << 27)) { set_error("rt_scene_add_mesh: more than 2^27 triangles"); return RT_ERR_INVALID; }
    for (uint32_t i = 0; i < 3 * nt; ++i)
        if (idx[i] >= nv) { set_error("rt_scene_add_mesh: vertex index out of range"); return RT_ERR_INVALID; }
    s->xyz.resize((base + nt) * 9);
    for (uint32_t i = 0; i < nt; ++i)
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k) s->xyz[(base + i) * 9 + v * 3 + k] = xyz[3 * (size_t)idx[3 * i + v] + k];
    s->tri_mat.resize(base + nt, material);
    // texture coordinates are kept once any mesh brings its own; until then
    // every triangle has the default (0,0), (1,0), (1,1)
    static const float kDefaultUv[6] = {0, 0, 1, 0, 1, 1};
    if (uv && !s->has_uv) {
        s->tri_uv.resize(base * 6);
        for (size_t i = 0; i < base; ++i) std::copy(kDefaultUv, kDefaultUv + 6, s->tri_uv.begin() + 6 * i);
        s->has_uv = true;
    }
    if (s->has_uv) {
        s->tri_uv.resize((base + nt) * 6);
        for (uint32_t i = 0; i < nt; ++i)
            std::copy(uv ? uv + 6 * (size_t)i : kDefaultUv, (uv ? uv + 6 * (size_t)i : kDefaultUv) + 6,
                      s->tri_uv.begin() + 6 * (base + i));
    }
    // corner normals likewise, once any mesh brings them (TriangleMeshData::
    // normals, shapes/Triangle.cu:379-418); NaN marks a triangle without
    const float nan = std::numeric_limits<float>::quiet_NaN();
    if (normals && !s->has_nrm) {
        s->tri_nrm.assign(base * 9, nan);
        s->has_nrm = true;
    }
    if (s->has_nrm) {
        s->tri_nrm.resize((base + nt) * 9, nan);
        if (normals) std::copy(normals, normals + 9 * (size_t)nt, s->tri_nrm.begin() + 9 * base);
    }
    if (first_id) *first_id = (uint32_t)base;
    s->committed = false;
    return RT_OK;
}

rt_status rt_scene_add_texture(rt_scene *s, const uint8_t *rgb, uint32_t width, uint32_t height, uint32_t *id)
{
    if (!s || !rgb) { set_error("rt_scene_add_texture: null argument"); return RT_ERR_INVALID; }
    if (width == 0 || height == 0 || (uint64_t)width * height > (1ull << 28)) {
        set_error("rt_scene_add_texture: size must be 1..2^28 texels");
        return RT_ERR_INVALID;
    }
    rt::TexDev t{};
    t.offset = (uint32_t)s->texels.size();
    if ((uint64_t)t.offset + (uint64_t)width * height > (1ull << 30)) {
        set_error("rt_scene_add_texture: more than 2^30 texels in the scene");
        return RT_ERR_UNSUPPORTED;
    }
    t.width = width;
    t.height = height;
    const size_t n = (size_t)width * height;
    s->texels.resize(t.offset + n);
    for (size_t i = 0; i < n; ++i)
        s->texels[t.offset + i] = (uint32_t)rgb[3 * i] | ((uint32_t)rgb[3 * i + 1] << 8) | ((uint32_t)rgb[3 * i + 2] << 16);
    s->textures.push_back(t);
    if (id) *id = (uint32_t)s->textures.size() - 1;
    s->committed = false;
    return RT_OK;
}

rt_status rt_scene_add_sphere(rt_scene *s, const float center[3], float radius, uint32_t material, int is_light)
{
    if (!s || !center) { set_error("rt_scene_add_sphere: null argument"); return RT_ERR_INVALID; }
    if (material >= s->mats.size()) { set_error("rt_scene_add_sphere: unknown material id"); return RT_ERR_INVALID; }
    if (!(radius > 0)) { set_error("rt_scene_add_sphere: radius must be > 0"); return RT_ERR_INVALID; }
    if (s->spheres.size() >= (size_t)rt::kMaxSpheres) { set_error("rt_scene_add_sphere: too many spheres"); return RT_ERR_UNSUPPORTED; }
    if (is_light && s->mats[material].type != RT_MAT_LIGHT) {
        set_error("rt_scene_add_sphere: a sphere light needs a RT_MAT_LIGHT material");
        return RT_ERR_INVALID;
    }
    rt::SphereDev d{};
    compose_sphere(center, radius, d);
    d.material = material;
    d.is_light = is_light ? 1u : 0u;
    d.transparent = s->mats[material].type == RT_MAT_DIELECTRIC ? 1u : 0u;
    s->spheres.push_back(d);
    s->committed = false;
    return RT_OK;
}

rt_status rt_scene_add_plane(rt_scene *s, const float corner[3], const float axis_x[3], const float axis_y[3],
                             const float normal[3], uint32_t material)
{
    if (!s || !corner || !axis_x || !axis_y || !normal) { set_error("rt_scene_add_plane: null argument"); return RT_ERR_INVALID; }
    if (material >= s->mats.size()) { set_error("rt_scene_add_plane: unknown material id"); return RT_ERR_INVALID; }
    if (s->planes.size() >= (size_t)rt::kMaxPlanes) { set_error("rt_scene_add_plane: too many planes"); return RT_ERR_UNSUPPORTED; }
    rt::PlaneDev d{};
    if (!compose_plane(corner, axis_x, axis_y, normal, d)) {
        set_error("rt_scene_add_plane: an axis is shorter than 1 or the axes are parallel (Plane truncates axis "
                  "lengths to integers)");
        return RT_ERR_INVALID;
    }
    d.material = material;
    d.transparent = s->mats[material].type == RT_MAT_DIELECTRIC ? 1u : 0u;
    s->planes.push_back(d);
    s->committed = false;
    return RT_OK;
}

rt_status rt_scene_add_point_light(rt_scene *s, const rt_point_light *l)
{
    if (!s || !l) { set_error("rt_scene_add_point_light: null argument"); return RT_ERR_INVALID; }
    if (s->lights.size() >= (size_t)rt::kMaxPointLights) { set_error("rt_scene_add_point_light: too many lights"); return RT_ERR_UNSUPPORTED; }
    s->lights.push_back(*l);
    s->committed = false;
    return RT_OK;
}

rt_status rt_scene_set_camera(rt_scene *s, const rt_camera *c)
{
    if (!s || !c) { set_error("rt_scene_set_camera: null argument"); return RT_ERR_INVALID; }
    if (c->width == 0 || c->height == 0 || !(c->frame_y > 0)) { set_error("rt_scene_set_camera: bad resolution or frame"); return RT_ERR_INVALID; }
    s->cam = *c;
    s->has_cam = true;
    compose_camera(*c, s->dev.cam);
    s->committed = false;
    return RT_OK;
}

rt_status rt_scene_set_bvh(rt_scene *s, const rt_bvh_opts *opts, uint32_t on_device)
{
    if (!s) { set_error("rt_scene_set_bvh: null scene"); return RT_ERR_INVALID; }
    rt_bvh_opts o{4, 32, 1.2f, 4};
    if (opts) o = *opts;
    if (o.width == 0) o.width = 4;
    if (o.max_leaf == 0) o.max_leaf = 4;
    if ((o.width != 2 && o.width != 4) || o.max_leaf > 8) {
        set_error("rt_scene_set_bvh: width must be 2 or 4 and max_leaf 1..8");
        return RT_ERR_INVALID;
    }
    s->bvh_opts = o;
    s->bvh_on_device = on_device != 0;
    s->committed = false;
    return RT_OK;
}

rt_status rt_scene_commit(rt_scene *s)
{
    if (!s) { set_error("rt_scene_commit: null scene"); return RT_ERR_INVALID; }
    if (!s->has_cam) { set_error("rt_scene_commit: no camera"); return RT_ERR_INVALID; }
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(s->ctx));
    RT_HIP(hipSetDevice(rt::ctx_device(s->ctx)));
    hipStream_t stream = (hipStream_t)rt::ctx_stream_ptr(s->ctx);
    free_device(s);
    const uint32_t nt = (uint32_t)s->tri_mat.size();
    std::vector<uint32_t> idx(3 * (size_t)nt);
    for (uint32_t i = 0; i < 3 * nt; ++i) idx[i] = i;
    rt_status st = RT_OK;
    if (s->bvh_on_device) {
        // the triangles go up once for the build; the scene's own copies
        // follow below
        float *dx = nullptr;
        uint32_t *di = nullptr;
        if (nt) {
            RT_HIP(hipMalloc(&dx, s->xyz.size() * 4));
            hipError_t e = hipMalloc(&di, idx.size() * 4);
            if (e == hipSuccess) e = hipMemcpyAsync(dx, s->xyz.data(), s->xyz.size() * 4, hipMemcpyHostToDevice, stream);
            if (e == hipSuccess) e = hipMemcpyAsync(di, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, stream);
            if (e != hipSuccess) {
                (void)hipFree(dx);
                if (di) (void)hipFree(di);
                set_error(std::string("rt_scene_commit: triangle upload: ") + hipGetErrorString(e));
                return RT_ERR_HIP;
            }
        }
        st = rt_bvh_build_device(s->ctx, dx, 3 * nt, di, nt, &s->bvh_opts, &s->bvh);
        if (dx) (void)hipFree(dx);
        if (di) (void)hipFree(di);
    } else {
        st = rt_bvh_build(s->ctx, s->xyz.data(), 3 * nt, idx.data(), nt, &s->bvh_opts, &s->bvh);
    }
    if (st != RT_OK) return st;
    std::vector<uint8_t> transparent(nt);
    for (uint32_t i = 0; i < nt; ++i) transparent[i] = s->mats[s->tri_mat[i]].type == RT_MAT_DIELECTRIC;
    if (nt) {
        st = rt_bvh_set_transparent(s->bvh, transparent.data());
        if (st != RT_OK) return st;
    }
    const size_t nmat = std::max<size_t>(1, s->mats.size()), npl = std::max<size_t>(1, s->lights.size());
    RT_HIP(hipMalloc(&s->d_tri_xyz, std::max<size_t>(1, s->xyz.size()) * 4));
    RT_HIP(hipMalloc(&s->d_tri_mat, std::max<size_t>(1, nt) * 4));
    RT_HIP(hipMalloc(&s->d_mat, nmat * sizeof(rt_material)));
    RT_HIP(hipMalloc(&s->d_pl, npl * sizeof(rt_point_light)));
    RT_HIP(hipMalloc(&s->d_plane, std::max<size_t>(1, s->planes.size()) * sizeof(rt::PlaneDev)));
    if (!s->planes.empty())
        RT_HIP(hipMemcpyAsync(s->d_plane, s->planes.data(), s->planes.size() * sizeof(rt::PlaneDev),
                              hipMemcpyHostToDevice, stream));
    for (const auto &m : s->mats)
        if (m.texture > s->textures.size()) {
            set_error("rt_scene_commit: a material names a texture the scene does not have");
            return RT_ERR_INVALID;
        }
    if (s->has_nrm && !s->tri_nrm.empty()) {
        // corner normals of every triangle (NaN first component: none)
        RT_HIP(hipMalloc(&s->d_tri_nrm, s->tri_nrm.size() * sizeof(float)));
        RT_HIP(hipMemcpyAsync(s->d_tri_nrm, s->tri_nrm.data(), s->tri_nrm.size() * sizeof(float),
                              hipMemcpyHostToDevice, stream));
    }
    if (!s->textures.empty()) {
        // texture coordinates are only read when the scene has textures
        RT_HIP(hipMalloc(&s->d_tex, s->textures.size() * sizeof(rt::TexDev)));
        RT_HIP(hipMalloc(&s->d_texels, s->texels.size() * sizeof(uint32_t)));
        RT_HIP(hipMemcpyAsync(s->d_tex, s->textures.data(), s->textures.size() * sizeof(rt::TexDev),
                              hipMemcpyHostToDevice, stream));
        RT_HIP(hipMemcpyAsync(s->d_texels, s->texels.data(), s->texels.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice, stream));
        if (s->has_uv && !s->tri_uv.empty()) {
            RT_HIP(hipMalloc(&s->d_tri_uv, s->tri_uv.size() * sizeof(float)));
            RT_HIP(hipMemcpyAsync(s->d_tri_uv, s->tri_uv.data(), s->tri_uv.size() * sizeof(float),
                                  hipMemcpyHostToDevice, stream));
        }
    }
    if (!s->xyz.empty()) RT_HIP(hipMemcpyAsync(s->d_tri_xyz, s->xyz.data(), s->xyz.size() * 4, hipMemcpyHostToDevice, stream));
    if (nt) RT_HIP(hipMemcpyAsync(s->d_tri_mat, s->tri_mat.data(), nt * 4, hipMemcpyHostToDevice, stream));
    if (!s->mats.empty()) RT_HIP(hipMemcpyAsync(s->d_mat, s->mats.data(), s->mats.size() * sizeof(rt_material), hipMemcpyHostToDevice, stream));
    if (!s->lights.empty()) RT_HIP(hipMemcpyAsync(s->d_pl, s->lights.data(), s->lights.size() * sizeof(rt_point_light), hipMemcpyHostToDevice, stream));
    RT_HIP(hipStreamSynchronize(stream));

    rt::SceneDev &d = s->dev;
    rt_bvh_info info;
    rt_bvh_get_info(s->bvh, &info);
    d.bvh = rt::bvh_device(s->bvh);
    d.tri_xyz = s->d_tri_xyz;
    d.tri_mat = s->d_tri_mat;
    d.mat = s->d_mat;
    d.pl = s->d_pl;
    d.nsph = (uint32_t)s->spheres.size();
    for (uint32_t j = 0; j < d.nsph; ++j) d.sph[j] = s->spheres[j];
    d.plane = s->d_plane;
    d.nplane = (uint32_t)s->planes.size();
    d.tex = s->d_tex;
    d.texels = s->d_texels;
    d.tri_uv = s->d_tri_uv;
    d.tri_nrm = s->d_tri_nrm;
    d.ntex = (uint32_t)s->textures.size();
    d.npl = (uint32_t)s->lights.size();
    uint32_t na = 0;
    for (uint32_t j = 0; j < d.nsph; ++j)
        if (s->spheres[j].is_light) d.area_light[na++] = j;
    d.nlights = d.npl + na;
    d.nmat = (uint32_t)s->mats.size();
    if (d.nlights > 32) { set_error("rt_scene_commit: more than 32 light slots"); return RT_ERR_UNSUPPORTED; }
    if (d.bvh.stack_cap > rt::kMaxTraversalStack) { set_error("rt_scene_commit: BVH too deep"); return RT_ERR_UNSUPPORTED; }
    s->committed = true;
    return RT_OK;
}

rt_bvh *rt_scene_bvh(rt_scene *s) { return s ? s->bvh : nullptr; }

rt_status rt_scene_trace_photons(rt_scene *s, const rt_photon_opts *opts, uint32_t capacity, float *location,
                                 float *direction, float *normal, double *radiance, int32_t *shape, uint32_t *key,
                                 uint32_t *count, void *stream_)
{
    const char *who = "rt_scene_trace_photons";
    if (count) *count = 0;
    if (!s || !opts || !count) { set_error(std::string(who) + ": null argument"); return RT_ERR_INVALID; }
    if (!s->committed) { set_error(std::string(who) + ": scene not committed"); return RT_ERR_INVALID; }
    if (opts->photons > (1u << 27)) { set_error(std::string(who) + ": at most 2^27 photon paths"); return RT_ERR_INVALID; }
    if (opts->max_depth == 0 || opts->max_depth > 32) { set_error(std::string(who) + ": max_depth must be 1..32"); return RT_ERR_INVALID; }
    if (capacity && (!location || !direction || !normal || !radiance || !shape)) {
        set_error(std::string(who) + ": null output array");
        return RT_ERR_INVALID;
    }
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(s->ctx));
    RT_HIP(hipSetDevice(rt::ctx_device(s->ctx)));
    hipStream_t stream = stream_ ? (hipStream_t)stream_ : (hipStream_t)rt::ctx_stream_ptr(s->ctx);
    return rt::trace_photons(s->dev, s->lights.data(), (uint32_t)s->lights.size(), *opts, capacity, location,
                             direction, normal, radiance, shape, key, count, stream, who);
}

}  // extern "C"

// accessors for batches (batch.hip)
namespace rt {
bool scene_committed(const rt_scene *s) { return s->committed; }
rt_ctx *scene_ctx(const rt_scene *s) { return s->ctx; }
const SceneDev &scene_device(const rt_scene *s) { return s->dev; }
uint64_t scene_stream_base(const rt_scene *s) { return s->stream_base; }
}  // namespace rt

extern "C" {

rt_status rt_scene_camera_ray(const rt_scene *s, uint32_t px, uint32_t py, float u0, float u1, float o[3], float d[3])
{
    if (!s || !s->has_cam || !o || !d) { set_error("rt_scene_camera_ray: null argument or no camera"); return RT_ERR_INVALID; }
    // same arithmetic as camera_ray() in rt_shade.h, on the host
    const rt::CameraDev &c = s->dev.cam;
    float fx = (float)px + u0, fy = (float)py + u1;
    fx = std::fmin(std::fmax(fx, 0.0f), (float)c.width);
    fy = std::fmin(std::fmax(fy, 0.0f), (float)c.height);
    const float *m = c.r2c;
    float xp = m[0] * fx + m[1] * fy + m[2] * 0.0f + m[3];
    float yp = m[4] * fx + m[5] * fy + m[6] * 0.0f + m[7];
    float zp = m[8] * fx + m[9] * fy + m[10] * 0.0f + m[11];
    float wp = m[12] * fx + m[13] * fy + m[14] * 0.0f + m[15];
    float inv = 1.0f / wp;
    F3 pc = f3(inv * xp, inv * yp, inv * zp);
    F3 dir = fnrm(pc);
    const float *w = c.c2w;
    const double me = (double)FLT_EPSILON * 0.5;
    const float eb3 = (float)((3 * me) / (1 - 3 * me));
    F3 oerr = f3(eb3 * std::fabs(w[3]), eb3 * std::fabs(w[7]), eb3 * std::fabs(w[11]));
    float ox = w[0] * 0.0f + w[1] * 0.0f + w[2] * 0.0f + w[3];
    float oy = w[4] * 0.0f + w[5] * 0.0f + w[6] * 0.0f + w[7];
    float oz = w[8] * 0.0f + w[9] * 0.0f + w[10] * 0.0f + w[11];
    float ow = w[12] * 0.0f + w[13] * 0.0f + w[14] * 0.0f + w[15];
    float oinv = 1.0f / ow;
    F3 org = f3(oinv * ox, oinv * oy, oinv * oz);
    F3 wd = f3(w[0] * dir.x + w[1] * dir.y + w[2] * dir.z, w[4] * dir.x + w[5] * dir.y + w[6] * dir.z,
               w[8] * dir.x + w[9] * dir.y + w[10] * dir.z);
    float ls = (wd.x * wd.x) + (wd.y * wd.y) + (wd.z * wd.z);
    if (ls > 0) {
        float dt = ((std::fabs(wd.x) * oerr.x) + (std::fabs(wd.y) * oerr.y) + (std::fabs(wd.z) * oerr.z)) / ls;
        org = f3(org.x + dt * wd.x, org.y + dt * wd.y, org.z + dt * wd.z);
    }
    o[0] = org.x; o[1] = org.y; o[2] = org.z;
    d[0] = wd.x; d[1] = wd.y; d[2] = wd.z;
    return RT_OK;
}

// Shared body of rt_render, rt_render_cached and rt_render_cache_frame.
// harvest_shift != kNoHarvest: one path in 2^harvest_shift ignores the cache
// and records its first cacheable vertex at cache_depth or deeper (with no
// cache, every path renders plain and those paths still record); *harvested
// receives the records of the (last) chunk, at most the wave capacity.
static rt_status render_frames(const char *fn, rt_scene *s, const rt_render_opts *opts, rt_cache *cache,
                               uint32_t cache_depth, double *accum, uint32_t *nvalid, uint32_t flags, void *stream_,
                               uint32_t harvest_shift = rt::kNoHarvest, uint32_t *harvested = nullptr,
                               rt::PhotonLaunch *photon = nullptr)
{
    const std::string name(fn);
    if (!s || !opts || !accum || !nvalid) { set_error(name + ": null argument"); return RT_ERR_INVALID; }
    if (!s->committed) { set_error(name + ": scene not committed"); return RT_ERR_INVALID; }
    if (flags & RT_FLAG_HOST_PTRS) { set_error(name + ": film buffers must be device pointers"); return RT_ERR_UNSUPPORTED; }
    if (flags & ~(uint32_t)(RT_FLAG_PROFILE | RT_FLAG_COUNT)) { set_error(name + ": unknown flag"); return RT_ERR_INVALID; }
    if (opts->max_depth == 0 || opts->max_depth > 0xffffu) { set_error(name + ": max_depth must be 1..65535"); return RT_ERR_INVALID; }
    if (cache && rt::cache_ctx(cache) != s->ctx) { set_error(name + ": cache and scene belong to different contexts"); return RT_ERR_INVALID; }
    if (cache && cache_depth >= 0xffffu) { set_error(name + ": cache_depth must be below 65535"); return RT_ERR_INVALID; }
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(s->ctx));
    RT_HIP(hipSetDevice(rt::ctx_device(s->ctx)));
    hipStream_t stream = stream_ ? (hipStream_t)stream_ : (hipStream_t)rt::ctx_stream_ptr(s->ctx);
    const uint32_t npix = s->cam.width * s->cam.height;
    const uint64_t total = (uint64_t)npix * opts->spp;
    // default wave: 2^24 paths (capped at the render's paths; ~10 GB of path
    // and shadow state at that size): on the 1080p procedural scene 2^21, 2^22,
    // 2^23 and 2^24 paths ran at 395, 448, 480 and 506 M samples/s
    uint32_t P = opts->path_capacity ? opts->path_capacity : (1u << 24);
    if (P > (1u << 26)) { set_error(name + ": path_capacity above 2^26"); return RT_ERR_INVALID; }
    if ((uint64_t)P > total) P = (uint32_t)std::max<uint64_t>(1, total);
    P = rt::stripe_capacity(P);
    const uint64_t R = std::min<uint64_t>(total, opts->chunk_paths ? opts->chunk_paths : rt::kChunkPaths);
    rt_status st = rt::ensure_buffers(s, P, fn);
    if (st == RT_OK) st = rt::ensure_result_buffer(s->wb, R, fn);
    if (st != RT_OK) return st;
    if (cache) {
        st = rt::ensure_cache_work(s, s->wb.capacity);
        if (st != RT_OK) return st;
    }
    if (photon) {
        st = rt::ensure_photon_rows(s, s->wb.capacity, *photon);
        if (st != RT_OK) return st;
        std::string err;
        hipError_t e = rt::photon_render_reserve(photon->map, s->wb.capacity, photon->k, photon->rgb, photon->dir, stream,
                                                 err);
        if (e != hipSuccess) {
            set_error(name + ": " + err + ": " + hipGetErrorString(e));
            return e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP;
        }
    }
    auto t0 = std::chrono::steady_clock::now();
    RT_HIP(hipMemsetAsync(s->wb.queues.stats, 0, rt::kNumStats * 8, stream));
    s->prof.events = (flags & RT_FLAG_PROFILE) != 0;
    s->prof.count = (flags & RT_FLAG_COUNT) != 0;
    s->prof.used = 0;
    rt::FilmDev film{accum, nvalid};
    rt::ChunkLaunch cl{};
    cl.scene = &s->dev;
    cl.paths = s->wb.paths;
    cl.shadow = s->wb.shadow;
    cl.queues = s->wb.queues;
    cl.capacity = std::min<uint32_t>(P, s->wb.capacity);
    cl.res = s->wb.res;
    cl.cache = cache ? &rt::cache_device(cache) : nullptr;
    cl.cwork = cache ? &s->cwork : nullptr;
    cl.batch = nullptr;
    cl.prof = (flags & (RT_FLAG_PROFILE | RT_FLAG_COUNT)) ? &s->prof : nullptr;
    cl.photon = photon;
    for (uint64_t first = 0; first < total; first += R) {
        rt::WaveParams wp{};
        wp.first_path = first;
        wp.count = (uint32_t)std::min<uint64_t>(R, total - first);
        wp.npix = npix;
        wp.spp_offset = opts->spp_offset;
        wp.seed = opts->seed;
        wp.max_depth = opts->max_depth;
        wp.rr_depth = opts->rr_depth;
        wp.stream_base = s->stream_base;
        wp.cache_depth = (cache || harvest_shift != rt::kNoHarvest) ? cache_depth : rt::kNoCache;
        wp.train = 0;
        wp.grid = opts->grid;
        wp.harvest_shift = harvest_shift;
        wp.harvest_cap = cl.capacity;
        wp.nee_scale = photon && photon->hybrid ? 0.5f : 1.0f;
        RT_HIP(rt::render_chunk(cl, film, wp, s->wb.host_counters, stream));
        if (harvested) {
            // the chunk's last counter read-back came after its final finalize
            const uint32_t k = s->wb.host_counters[rt::kCTrain * rt::kCtrStride];
            *harvested = std::min(k, cl.capacity);
        }
    }
    unsigned long long counts[rt::kNumStats];
    RT_HIP(hipMemcpyAsync(counts, s->wb.queues.stats, sizeof counts, hipMemcpyDeviceToHost, stream));
    RT_HIP(hipStreamSynchronize(stream));
    s->stats = rt_render_stats{};
    s->stats.paths = total;
    s->stats.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rt::fill_stats(s->stats, counts, s->prof, flags, name.c_str());
}

rt_status rt_render(rt_scene *s, const rt_render_opts *opts, double *accum, uint32_t *nvalid, uint32_t flags,
                    void *stream)
{
    return render_frames("rt_render", s, opts, nullptr, 0, accum, nvalid, flags, stream);
}

rt_status rt_render_photon(rt_scene *s, const rt_render_opts *opts, rt_photon_map *map, const double *photon_rgb,
                           const float *photon_dir, const rt_photon_render_opts *popts, double *accum,
                           uint32_t *nvalid, uint32_t flags, void *stream)
{
    if (!s || !map || !popts) { set_error("rt_render_photon: null argument"); return RT_ERR_INVALID; }
    if (popts->mode != RT_RENDER_PHOTON_MAPPING && popts->mode != RT_RENDER_HYBRID) {
        set_error("rt_render_photon: mode must be RT_RENDER_PHOTON_MAPPING or RT_RENDER_HYBRID");
        return RT_ERR_INVALID;
    }
    if (popts->k == 0 || popts->k > 1024) { set_error("rt_render_photon: k must be 1..1024"); return RT_ERR_INVALID; }
    if (!(popts->radius > 0.0f)) { set_error("rt_render_photon: radius must be positive"); return RT_ERR_INVALID; }
    if (rt::photon_map_ctx(map) != s->ctx) { set_error("rt_render_photon: map and scene belong to different contexts"); return RT_ERR_INVALID; }
    uint32_t n = 0, depth = 0;
    rt_status st = rt_photon_map_get_info(map, &n, &depth);
    if (st != RT_OK) return st;
    if (n && (!photon_rgb || !photon_dir)) { set_error("rt_render_photon: null photon arrays"); return RT_ERR_INVALID; }
    rt::PhotonLaunch ph{};
    ph.map = map;
    ph.rgb = photon_rgb;
    ph.dir = photon_dir;
    ph.k = popts->k;
    ph.radius = popts->radius;
    ph.hybrid = popts->mode == RT_RENDER_HYBRID ? 1u : 0u;
    return render_frames("rt_render_photon", s, opts, nullptr, 0, accum, nvalid, flags, stream, rt::kNoHarvest,
                         nullptr, &ph);
}

rt_status rt_render_cached(rt_scene *s, const rt_render_opts *opts, rt_cache *cache, uint32_t cache_depth,
                           double *accum, uint32_t *nvalid, uint32_t flags, void *stream)
{
    if (!cache) { set_error("rt_render_cached: null cache"); return RT_ERR_INVALID; }
    return render_frames("rt_render_cached", s, opts, cache, cache_depth, accum, nvalid, flags, stream);
}

rt_status rt_render_cache_frame(rt_scene *s, const rt_render_opts *opts, rt_cache *cache,
                                const rt_cache_frame_opts *copts, double *accum, uint32_t *nvalid,
                                rt_cache_frame_stats *stats, uint32_t flags, void *stream_)
{
    if (stats) *stats = rt_cache_frame_stats{};
    if (!cache || !copts) { set_error("rt_render_cache_frame: null cache or options"); return RT_ERR_INVALID; }
    if (copts->cache_depth >= 0xffffu) { set_error("rt_render_cache_frame: cache_depth must be below 65535"); return RT_ERR_INVALID; }
    if (s && rt::cache_ctx(cache) != s->ctx) { set_error("rt_render_cache_frame: cache and scene belong to different contexts"); return RT_ERR_INVALID; }
    if (!s) { set_error("rt_render_cache_frame: null scene"); return RT_ERR_INVALID; }
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(s->ctx));
    RT_HIP(hipSetDevice(rt::ctx_device(s->ctx)));
    hipStream_t stream = stream_ ? (hipStream_t)stream_ : (hipStream_t)rt::ctx_stream_ptr(s->ctx);
    // a cache reset after diverging sits this frame out: plain path tracing,
    // every path harvested for its fresh parameters
    const bool fallback = rt::cache_take_fallback(cache);
    const bool plain = copts->plain != 0 || fallback;
    const uint32_t shift = plain ? 0u : (copts->harvest_shift >= 32 ? rt::kNoHarvest : copts->harvest_shift);
    uint32_t harvested = 0;
    rt_status st = render_frames("rt_render_cache_frame", s, opts, plain ? nullptr : cache, copts->cache_depth, accum,
                                 nvalid, flags, stream_, shift, &harvested);
    if (st != RT_OK) return st;
    uint32_t done = 0;
    float loss = std::numeric_limits<float>::quiet_NaN();
    bool reset = false;
    if (shift != rt::kNoHarvest) {
        st = rt::cache_train_harvest(cache, s->wb.paths.cin, s->wb.paths.ttgt, harvested, copts->train_steps, stream,
                                     &done, &loss, &reset);
        if (st != RT_OK) return st;
    }
    if (stats) {
        stats->samples = harvested;
        stats->steps = done;
        stats->loss = loss;
        stats->fallback = fallback ? 1u : 0u;
        stats->reset = reset ? 1u : 0u;
    }
    return RT_OK;
}

rt_status rt_cache_train_scene(rt_cache *c, rt_scene *s, const rt_render_opts *opts, uint32_t cache_depth,
                               uint32_t n_paths, float *loss, uint32_t *n_samples, void *stream_)
{
    if (n_samples) *n_samples = 0;
    if (!c || !s || !opts) { set_error("rt_cache_train_scene: null argument"); return RT_ERR_INVALID; }
    if (!s->committed) { set_error("rt_cache_train_scene: scene not committed"); return RT_ERR_INVALID; }
    if (rt::cache_ctx(c) != s->ctx) { set_error("rt_cache_train_scene: cache and scene belong to different contexts"); return RT_ERR_INVALID; }
    if (opts->max_depth == 0 || opts->max_depth > 0xffffu) { set_error("rt_cache_train_scene: max_depth must be 1..65535"); return RT_ERR_INVALID; }
    if (cache_depth >= 0xffffu) { set_error("rt_cache_train_scene: cache_depth must be below 65535"); return RT_ERR_INVALID; }
    if (n_paths == 0 || n_paths > (1u << 26)) { set_error("rt_cache_train_scene: n_paths must be 1..2^26"); return RT_ERR_INVALID; }
    std::lock_guard<std::recursive_mutex> g(rt::ctx_lock(s->ctx));
    RT_HIP(hipSetDevice(rt::ctx_device(s->ctx)));
    hipStream_t stream = stream_ ? (hipStream_t)stream_ : (hipStream_t)rt::ctx_stream_ptr(s->ctx);
    const uint32_t P = rt::stripe_capacity(n_paths);
    rt_status st = rt::ensure_buffers(s, P, "rt_cache_train_scene");
    if (st == RT_OK) st = rt::ensure_cache_work(s, s->wb.capacity);
    if (st != RT_OK) return st;
    RT_HIP(hipMemsetAsync(s->wb.queues.stats, 0, rt::kNumStats * 8, stream));
    rt::WaveParams wp{};
    wp.first_path = (uint64_t)opts->spp_offset * n_paths;
    wp.count = n_paths;
    wp.npix = s->cam.width * s->cam.height;
    wp.spp_offset = 0;
    wp.seed = opts->seed;
    wp.max_depth = opts->max_depth;
    wp.rr_depth = opts->rr_depth;
    wp.stream_base = s->stream_base;
    wp.cache_depth = cache_depth;
    wp.train = 1;
    rt::FilmDev film{nullptr, nullptr};
    rt::ChunkLaunch cl{};
    cl.scene = &s->dev;
    cl.paths = s->wb.paths;
    cl.shadow = s->wb.shadow;
    cl.queues = s->wb.queues;
    cl.capacity = P;  // a slot for every training path: no read-back needed
    cl.cache = &rt::cache_device(c);  // training paths end in the cache after their suffix
    cl.cwork = &s->cwork;
    // no host read-back: every training path starts at once and ends within max_depth iterations
    RT_HIP(rt::render_chunk(cl, film, wp, nullptr, stream));
    // the sample count stays on the device: no host round trip unless the
    // caller asks for the loss or the count
    const uint32_t *k_dev = rt::scount(s->wb.queues, rt::kCTrain);
    float *loss_dev = nullptr;
    st = rt::cache_train_rows_dev(c, s->wb.paths.cin, s->wb.paths.ttgt, n_paths, k_dev, stream, &loss_dev);
    if (st != RT_OK) return st;
    if (loss || n_samples) {
        uint32_t k = 0;
        float l = 0;
        RT_HIP(hipMemcpyAsync(&k, k_dev, 4, hipMemcpyDeviceToHost, stream));
        RT_HIP(hipMemcpyAsync(&l, loss_dev, 4, hipMemcpyDeviceToHost, stream));
        RT_HIP(hipStreamSynchronize(stream));
        const uint32_t n = std::min(std::min(k, n_paths), rt::cache_batch(c));
        if (n_samples) *n_samples = n;
        if (loss) *loss = n ? l / (3.0f * (float)n) : 0.0f;
    }
    return RT_OK;
}

rt_status rt_render_get_stats(const rt_scene *s, rt_render_stats *st)
{
    if (!s || !st) { set_error("rt_render_get_stats: null argument"); return RT_ERR_INVALID; }
    *st = s->stats;
    return RT_OK;
}

}  // extern "C"
