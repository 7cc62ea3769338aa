// photon_trace.hip -- the photon pre-pass on the GPU: Scene::generatePhotons
// and computePhotons (integration/Scene.cu:989-1021,
// integration/photonDispersion.cu:26-192).
//
// Each photon path picks a point light with probability proportional to its
// power (Scene::selectOneRandomLight, Scene.cu:427-455), leaves it in a
// uniformly sampled direction (PointLight::sampleLight, light/PointLight.cu:
// 68-95) carrying color * intensity * 4 pi / n_l / p_l, where n_l is the
// number of photons that picked light l and p_l its selection probability,
// and bounces through the scene for at most max_depth intersections:
//   * at a scattering hit (Lambertian, Trowbridge, plain) a photon is stored
//     with the hit point, the incoming direction, the shading normal, the
//     radiance totalLight * pathLightFactor and the shape id (analytic shape
//     k: k, in sphere-then-plane order; triangle t: shapes + t, as
//     Scene::intersectShapes numbers them, Scene.cu:209-227);
//   * the material's sampleDirection gives the bounce; emission of the hit
//     (light: kd * intensity, plain: ke) is added to totalLight times the
//     throughput; a black weight ends the path;
//   * Russian roulette at scattering hits deeper than rr_depth
//     (pathTracing.cu:81-102, photon_first_pass).
// One lane traces one photon path (the pass is ~1M short paths, a small
// fraction of a frame), with the renderer's traversal, surface and BSDF code.
// Photons are appended with a key (path index << 5 | bounce) and sorted by
// it, so the output order is deterministic: by photon path, then bounce.
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>

#include <cmath>
#include <string>
#include <vector>

#include "rt_device.h"
#include "rt_internal.h"
#include "rt_photon_trace.h"
#include "rt_render.h"
#include "rt_shade.h"
#include "rt_surface.h"

namespace rt {

namespace {

constexpr int kPhBlock = kTravBlock;  // the LDS stack column stride assumes this block size
constexpr uint32_t kKeyShift = 5;     // bounce index bits of a photon key

struct PhotonTraceArgs {
    const double *cdf;       // nl cumulative selection probabilities
    const double *light_rgb; // nl * 3 photon radiance per light (after the counts are known)
    uint32_t nl, n, seed, max_depth, rr_depth;
    uint32_t nshapes;        // spheres + planes: the first triangle's shape id
    // staging (append order)
    float *pos, *dir, *nrm;  // 3 per photon
    double *rgb;             // 3 per photon
    int32_t *shape;
    uint32_t *key;
    uint32_t *counter;
    uint32_t cap;
};

// The photon path's generator: stream i of the seed's state.
__device__ __forceinline__ Pcg32 photon_rng(uint32_t seed, uint32_t i)
{
    return pcg32_seed(((uint64_t)seed << 32) ^ 0x9E3779B97F4A7C15ull, (uint64_t)i);
}

// First light whose cumulative probability reaches u (the reference's
// running subtraction); rounding past the end falls back to the last light
// with a positive probability.
__device__ __forceinline__ uint32_t select_light(const double *cdf, uint32_t nl, float u)
{
    for (uint32_t l = 0; l < nl; ++l)
        if ((double)u <= cdf[l]) return l;
    uint32_t last = nl - 1;
    while (last > 0 && cdf[last] == cdf[last - 1]) --last;
    return last;
}

// Photons per light.  Most lanes of a wavefront pick the same few lights, so
// each distinct light of the wavefront takes one atomic (a same-address
// atomic per photon serialises: 1M of them took 11.9 ms).
__global__ __launch_bounds__(256) void k_photon_select(PhotonTraceArgs a, uint32_t *counts)
{
    const uint32_t lane = __lane_id();
    for (uint32_t i0 = blockIdx.x * 256; i0 < a.n; i0 += gridDim.x * 256) {
        const uint32_t i = i0 + threadIdx.x;
        uint32_t l = 0xffffffffu;
        if (i < a.n) {
            Pcg32 g = photon_rng(a.seed, i);
            l = select_light(a.cdf, a.nl, pcg32_float(g));
        }
        uint64_t left = __ballot(l != 0xffffffffu);
        while (left) {
            const uint32_t leader = (uint32_t)(__ffsll((unsigned long long)left) - 1);
            const uint32_t ll = (uint32_t)__shfl((int)l, (int)leader);
            const uint64_t same = __ballot(l == ll);
            if (lane == leader) atomicAdd(&counts[ll], (uint32_t)__popcll(same));
            left &= ~same;
        }
    }
}

__device__ __forceinline__ bool scattering(uint32_t type)
{
    return type == RT_MAT_LAMBERTIAN || type == RT_MAT_TROWBRIDGE || type == RT_MAT_TROWBRIDGE_BASIC ||
           type == RT_MAT_PLAIN;
}

template <bool WIDE>
__global__ __launch_bounds__(kPhBlock) void k_photon_trace(SceneDev sc, PhotonTraceArgs a, int cap)
{
    extern __shared__ int32_t s_stack[];
    StackTail<PathStack> s_tail;
    const PathStack stack{s_stack + threadIdx.x, s_tail};
    const float4 *nodes = (const float4 *)sc.bvh.nodes;
    const float4 *tris = (const float4 *)sc.bvh.tris;
    uint32_t n_node = 0, n_tri = 0;
    for (uint32_t i = blockIdx.x * kPhBlock + threadIdx.x; i < a.n; i += gridDim.x * kPhBlock) {
        Pcg32 g = photon_rng(a.seed, i);
        const uint32_t l = select_light(a.cdf, a.nl, pcg32_float(g));
        // uniformSampleSphere (sampling/Sampler.cu:47-52)
        const float u0 = pcg32_float(g), u1 = pcg32_float(g);
        const float phi = 2.0f * kPi * u0, theta = acosf(2.0f * u1 - 1.0f);
        const rt_point_light &pl = sc.pl[l];
        V3 o = mk(pl.pos[0], pl.pos[1], pl.pos[2]);
        V3 d = mk(sinf(theta) * cosf(phi), sinf(theta) * sinf(phi), cosf(theta));
        double tl[3] = {a.light_rgb[3 * l], a.light_rgb[3 * l + 1], a.light_rgb[3 * l + 2]};
        V3 beta = mk(1, 1, 1);
        for (uint32_t depth = 0; depth < a.max_depth; ++depth) {
            // closest hit: the tree's triangle, then the analytic shapes
            // (strict <), as the renderer's extend and classify stages
            TravState s;
            trav_begin(s, o.x, o.y, o.z, d.x, d.y, d.z, INFINITY, false, 0);
            if (sc.bvh.nodes)
                while (!trav_step_any<WIDE, false, false>(s, nodes, tris, stack, cap, n_node, n_tri)) {
                }
            int32_t hid = s.hit;
            float th = s.thit;
            for (uint32_t k = 0; k < sc.nsph; ++k) {
                V3 so, sd;
                const float ts = sphere_core(sc.sph[k], o, d, INFINITY, so, sd);
                if (ts >= 0 && ts < th) { th = ts; hid = -(int32_t)k - 2; }
            }
            for (uint32_t k = 0; k < sc.nplane; ++k) {
                float u, v, dz;
                const float tp = plane_core(sc.plane[k], o, d, INFINITY, u, v, dz);
                if (tp > 0 && tp < th) { th = tp; hid = -(int32_t)(sc.nsph + k) - 2; }
            }
            if (hid == -1) break;
            Surface su{};
            V3 dd;
            surface_of(sc, make_float4(o.x, o.y, o.z, d.x), make_float4(d.y, d.z, th, __int_as_float(hid)), su, dd);
            rt_material m = sc.mat[su.mat];
            if (sc.ntex) apply_texture(sc, su, m);
            const bool scat = scattering(m.type);
            // one append per wavefront over the lanes storing a photon now
            const uint64_t smask = __ballot(scat);
            if (scat) {
                const uint32_t lane = __lane_id();
                const uint32_t leader = (uint32_t)(__ffsll((unsigned long long)smask) - 1);
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(a.counter, (uint32_t)__popcll(smask));
                base = (uint32_t)__shfl((int)base, (int)leader);
                const uint32_t pos = base + (uint32_t)__popcll(smask & ((1ull << lane) - 1ull));
                if (pos < a.cap) {
                    a.pos[3 * (size_t)pos] = su.p.x; a.pos[3 * (size_t)pos + 1] = su.p.y; a.pos[3 * (size_t)pos + 2] = su.p.z;
                    a.dir[3 * (size_t)pos] = d.x; a.dir[3 * (size_t)pos + 1] = d.y; a.dir[3 * (size_t)pos + 2] = d.z;
                    a.nrm[3 * (size_t)pos] = su.n.x; a.nrm[3 * (size_t)pos + 1] = su.n.y; a.nrm[3 * (size_t)pos + 2] = su.n.z;
                    a.rgb[3 * (size_t)pos] = tl[0] * (double)beta.x;
                    a.rgb[3 * (size_t)pos + 1] = tl[1] * (double)beta.y;
                    a.rgb[3 * (size_t)pos + 2] = tl[2] * (double)beta.z;
                    a.shape[pos] = hid >= 0 ? (int32_t)(a.nshapes + (uint32_t)hid) : -hid - 2;
                    a.key[pos] = (i << kKeyShift) | depth;
                }
            }
            // SurfaceInteraction::sampleRay in the shading frame
            V3 t, b;
            onb(su.n, t, b);
            const V3 wow = neg(d);
            const V3 wo = mk(dot(wow, t), dot(wow, b), dot(wow, su.n));
            V3 wi;
            float pdf;
            bool delta;
            V3 wgt;
            if (m.type == RT_MAT_LAMBERTIAN) wgt = bsdf_sample<RT_MAT_LAMBERTIAN>(m, wo, g, wi, pdf, delta);
            else if (m.type == RT_MAT_DIELECTRIC || m.type == RT_MAT_MIRROR)
                wgt = bsdf_sample<RT_MAT_DIELECTRIC>(m, wo, g, wi, pdf, delta);
            else if (m.type == RT_MAT_TROWBRIDGE || m.type == RT_MAT_TROWBRIDGE_BASIC)
                wgt = bsdf_sample<RT_MAT_TROWBRIDGE>(m, wo, g, wi, pdf, delta);
            else wgt = bsdf_sample<RT_MAT_LIGHT>(m, wo, g, wi, pdf, delta);
            // emission of the hit into totalLight (photon rays are never
            // camera rays: a plain surface gives ke)
            if (m.type == RT_MAT_LIGHT) {
                const float I = m.intensity;
                tl[0] += (double)beta.x * (double)(m.kd[0] * I);
                tl[1] += (double)beta.y * (double)(m.kd[1] * I);
                tl[2] += (double)beta.z * (double)(m.kd[2] * I);
            } else if (m.type == RT_MAT_PLAIN) {
                tl[0] += (double)beta.x * (double)m.ke[0];
                tl[1] += (double)beta.y * (double)m.ke[1];
                tl[2] += (double)beta.z * (double)m.ke[2];
            }
            beta = hmul(beta, wgt);
            if (black(wgt)) break;
            if (scat && depth > a.rr_depth) {
                const float q = fmaxf(0.05f, 1 - luminance(beta));
                if (pcg32_float(g) < q) break;
                const float sc1 = 1 - q;
                beta = mk(beta.x / sc1, beta.y / sc1, beta.z / sc1);
            }
            const V3 ww = add(add(mul(t, wi.x), mul(b, wi.y)), mul(su.n, wi.z));
            o = spawn_origin(su.p, su.perr, su.n, ww);
            d = ww;
        }
    }
}

// The sorted output: entry j takes staging record perm[j].
__global__ __launch_bounds__(256) void k_photon_scatter(PhotonTraceArgs a, const uint32_t *perm, uint32_t m,
                                                        float *pos, float *dir, float *nrm, double *rgb,
                                                        int32_t *shape)
{
    for (uint32_t j = blockIdx.x * 256 + threadIdx.x; j < m; j += gridDim.x * 256) {
        const uint32_t e = perm[j];
        for (int c = 0; c < 3; ++c) {
            pos[3 * (size_t)j + c] = a.pos[3 * (size_t)e + c];
            dir[3 * (size_t)j + c] = a.dir[3 * (size_t)e + c];
            nrm[3 * (size_t)j + c] = a.nrm[3 * (size_t)e + c];
            rgb[3 * (size_t)j + c] = a.rgb[3 * (size_t)e + c];
        }
        shape[j] = a.shape[e];
    }
}

__global__ __launch_bounds__(256) void k_iota(uint32_t *v, uint32_t n)
{
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) v[i] = i;
}

uint32_t blocks_for(uint64_t n, uint32_t block, uint32_t cap)
{
    const uint64_t b = (n + block - 1) / block;
    return (uint32_t)(b < 1 ? 1 : (b > cap ? cap : b));
}

struct DevMem {
    void *p = nullptr;
    ~DevMem() { if (p) (void)hipFree(p); }
};

}  // namespace

void photon_light_table(const rt_point_light *lights, uint32_t nl, std::vector<double> &p, std::vector<double> &cdf)
{
    p.assign(nl, 0.0);
    cdf.assign(nl, 0.0);
    double total = 0;
    for (uint32_t l = 0; l < nl; ++l) {
        const rt_point_light &L = lights[l];
        // sum(PointLight::power()) = sum(color) * intensity * 4 pi
        p[l] = ((double)L.color[0] + (double)L.color[1] + (double)L.color[2]) * (double)L.intensity * (4.0 * M_PI);
        total += p[l];
    }
    double run = 0;
    for (uint32_t l = 0; l < nl; ++l) {
        p[l] = total > 0 ? p[l] / total : 0.0;
        run += p[l];
        cdf[l] = run;
    }
}

rt_status trace_photons(const SceneDev &sc, const rt_point_light *lights, uint32_t nl, const rt_photon_opts &o,
                        uint32_t capacity, float *location, float *direction, float *normal, double *radiance,
                        int32_t *shape, uint32_t *key, uint32_t *count, hipStream_t s, const char *who)
{
    *count = 0;
    if (nl == 0 || o.photons == 0) return RT_OK;  // generatePhotons runs only with lights
    std::vector<double> p, cdf;
    photon_light_table(lights, nl, p, cdf);
    if (!(cdf[nl - 1] > 0)) {
        set_error(std::string(who) + ": the point lights carry no power");
        return RT_ERR_INVALID;
    }
    auto fail = [&](hipError_t e, const char *what) {
        set_error(std::string(who) + ": " + what + ": " + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP;
    };
    // staging for capacity photons: 9 floats, 3 doubles, shape, key, and
    // the sort's second key / value arrays and permutation
    const size_t C = capacity ? capacity : 1;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t r = off; off += (bytes + 255) & ~(size_t)255; return r; };
    const size_t o_cdf = take(nl * sizeof(double)), o_rgbl = take(3 * nl * sizeof(double)),
                 o_cnt = take(nl * sizeof(uint32_t) + sizeof(uint32_t)), o_pos = take(C * 12), o_dir = take(C * 12),
                 o_nrm = take(C * 12), o_rgb = take(C * 24), o_shape = take(C * 4), o_key = take(C * 4),
                 o_key2 = take(C * 4), o_val = take(C * 4), o_val2 = take(C * 4);
    const uint32_t kbits = 32u - (uint32_t)__builtin_clz((uint32_t)(((uint64_t)o.photons << kKeyShift) - 1) | 1u);
    size_t sort_bytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const uint32_t *)nullptr,
                                                      (uint32_t *)nullptr, (const uint32_t *)nullptr,
                                                      (uint32_t *)nullptr, (int)C, 0, (int)kbits, s);
    if (e != hipSuccess) return fail(e, "sort sizing");
    const size_t o_sort = take(sort_bytes);
    DevMem mem;
    e = hipMalloc(&mem.p, off);
    if (e != hipSuccess) return fail(e, "photon work space");
    char *base = (char *)mem.p;
    PhotonTraceArgs a{};
    a.cdf = (const double *)(base + o_cdf);
    a.light_rgb = (const double *)(base + o_rgbl);
    a.nl = nl;
    a.n = o.photons;
    a.seed = o.seed;
    a.max_depth = o.max_depth;
    a.rr_depth = o.rr_depth;
    a.nshapes = sc.nsph + sc.nplane;
    a.pos = (float *)(base + o_pos);
    a.dir = (float *)(base + o_dir);
    a.nrm = (float *)(base + o_nrm);
    a.rgb = (double *)(base + o_rgb);
    a.shape = (int32_t *)(base + o_shape);
    a.key = (uint32_t *)(base + o_key);
    uint32_t *counts = (uint32_t *)(base + o_cnt);
    a.counter = counts + nl;
    a.cap = capacity;
    if ((e = hipMemcpyAsync(base + o_cdf, cdf.data(), nl * sizeof(double), hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemsetAsync(counts, 0, (nl + 1) * sizeof(uint32_t), s)) != hipSuccess)
        return fail(e, "upload");
    hipLaunchKernelGGL(k_photon_select, dim3(blocks_for(o.photons, 256, 4096)), dim3(256), 0, s, a, counts);
    std::vector<uint32_t> n_l(nl);
    if ((e = hipGetLastError()) != hipSuccess ||
        (e = hipMemcpyAsync(n_l.data(), counts, nl * sizeof(uint32_t), hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
        return fail(e, "light selection");
    // PointLight::sampleLight(sampler, n_l) then / lightSelectionPdf(l)
    // (Scene.cu:1005-1012): color * intensity * 4 pi / n_l / p_l
    std::vector<double> rgbl(3 * nl, 0.0);
    for (uint32_t l = 0; l < nl; ++l)
        for (int c = 0; c < 3; ++c)
            if (n_l[l])
                rgbl[3 * l + c] = (double)lights[l].color[c] * (double)lights[l].intensity * (4.0 * M_PI) /
                                  (double)n_l[l] / p[l];
    if ((e = hipMemcpyAsync(base + o_rgbl, rgbl.data(), rgbl.size() * sizeof(double), hipMemcpyHostToDevice, s)) !=
        hipSuccess)
        return fail(e, "upload");
    const int cap = (int)sc.bvh.stack_cap;
    const size_t lds = extend_lds_bytes(sc.bvh.stack_cap);
    const dim3 grid(blocks_for(o.photons, kPhBlock, 2048));
    if (sc.bvh.wide) hipLaunchKernelGGL(k_photon_trace<true>, grid, dim3(kPhBlock), lds, s, sc, a, cap);
    else hipLaunchKernelGGL(k_photon_trace<false>, grid, dim3(kPhBlock), lds, s, sc, a, cap);
    uint32_t m = 0;
    if ((e = hipGetLastError()) != hipSuccess ||
        (e = hipMemcpyAsync(&m, a.counter, sizeof(uint32_t), hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
        return fail(e, "photon tracing");
    *count = m;
    if (m > capacity) {
        set_error(std::string(who) + ": " + std::to_string(m) + " photons stored, capacity " +
                  std::to_string(capacity));
        return RT_ERR_INVALID;
    }
    if (m == 0) return RT_OK;
    uint32_t *val = (uint32_t *)(base + o_val), *val2 = (uint32_t *)(base + o_val2);
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(m, 256, 4096)), dim3(256), 0, s, val, m);
    e = hipcub::DeviceRadixSort::SortPairs(base + o_sort, sort_bytes, a.key, (uint32_t *)(base + o_key2), val, val2,
                                           (int)m, 0, (int)kbits, s);
    if (e != hipSuccess) return fail(e, "sort");
    hipLaunchKernelGGL(k_photon_scatter, dim3(blocks_for(m, 256, 4096)), dim3(256), 0, s, a, val2, m, location,
                       direction, normal, radiance, shape);
    if (key && (e = hipMemcpyAsync(key, base + o_key2, (size_t)m * sizeof(uint32_t), hipMemcpyDeviceToDevice, s)) !=
                   hipSuccess)
        return fail(e, "photon keys");
    if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(s)) != hipSuccess)
        return fail(e, "photon output");
    return RT_OK;
}

}  // namespace rt
