// render.hip -- wavefront path tracer for gfx950.
//
// A wave of up to P paths lives in structure-of-arrays state in HBM.  Per
// bounce:
//   extend   closest hit (BVH + analytic spheres) for every active path,
//            then material classification into per-class queues with
//            wavefront ballots (one atomic per 64 lanes per queue);
//   shade<M> one kernel per material class (Lambertian, dielectric, GGX,
//            light, miss), so a wavefront runs a single BSDF: resolves the
//            previous bounce's next-event estimate, emits this bounce's
//            shadow rays, samples the BSDF, applies Russian roulette and
//            appends survivors to the next active queue;
//   connect  any hit for every shadow ray; occluded rays clear their light
//            bit in the path's NEE mask;
// and a final pass adds finished paths to the FP64 film.
//
// The estimator is the reference's (integration/pathTracing.cu:33-123,
// 173-205, 284-384; integration/Scene.cu:165-335, 468-585): emission at
// every hit, next-event estimation at scattering materials over every point
// light with weight 1 / (1 + pdf_bsdf^2) (Scene.cu:331, pdf_light = 1),
// Russian roulette at scattering materials beyond rr_depth with
// q = max(0.05, 1 - luminance(beta)) (pathTracing.cu:91-102).  Next-event
// contributions are summed per path in light order, so the result does not
// depend on the order in which shadow rays complete.
//
// Every queue kernel is a grid-stride loop over counts read from device
// memory, so a wave is a fixed launch sequence with no host round trip.
// Path slots, queues and their counters are striped (rt_render.h): a
// workgroup serves stripe blockIdx % kStripes, so the appends of the whole
// chip spread over 64 counters instead of queueing on one address.
#include "rt_device.h"
#include "rt_render.h"
#include "rt_shade.h"
#include "rt_photon_render.h"
#include "rt_surface.h"
#include "rt_cache.h"

namespace rt {

constexpr int kBlock = 128;
constexpr int kGrid = 2048;  // 8 workgroups per CU on 256 CUs

__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool pred)
{
    const uint64_t mask = __ballot(pred);
    if (mask == 0) return 0;
    const uint32_t lane = __lane_id();
    const uint32_t leader = (uint32_t)(__ffsll((unsigned long long)mask) - 1);
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, (int)leader);
    return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
}

// Exclusive prefix sum over a wavefront's lanes; `total` is the full sum.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t &total)
{
    const uint32_t lane = __lane_id();
    uint32_t incl = v;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t t = (uint32_t)__shfl_up((int)incl, off);
        if (lane >= off) incl += t;
    }
    total = (uint32_t)__shfl((int)incl, 63);
    return incl - v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off);
    return v;
}

// The stripe a queue kernel's workgroup serves, and its share of the
// stripe's items: [j0 + thread, ...) with step `step` (every stripe has at
// least one workgroup when the grid has at least kStripes).
struct StripeSpan {
    uint32_t s, j0, step;
};

__device__ __forceinline__ StripeSpan stripe_span()
{
    const uint32_t s = blockIdx.x % kStripes;
    const uint32_t nb = (gridDim.x - s + kStripes - 1) / kStripes;
    return {s, (blockIdx.x / kStripes) * kBlock, nb * kBlock};
}

// The record of a path slot and its words.
__device__ __forceinline__ float4 *path_rec(const PathsDev &p, uint32_t slot)
{
    return p.rec + (size_t)slot * kPathQuads;
}
__device__ __forceinline__ uint32_t *path_word(const PathsDev &p, uint32_t slot, uint32_t quad, uint32_t w)
{
    return (uint32_t *)(path_rec(p, slot) + quad) + w;
}
__device__ __forceinline__ uint32_t bits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float as_f(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ float as_f(int32_t i) { return __int_as_float(i); }

// The scene of path `slot`: the kernel's scene, or the batch table entry.
template <bool B>
__device__ __forceinline__ const SceneDev &scene_of(const SceneDev &sc0, const BatchDev &bt, const PathsDev &p,
                                                    uint32_t slot)
{
    if constexpr (B) return bt.scenes[*path_word(p, slot, kPqScene, 0)];
    else return sc0;
}

// Read-only data at a wavefront-uniform address, read through the constant
// address space so the compiler issues scalar loads: the scene record of a
// wavefront whose paths all belong to one scene of a batch, and what that
// record points to at uniform indices (point lights, light materials).
// The empty asm statement keeps the constant-space pointer (in scalar
// registers) opaque, so the casts are not folded back into a global pointer;
// ptr must be wavefront-uniform.
template <typename T>
__device__ __forceinline__ const T &const_ref(const T *ptr)
{
    const __attribute__((address_space(4))) T *cp = (const __attribute__((address_space(4))) T *)ptr;
    asm("" : "+s"(cp));
    return *(const T *)cp;
}
template <bool UNI, typename T>
__device__ __forceinline__ const T &uniform_ref(const T *ptr)
{
    if constexpr (UNI) return const_ref(ptr);
    else return *ptr;
}

template <bool B>
__device__ __forceinline__ uint64_t stream_base_of(const SceneDev &sc, const WaveParams &wp)
{
    if constexpr (B) return sc.stream_base;
    else return wp.stream_base;
}

// PCG32 generator of a path from its kPqRng quad (state, pixel)
__device__ __forceinline__ Pcg32 path_rng(float4 q, uint64_t stream_base)
{
    Pcg32 g;
    g.state = (uint64_t)bits(q.x) | ((uint64_t)bits(q.y) << 32);
    g.inc = ((stream_base + bits(q.z)) << 1u) | 1u;
    return g;
}

// ---------------------------------------------------------------------
// Counter resets.  Per iteration: the next live list, the material queues,
// the shadow list, the terminated list, the claim counters of the
// persistent kernels and the cache queries; at a chunk's start: all.
// Workgroup L < kNumLists clears list L's stripe counters (one lane per
// stripe), the last workgroup the scalar counters.
__global__ __launch_bounds__(kStripes) void k_reset(QueuesDev q, uint32_t cur, uint32_t chunk_start)
{
    const uint32_t L = blockIdx.x, t = threadIdx.x;
    if (L < kNumLists) {
        const bool clear = chunk_start || L == kLActive0 + (cur ^ 1u) || (L >= kLQueue && L <= kLTerm) ||
                           L == kLClaimExt || L == kLClaimCon;
        if (clear) *lcount(q, L, t) = 0;
    } else if (t < kNumCounters && (chunk_start || t == kCCacheQ)) {
        *scount(q, t) = 0;
    }
}

// A new path in `slot`: global path index g (chunk index k).  Within its
// scene, g is sample * npix + pixel (a batch concatenates its scenes' index
// ranges).  Camera ray and PCG32 stream as the single-scene renderer.
template <bool B>
__device__ __forceinline__ void init_path(const SceneDev &sc0, const BatchDev &bt, const PathsDev &p,
                                          const WaveParams &wp, uint32_t slot, uint64_t g, uint32_t k)
{
    float4 *rec = path_rec(p, slot);
    uint32_t sid = 0;
    if constexpr (B) {
        // the hint is the scene at the start of g's 2^16-path block; scenes
        // hold at least one path, so a short forward walk ends at g's scene
        sid = bt.hint[g >> kHintShift];
        while (sid + 1 < bt.n && bt.bases[sid + 1] <= g) ++sid;
        g -= bt.bases[sid];
        rec[kPqScene] = make_float4(as_f(sid), as_f(bt.root[sid]), 0, 0);
    }
    const SceneDev &sc = B ? bt.scenes[sid] : sc0;
    const uint32_t npix = B ? sc.cam.width * sc.cam.height : wp.npix;
    const uint32_t seed = B ? sc.seed : wp.seed;
    const uint64_t stream_base = stream_base_of<B>(sc, wp);
    const uint32_t first_sample = wp.spp_offset_dev ? *wp.spp_offset_dev : wp.spp_offset;
    uint32_t sample = first_sample + (uint32_t)(g / npix);
    uint32_t pix = (uint32_t)(g % npix);
    if (!B && wp.train) {
        // training paths: pseudo-random pixels, sample indices disjoint
        // from the rendering samples (top bit set)
        Pcg32 h = pcg32_seed(g, 0x7a11u);
        pix = pcg32_next(h) % npix;
        sample = 0x80000000u | (uint32_t)g;
    }
    const uint32_t px = pix % sc.cam.width, py = pix / sc.cam.width;
    const uint64_t initstate = ((uint64_t)seed << 32) ^ (uint64_t)sample ^ 0x853c49e6748fea9bULL;
    Pcg32 r = pcg32_seed(initstate, stream_base + pix);
    const float u0 = pcg32_float(r), u1 = pcg32_float(r);
    V3 o, d;
    camera_ray(sc.cam, px, py, u0, u1, o, d);
    // depth 0; a camera ray counts as delta for MIS (bit 16) and is a camera
    // ray (bit 17, Ray::cameraRay); harvest paths carry bit 18
    uint32_t harvest = 0;
    if (!B && wp.harvest_shift != kNoHarvest) {
        // Fibonacci hash of the global path index: its top harvest_shift bits
        // are zero for one path in 2^harvest_shift
        const uint32_t h = (uint32_t)((g * 0x9E3779B97F4A7C15ull) >> 32);
        harvest = (wp.harvest_shift == 0 || (h >> (32u - wp.harvest_shift)) == 0) ? kMetaHarvest : 0u;
    }
    const uint32_t meta = (1u << 16) | (1u << 17) | harvest;
    // the whole record but the scene quad, so the line is written in full
    rec[kPqRay] = make_float4(o.x, o.y, o.z, d.x);
    rec[kPqHit] = make_float4(d.y, d.z, INFINITY, as_f(-1));
    rec[kPqBeta] = make_float4(1, 1, 1, as_f(meta));
    rec[kPqL] = make_float4(0, 0, 0, as_f(k));
    rec[kPqRng] = make_float4(as_f((uint32_t)r.state), as_f((uint32_t)(r.state >> 32)), as_f(pix), as_f(0u));
    rec[kPqPrev] = make_float4(0, o.x, o.y, o.z);
    rec[kPqNee] = make_float4(0, 0, 0, as_f(0u));
}

// Path regeneration, per stripe.  The free slots of stripe s (all of them at
// a chunk's start, else the slots whose paths ended this iteration, compact
// in the stripe's terminated list) take new paths and are appended to the
// stripe's next live list after its surviving paths.  At a chunk's start
// stripe s takes chunk paths s, s + 64, s + 128, ... (balanced stripes);
// later, the stripes' free slots in stripe order take the chunk's next
// paths in order.  No atomics: k_regen_commit advances the counters after.
// Called by whole wavefronts (lane j reads stripe j's free count).
struct RegenPlan {
    uint32_t next;   // chunk paths handed out before this regeneration
    uint32_t n_new;  // paths handed out now (all stripes)
    uint32_t take;   // paths stripe s takes
    uint32_t pre;    // chunk index of its first path - next (later iterations)
};

__device__ __forceinline__ RegenPlan regen_plan(const QueuesDev &q, const WaveParams &wp, uint32_t init,
                                                uint32_t capacity, uint32_t s)
{
    RegenPlan r;
    const uint32_t count = wp.chunk_dev ? (uint32_t)wp.chunk_dev[1] : wp.count;
    r.next = *scount(q, kCNext);
    const uint32_t left = r.next < count ? count - r.next : 0u;
    if (init) {
        r.n_new = capacity < left ? capacity : left;
        r.take = r.n_new > s ? (r.n_new - s + kStripes - 1) / kStripes : 0u;
        r.pre = 0;
    } else {
        const uint32_t nfree = *lcount(q, kLTerm, __lane_id());
        uint32_t total = 0;
        const uint32_t pre = wave_excl_scan(nfree, total);
        r.n_new = total < left ? total : left;
        r.pre = (uint32_t)__shfl((int)pre, (int)s);
        const uint32_t f = (uint32_t)__shfl((int)nfree, (int)s);
        r.take = left > r.pre ? (left - r.pre < f ? left - r.pre : f) : 0u;
    }
    return r;
}

template <bool B>
__global__ __launch_bounds__(kBlock) void k_regen(SceneDev sc0, BatchDev bt, PathsDev p, QueuesDev q, WaveParams wp,
                                                  uint32_t nxt, uint32_t init, uint32_t capacity)
{
    const uint64_t first = wp.chunk_dev ? wp.chunk_dev[0] : wp.first_path;
    const StripeSpan sp = stripe_span();
    const uint32_t s = sp.s, cs = q.cs;
    const RegenPlan r = regen_plan(q, wp, init, capacity, s);
    const uint32_t n_live = *lcount(q, kLActive0 + nxt, s);
    for (uint32_t j = sp.j0 + threadIdx.x; j < r.take; j += sp.step) {
        const uint32_t slot = init ? s * cs + j : q.term[s * cs + j];
        const uint32_t k = r.next + (init ? j * kStripes + s : r.pre + j);
        init_path<B>(sc0, bt, p, wp, slot, first + k, k);
        q.active[nxt][s * cs + n_live + j] = slot;
    }
}

// One wavefront, lane s = stripe s: the stripes' live counts grow by what
// they took; the next unclaimed path and the total live count (for the
// host's end-of-chunk test) are updated.
__global__ __launch_bounds__(kStripes) void k_regen_commit(QueuesDev q, WaveParams wp, uint32_t nxt, uint32_t init,
                                                           uint32_t capacity)
{
    const uint32_t s = threadIdx.x;
    const RegenPlan r = regen_plan(q, wp, init, capacity, s);
    uint32_t *live = lcount(q, kLActive0 + nxt, s);
    const uint32_t n = *live + r.take;
    const uint32_t total = wave_sum(n);
    *live = n;
    if (s == 0) {
        *scount(q, kCNext) = r.next + r.n_new;
        *scount(q, kCLive) = total;
    }
}

// ---------------------------------------------------------------------
// extend: closest hit + classification
// ---------------------------------------------------------------------
// A vertex the radiance cache may stand for: diffuse, or glossy with GGX
// alpha >= 0.2 (roughness = alpha^2).  Specular and sharper glossy vertices
// keep tracing, so the cache only ever answers smooth view dependence.
__device__ __forceinline__ bool cacheable(const rt_material &m)
{
    return m.type == RT_MAT_LAMBERTIAN ||
           ((m.type == RT_MAT_TROWBRIDGE || m.type == RT_MAT_TROWBRIDGE_BASIC) && m.roughness >= 0.04f);
}

// Photon mapping: the scattering hits whose photon estimate can be nonzero
// (Lambertian and Trowbridge; a plain surface reflects no photons, so it
// keeps its emissive class and ends there as well).
__device__ __forceinline__ bool photon_vertex(const rt_material &m)
{
    return m.type == RT_MAT_LAMBERTIAN || m.type == RT_MAT_TROWBRIDGE || m.type == RT_MAT_TROWBRIDGE_BASIC;
}

__device__ __forceinline__ uint32_t classify(const SceneDev &sc, int32_t hid, bool cache_depth_reached,
                                             bool photon = false)
{
    if (hid == -1) return kQMiss;
    const uint32_t k = (uint32_t)(-hid - 2);
    const uint32_t mi = hid >= 0 ? sc.tri_mat[hid] : k < sc.nsph ? sc.sph[k].material : sc.plane[k - sc.nsph].material;
    if (cache_depth_reached && cacheable(sc.mat[mi])) return kQCache;
    if (photon && photon_vertex(sc.mat[mi])) return kQCache;
    const uint32_t t = sc.mat[mi].type;
    if (t == RT_MAT_LAMBERTIAN) return kQLambert;
    // the specular non-scattering class: dielectric and mirror
    if (t == RT_MAT_DIELECTRIC || t == RT_MAT_MIRROR) return kQDielectric;
    // the emissive terminal class: light and plain (both end the path)
    if (t == RT_MAT_LIGHT || t == RT_MAT_PLAIN) return kQLight;
    return kQGgx;
}

// Persistent closest hit against the BVH: lanes claim active paths (their
// workgroup's stripe first, then any stripe with paths left) and walk them
// one node visit per step (rt_device.h), so a few long rays do not hold the
// rest of the wavefront idle.  The analytic shapes and the material
// classification follow in k_classify, which keeps this loop's register
// footprint at the traversal's.  STATS: the counting variant (node visits and
// triangle tests summed into the stats words), never the timed kernel.
template <bool B, bool WIDE, bool STATS>
__global__ __launch_bounds__(kBlock) void k_extend(SceneDev sc0, BatchDev bt, PathsDev p, QueuesDev q, uint32_t cur,
                                                   int cap)
{
    extern __shared__ int32_t s_stack[];
    StackTail<PathStack> s_tail;
    const PathStack stack{s_stack + threadIdx.x, s_tail};
    const uint32_t *active = q.active[cur];
    const uint32_t cs = q.cs;
    auto count = [&](uint32_t j) { return *lcount(q, kLActive0 + cur, j); };
    StripeCursor c;
    cursor_init(c, blockIdx.x % kStripes, count);
    TravState s;
    bool has = false;
    uint32_t slot = 0, n_node = 0, n_tri = 0;
    // a batch traverses its pooled trees from each path's scene root
    const float4 *nodes = (const float4 *)(B ? bt.nodes : sc0.bvh.nodes);
    const float4 *tris = (const float4 *)(B ? bt.tris : sc0.bvh.tris);
    for (;;) {
        uint32_t stripe = 0;
        const uint32_t r = wave_claim(c, lcount(q, kLClaimExt, 0), count, has, stripe);
        if (r != 0xffffffffu) {
            slot = active[stripe * cs + r];
            const float4 *rec = path_rec(p, slot);
            const int32_t root = B ? (int32_t)*path_word(p, slot, kPqScene, 1) : 0;
            const float4 a = rec[kPqRay];
            const float2 b = *(const float2 *)(rec + kPqHit);
            trav_begin(s, a.x, a.y, a.z, a.w, b.x, b.y, INFINITY, false, root);
            has = true;
        }
        if (__ballot(has) == 0) {
            if (cursor_finished(c)) break;
            continue;
        }
        if (has && trav_step_any<WIDE, false, STATS>(s, nodes, tris, stack, cap, n_node, n_tri)) {
            ((float2 *)(path_rec(p, slot) + kPqHit))[1] = make_float2(s.thit, as_f(s.hit));
            has = false;
        }
    }
    if (STATS) {  // the loop ends wavefront-uniformly
        const uint32_t a = wave_sum(n_node), b = wave_sum(n_tri);
        if (__lane_id() == 0) {
            atomicAdd(&q.stats[kStatExtNodes], (unsigned long long)a);
            atomicAdd(&q.stats[kStatExtTris], (unsigned long long)b);
        }
    }
}

// Analytic shapes after the BVH triangle, strict < (Scene.cu:221-227,
// Sphere.cu:321), then material classification into the stripe's
// per-class queues with wavefront ballots (one atomic per 64 lanes per
// queue, on the stripe's counter).
template <bool B>
__global__ __launch_bounds__(kBlock) void k_classify(SceneDev sc0, BatchDev bt, PathsDev p, QueuesDev q, uint32_t cur,
                                                     uint32_t cache_depth, uint32_t need_record, uint32_t photon)
{
    const StripeSpan sp = stripe_span();
    const uint32_t cs = q.cs, base = sp.s * cs;
    const uint32_t n = *lcount(q, kLActive0 + cur, sp.s);
    // every lane of a wavefront runs the same number of iterations so the
    // ballots below see the whole wavefront
    for (uint32_t j0 = sp.j0; j0 < n; j0 += sp.step) {
        const uint32_t j = j0 + threadIdx.x;
        const bool live = j < n;
        const uint32_t slot = live ? q.active[cur][base + j] : 0;
        // a batch wavefront whose paths share one scene reads its record with
        // scalar loads (the ray-shape tests loop over that scene's shapes)
        auto body = [&](const SceneDev &sc, auto uni) -> uint32_t {
            constexpr bool UNI = decltype(uni)::value;
                float4 *rec = path_rec(p, slot);
                const float4 qh = rec[kPqHit];
                int32_t hid = (int32_t)bits(qh.w);
                float th = qh.z;
                if (sc.nsph || sc.nplane) {
                    const float4 qr = rec[kPqRay];
                    const V3 o = mk(qr.x, qr.y, qr.z), d = mk(qr.w, qh.x, qh.y);
                    bool changed = false;
                    for (uint32_t k = 0; k < sc.nsph; ++k) {
                        V3 so, sd;
                        const float ts = sphere_core(sc.sph[k], o, d, INFINITY, so, sd);
                        if (ts >= 0 && ts < th) {
                            th = ts;
                            hid = -(int32_t)k - 2;
                            changed = true;
                        }
                    }
                    for (uint32_t k = 0; k < sc.nplane; ++k) {
                        float u, v, dz;
                        const float tp = plane_core(uniform_ref<UNI>(sc.plane + k), o, d, INFINITY, u, v, dz);
                        if (tp > 0 && tp < th) {  // strict (Plane.cu:88-98)
                            th = tp;
                            hid = -(int32_t)(sc.nsph + k) - 2;
                            changed = true;
                        }
                    }
                    if (changed) ((float2 *)(rec + kPqHit))[1] = make_float2(th, as_f(hid));
                }
                // training paths end in the cache only after their training record
                const uint32_t meta = *path_word(p, slot, kPqBeta, 3);
                const bool deep = cache_depth != kNoCache && (meta & 0xffffu) >= cache_depth &&
                                  (!need_record || *path_word(p, slot, kPqNee, 3) != 0) && !(meta & kMetaHarvest);
                return classify(sc, hid, deep, photon == 1u);
        };
        uint32_t cls = 0xffu;
        if constexpr (B) {
            const uint32_t sid = *path_word(p, slot, kPqScene, 0);
            const uint32_t s0 = __builtin_amdgcn_readfirstlane(sid);
            if (__ballot(live && sid != s0) == 0) {
                const SceneDev &su = const_ref(bt.scenes + (s0 < bt.n ? s0 : 0u));
                if (live) cls = body(su, std::true_type{});
            } else if (live) {
                cls = body(bt.scenes[sid], std::false_type{});
            }
        } else if (live) {
            cls = body(sc0, std::true_type{});
        }
#pragma unroll
        for (uint32_t k = 0; k < kNumQ; ++k) {
            const uint32_t pos = wave_append(lcount(q, kLQueue + k, sp.s), cls == k);
            if (cls == k) q.q[k][base + pos] = slot;
        }
        if (photon == 2u) {
            // hybrid: the scattering hits are shaded as usual and also queried
            const bool pq = cls == kQLambert || cls == kQGgx;
            const uint32_t pos = wave_append(lcount(q, kLQueue + kQCache, sp.s), pq);
            if (pq) q.q[kQCache][base + pos] = slot;
        }
    }
}

// ---------------------------------------------------------------------
// shading
// ---------------------------------------------------------------------
// The radiance L after adding the previous bounce's unoccluded next-event
// contributions (mask: their light slots), in light order, times the
// throughput nb they were computed with.
__device__ __forceinline__ V3 resolved_L(const PathsDev &p, uint32_t slot, uint32_t mask, V3 L, V3 nb)
{
    if (mask == 0) return L;
    V3 Ld = mk(0, 0, 0);
    const float *c = p.nee_c + (size_t)slot * p.nee_stride * 3;
    while (mask) {
        const uint32_t l = __ffs(mask) - 1;
        mask &= mask - 1;
        Ld = add(Ld, mk(c[3 * l], c[3 * l + 1], c[3 * l + 2]));
    }
    return mk(L.x + nb.x * Ld.x, L.y + nb.y * Ld.y, L.z + nb.z * Ld.z);
}

// resolved_L in place: the record's radiance takes the contributions and its
// mask is cleared.
__device__ __forceinline__ void resolve_nee(const PathsDev &p, uint32_t slot)
{
    float4 *rec = path_rec(p, slot);
    const uint32_t mask = bits(rec[kPqRng].w);
    if (mask == 0) return;
    const float4 qL = rec[kPqL], qn = rec[kPqNee];
    const V3 L = resolved_L(p, slot, mask, mk(qL.x, qL.y, qL.z), mk(qn.x, qn.y, qn.z));
    rec[kPqL] = make_float4(L.x, L.y, L.z, qL.w);
    *path_word(p, slot, kPqRng, 3) = 0;
}

// Shadow-ray occlusion by the analytic shapes (Scene.cu:108-150): opaque
// spheres other than `skip` (the sampled light) and opaque planes hit at
// t <= tmax.
__device__ __forceinline__ bool spheres_occlude(const SceneDev &sc, V3 o, V3 d, float tmax, int32_t skip)
{
    for (uint32_t j = 0; j < sc.nsph; ++j) {
        if ((int32_t)j == skip || sc.sph[j].transparent) continue;
        V3 so, sd;
        const float ts = sphere_core(sc.sph[j], o, d, INFINITY, so, sd);
        if (ts >= 0 && !(ts > tmax)) return true;
    }
    for (uint32_t k = 0; k < sc.nplane; ++k) {
        if (sc.plane[k].transparent) continue;
        float u, v, dz;
        const float tp = plane_core(sc.plane[k], o, d, INFINITY, u, v, dz);
        if (tp > 0 && !(tp > tmax)) return true;
    }
    return false;
}

// Radiance-cache query record (12 floats): position, the normal facing the
// incoming ray, the direction back along it, a roughness feature (diffuse 1,
// GGX alpha), padding.
__device__ __forceinline__ void write_record(float *r, const Surface &s, V3 d, const rt_material &m)
{
    r[0] = s.p.x; r[1] = s.p.y; r[2] = s.p.z;
    r[3] = s.n.x; r[4] = s.n.y; r[5] = s.n.z;
    r[6] = -d.x; r[7] = -d.y; r[8] = -d.z;
    r[9] = m.type == RT_MAT_LAMBERTIAN ? 1.0f : sqrtf(m.roughness);
    r[10] = 0; r[11] = 0;
}

template <uint32_t CLS, bool B>
__global__ __launch_bounds__(kBlock) void k_shade(SceneDev sc0, BatchDev bt, PathsDev p, ShadowDev sh, QueuesDev q,
                                                  WaveParams wp, uint32_t cur)
{
    constexpr bool kScatter = CLS == kQLambert || CLS == kQGgx;
    const StripeSpan sp = stripe_span();
    const uint32_t cs = q.cs, base = sp.s * cs;
    const uint32_t n = *lcount(q, kLQueue + CLS, sp.s);
    const uint32_t *queue = q.q[CLS] + base;
    const size_t sbase = (size_t)base * p.nee_stride;  // the stripe's shadow rays
    const uint32_t nxt = cur ^ 1u;
    // Every lane of a wavefront runs the same iterations (and the same light
    // loops: kernel arguments, or in a batch the wavefront's maximum) so the
    // ballots in wave_append see whole wavefronts.
    for (uint32_t j0 = sp.j0; j0 < n; j0 += sp.step) {
        const uint32_t i = j0 + threadIdx.x;
        const bool live = i < n;
        const uint32_t slot = live ? queue[i] : 0;
        // a batch wavefront whose paths share one scene reads that scene's
        // record with scalar loads; a mixed one per lane
        auto body = [&](const SceneDev &sc, auto uni) {
            constexpr bool UNI = decltype(uni)::value;
            float4 *rec = path_rec(p, slot);
            bool alive = false;
            Surface s{};
            V3 d = mk(0, 0, 1), t = mk(1, 0, 0), b = mk(0, 1, 0), wo = mk(0, 0, 1);
            rt_material m{};
            Pcg32 g{0, 1};
            uint32_t meta = 0, depth = 0, tvalid = 0, emit = 0;
            V3 beta = mk(0, 0, 0), L = mk(0, 0, 0);
            float4 qh{}, qL{}, qg{};
            if (live) {
                // the previous bounce's next-event contributions first
                qL = rec[kPqL];
                qg = rec[kPqRng];
                const float4 qn = rec[kPqNee];
                L = resolved_L(p, slot, bits(qg.w), mk(qL.x, qL.y, qL.z), mk(qn.x, qn.y, qn.z));
                tvalid = bits(qn.w);
                if (CLS != kQMiss) {
                    const float4 qr = rec[kPqRay], qb = rec[kPqBeta];
                    qh = rec[kPqHit];
                    surface_of(sc, qr, qh, s, d);
                    m = sc.mat[s.mat];
                    if (sc.ntex) apply_texture(sc, s, m);
                    g = path_rng(qg, stream_base_of<B>(sc, wp));
                    meta = bits(qb.w);
                    depth = meta & 0xffffu;
                    beta = mk(qb.x, qb.y, qb.z);
                    onb(s.n, t, b);
                    const V3 wow = neg(d);
                    wo = mk(dot(wow, t), dot(wow, b), dot(wow, s.n));
                    // training paths remember their first cacheable vertex at or
                    // beyond the cache depth: the radiance gathered from here on,
                    // divided by the throughput on arrival, is the cache's target
                    if (!B && kScatter && (wp.train || (meta & kMetaHarvest)) && depth >= wp.cache_depth &&
                        tvalid == 0 && cacheable(m)) {
                        write_record(p.tin + (size_t)slot * 12, s, d, m);
                        p.tLb[3 * (size_t)slot] = L.x; p.tLb[3 * (size_t)slot + 1] = L.y; p.tLb[3 * (size_t)slot + 2] = L.z;
                        p.tB[3 * (size_t)slot] = beta.x; p.tB[3 * (size_t)slot + 1] = beta.y; p.tB[3 * (size_t)slot + 2] = beta.z;
                        tvalid = 1;
                    }
                }
            }
            if (kScatter) {
                // next-event estimation over every point light (Scene.cu:468-525)
                // the shadow rays' tree root (batches)
                const int32_t proot = B && live ? (int32_t)*path_word(p, slot, kPqScene, 1) : 0;
                float *c = p.nee_c + (size_t)slot * p.nee_stride * 3;
                const uint32_t npl = live ? sc.npl : 0u, nal = live ? sc.nlights - sc.npl : 0u;
                const uint32_t npl_w = UNI ? sc.npl : wave_max(npl);
                const uint32_t nal_w = UNI ? sc.nlights - sc.npl : wave_max(nal);
                for (uint32_t l = 0; l < npl_w; ++l) {
                    bool e = false;
                    V3 Lw = mk(0, 0, 1), o2 = mk(0, 0, 0);
                    float dist = 0;
                    if (live && l < npl) {
                        const rt_point_light &pl = uniform_ref<UNI>(sc.pl + l);
                        const V3 lp = mk(pl.pos[0], pl.pos[1], pl.pos[2]);
                        const V3 tol = sub(lp, s.p);
                        Lw = nrm(tol);
                        const float d2 = dot(tol, tol);
                        const float I = pl.intensity;
                        const V3 Li = mk(pl.color[0] * I * (1 / d2), pl.color[1] * I * (1 / d2), pl.color[2] * I * (1 / d2));
                        if (dot(Lw, s.n) > 0) {  // isOrientedTo (SurfaceInteraction.cu:283)
                            const V3 wi = mk(dot(Lw, t), dot(Lw, b), dot(Lw, s.n));
                            const V3 f = bsdf_f_of<CLS == kQLambert>(m, wo, wi);
                            const float pf = bsdf_pdf_of<CLS == kQLambert>(m, wo, wi);
                            if (!black(f) && !black(Li)) {
                                const float w = 1.0f / (1.0f + pf * pf);  // powerHeuristic(1, 1, 1, pf)
                                c[3 * l] = f.x * Li.x * w * wp.nee_scale;
                                c[3 * l + 1] = f.y * Li.y * w * wp.nee_scale;
                                c[3 * l + 2] = f.z * Li.z * w * wp.nee_scale;
                                dist = len(tol);
                                o2 = spawn_origin(s.p, s.perr, s.n, Lw);
                                e = !spheres_occlude(sc, o2, Lw, dist, -1);
                            }
                        }
                    }
                    const size_t pos = sbase + wave_append(lcount(q, kLShadow, sp.s), e);
                    if (e) {
                        emit |= 1u << l;
                        sh.ox[pos] = o2.x; sh.oy[pos] = o2.y; sh.oz[pos] = o2.z;
                        sh.dx[pos] = Lw.x; sh.dy[pos] = Lw.y; sh.dz[pos] = Lw.z;
                        sh.tmax[pos] = dist;
                        sh.slot[pos] = (slot << 5) | l;
                        if constexpr (B) sh.root[pos] = proot;
                    }
                }
                for (uint32_t a = 0; a < nal_w; ++a) {
                    uint32_t j = 0, l = 0;
                    bool e = false;
                    V3 Lw = mk(0, 0, 1), o2 = mk(0, 0, 0);
                    float dist = 0;
                    if (live && a < nal) {
                        j = sc.area_light[a];
                        l = sc.npl + a;
                        const float a0 = pcg32_float(g), a1 = pcg32_float(g);
                        float plp;
                        if (sphere_light_sample(sc.sph[j], s.p, a0, a1, Lw, dist, plp) && dot(Lw, s.n) > 0) {
                            const V3 wi = mk(dot(Lw, t), dot(Lw, b), dot(Lw, s.n));
                            const V3 f = bsdf_f_of<CLS == kQLambert>(m, wo, wi);
                            if (!black(f)) {
                                const float pf = bsdf_pdf_of<CLS == kQLambert>(m, wo, wi);
                                const rt_material &lm = uniform_ref<UNI>(sc.mat + sc.sph[j].material);
                                const float w = plp * plp / (plp * plp + pf * pf);
                                const float k = w / plp;
                                c[3 * l] = f.x * lm.kd[0] * lm.intensity * k * wp.nee_scale;
                                c[3 * l + 1] = f.y * lm.kd[1] * lm.intensity * k * wp.nee_scale;
                                c[3 * l + 2] = f.z * lm.kd[2] * lm.intensity * k * wp.nee_scale;
                                o2 = spawn_origin(s.p, s.perr, s.n, Lw);
                                e = !spheres_occlude(sc, o2, Lw, dist * 0.9999f, (int32_t)j);
                            }
                        }
                    }
                    const size_t pos = sbase + wave_append(lcount(q, kLShadow, sp.s), e);
                    if (e) {
                        emit |= 1u << l;
                        sh.ox[pos] = o2.x; sh.oy[pos] = o2.y; sh.oz[pos] = o2.z;
                        sh.dx[pos] = Lw.x; sh.dy[pos] = Lw.y; sh.dz[pos] = Lw.z;
                        sh.tmax[pos] = dist * 0.9999f;
                        sh.slot[pos] = (slot << 5) | l;
                        if constexpr (B) sh.root[pos] = proot;
                    }
                }
                // the throughput these estimates were made with
                if (live) rec[kPqNee] = make_float4(beta.x, beta.y, beta.z, as_f(tvalid));
            }
            if (live && CLS != kQMiss) {
                if (CLS == kQLight && m.type == RT_MAT_PLAIN) {
                    // PlainMaterial::getEmittedRadiance (PlainMaterial.cu:62-75):
                    // kd towards camera rays, ke towards the others
                    const float *e = ((meta >> 17) & 1u) ? m.kd : m.ke;
                    L = add(L, mk(beta.x * e[0], beta.y * e[1], beta.z * e[2]));
                } else if (CLS == kQLight) {
                    // getEmittedRadiance (LightMaterial.cu:20-23), MIS-weighted
                    // against the light's own next-event sampling
                    float w = 1;
                    const int32_t hid = (int32_t)bits(qh.w);
                    if (hid <= -2 && (uint32_t)(-hid - 2) < sc.nsph && sc.sph[-hid - 2].is_light && !((meta >> 16) & 1u)) {
                        const float4 qp = rec[kPqPrev];
                        const float plp = sphere_light_pdf(sc.sph[-hid - 2], mk(qp.y, qp.z, qp.w));
                        const float pp = qp.x;
                        w = pp * pp / (pp * pp + plp * plp);
                    }
                    const float I = m.intensity;
                    L = add(L, mk(beta.x * m.kd[0] * I * w, beta.y * m.kd[1] * I * w, beta.z * m.kd[2] * I * w));
                }
                V3 wi;
                float pdf;
                bool delta;
                constexpr uint32_t MT = CLS == kQLambert ? RT_MAT_LAMBERTIAN
                                        : CLS == kQDielectric ? RT_MAT_DIELECTRIC
                                        : CLS == kQGgx ? RT_MAT_TROWBRIDGE : RT_MAT_LIGHT;
                const V3 wgt = bsdf_sample<MT>(m, wo, g, wi, pdf, delta);
                beta = hmul(beta, wgt);
                alive = !black(wgt);
                if (alive && kScatter && depth > wp.rr_depth) {
                    const float qq = fmaxf(0.05f, 1 - luminance(beta));
                    if (pcg32_float(g) < qq) {
                        alive = false;
                    } else {
                        const float sc1 = 1 - qq;
                        beta = mk(beta.x / sc1, beta.y / sc1, beta.z / sc1);
                    }
                }
                if (alive && depth + 1 >= wp.max_depth) alive = false;
                uint32_t meta_out = meta;
                if (alive) {
                    const V3 ww = add(add(mul(t, wi.x), mul(b, wi.y)), mul(s.n, wi.z));
                    const V3 no = spawn_origin(s.p, s.perr, s.n, ww);
                    rec[kPqRay] = make_float4(no.x, no.y, no.z, ww.x);
                    rec[kPqHit] = make_float4(ww.y, ww.z, qh.z, qh.w);
                    rec[kPqPrev] = make_float4(pdf, s.p.x, s.p.y, s.p.z);
                    // a camera ray stays one through specular hits (dielectric,
                    // mirror, Trowbridge: SurfaceInteraction.cu:252-254)
                    const uint32_t camera = (CLS != kQLambert) ? (meta >> 17) & 1u : 0u;
                    meta_out = (depth + 1) | ((delta ? 1u : 0u) << 16) | (camera << 17) | (meta & kMetaHarvest);
                }
                rec[kPqBeta] = make_float4(beta.x, beta.y, beta.z, as_f(meta_out));
                qg.x = as_f((uint32_t)g.state);
                qg.y = as_f((uint32_t)(g.state >> 32));
            }
            if (live) {
                // radiance, generator state and the new next-event mask (none
                // outside the scattering classes: the old one is resolved)
                rec[kPqL] = make_float4(L.x, L.y, L.z, qL.w);
                rec[kPqRng] = make_float4(qg.x, qg.y, qg.z, as_f(emit));
            }
            const uint32_t ppos = wave_append(lcount(q, kLActive0 + nxt, sp.s), live && alive);
            if (live && alive) q.active[nxt][base + ppos] = slot;
            const uint32_t tpos = wave_append(lcount(q, kLTerm, sp.s), live && !alive);
            if (live && !alive) q.term[base + tpos] = slot;
        };
        if constexpr (B) {
            const uint32_t sid = *path_word(p, slot, kPqScene, 0);
            const uint32_t s0 = __builtin_amdgcn_readfirstlane(sid);
            if (__ballot(live && sid != s0) == 0) body(const_ref(bt.scenes + (s0 < bt.n ? s0 : 0u)), std::true_type{});
            else body(bt.scenes[sid], std::false_type{});
        } else {
            body(sc0, std::true_type{});
        }
    }
}

// Cache vertices: the path ends here.  Its previous next-event estimate is
// resolved and its query record is written to the iteration's query rows;
// the cache's answer times the throughput on arrival is added after the
// bounce loop (k_cache_apply).  The rows are dense in stripe order: stripe
// s's queries start at the sum of the lower stripes' cache-queue counts,
// so no counter is shared.
__global__ __launch_bounds__(kBlock) void k_cache_vertex(SceneDev sc, PathsDev p, QueuesDev q)
{
    const StripeSpan sp = stripe_span();
    const uint32_t base = sp.s * q.cs;
    const uint32_t n = *lcount(q, kLQueue + kQCache, sp.s);
    uint32_t total = 0;
    const uint32_t pre = wave_excl_scan(*lcount(q, kLQueue + kQCache, __lane_id()), total);
    const uint32_t row0 = (uint32_t)__shfl((int)pre, (int)sp.s);
    if (blockIdx.x == 0 && threadIdx.x == 0) *scount(q, kCCacheQ) = total;
    for (uint32_t j0 = sp.j0; j0 < n; j0 += sp.step) {
        const uint32_t i = j0 + threadIdx.x;
        const bool live = i < n;
        const uint32_t slot = live ? q.q[kQCache][base + i] : 0;
        Surface s{};
        V3 d = mk(0, 0, 1);
        if (live) {
            resolve_nee(p, slot);
            const float4 *rec = path_rec(p, slot);
            surface_of(sc, rec[kPqRay], rec[kPqHit], s, d);
        }
        const uint32_t k = row0 + i;
        if (live) {
            write_record(p.cqin + (size_t)k * 12, s, d, sc.mat[s.mat]);
            p.cslot[k] = slot;
        }
        const uint32_t tpos = wave_append(lcount(q, kLTerm, sp.s), live);
        if (live) q.term[base + tpos] = slot;
    }
}

// Photon vertices (photon mapping): the path ends at its first scattering
// hit.  Its query -- the hit point, the shading normal, the direction back
// along the ray, the material with the textured kd and the shape id in the
// photons' numbering (spheres, planes, then triangles) -- goes to the
// iteration's query rows (dense in stripe order, like the cache's), and
// k_photon_apply adds the throughput times the map's estimate
// (getPhotonMappingContribution, Scene.cu:724-746, added in
// computeDirectLightning, pathTracing.cu:186-204).  Lambertian and
// Trowbridge surfaces emit nothing, so nothing else is added here.
template <bool HYBRID>
__global__ __launch_bounds__(kBlock) void k_photon_vertex(SceneDev sc, PathsDev p, QueuesDev q, PhotonLaunch ph)
{
    const StripeSpan sp = stripe_span();
    const uint32_t base = sp.s * q.cs;
    const uint32_t n = *lcount(q, kLQueue + kQCache, sp.s);
    uint32_t total = 0;
    const uint32_t pre = wave_excl_scan(*lcount(q, kLQueue + kQCache, __lane_id()), total);
    const uint32_t row0 = (uint32_t)__shfl((int)pre, (int)sp.s);
    if (blockIdx.x == 0 && threadIdx.x == 0) *scount(q, kCCacheQ) = total;
    for (uint32_t j0 = sp.j0; j0 < n; j0 += sp.step) {
        const uint32_t i = j0 + threadIdx.x;
        const bool live = i < n;
        const uint32_t slot = live ? q.q[kQCache][base + i] : 0;
        if (live) {
            if (!HYBRID) resolve_nee(p, slot);
            const float4 *rec = path_rec(p, slot);
            Surface s{};
            V3 d = mk(0, 0, 1);
            surface_of(sc, rec[kPqRay], rec[kPqHit], s, d);
            const int32_t hid = (int32_t)bits(rec[kPqHit].w);
            const uint32_t k = row0 + i;
            if (HYBRID) {
                const float4 qb = rec[kPqBeta];
                ph.qbeta[3 * (size_t)k] = qb.x; ph.qbeta[3 * (size_t)k + 1] = qb.y; ph.qbeta[3 * (size_t)k + 2] = qb.z;
            }
            ph.qp[3 * (size_t)k] = s.p.x; ph.qp[3 * (size_t)k + 1] = s.p.y; ph.qp[3 * (size_t)k + 2] = s.p.z;
            ph.qn[3 * (size_t)k] = s.n.x; ph.qn[3 * (size_t)k + 1] = s.n.y; ph.qn[3 * (size_t)k + 2] = s.n.z;
            ph.qv[3 * (size_t)k] = -d.x; ph.qv[3 * (size_t)k + 1] = -d.y; ph.qv[3 * (size_t)k + 2] = -d.z;
            rt_material m = sc.mat[s.mat];
            apply_texture(sc, s, m);
            ph.qm[k] = m;
            ph.qs[k] = hid >= 0 ? (int32_t)(sc.nsph + sc.nplane) + hid : -hid - 2;
            ph.qslot[k] = slot;
        }
        if (!HYBRID) {
            const uint32_t tpos = wave_append(lcount(q, kLTerm, sp.s), live);
            if (live) q.term[base + tpos] = slot;
        }
    }
}

// L += beta * estimate for this iteration's photon queries: photon mapping,
// the paths ended at their photon vertex (beta is the record's) and are
// finalized right after; hybrid, L += beta on arrival * estimate / 2 after
// the vertex was shaded (its next-event terms are halved in k_shade).
template <bool HYBRID>
__global__ __launch_bounds__(kBlock) void k_photon_apply(PathsDev p, QueuesDev q, PhotonLaunch ph)
{
    const uint32_t n = *scount(q, kCCacheQ);
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
        float4 *rec = path_rec(p, ph.qslot[k]);
        const double *c = ph.qrgb + 3 * (size_t)k;
        const float4 qL = rec[kPqL];
        if (HYBRID) {
            const float *b = ph.qbeta + 3 * (size_t)k;
            rec[kPqL] = make_float4(qL.x + b[0] * ((float)c[0] * 0.5f), qL.y + b[1] * ((float)c[1] * 0.5f),
                                    qL.z + b[2] * ((float)c[2] * 0.5f), qL.w);
        } else {
            const float4 qb = rec[kPqBeta];
            rec[kPqL] = make_float4(qL.x + qb.x * (float)c[0], qL.y + qb.y * (float)c[1], qL.z + qb.z * (float)c[2],
                                    qL.w);
        }
    }
}

// ---------------------------------------------------------------------
// connect: any hit for shadow rays (Scene.cu:108-150, 238-276)
// ---------------------------------------------------------------------
// Persistent any hit of the shadow queue against the BVH (the analytic
// shapes were tested when the shadow ray was made, in k_shade); an occluded
// ray clears its light bit in the path's NEE mask.  STATS: the counting
// variant (node visits, triangle tests and occluded rays into the stats words).
template <bool B, bool WIDE, bool STATS>
__global__ __launch_bounds__(kBlock) void k_connect(SceneDev sc0, BatchDev bt, PathsDev p, ShadowDev sh, QueuesDev q,
                                                    int cap)
{
    extern __shared__ int32_t s_stack[];
    StackTail<PathStack> s_tail;
    const PathStack stack{s_stack + threadIdx.x, s_tail};
    const size_t sstride = (size_t)q.cs * p.nee_stride;  // shadow rays per stripe
    auto count = [&](uint32_t j) { return *lcount(q, kLShadow, j); };
    StripeCursor c;
    cursor_init(c, blockIdx.x % kStripes, count);
    TravState s;
    bool has = false;
    uint32_t sl = 0, n_node = 0, n_tri = 0, n_occ = 0;
    const float4 *nodes = (const float4 *)(B ? bt.nodes : sc0.bvh.nodes);
    const float4 *tris = (const float4 *)(B ? bt.tris : sc0.bvh.tris);
    for (;;) {
        uint32_t stripe = 0;
        const uint32_t r = wave_claim(c, lcount(q, kLClaimCon, 0), count, has, stripe);
        if (r != 0xffffffffu) {
            const size_t e = stripe * sstride + r;
            sl = sh.slot[e];
            const int32_t root = B ? sh.root[e] : 0;
            trav_begin(s, sh.ox[e], sh.oy[e], sh.oz[e], sh.dx[e], sh.dy[e], sh.dz[e], sh.tmax[e], true, root);
            has = true;
        }
        if (__ballot(has) == 0) {
            if (cursor_finished(c)) break;
            continue;
        }
        if (has && trav_step_any<WIDE, true, STATS>(s, nodes, tris, stack, cap, n_node, n_tri)) {
            if (s.occ) atomicAnd(path_word(p, sl >> 5, kPqRng, 3), ~(1u << (sl & 31u)));
            if (STATS) n_occ += s.occ ? 1u : 0u;
            has = false;
        }
    }
    if (STATS) {  // the loop ends wavefront-uniformly
        const uint32_t a = wave_sum(n_node), b = wave_sum(n_tri), o = wave_sum(n_occ);
        if (__lane_id() == 0) {
            atomicAdd(&q.stats[kStatShNodes], (unsigned long long)a);
            atomicAdd(&q.stats[kStatShTris], (unsigned long long)b);
            atomicAdd(&q.stats[kStatOccluded], (unsigned long long)o);
        }
    }
}

// ---------------------------------------------------------------------
// finalize: paths that ended this iteration take their last next-event
// contributions.  Rendering: the radiance goes to res[path index].
// Training: a path that passed a cacheable vertex at the cache depth yields
// the target (L_end - L_arrival) / beta_arrival per channel (0 where beta is
// 0), compacted with its record into cin / ttgt; non-finite targets are
// dropped.
// ---------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_finalize(PathsDev p, QueuesDev q, float *res, uint32_t train,
                                                    uint32_t harvest, uint32_t cap)
{
    const StripeSpan sp = stripe_span();
    const uint32_t n = *lcount(q, kLTerm, sp.s);
    const uint32_t *term = q.term + sp.s * q.cs;
    for (uint32_t j0 = sp.j0; j0 < n; j0 += sp.step) {
        const uint32_t i = j0 + threadIdx.x;
        const bool live = i < n;
        const uint32_t slot = live ? term[i] : 0;
        // the slot is refilled next, so the resolved radiance stays in registers
        V3 L = mk(0, 0, 0);
        uint32_t tvalid = 0;
        float4 qL{};
        if (live) {
            const float4 *rec = path_rec(p, slot);
            qL = rec[kPqL];
            const float4 qn = rec[kPqNee];
            L = resolved_L(p, slot, bits(rec[kPqRng].w), mk(qL.x, qL.y, qL.z), mk(qn.x, qn.y, qn.z));
            tvalid = bits(qn.w);
        }
        if (!train && live) {
            float *r = res + 3 * (size_t)bits(qL.w);
            r[0] = L.x; r[1] = L.y; r[2] = L.z;
        }
        if (!train && !harvest) continue;  // wavefront-uniform (kernel arguments)
        bool ok = false;
        float tg[3] = {0, 0, 0};
        if (live && tvalid) {
            const float Le[3] = {L.x, L.y, L.z};
            ok = true;
            for (int c = 0; c < 3; ++c) {
                const float b = p.tB[3 * (size_t)slot + c];
                tg[c] = b > 0 ? (Le[c] - p.tLb[3 * (size_t)slot + c]) / b : 0.0f;
                ok = ok && isfinite(tg[c]);
            }
        }
        const uint32_t k = wave_append(scount(q, kCTrain), ok);
        if (ok && k < cap) {
            const float4 *src = (const float4 *)(p.tin + (size_t)slot * 12);
            float4 *dst = (float4 *)(p.cin + (size_t)k * 12);
            dst[0] = src[0]; dst[1] = src[1]; dst[2] = src[2];
            p.ttgt[3 * (size_t)k] = fmaxf(tg[0], 0.0f);
            p.ttgt[3 * (size_t)k + 1] = fmaxf(tg[1], 0.0f);
            p.ttgt[3 * (size_t)k + 2] = fmaxf(tg[2], 0.0f);
        }
    }
}

// L += beta * cache(record) for this iteration's cache queries; the paths
// ended at their cache vertex and are finalized right after
__global__ __launch_bounds__(kBlock) void k_cache_apply(PathsDev p, QueuesDev q)
{
    const uint32_t n = *scount(q, kCCacheQ);
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
        const uint32_t slot = p.cslot[k];
        const float *c = p.crgb + 3 * (size_t)k;
        float4 *rec = path_rec(p, slot);
        const float4 qb = rec[kPqBeta], qL = rec[kPqL];
        rec[kPqL] = make_float4(qL.x + qb.x * c[0], qL.y + qb.y * c[1], qL.z + qb.z * c[2], qL.w);
    }
}

// film: FP64 accumulation with the reference's NaN/Inf sentinel bookkeeping
// (pathTracing.cu:474-486): only finite samples are summed and counted (the
// consumer rescales by samples / nvalid).  One thread per pixel sums that
// pixel's samples in sample order, so the result is bit-reproducible and
// needs no atomics.  Chunk paths [lo, hi) belong to one scene of npix
// pixels; path lo is that scene's pixel pix0.
__global__ __launch_bounds__(kBlock) void k_film(const float *res, FilmDev film, uint32_t lo, uint32_t hi,
                                                 uint32_t npix, uint32_t pix0)
{
    const uint32_t span = hi - lo;
    const uint32_t npx = span < npix ? span : npix;
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < npx; t += gridDim.x * kBlock) {
        double sx = 0, sy = 0, sz = 0;
        uint32_t nv = 0;
        for (uint32_t k = lo + t; k < hi; k += npix) {
            const float lx = res[3 * (size_t)k], ly = res[3 * (size_t)k + 1], lz = res[3 * (size_t)k + 2];
            if (isfinite(lx) && isfinite(ly) && isfinite(lz)) {
                sx += (double)lx;
                sy += (double)ly;
                sz += (double)lz;
                ++nv;
            }
        }
        uint32_t px = pix0 + t;
        if (px >= npix) px -= npix;
        film.accum[3 * (size_t)px] += sx;
        film.accum[3 * (size_t)px + 1] += sy;
        film.accum[3 * (size_t)px + 2] += sz;
        film.nvalid[px] += nv;
    }
}

// Ray statistics of an iteration (one wavefront, lane s = stripe s).
__global__ __launch_bounds__(kStripes) void k_count(QueuesDev q, uint32_t cur)
{
    const uint32_t s = threadIdx.x;
    const uint32_t ext = wave_sum(*lcount(q, kLActive0 + cur, s));
    const uint32_t shd = wave_sum(*lcount(q, kLShadow, s));
    if (s == 0) {
        q.stats[0] += ext;
        q.stats[1] += shd;
    }
}

__global__ void k_set_u32(uint32_t *p, uint32_t v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

hipError_t set_u32(uint32_t *p, uint32_t v, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, s, p, v);
    return hipGetLastError();
}

__global__ void k_set_u64(uint64_t *p, uint64_t v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

hipError_t set_u64(uint64_t *p, uint64_t v, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, p, v);
    return hipGetLastError();
}

size_t extend_lds_bytes(uint32_t stack_cap)
{
    return stack_lds_bytes<PathStack>(stack_cap, kBlock);
}

// Workgroups of a persistent traversal launch: as many as fit on the chip
// (LDS-limited, at most 16 per CU).
static dim3 persistent_grid(size_t lds)
{
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess || cus <= 0) cus = 256;
    }
    size_t per_cu = lds ? (160u * 1024u) / lds : 16;
    per_cu = per_cu > 16 ? 16 : (per_cu < 1 ? 1 : per_cu);
    return dim3((unsigned)(cus * per_cu));
}

// Workgroups of the queue kernels: every stripe needs at least one.
static dim3 queue_grid(const WaveParams &wp)
{
    const uint32_t g = wp.grid ? wp.grid : (uint32_t)kGrid;
    return dim3(g < kStripes ? kStripes : g);
}

// The chunk's queues with the stripe size of its capacity.
static hipError_t chunk_queues(const ChunkLaunch &c, QueuesDev &q)
{
    if (c.capacity == 0 || c.capacity % kStripes) return hipErrorInvalidValue;
    q = c.queues;
    q.cs = c.capacity / kStripes;
    return hipSuccess;
}

hipError_t chunk_begin(const ChunkLaunch &c, const WaveParams &wp, hipStream_t s)
{
    QueuesDev q;
    hipError_t e = chunk_queues(c, q);
    if (e != hipSuccess) return e;
    const dim3 blk(kBlock);
    const dim3 grid = queue_grid(wp);
    const SceneDev &sc = *c.scene;
    const BatchDev bt = c.batch ? c.batch->dev : BatchDev{nullptr, nullptr, 0, nullptr, nullptr, nullptr, nullptr};
    hipLaunchKernelGGL(k_reset, dim3(kNumLists + 1), dim3(kStripes), 0, s, q, 0u, 1u);
    if (c.batch) hipLaunchKernelGGL(k_regen<true>, grid, blk, 0, s, sc, bt, c.paths, q, wp, 0u, 1u, c.capacity);
    else hipLaunchKernelGGL(k_regen<false>, grid, blk, 0, s, sc, bt, c.paths, q, wp, 0u, 1u, c.capacity);
    hipLaunchKernelGGL(k_regen_commit, dim3(1), dim3(kStripes), 0, s, q, wp, 0u, 1u, c.capacity);
    return hipGetLastError();
}

hipError_t chunk_iterate(const ChunkLaunch &c, const WaveParams &wp, uint32_t iterations, hipStream_t s)
{
    const bool B = c.batch != nullptr;
    const SceneDev &sc = *c.scene;
    const BatchDev bt = B ? c.batch->dev : BatchDev{nullptr, nullptr, 0, nullptr, nullptr, nullptr, nullptr};
    const uint32_t stack_cap = B ? c.batch->stack_cap : sc.bvh.stack_cap;
    const bool wide = sc.bvh.wide != 0;  // a batch's scenes share one layout
    const int cap = (int)stack_cap;
    const size_t lds = extend_lds_bytes(stack_cap);
    const dim3 blk(kBlock);
    const dim3 grid = queue_grid(wp);
    const dim3 pgrid = persistent_grid(lds);
    const PathsDev &paths = c.paths;
    const ShadowDev &shadow = c.shadow;
    QueuesDev q;
    hipError_t qe = chunk_queues(c, q);
    if (qe != hipSuccess) return qe;
    // Rendering paths end at their first cacheable vertex at the cache depth.
    // Training paths record there, trace kTrainSuffix more bounces and end
    // at the next cacheable vertex, whose cache answer completes their
    // target (self-training on the cache's own prediction).
    const bool lookups = !B && wp.cache_depth != kNoCache && c.cache && c.cwork;
    const bool photon = !B && c.photon != nullptr;
    if (photon && (lookups || wp.train || wp.harvest_shift != kNoHarvest)) return hipErrorInvalidValue;
    // a cache depth without lookups: only a plain wave that harvests records
    if (!wp.train && wp.cache_depth != kNoCache && !lookups && wp.harvest_shift == kNoHarvest)
        return hipErrorInvalidValue;
    if (B && wp.train) return hipErrorInvalidValue;
    if (lookups && c.cwork->rows < c.capacity) return hipErrorInvalidValue;
    if (iterations & 1u) return hipErrorInvalidValue;
    const uint32_t ext_depth = !lookups ? kNoCache : (wp.train ? wp.cache_depth + kTrainSuffix : wp.cache_depth);
    const uint32_t need_record = wp.train ? 1u : 0u;
    const bool cnt = c.prof && c.prof->count;
    uint32_t cur = 0;
    // a traversal launch, bracketed by an event pair when profiling
    auto timed = [&](uint8_t kind, auto launch) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        const bool ev = c.prof && c.prof->events && c.prof->next(kind, e0, e1);
        if (ev) (void)hipEventRecord(e0, s);
        launch();
        if (ev) (void)hipEventRecord(e1, s);
    };
    auto extend = [&]() {
        auto go = [&](auto kern) { hipLaunchKernelGGL(kern, pgrid, blk, lds, s, sc, bt, paths, q, cur, cap); };
        if (B && wide) cnt ? go(k_extend<true, true, true>) : go(k_extend<true, true, false>);
        else if (B) cnt ? go(k_extend<true, false, true>) : go(k_extend<true, false, false>);
        else if (wide) cnt ? go(k_extend<false, true, true>) : go(k_extend<false, true, false>);
        else cnt ? go(k_extend<false, false, true>) : go(k_extend<false, false, false>);
    };
    auto connect = [&]() {
        auto go = [&](auto kern) { hipLaunchKernelGGL(kern, pgrid, blk, lds, s, sc, bt, paths, shadow, q, cap); };
        if (B && wide) cnt ? go(k_connect<true, true, true>) : go(k_connect<true, true, false>);
        else if (B) cnt ? go(k_connect<true, false, true>) : go(k_connect<true, false, false>);
        else if (wide) cnt ? go(k_connect<false, true, true>) : go(k_connect<false, true, false>);
        else cnt ? go(k_connect<false, false, true>) : go(k_connect<false, false, false>);
    };
    for (uint32_t it = 0; it < iterations; ++it) {
        const uint32_t nxt = cur ^ 1u;
        hipLaunchKernelGGL(k_reset, dim3(kNumLists + 1), dim3(kStripes), 0, s, q, cur, 0u);
        timed(0, extend);
        if (B) {
            hipLaunchKernelGGL(k_classify<true>, grid, blk, 0, s, sc, bt, paths, q, cur, ext_depth, need_record, 0u);
            hipLaunchKernelGGL((k_shade<kQLambert, true>), grid, blk, 0, s, sc, bt, paths, shadow, q, wp, cur);
            hipLaunchKernelGGL((k_shade<kQDielectric, true>), grid, blk, 0, s, sc, bt, paths, shadow, q, wp, cur);
            hipLaunchKernelGGL((k_shade<kQGgx, true>), grid, blk, 0, s, sc, bt, paths, shadow, q, wp, cur);
            hipLaunchKernelGGL((k_shade<kQLight, true>), grid, blk, 0, s, sc, bt, paths, shadow, q, wp, cur);
            hipLaunchKernelGGL((k_shade<kQMiss, true>), grid, blk, 0, s, sc, bt, paths, shadow, q, wp, cur);
        } else {
            hipLaunchKernelGGL(k_classify<false>, grid, blk, 0, s, sc, bt, paths, q, cur, ext_depth, need_record,
                               photon ? (c.photon->hybrid ? 2u : 1u) : 0u);
            // hybrid queries record the throughput on arrival, before shading
            if (photon && c.photon->hybrid)
                hipLaunchKernelGGL(k_photon_vertex<true>, grid, blk, 0, s, sc, paths, q, *c.photon);
            hipLaunchKernelGGL((k_shade<kQLambert, false>), grid, blk, 0, s, sc, bt, paths, shadow, q, wp, cur);
            hipLaunchKernelGGL((k_shade<kQDielectric, false>), grid, blk, 0, s, sc, bt, paths, shadow, q, wp, cur);
            hipLaunchKernelGGL((k_shade<kQGgx, false>), grid, blk, 0, s, sc, bt, paths, shadow, q, wp, cur);
            hipLaunchKernelGGL((k_shade<kQLight, false>), grid, blk, 0, s, sc, bt, paths, shadow, q, wp, cur);
            hipLaunchKernelGGL((k_shade<kQMiss, false>), grid, blk, 0, s, sc, bt, paths, shadow, q, wp, cur);
        }
        if (lookups) hipLaunchKernelGGL(k_cache_vertex, grid, blk, 0, s, sc, paths, q);
        if (photon && !c.photon->hybrid)
            hipLaunchKernelGGL(k_photon_vertex<false>, grid, blk, 0, s, sc, paths, q, *c.photon);
        hipLaunchKernelGGL(k_count, dim3(1), dim3(kStripes), 0, s, q, cur);
        timed(1, connect);
        if (lookups) {
            // cin rows are this iteration's queries (training records are
            // compacted into cin only by k_finalize, after this)
            hipError_t e = cache_infer(*c.cache, *c.cwork, paths.cqin, c.capacity, paths.crgb, s,
                                       scount(q, kCCacheQ));
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_cache_apply, grid, blk, 0, s, paths, q);
        }
        if (photon) {
            const PhotonLaunch &ph = *c.photon;
            hipError_t e = photon_gather_render(ph.map, ph.qp, ph.qn, ph.qv, ph.qm, ph.qs, c.capacity,
                                                scount(q, kCCacheQ), ph.k, ph.radius, ph.qrgb, s);
            if (e != hipSuccess) return e;
            if (ph.hybrid) hipLaunchKernelGGL(k_photon_apply<true>, grid, blk, 0, s, paths, q, ph);
            else hipLaunchKernelGGL(k_photon_apply<false>, grid, blk, 0, s, paths, q, ph);
        }
        const uint32_t harvest = wp.harvest_shift != kNoHarvest ? 1u : 0u;
        const uint32_t cap = wp.train ? c.capacity : (harvest ? wp.harvest_cap : 0u);
        hipLaunchKernelGGL(k_finalize, grid, blk, 0, s, paths, q, c.res, wp.train, harvest, cap);
        if (B) hipLaunchKernelGGL(k_regen<true>, grid, blk, 0, s, sc, bt, paths, q, wp, nxt, 0u, c.capacity);
        else hipLaunchKernelGGL(k_regen<false>, grid, blk, 0, s, sc, bt, paths, q, wp, nxt, 0u, c.capacity);
        hipLaunchKernelGGL(k_regen_commit, dim3(1), dim3(kStripes), 0, s, q, wp, nxt, 0u, c.capacity);
        cur = nxt;
    }
    return hipGetLastError();
}

hipError_t chunk_film(const ChunkLaunch &c, const FilmDev &film, const WaveParams &wp, hipStream_t s)
{
    if (wp.train) return hipSuccess;
    const dim3 blk(kBlock);
    const dim3 grid = queue_grid(wp);
    if (!c.batch) {
        hipLaunchKernelGGL(k_film, grid, blk, 0, s, c.res, film, 0u, wp.count, wp.npix,
                           (uint32_t)(wp.first_path % wp.npix));
    } else {
        // one film segment per scene that has paths in this chunk
        const BatchLaunch &b = *c.batch;
        const uint64_t w0 = wp.first_path, w1 = wp.first_path + wp.count;
        for (uint32_t sidx = 0; sidx < b.dev.n; ++sidx) {
            const uint64_t lo = b.bases[sidx] > w0 ? b.bases[sidx] : w0;
            const uint64_t hi = b.bases[sidx + 1] < w1 ? b.bases[sidx + 1] : w1;
            if (lo >= hi) continue;
            const uint32_t np = b.npix[sidx];
            hipLaunchKernelGGL(k_film, grid, blk, 0, s, c.res, b.films[sidx], (uint32_t)(lo - w0), (uint32_t)(hi - w0),
                               np, (uint32_t)((lo - b.bases[sidx]) % np));
        }
    }
    return hipGetLastError();
}

hipError_t render_chunk(const ChunkLaunch &c, const FilmDev &film, const WaveParams &wp, uint32_t *host_counters,
                        hipStream_t s)
{
    constexpr uint32_t kBlockIters = 4;
    hipError_t e = chunk_begin(c, wp, s);
    if (e != hipSuccess) return e;
    if (!host_counters) {
        // no read-back: every path is generated at the start (count <=
        // capacity) and needs at most max_depth iterations
        if (wp.count > c.capacity) return hipErrorInvalidValue;
        e = chunk_iterate(c, wp, (wp.max_depth + 1u) & ~1u, s);
        if (e != hipSuccess) return e;
        return chunk_film(c, film, wp, s);
    }
    // While unclaimed paths remain every slot is busy, and a slot holds a
    // path for at most max_depth iterations; the drain takes max_depth more.
    const uint64_t rounds = ((uint64_t)wp.count + c.capacity - 1) / c.capacity + 1;
    const uint64_t max_iters = rounds * wp.max_depth + kBlockIters;
    bool done = false;
    for (uint64_t it = 0; it < max_iters && !done; it += kBlockIters) {
        e = chunk_iterate(c, wp, kBlockIters, s);
        if (e != hipSuccess) return e;
        e = hipMemcpyAsync(host_counters, c.queues.counters, kCounterWords * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           s);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        done = host_counters[kCNext * kCtrStride] >= wp.count && host_counters[kCLive * kCtrStride] == 0;
    }
    if (!done) return hipErrorUnknown;  // cannot happen: the bound above holds
    return chunk_film(c, film, wp, s);
}

}  // namespace rt
