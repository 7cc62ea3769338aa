// rt_render.h -- device buffers and launch interface of the wavefront path
// tracer (render.hip), used by the host scene code (scene.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#include "rt_scene_types.h"
#include "rt_trace.h"

namespace rt {

// Queue classes filled by the extend stage (material-sorted shading).  kQCache
// takes every hit at the radiance cache's depth when a cache is attached.
enum : uint32_t { kQLambert = 0, kQDielectric = 1, kQGgx = 2, kQLight = 3, kQMiss = 4, kQCache = 5, kNumQ = 6 };

// Striped work lists.  Every list is cut into kStripes stripes of cs path
// slots (rt_device.h): stripe s of list L keeps its item count at
// sctr[(L * kStripes + s) * kCtrStride] and its items at [s * cs, (s + 1) * cs)
// of the list's array (the shadow list: cs * nee_stride entries per stripe).
// Path slots are striped the same way, and a queue kernel's workgroup works
// on stripe blockIdx % kStripes only, so what a stripe's paths produce stays
// in that stripe and can never overflow it.  kLClaimExt / kLClaimCon are the
// claim counters of the persistent extend / connect kernels.
enum : uint32_t {
    kLActive0 = 0, kLActive1 = 1, kLQueue = 2, /* 2..7 */ kLShadow = 8, kLTerm = 9, kLClaimExt = 10,
    kLClaimCon = 11, kNumLists = 12
};
// Scalar device counters, kCtrStride words apart: the chunk's next unclaimed
// path, the live paths after regeneration (summed over stripes, for the
// host's end-of-chunk test), the cache queries of an iteration and the
// training records of a chunk.
enum : uint32_t { kCNext = 0, kCLive = 1, kCCacheQ = 2, kCTrain = 3, kNumCounters = 4 };
constexpr uint32_t kCounterWords = kNumCounters * kCtrStride;
constexpr uint32_t kStripedWords = kNumLists * kStripes * kCtrStride;
constexpr uint32_t kNoCache = 0xffffffffu;
// bounces a training path traces past its training record before it ends
// in the cache
constexpr uint32_t kTrainSuffix = 2;

// Scene as the kernels see it (passed by value).
struct SceneDev {
    DevBvh bvh;
    const float *tri_xyz;      // 9 floats per triangle id
    const uint32_t *tri_mat;   // material per triangle id
    const rt_material *mat;
    const rt_point_light *pl;
    SphereDev sph[kMaxSpheres];
    const PlaneDev *plane;  // nplane planes, tested after the spheres
    const TexDev *tex;      // ntex textures over `texels`
    const uint32_t *texels;
    const float *tri_uv;    // 6 floats per triangle, or null: the default (0,0), (1,0), (1,1)
    const float *tri_nrm;   // 9 floats per triangle (corner normals, NaN: none), or null
    CameraDev cam;
    uint32_t nsph, npl, nlights;  // nlights = npl + area lights (NEE slots per path)
    uint32_t area_light[kMaxSpheres];  // sphere index of area light a
    uint32_t nmat;
    uint32_t nplane;
    uint32_t ntex;
    // batched rendering: the scene's own PCG seed and stream offset (a
    // single-scene wave takes both from WaveParams)
    uint32_t seed;
    uint64_t stream_base;
};

// Scenes of a batch as the kernels see them: every path carries its scene
// index; scene s owns global path indices [bases[s], bases[s + 1]).
struct BatchDev {
    const SceneDev *scenes;
    const uint64_t *bases;  // n + 1 entries
    uint32_t n;
    // the scenes' trees pooled into one node array and one triangle array
    // (child and leaf references rebased), so traversal reads uniform base
    // pointers; scene s starts at node root[s]
    const void *nodes;
    const void *tris;
    const uint32_t *root;
    // scene of global path index g << kHintShift: where a search for the
    // scene of g starts (it moves forward through bases from there)
    const uint32_t *hint;
};
constexpr uint32_t kHintShift = 16;

// Structure-of-arrays path state (capacity P).
// Path state: one 128-byte record (kPathQuads float4, one cache line) per
// slot.  The kernels of a bounce reach paths through queues whose slots are
// scattered, so a path's fields share a line: a kernel that touches a path
// moves whole lines instead of a 32- or 64-byte sector per 4-byte field.
// Quads of a record (words 4q .. 4q+3):
//   kPqRay   origin xyz, direction x
//   kPqHit   direction yz, hit t, hit id (int bits: >= 0 triangle, -1 miss,
//            <= -2 analytic shape -(k+2): sphere k < nsph, else plane k - nsph)
//   kPqBeta  throughput xyz, meta (depth | prev_delta << 16 | camera << 17 | harvest << 18)
//   kPqL     radiance xyz, path index within the chunk (result slot)
//   kPqRng   PCG32 state (low, high word), pixel, NEE mask (unoccluded light
//            slots of the last next-event estimation)
//   kPqPrev  previous BSDF pdf, previous vertex xyz (MIS of sphere lights)
//   kPqNee   throughput at the last next-event estimation xyz, training
//            record valid
//   kPqScene scene of the path, root node of its scene in the batch's pooled
//            tree (batched waves), two unused words
enum : uint32_t {
    kPqRay = 0, kPqHit = 1, kPqBeta = 2, kPqL = 3, kPqRng = 4, kPqPrev = 5, kPqNee = 6, kPqScene = 7,
    kPathQuads = 8
};

struct PathsDev {
    float4 *rec;                // kPathQuads per slot
    float *nee_c;               // P * nee_stride * 3: per-light f * Li * w
    uint32_t nee_stride;        // light slots per path (max over a batch's scenes)
    // radiance cache: query records (P x 12) and answers (P x 3) of one
    // iteration, in queue order, and the path slot of each query
    float *cqin, *crgb;
    uint32_t *cslot;
    float *cin;                 // compacted training records (P x 12)
    // training paths: the record at the cache depth, L and beta on arrival
    float *tin;                 // P x 12
    float *tLb, *tB;            // P x 3 each
    float *ttgt;               // compacted targets (P x 3); records compacted into cin
};

struct ShadowDev {
    float *ox, *oy, *oz, *dx, *dy, *dz, *tmax;
    uint32_t *slot;             // path slot << 5 | light slot
    int32_t *root;              // batches: root node of the ray's scene in the pooled tree
};

struct QueuesDev {
    uint32_t *active[2];
    uint32_t *q[kNumQ];
    uint32_t *term;
    uint32_t *sctr;             // kStripedWords: striped list counters
    uint32_t *counters;         // kCounterWords: scalar counters
    unsigned long long *stats;  // kNumStats words (kStat*)
    uint32_t cs;                // path slots per stripe (capacity / kStripes)
};

__host__ __device__ inline uint32_t *lcount(const QueuesDev &q, uint32_t list, uint32_t stripe)
{
    return q.sctr + ((size_t)list * kStripes + stripe) * kCtrStride;
}
__host__ __device__ inline uint32_t *scount(const QueuesDev &q, uint32_t id)
{
    return q.counters + (size_t)id * kCtrStride;
}
// Path slots of a wave of `requested` paths: a whole number per stripe.
inline uint32_t stripe_capacity(uint32_t requested)
{
    const uint32_t r = requested ? requested : 1u;
    return (r + kStripes - 1) / kStripes * kStripes;
}

struct FilmDev {
    double *accum;              // W*H*3
    uint32_t *nvalid;           // W*H
};

struct WaveParams {
    uint64_t first_path;        // global path index of the chunk's first path
    uint32_t count;             // paths in the chunk
    uint32_t npix;              // W*H
    uint32_t spp_offset, seed, max_depth, rr_depth;
    uint64_t stream_base;       // PCG stream offset (scene id for batches)
    uint32_t cache_depth;       // hits at this depth end in the cache (kNoCache: off)
    uint32_t train;             // 1: training wave (random pixels, no film, records at cache_depth)
    uint32_t grid;              // workgroups of the queue kernels (0: default)
    const uint32_t *spp_offset_dev;  // if set, the first sample index is read here (graph replays)
    const uint64_t *chunk_dev;  // if set, {first_path, count} are read here (graph replays)
    // harvesting (rendering waves): one path in 2^harvest_shift (by a hash of
    // its global index; kNoHarvest: none) ignores the cache, and its first
    // cacheable vertex at cache_depth or deeper becomes a training record;
    // k_finalize compacts up to harvest_cap records into cin / ttgt
    uint32_t harvest_shift = 0xffffffffu;  // kNoHarvest
    uint32_t harvest_cap = 0;
    float nee_scale = 1.0f;     // factor of every next-event term (RENDER_MODE HYBRID: 0.5)
};
constexpr uint32_t kNoHarvest = 0xffffffffu;
// path meta word: depth (bits 0-15), previous bounce delta (16), camera ray
// (17), harvest path (18)
constexpr uint32_t kMetaHarvest = 1u << 18;

// Host view of a batch for one wave launch: the device tables plus, per
// scene, the path-index base (n + 1 entries), pixel count and film.
struct BatchLaunch {
    BatchDev dev;
    const uint64_t *bases;
    const uint32_t *npix;
    const FilmDev *films;
    uint32_t stack_cap;         // largest traversal stack of the batch's trees (LDS)
};

struct CacheDev;
struct CacheWork;

// Photon mapping in the bounce loop (integration/Scene.cu:912-967): with
// RENDER_MODE PHOTON_MAPPING a path ends at its first scattering hit
// (pathTracing.cu:84-88) and its estimate there is the photon map's
// density estimate instead of next-event estimation; with HYBRID every
// scattering hit takes half of each and the path goes on.  The map (kd-tree of
// photon locations and shapes), each photon's radiance and direction, the
// search's k and radius, and per wave slot the query buffers.
struct PhotonLaunch {
    rt_photon_map *map;
    const double *rgb;          // 3 per photon
    const float *dir;           // 3 per photon
    uint32_t k;
    float radius;
    float *qp, *qn, *qv;        // query point, shading normal, direction to the viewer (3 each per slot)
    rt_material *qm;            // query material (textured kd)
    int32_t *qs;                // query shape id (the photons' numbering)
    uint32_t *qslot;            // the query's path slot
    double *qrgb;               // the estimate (3 per query)
    float *qbeta;               // hybrid: the throughput on arrival (3 per query)
    uint32_t hybrid;            // RENDER_MODE HYBRID: (estimate + next-event) / 2, the path goes on
};

// Optional measurement of a frame's traversal kernels.  With events, every
// k_extend / k_connect launch is bracketed by a pair of HIP events on the
// launch stream (pairs are reused across frames; the host sums their elapsed
// times after the frame).  With count, the kernels run their counting
// variants, which add node visits, triangle tests and occluded shadow rays to
// the queues' stats words 2..6 (never the timed kernels).
struct KernelProfile {
    bool events = false;
    bool count = false;
    std::vector<hipEvent_t> ev;    // pairs: before, after
    std::vector<uint8_t> kind;     // per pair: 0 extend, 1 connect
    size_t used = 0;               // pairs recorded this frame
    // the next pair for a launch of `k` (created on first use); false if
    // an event could not be created
    bool next(uint8_t k, hipEvent_t &a, hipEvent_t &b)
    {
        if (used == kind.size()) {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (hipEventCreate(&e0) != hipSuccess) return false;
            if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return false; }
            ev.push_back(e0);
            ev.push_back(e1);
            kind.push_back(k);
        }
        kind[used] = k;
        a = ev[2 * used];
        b = ev[2 * used + 1];
        ++used;
        return true;
    }
    ~KernelProfile()
    {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};
// Stats words of QueuesDev::stats: rays, then the counting variants' sums.
enum : uint32_t {
    kStatExt = 0, kStatShadow = 1, kStatExtNodes = 2, kStatExtTris = 3, kStatShNodes = 4, kStatShTris = 5,
    kStatOccluded = 6, kNumStats = 8
};

// Streaming renderer.  A chunk of wp.count paths (global indices
// wp.first_path ...) runs through `capacity` path slots: every iteration
// extends, shades and connects the live paths, writes the radiance of the
// paths that ended to res[path index] and refills their slots with the
// chunk's next paths, so the slots stay full until the chunk runs dry.
// The film pass then sums res per pixel in sample order (bit-reproducible).
//
// Everything a stage launches is kernels only (no memset, copy or host
// synchronisation), so stages can be captured into graphs; the chunk's
// first path and count are read from device memory (chunk_dev) so one
// captured iteration block serves every chunk.
struct ChunkLaunch {
    const SceneDev *scene;      // single scene (batch == nullptr)
    PathsDev paths;
    ShadowDev shadow;
    QueuesDev queues;
    uint32_t capacity;
    float *res;                 // 3 floats per chunk path
    const CacheDev *cache;      // rendering with cache lookups (single scene)
    const CacheWork *cwork;
    const BatchLaunch *batch;
    KernelProfile *prof;        // optional (never under graph capture)
    const PhotonLaunch *photon; // photon mapping (single scene, no cache)
};
// chunk start: clears every counter and fills all slots
hipError_t chunk_begin(const ChunkLaunch &c, const WaveParams &wp, hipStream_t s);
// `iterations` bounce iterations (even, so the live list ends where it began)
hipError_t chunk_iterate(const ChunkLaunch &c, const WaveParams &wp, uint32_t iterations, hipStream_t s);
// film pass of a finished chunk (FP64 sums; skipped for training)
hipError_t chunk_film(const ChunkLaunch &c, const FilmDev &film, const WaveParams &wp, hipStream_t s);

// Host driver: begin, iterate in blocks of 4 until the chunk's paths are
// all generated and none is live (one small read-back per block), film.
// host_counters: pinned host memory for kCounterWords words.  c.capacity
// must be a multiple of kStripes.
hipError_t render_chunk(const ChunkLaunch &c, const FilmDev &film, const WaveParams &wp, uint32_t *host_counters,
                        hipStream_t s);

size_t extend_lds_bytes(uint32_t stack_cap);

// *p = v from a one-thread kernel (for values a captured graph reads).
hipError_t set_u32(uint32_t *p, uint32_t v, hipStream_t s);
hipError_t set_u64(uint64_t *p, uint64_t v, hipStream_t s);

}  // namespace rt
