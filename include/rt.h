/*
 * rt.h -- C ABI of the MI355X-native path-tracing hot path.
 *
 * Plain C types only (pointers, sizes, enums, POD structs).  Every entry
 * point returns an rt_status; on failure rt_last_error() returns a
 * thread-local message.  No exception crosses this boundary.
 *
 * Pointer convention: ray, hit and image buffers passed to rt_trace_* and
 * rt_render* are DEVICE pointers (HBM) unless RT_FLAG_HOST_PTRS is set, in
 * which case they are host pointers and the call stages them through
 * device memory and blocks until the results are on the host (the
 * reference's BVH::intersectRays contract).
 *
 * Streams: a stream argument is a hipStream_t passed as void*; NULL means
 * the context's own stream.  Calls on one rt_ctx from several host threads
 * are serialised by the context's lock (the reference reaches
 * BVH::intersectRays concurrently from per-light threads,
 * integration/Scene.cu:557-574).
 */
#ifndef RT_AMD_H
#define RT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    RT_OK = 0,
    RT_ERR_INVALID = 1,     /* bad argument, size or state */
    RT_ERR_HIP = 2,         /* a HIP runtime call failed */
    RT_ERR_OOM = 3,         /* device allocation failed */
    RT_ERR_UNSUPPORTED = 4, /* feature not built for this scene */
    RT_ERR_NO_DEVICE = 5    /* no HIP device visible */
} rt_status;

enum {
    RT_FLAG_HOST_PTRS = 1u,  /* ray / hit buffers are host memory; the call blocks */
    RT_FLAG_GRAPH = 2u,      /* rt_batch_render: capture the frame into a hipGraph and replay it */
    RT_FLAG_SHADOW = 4u,     /* rt_trace_closest: closest OPAQUE hit (transparent triangles skipped) */
    RT_FLAG_PROFILE = 8u,    /* rt_render*, rt_batch_render: time every traversal launch with HIP events */
    RT_FLAG_COUNT = 16u      /* rt_render*, rt_batch_render: counting traversal kernels (node visits, ...) */
};

/* Thread-local message of the last failing call on this thread. */
const char *rt_last_error(void);
/* ABI version (major << 16 | minor). */
uint32_t rt_version(void);

/* ------------------------------------------------------------------ */
/* Context                                                              */
/* ------------------------------------------------------------------ */
typedef struct rt_ctx rt_ctx;

rt_status rt_ctx_create(int device, rt_ctx **out);
rt_status rt_ctx_destroy(rt_ctx *ctx);
rt_status rt_ctx_synchronize(rt_ctx *ctx);
/* The context's stream (hipStream_t as void*). */
void *rt_ctx_stream(rt_ctx *ctx);

/* Device memory helpers for C callers that do not own an allocator. */
rt_status rt_malloc(rt_ctx *ctx, size_t bytes, void **out);
rt_status rt_free(rt_ctx *ctx, void *ptr);
rt_status rt_memcpy_h2d(rt_ctx *ctx, void *dst, const void *src, size_t bytes);
rt_status rt_memcpy_d2h(rt_ctx *ctx, void *dst, const void *src, size_t bytes);

/* ------------------------------------------------------------------ */
/* BVH over a triangle soup                                             */
/* ------------------------------------------------------------------ */
/* Replaces the reference's BVH(meshes, initTriangles, meshData, cpuTris)
 * constructor (cuda_libraries/BVH.cu:314-341).  The reference builds a
 * Morton-pairing tree with one triangle per leaf; this build is a binned-SAH
 * two-wide tree with up to max_leaf triangles per leaf whose nodes hold
 * both child boxes (64-byte records, one cache line per visit).  Triangle
 * ids are the index of the triangle in idx (the reference's concatenated
 * triangle list, BVH.cu:288-291).  xyz and idx are HOST arrays.  With
 * ctx == NULL the tree is built on the host only (rt_bvh_get_info and
 * rt_bvh_export work; traces return RT_ERR_INVALID). */
typedef struct {
    uint32_t max_leaf;   /* 1..8, default 4 */
    uint32_t bins;       /* SAH bins per axis, 4..64, default 32 */
    float    trav_cost;  /* SAH traversal cost relative to a triangle test, default 1.2 */
    uint32_t width;      /* traversal layout: 4 = four-wide nodes with 8-bit quantized child
                            boxes (default, also for 0), 2 = two-wide nodes with FP32 boxes */
} rt_bvh_opts;

typedef struct {
    uint32_t num_tris;
    uint32_t num_nodes;   /* internal nodes (64 B each) */
    uint32_t num_leaves;
    uint32_t max_depth;   /* internal levels on the longest root-leaf path */
    uint64_t node_bytes;
    uint64_t tri_bytes;   /* 48 B per triangle */
    double   build_ms;
    double   sah_cost;
    uint32_t width;       /* layout the traces use */
    uint32_t num_nodes4;  /* four-wide nodes (64 B each) */
    uint32_t depth4;      /* levels of the four-wide tree */
    uint32_t stack_cap;   /* traversal stack entries per lane for the layout in use */
} rt_bvh_info;

typedef struct rt_bvh rt_bvh;

rt_status rt_bvh_build(rt_ctx *ctx, const float *xyz, uint32_t nv, const uint32_t *idx,
                       uint32_t nt, const rt_bvh_opts *opts, rt_bvh **out);
/* The same tree object built on the GPU, from DEVICE arrays on ctx's device
 * (xyz: nv x 3 floats, idx: nt x 3 vertex indices), ordered on ctx's stream;
 * returns when the tree is ready.  This is the reference's own arrangement:
 * its BVH constructor (cuda_libraries/BVH.cu:314-341) takes the triangles
 * already on the device and builds there (Morton codes, sort, pairing,
 * BVH.cu:136-151).  Here: 30-bit Morton codes of the triangle centres, a
 * radix sort, a binary radix tree over the sorted codes (one internal node
 * per adjacent pair of leaves, ties split by index), boxes and SAH costs
 * bottom-up; a subtree of at most max_leaf triangles becomes a leaf when
 * (triangles x area) <= (trav_cost x area + its children's costs); then
 * the same 64-byte node records, padding and four-wide collapse as
 * rt_bvh_build.  opts->bins is ignored.  The same input always gives the
 * same records.  The tree's SAH cost is higher than the binned-SAH build's
 * (rt_bvh_info.sah_cost), so traces are somewhat slower on it; the build
 * takes milliseconds instead of hundreds.  Everything rt_bvh_build's tree
 * supports works on it; export4 of a two-wide build copies nothing
 * (num_nodes4 = 0).  Invalid indices or non-finite coordinates return
 * RT_ERR_INVALID. */
rt_status rt_bvh_build_device(rt_ctx *ctx, const float *xyz, uint32_t nv, const uint32_t *idx,
                              uint32_t nt, const rt_bvh_opts *opts, rt_bvh **out);
rt_status rt_bvh_destroy(rt_bvh *bvh);
rt_status rt_bvh_get_info(const rt_bvh *bvh, rt_bvh_info *info);
/* Per-triangle transparency (host array of nt bytes, indexed by triangle
 * id): transparent triangles are skipped by rt_trace_any, as the reference
 * skips them for shadow rays (shapes/Triangle.cu:439-440). */
rt_status rt_bvh_set_transparent(rt_bvh *bvh, const uint8_t *flags);
/* Copy the device tree to host arrays of num_nodes*16 and num_tris*12
 * floats: 64-byte nodes (child 0 box x/y, child 1 box x/y, both z ranges,
 * two child links; a link < 0 is a leaf ~(first << 3 | count - 1)) and
 * 48-byte triangles (v0 xyz + id, v1 xyz + flags, v2 xyz + 0), as the
 * oracle's counting traversal reads them. */
rt_status rt_bvh_export(const rt_bvh *bvh, float *nodes, float *tris);
/* Copy the four-wide tree (num_nodes4 * 16 floats): per node the origin,
 * a word of scale exponents and slot-valid bits, four child links and the
 * 8-bit quantized child bounds (lo_x, hi_x, lo_y, hi_y, lo_z, hi_z words,
 * byte i = child i), followed by two zero words. */
rt_status rt_bvh_export4(const rt_bvh *bvh, float *nodes4);

/* ------------------------------------------------------------------ */
/* Ray queries                                                          */
/* ------------------------------------------------------------------ */
/* Structure-of-arrays rays: origin, direction, tMax (FP32).  tmax may be
 * NULL (+infinity). */
typedef struct {
    const float *ox, *oy, *oz;
    const float *dx, *dy, *dz;
    const float *tmax;
} rt_rays;

/* Closest hit.  Replaces BVH::intersectRays (cuda_libraries/BVH.cu:541-584)
 * and its per-ray BVH::intersectRay (BVH.cu:355-448) with the
 * Triangle::intersects arithmetic of shapes/Triangle.cu:433-643 done
 * bit-for-bit in FP32.  prim[i] = triangle id or -1; t[i] = hit distance
 * (+inf on a miss); t may be NULL.
 *
 * flags: RT_FLAG_HOST_PTRS (host buffers, blocking) and RT_FLAG_SHADOW.
 * RT_FLAG_SHADOW is BVH::intersectRays(rays, ids, checkShadows = true): the
 * closest hit among the triangles that are not transparent
 * (rt_bvh_set_transparent), as Triangle::intersects returns false for a
 * transparent triangle under checkShadow (Triangle.cu:439-440).  Its caller
 * Scene::intersects re-tests triangles[id] for t (Scene.cu:108-150) and
 * Scene::isVisible compares that t with the light distance (Scene.cu:238-276);
 * t[i] here is that same distance.  Any other flag bit is RT_ERR_INVALID. */
rt_status rt_trace_closest(rt_bvh *bvh, const rt_rays *rays, uint32_t n, int32_t *prim,
                           float *t, uint32_t flags, void *stream);

/* Any hit (shadow rays).  Replaces the shadow use of the same call,
 * Scene::intersects(..., checkShadow=true) (integration/Scene.cu:108-150)
 * followed by the visibility test of Scene::isVisible (Scene.cu:262-273):
 * occluded[i] = 1 iff a non-transparent triangle is accepted at
 * t <= tmax[i] (tmax = distance to the light). */
rt_status rt_trace_any(rt_bvh *bvh, const rt_rays *rays, uint32_t n, uint8_t *occluded,
                       uint32_t flags, void *stream);

/* Counting run of the same traversal over device rays: total node visits
 * (64 B each) and triangle tests (48 B each).  Prices the algorithmic bytes
 * of a launch; it is a separate, slower kernel variant, never the timed one. */
rt_status rt_trace_stats(rt_bvh *bvh, const rt_rays *rays, uint32_t n, int any,
                         uint64_t *node_visits, uint64_t *tri_tests);

/* Time the dominant kernel of the last rt_trace_* call on this BVH:
 * elapsed milliseconds between HIP events recorded around that kernel
 * on the stream it was launched on. */
rt_status rt_bvh_last_kernel_ms(const rt_bvh *bvh, float *ms);

/* ------------------------------------------------------------------ */
/* Scene and path-traced rendering                                      */
/* ------------------------------------------------------------------ */
/* Material types: the reference's material classes on the hot path. */
typedef enum {
    RT_MAT_LAMBERTIAN = 0,       /* materials/LambertianMaterial.cu */
    RT_MAT_DIELECTRIC = 1,       /* materials/DielectricMaterial.cu, GlassMaterial.cu (ior 1.5) */
    RT_MAT_TROWBRIDGE = 2,       /* materials/TrowbridgeMaterial.cu (GGX, conductor Fresnel) */
    RT_MAT_TROWBRIDGE_BASIC = 3, /* materials/TrowbridgeBasicMaterial.cu (GGX, F = 1) */
    RT_MAT_LIGHT = 4,            /* materials/LightMaterial.cu (emissive) */
    RT_MAT_MIRROR = 5,           /* materials/MirrorMaterial.cu (PerfectReflectionBSDF, weight kd) */
    RT_MAT_PLAIN = 6             /* materials/PlainMaterial.cu (emits kd to camera rays, ke to others; absorbs) */
} rt_material_type;

typedef struct {
    uint32_t type;      /* rt_material_type */
    float kd[3];        /* albedo / colour (kd, ks = kt, light colour, mirror colour) */
    float ior;          /* dielectric index of refraction */
    float roughness;    /* Trowbridge: alpha = sqrt(roughness) (TrowbridgeBSDF.cu:56-59) */
    float eta[3], k[3]; /* Trowbridge conductor Fresnel */
    float intensity;    /* light: emitted radiance = kd * intensity */
    float ke[3];        /* plain: radiance emitted towards rays that are not camera rays */
    uint32_t texture;   /* 0: none; k + 1: the scene's texture k, multiplying kd of a Lambertian
                         * (LambertianBSDF.cu:22-74) or the camera-ray kd of a plain material */
} rt_material;

/* PointLight(translate(pos), color, intensity) (light/PointLight.cu:15-66). */
typedef struct {
    float pos[3];
    float color[3];
    float intensity;
} rt_point_light;

/* The .obejota camera (cuda_libraries/FileImporting.cu:404-436): lookAt
 * from pos to lookat with up, cameraToScreen = translate(cts), vertical
 * frame size frame_y, resolution width x height. */
typedef struct {
    float pos[3], lookat[3], up[3], cts[3];
    float frame_y;
    uint32_t width, height;
} rt_camera;

typedef struct {
    uint32_t spp;         /* samples per pixel in this call */
    uint32_t spp_offset;  /* index of the first sample (progressive rendering) */
    uint32_t max_depth;   /* MAX_RAY_RECURSION_DEPTH (types.h:41) */
    uint32_t rr_depth;    /* Russian roulette when depth > rr_depth (pathTracing.cu:91), 3 */
    uint32_t seed;        /* PCG32 seed */
    uint32_t path_capacity; /* paths in flight per wave, 0 = default (2^24, capped at the render's paths;
                               ~600 B of device memory per path) */
    uint32_t grid;          /* workgroups of the queue kernels, 0 = default (2048) */
    uint32_t chunk_paths;   /* paths per streamed chunk (12 B of result buffer each), 0 = default (2^28) */
} rt_render_opts;

typedef struct rt_scene rt_scene;

/* Replaces Scene::build(shapes, lights, triangles, meshes)
 * (integration/Scene.cu:56-95): shapes are added one by one, then
 * rt_scene_commit builds the BVH and uploads everything. */
rt_status rt_scene_create(rt_ctx *ctx, rt_scene **out);
rt_status rt_scene_destroy(rt_scene *scene);
rt_status rt_scene_add_material(rt_scene *scene, const rt_material *m, uint32_t *id);
/* A triangle mesh (TriangleMesh + Triangle list, shapes/Triangle.cu:14-100);
 * its triangles get consecutive ids after the ones already added; first_id
 * (may be NULL) receives the first. */
rt_status rt_scene_add_mesh(rt_scene *scene, const float *xyz, uint32_t nv, const uint32_t *idx,
                            uint32_t nt, uint32_t material, uint32_t *first_id);
/* The same with texture coordinates (TriangleMesh::textureCoords,
 * shapes/Triangle.cu:134-152): uv holds 6 floats per triangle (u, v of its
 * three corners); NULL gives every triangle the reference's default
 * (0,0), (1,0), (1,1). */
rt_status rt_scene_add_mesh_uv(rt_scene *scene, const float *xyz, uint32_t nv, const uint32_t *idx, uint32_t nt,
                               const float *uv, uint32_t material, uint32_t *first_id);
/* The same with vertex normals as well (TriangleMesh::normals with each
 * face's normal indices, shapes/Triangle.cu:14-100): normals holds 9 floats
 * per triangle, the normals of its three corners (a NaN first component: a
 * triangle whose face names no normals); uv may be NULL.  A triangle with
 * normals shades with the interpolated normal normalize(b0 n0 + b1 n1 +
 * b2 n2), not turned towards the ray, and its hit point moves 0.15 along it
 * (Triangle.cu:379-418). */
rt_status rt_scene_add_mesh_normals(rt_scene *scene, const float *xyz, uint32_t nv, const uint32_t *idx,
                                    uint32_t nt, const float *uv, const float *normals, uint32_t material,
                                    uint32_t *first_id);
/* A texture (textures/Texture.cu): width x height texels of 3 bytes (R, G,
 * B), row y = the reference's pixelMatrix[y] (rows in BMP file order), read
 * nearest-texel at (u * width, v * height) and scaled by 1/255.  id (may be
 * NULL) receives k for rt_material.texture = k + 1. */
rt_status rt_scene_add_texture(rt_scene *scene, const uint8_t *rgb, uint32_t width, uint32_t height, uint32_t *id);
/* Sphere(center, radius) (shapes/Sphere.cu:233-245).  is_light = 1 also
 * samples it as a sphere area light (next-event estimation with MIS); its
 * material must be RT_MAT_LIGHT. */
rt_status rt_scene_add_sphere(rt_scene *scene, const float center[3], float radius, uint32_t material,
                              int is_light);
/* Plane(corner, axis_x, axis_y, normal) (shapes/Plane.cu:27-48): the
 * parallelogram corner + [0,1] axis_x + [0,1] axis_y, where the axis lengths
 * are first truncated to integers as the reference does (an axis shorter
 * than 1 is refused); normal picks the side of cross(axis_x, axis_y).
 * Intersected after the spheres with the reference's Plane::intersect. */
rt_status rt_scene_add_plane(rt_scene *scene, const float corner[3], const float axis_x[3], const float axis_y[3],
                             const float normal[3], uint32_t material);
rt_status rt_scene_add_point_light(rt_scene *scene, const rt_point_light *light);
rt_status rt_scene_set_camera(rt_scene *scene, const rt_camera *cam);
/* The BVH rt_scene_commit builds (no counterpart in the reference, whose
 * Scene::build always builds its GPU Morton tree, Scene.cu:56-95): opts as
 * for rt_bvh_build (NULL: the defaults), and on_device != 0 builds it on
 * the GPU with rt_bvh_build_device instead of the host's binned SAH
 * (milliseconds instead of seconds for millions of triangles; traces a few
 * per cent slower).  Takes effect at the next commit. */
rt_status rt_scene_set_bvh(rt_scene *scene, const rt_bvh_opts *opts, uint32_t on_device);
rt_status rt_scene_commit(rt_scene *scene);
/* The scene's triangle BVH (valid after commit; owned by the scene). */
rt_bvh *rt_scene_bvh(rt_scene *scene);

/* Photon pre-pass: generatePhotons + computePhotons (integration/
 * photonDispersion.cu:26-192, Scene.cu:989-1021) on the GPU.  Each of
 * `photons` paths picks a point light with probability proportional to its
 * power (Scene::selectOneRandomLight, Scene.cu:427-455) and leaves it in a
 * uniform direction (PointLight::sampleLight, light/PointLight.cu:68-95)
 * carrying color * intensity * 4 pi / n_l / p_l (n_l: photons that picked
 * light l, p_l: its selection probability, Scene.cu:1005-1012).  Each
 * intersection with a scattering material (Lambertian, Trowbridge, plain)
 * stores a photon; the material's sampleDirection continues the path, the
 * hit's emission is added to the carried light, a black weight ends it and
 * Russian roulette applies past rr_depth (photonDispersion.cu:26-75,
 * pathTracing.cu:62-123), for at most max_depth intersections.  Random
 * numbers: photon path i draws from its own PCG32 stream (light, two for the
 * direction, then the materials' draws and the roulette in path order).
 * Outputs, device arrays of capacity entries: location (3 floats), direction
 * of the ray that arrived (3 floats), shading normal (3 floats), radiance
 * (3 doubles, totalLight * pathLightFactor) and shape id (analytic shape k:
 * k, spheres then planes; triangle t: spheres + planes + t, as
 * Scene::intersectShapes numbers them, Scene.cu:209-227); key (may be NULL)
 * receives path << 5 | bounce.  Photons are ordered by path, then bounce
 * (ascending key).  *count (host) receives their number; more
 * than capacity gives RT_ERR_INVALID with *count set, so the caller can
 * grow.  A scene without point lights stores none (generatePhotons runs only
 * with lights).  Blocking: returns when the photons are written. */
typedef struct {
    uint32_t photons;    /* photon paths, 1 .. 2^27 (N_PHOTON_SAMPLES = 1M, types.h:57) */
    uint32_t max_depth;  /* intersections per path, 1 .. 32 (MAX_PHOTON_RECURSION_DEPTH = 30, types.h:61) */
    uint32_t rr_depth;   /* Russian roulette after this depth (3, pathTracing.cu:91) */
    uint32_t seed;
} rt_photon_opts;
rt_status rt_scene_trace_photons(rt_scene *scene, const rt_photon_opts *opts, uint32_t capacity, float *location,
                                 float *direction, float *normal, double *radiance, int32_t *shape, uint32_t *key,
                                 uint32_t *count, void *stream);

/* Path-trace spp samples per pixel and ADD them into accum (width*height*3
 * doubles, per-pixel radiance sums) and nvalid (width*height counts of
 * finite samples).  Replaces PathTracerIntegrator::render
 * (integration/PathTracerIntegrator.cu:140-212) with computeSamples
 * (pathTracing.cu:284-414); the film convention is Film::addSample
 * (camera/Film.cu:111-118): the reference's pixel sum is
 * accum * (samples / nvalid) (pathTracing.cu:474-486). */
rt_status rt_render(rt_scene *scene, const rt_render_opts *opts, double *accum, uint32_t *nvalid,
                    uint32_t flags, void *stream);

/* Camera ray of pixel (px, py) with film jitter (u0, u1), as the renderer's
 * generate stage computes it (host copy for checking). */
rt_status rt_scene_camera_ray(const rt_scene *scene, uint32_t px, uint32_t py, float u0, float u1,
                              float o[3], float d[3]);

/* Statistics of the last rt_render call.  The rest of the fields are filled
 * by frames rendered with the measurement flags:
 *   RT_FLAG_PROFILE: every launch of the closest-hit (extension rays) and
 *     any-hit (shadow rays) traversal kernels is bracketed by HIP events on
 *     its stream; *_ms is their summed elapsed time over the frame.
 *   RT_FLAG_COUNT: those kernels run their counting variants (slower, never
 *     the timed ones) and sum node visits (64 B each), triangle tests (48 B
 *     each) and shadow rays found occluded.  Images are unchanged. */
typedef struct {
    uint64_t paths, extension_rays, shadow_rays;
    double ms;             /* wall time of the call (device-synchronised) */
    double extend_ms, connect_ms;
    uint32_t extend_launches, connect_launches;
    uint64_t ext_node_visits, ext_tri_tests;
    uint64_t shadow_node_visits, shadow_tri_tests, shadow_occluded;
} rt_render_stats;
rt_status rt_render_get_stats(const rt_scene *scene, rt_render_stats *stats);

/* ------------------------------------------------------------------ */
/* Film output: tone mapping and image files                            */
/* ------------------------------------------------------------------ */
/* ToneMappingType (camera/Film.cu:13-21), in the reference's order. */
typedef enum {
    RT_TONE_EQUALIZE = 0,       /* Film.cu:156-162  pixel / max channel */
    RT_TONE_REINHARD = 1,       /* Film.cu:164-180  extended Reinhard on luminance, white = max luminance */
    RT_TONE_REINHARD_JODIE = 2, /* Film.cu:182-193 */
    RT_TONE_ACES_FILMIC = 3,    /* Film.cu:224-247  fitted ACES RRT+ODT */
    RT_TONE_UNCHARTED2 = 4,     /* Film.cu:195-222  filmic curve, exposure 2, white 11.2 (no clamp) */
    RT_TONE_GAMMA_CLAMP = 5     /* Film.cu:250-264  (pixel / max channel)^(1/2.2) */
} rt_tone_mapping;

/* Film::writeToBMP (camera/BMP_Format.cu, 24-bit, rows bottom-up and padded
 * to 4 bytes) and Film::writeToPPM (binary P6). */
typedef enum { RT_IMAGE_BMP = 0, RT_IMAGE_PPM = 1 } rt_image_format;

/* Bytes of the image's pixel array (0 for an unknown format or a side
 * outside 1..32768). */
size_t rt_image_pixel_bytes(rt_image_format format, uint32_t width, uint32_t height);

/* Writes the file header that precedes the pixel array into out (at most cap
 * bytes; out may be null) and returns its length: BMP 54 bytes as
 * BitmapImage::write_image emits them (BMP_Format.cu), including its
 * Image_Size field of (width + 4 - width % 4) * height; PPM
 * "P6\n<w> <h>\n255\n" (Film.cu:339-350).  0 for a bad format or size. */
size_t rt_image_header(rt_image_format format, uint32_t width, uint32_t height, uint8_t *out, size_t cap);

/* Tone-maps a film and writes its image pixel array into `pixels` (device
 * memory, rt_image_pixel_bytes bytes, 4-byte aligned), replacing
 * Film::toneMapBMP / toneMapPPM (camera/Film.cu:285-337) with
 * toneMaxPixel (Film.cu:266-283) and the maxima of Film.cu:120-154.
 *   sums     device, width*height*3 doubles: the pixel sums
 *            (Film::pixelLight), or, when nvalid is not null, rt_render's
 *            valid-sample sums, whose pixel sum is sums * samples / nvalid
 *   samples  the film's sample count (samplesLoadedFromFile +
 *            samplesComputedPixel0), 1..2^53
 * Bytes are (unsigned char)(value * 255) as the reference computes them on
 * x86-64: values outside [0, 1] are not clamped by the equalize, uncharted2
 * and gamma operators, and wrap modulo 256 (NaN gives 0).  Enqueued on
 * `stream` (null: the context's stream); the call returns without waiting.
 * Calls on one context share its partial-maxima buffer: issue them on one
 * stream. */
rt_status rt_film_tonemap(rt_ctx *ctx, const double *sums, const uint32_t *nvalid, uint32_t width, uint32_t height,
                          uint64_t samples, rt_tone_mapping mode, rt_image_format format, uint8_t *pixels,
                          size_t pixel_bytes, void *stream);

/* rt_film_tonemap without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, checks and output as rt_film_tonemap; the partial maxima live in
 * a file-scope device array of 1025 pairs of doubles (16.4 KB) that each GPU
 * keeps for the life of the process, instead of the context's scratch that
 * rt_film_tonemap allocates on its first call.  A film whose max pass would
 * need more partial pairs than the array holds is refused (RT_ERR_INVALID; no
 * film within 32768 x 32768 does).  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other contexts, other
 * threads) race; order them. */
rt_status rt_film_tonemap_noalloc(rt_ctx *ctx, const double *sums, const uint32_t *nvalid, uint32_t width,
                                  uint32_t height, uint64_t samples, rt_tone_mapping mode, rt_image_format format,
                                  uint8_t *pixels, size_t pixel_bytes, void *stream);

/* ------------------------------------------------------------------ */
/* Photon map: kd-tree and k-nearest search                             */
/* ------------------------------------------------------------------ */
/* The photon map of Scene::buildPhotonStructures / search_nearest
 * (integration/Scene.cu:704-722, 1102-1114) over nn::KDTree<Photon, 3,
 * PhotonPosition> (cuda_libraries/kdtree.h). */
typedef struct rt_photon_map rt_photon_map;

/* Builds the kd-tree of n photon locations (xyz: host, 3n floats) as
 * KDTree's constructor does (median split on the axis of largest extent) and
 * uploads it.  shape: host, n shape ids (Photon::shapeId, -1 .. 2^30-2), or
 * null for all -1.  Photons are referred to by their index in xyz. */
rt_status rt_photon_map_build(rt_ctx *ctx, const float *xyz, const int32_t *shape, uint32_t n,
                              rt_photon_map **out);
rt_status rt_photon_map_destroy(rt_photon_map *map);
rt_status rt_photon_map_get_info(const rt_photon_map *map, uint32_t *n, uint32_t *depth);

/* KDTree::nearest_neighbors(p, k, max_distance, shape) for nq query points
 * (device buffers).  queries: 3*nq floats; query_shape: nq ids (-1: any
 * shape) or null.  Outputs per query: count[q] <= k photons, their indices
 * index[q*k ..] and float distances distance[q*k ..] in the reference's
 * result order: insertion order below k, then the layout that libstdc++'s
 * std::make_heap / push_heap / pop_heap leave (kdtree.h:123-130), ties
 * included.  The set is the k nearest photons strictly within max_distance
 * (on the query's shape).  k: 1..1024, nq*k <= 2^31.  Enqueued on stream
 * (null: the context's stream).  The map keeps one work space (query sort
 * keys, permutation) that every call on it uses: issue the calls on one map
 * from one stream, or order them with events. */
rt_status rt_photon_map_knn(rt_photon_map *map, const float *queries, const int32_t *query_shape, uint32_t nq,
                            uint32_t k, float max_distance, uint32_t *count, uint32_t *index, float *distance,
                            void *stream);

/* Photon-mapping density estimate on Lambertian surfaces:
 * Scene::getPhotonMappingContribution's combinePhotonContributions
 * (integration/Scene.cu:607-634, 724-746) for nq query points at once.  Per
 * query: the k nearest photons within max_distance (on query_shape[q] unless
 * it is -1 or query_shape is null), then
 *   out = sum(photon_rgb * kd / pi * |cos(normal, -photon_dir)|) / (pi * d_max^2)
 * with d_max the farthest photon found (0 photons give 0).  photon_rgb
 * (Photon::radiance, n*3 doubles) and photon_dir (Photon::ray.direction, n*3
 * floats) are indexed like the map's photons; points, normals (the shading
 * normals) and kd are nq*3 floats; out is nq*3 doubles; all device memory.
 * Photons are summed in the search's result order, as the reference sums its
 * vector.  Shares the map's work space and result buffers with
 * rt_photon_map_knn: one stream per map. */
rt_status rt_photon_gather_lambert(rt_photon_map *map, const double *photon_rgb, const float *photon_dir,
                                   const float *points, const float *normals, const float *kd,
                                   const int32_t *query_shape, uint32_t nq, uint32_t k, float max_distance,
                                   double *out, void *stream);

/* The same density estimate for any material of the scene vocabulary:
 * combinePhotonContributions (Scene.cu:607-634) with
 * SurfaceInteraction::lightScattered(-photon.ray.direction) (transformations/
 * SurfaceInteraction.cu:278-292), i.e. the material's lightReflected
 * (f * |cos wi|) for the view direction wo and wi = -photon_dir, both taken
 * in the shading frame of normals[q]:
 *   out = sum(photon_rgb * lightReflected(wo, -photon_dir)) / (pi * d_max^2)
 * Lambertian and Trowbridge materials scatter photons; plain surfaces reflect
 * nothing (lightReflected = 0) and the other materials are not scattering
 * (getPhotonMappingContribution gives them 0, Scene.cu:724-746).  views: nq*3
 * floats, the direction towards the viewer (-ray direction); materials: the
 * material of each query (nq entries, device, rt_material; textures are not
 * applied: pass the textured kd).  Other arguments as
 * rt_photon_gather_lambert. */
rt_status rt_photon_gather(rt_photon_map *map, const double *photon_rgb, const float *photon_dir,
                           const float *points, const float *normals, const float *views,
                           const rt_material *materials, const int32_t *query_shape, uint32_t nq, uint32_t k,
                           float max_distance, double *out, void *stream);

/* Subsurface scattering from the photon map: Scene::combinePhotonsSubsurface
 * (integration/Scene.cu:667-702) over search_nearest(p, radius, unlimited,
 * shape) (Scene.cu:806-814) with the dipole of diffusionTermBSSRDF and
 * impPaper_dipoleDiffusionAproximation (cuda_libraries/geometricMath.cu:
 * 451-522).  Per query: every photon strictly within radius of points[q] on
 * query_shape[q] (all shapes if it is -1 or query_shape is null), in the
 * search's insertion order, then
 *   out = sum(photon_rgb * Sd(p, photon) * kd) / (2 pi d_max^2)
 * where Sd = 1/pi * Ft(cos(photon_nrm, -photon_dir)) * Rd(|p - photon|) *
 * Ft(cos(normals[q], views[q])), Ft = 1 - fresnelDielectric(cos, 1, eta).
 * Like the reference, a query that finds no photon gets 0 / 0 = NaN.
 * photon_nrm: the photons' surface normals (n*3 floats); views: nq*3 floats,
 * -camera ray direction; params: the material's SSMaterial parameters.
 * Device buffers; shares the map's work space (one stream per map). */
typedef struct {
    float kd[3];       /* diffuse colour (SSMaterial::kd) */
    float sigma_s[3];  /* scattering coefficient */
    float sigma_a[3];  /* absorption coefficient */
    float g;           /* phase function asymmetry */
    float eta;         /* index of refraction */
} rt_sss_params;
rt_status rt_photon_subsurface(rt_photon_map *map, const double *photon_rgb, const float *photon_dir,
                               const float *photon_nrm, const float *points, const float *normals,
                               const float *views, const int32_t *query_shape, uint32_t nq, float radius,
                               const rt_sss_params *params, double *out, void *stream);

/* Photon mapping in the bounce loop: the reference's RENDER_MODE
 * PHOTON_MAPPING (cuda_libraries/types.h:51; Scene::sampleDirectLight,
 * integration/Scene.cu:912-967; finishStep, pathTracing.cu:84-88).  Paths
 * start and bounce as in rt_render (same seeds, camera, specular bounces,
 * emission), but there is no next-event estimation: at its first
 * scattering hit (Lambertian, Trowbridge, plain) a path adds its throughput
 * times getPhotonMappingContribution -- the map's k nearest photons within
 * radius on the hit's shape, through combinePhotonContributions, as
 * rt_photon_gather computes it -- and ends.  The map holds the photons of
 * rt_scene_trace_photons (or any others: locations and shape ids, with
 * shape k = sphere k, then planes, then triangles); photon_rgb and
 * photon_dir are their radiance (3 doubles) and direction (3 floats) in the
 * map's input order, device memory.  The film, flags and stream are as for
 * rt_render; the map must be on the scene's context, and its work space
 * serves this call (one render or search per map at a time).  Each
 * iteration searches every path that ended there (k results per path slot
 * of work space: 8 * k bytes per slot).  RT_RENDER_HYBRID (RENDER_MODE
 * HYBRID): every scattering hit adds its throughput times (photon estimate
 * + next-event estimate) / 2 (Scene.cu:963-967) and the path goes on as in
 * rt_render. */
enum { RT_RENDER_PATH_TRACING = 0, RT_RENDER_PHOTON_MAPPING = 1, RT_RENDER_HYBRID = 2 };
typedef struct {
    uint32_t mode;   /* RT_RENDER_PHOTON_MAPPING or RT_RENDER_HYBRID */
    uint32_t k;      /* photons per estimate, 1..1024 (MAX_NEIGHBOURS, types.h:60: 100) */
    float radius;    /* search radius (PHOTON_SEARCH_RADIUS, types.h:58: 100) */
} rt_photon_render_opts;

rt_status rt_render_photon(rt_scene *scene, const rt_render_opts *opts, rt_photon_map *map,
                           const double *photon_rgb, const float *photon_dir,
                           const rt_photon_render_opts *popts, double *accum, uint32_t *nvalid,
                           uint32_t flags, void *stream);

/* ------------------------------------------------------------------ */
/* Batches: many scenes in one launch sequence                          */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference, which renders one scene per process
 * (PathTracerIntegrator::render, integration/PathTracerIntegrator.cu:146-210;
 * several GPUs = several processes, boxMulti.cu:174-190).  A batch renders
 * n committed scenes of any sizes together: their samples form one index
 * space, waves of paths from all scenes share the ray and material queues,
 * and each path carries its scene (own BVH, materials, lights, camera).
 * Per scene, the image equals rt_render's with the same seed, sample offset
 * and options (up to FP64 summation order across wave boundaries).  Scenes
 * must stay alive and committed while the batch exists. */
typedef struct rt_batch rt_batch;

typedef struct {
    uint32_t spp, spp_offset, max_depth, rr_depth;
    uint32_t path_capacity;  /* paths per wave, 0 = the autotuned or default value */
    uint32_t grid;           /* workgroups of the queue kernels, 0 = autotuned or default */
    uint32_t chunk_paths;    /* paths per streamed chunk, 0 = default (2^28) */
} rt_batch_opts;

rt_status rt_batch_create(rt_ctx *ctx, rt_scene *const *scenes, uint32_t n, rt_batch **out);
rt_status rt_batch_destroy(rt_batch *batch);

/* Renders opts->spp samples per pixel of every scene, accumulating into the
 * device films accum[i] (W_i*H_i*3 FP64 sums) and nvalid[i] (W_i*H_i valid
 * counts).  seeds: n PCG seeds (NULL: 1 for every scene).  RT_FLAG_GRAPH:
 * the frame's launch sequence is captured into a hipGraph on the first call
 * and replayed by later calls with the same spp, depths, wave size, grid and
 * films; the sample offset and seeds are read from device memory at replay,
 * so each call renders its own samples.  A change of any captured option
 * recaptures.  Blocks until the frame is done. */
rt_status rt_batch_render(rt_batch *batch, const rt_batch_opts *opts, const uint32_t *seeds, double *const *accum,
                          uint32_t *const *nvalid, uint32_t flags, void *stream);

/* Autotuner: renders opts->spp samples per pixel (clamped to 1..4) of every
 * scene (into scratch films) for each candidate wave size {2^19 .. 2^24,
 * capped at the trial frame's paths} x grid {1024, 2048, 4096, 8192}, twice
 * each, and keeps the fastest pair as the batch's default.  entries (may be NULL) receives up to
 * max_entries (path_capacity, grid, ms) triples in the order tried. */
typedef struct {
    uint32_t path_capacity, grid;
    float ms;  /* best of the two runs of one trial frame */
} rt_tune_entry;
rt_status rt_batch_autotune(rt_batch *batch, const rt_batch_opts *opts, rt_tune_entry *best, rt_tune_entry *entries,
                            uint32_t max_entries, uint32_t *n_entries);
rt_status rt_batch_get_stats(const rt_batch *batch, rt_render_stats *stats);

/* ------------------------------------------------------------------ */
/* Neural radiance cache                                                */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference (whose only learned component is the
 * per-shape polynomial regression of sampling/ShapeLightDistribution.cu:
 * 20-345, run in its MACHINE_LEARNING mode).  A multiresolution hash-grid
 * encoding of position plus normal, view direction and roughness feeds a
 * 4x64 ReLU MLP predicting outgoing RGB radiance; BF16 matrix-core matmuls
 * with FP32 accumulation, FP32 master weights and Adam. */
typedef struct {
    uint32_t levels;          /* hash-grid levels, 1..16 */
    uint32_t log2_table;      /* entries per level = 2^log2_table */
    uint32_t base_res;        /* coarsest level resolution */
    float per_level_scale;    /* resolution growth per level */
    float lr, beta1, beta2, eps;  /* Adam */
    uint32_t batch;           /* largest training batch (and query chunk) */
    uint32_t seed;            /* parameter initialisation */
    float aabb_lo[3], aabb_hi[3];  /* positions are normalised to this box */
} rt_cache_opts;

typedef struct {
    uint32_t levels, table_size;
    uint32_t res[16], dense[16];
    uint64_t mlp_params, grid_params;
    uint32_t steps;
} rt_cache_info;

typedef struct rt_cache rt_cache;

rt_status rt_cache_create(rt_ctx *ctx, const rt_cache_opts *opts, rt_cache **out);
rt_status rt_cache_destroy(rt_cache *cache);
rt_status rt_cache_get_info(const rt_cache *cache, rt_cache_info *info);
/* Queries: n records of 12 floats (position 3, normal 3, view direction 3,
 * roughness 1, 2 unused), device pointers; rgb: n x 3 floats (>= 0). */
rt_status rt_cache_query(rt_cache *cache, const float *in, uint32_t n, float *rgb, void *stream);
/* One Adam step on a batch (n <= opts.batch) of records and RGB targets
 * (device).  loss (host, may be NULL) receives the batch's mean relative L2
 * loss and makes the call blocking. */
rt_status rt_cache_train_step(rt_cache *cache, const float *in, const float *target, uint32_t n, float *loss,
                              void *stream);
/* Gradients of the batch's mean loss without an update (host outputs of
 * mlp_params and grid_params floats); for checking against a CPU model. */
rt_status rt_cache_gradients(rt_cache *cache, const float *in, const float *target, uint32_t n,
                             float *mlp_grad, float *grid_grad, float *loss);
/* FP32 parameters (host arrays of mlp_params / grid_params floats).  MLP
 * layout: 5 matrices W_l[in][out] row-major at offsets 0, 4096, 8192,
 * 12288, 16384; the last is 64 x 16 with columns 3..15 unused.  Grid:
 * [levels][table][2]. */
rt_status rt_cache_get_params(const rt_cache *cache, float *mlp, float *grid);
rt_status rt_cache_set_params(rt_cache *cache, const float *mlp, const float *grid);

/* The radiance cache inside the path tracer.  No counterpart in the
 * reference (its learned mode is a per-shape polynomial fit,
 * integration/Denoiser.cu, sampling/ShapeLightDistribution.cu:84-272); these
 * extend PathTracerIntegrator::render (integration/PathTracerIntegrator.cu:
 * 146-210) as the plain rt_render does.
 *
 * A cache vertex is the first hit, at path depth >= cache_depth (camera hit =
 * depth 0), on a diffuse surface or a GGX surface with roughness >= 0.04
 * (alpha >= 0.2).  Specular and sharper glossy hits keep tracing.
 *
 * rt_render_cached: as rt_render, but every path ends at its cache vertex
 * (after the emission and next-event estimate of the vertices before it) and
 * takes beta * cache(vertex).  The cache and the scene share one context.
 *
 * rt_cache_train_scene: one online training step.  Traces n_paths training
 * paths (pseudo-random pixels; sample indices disjoint from rendering ones;
 * opts->spp_offset selects the frame's set, opts->spp is ignored) with the
 * plain estimator.  A path records its first cache vertex, traces two more
 * bounces and ends at its next cache vertex, where the cache's own answer
 * stands for the rest of the path (self-training, as in neural radiance
 * caching).  The radiance gathered from the recorded vertex on, divided by
 * the throughput on arrival there, is the target, and one Adam step runs on
 * up to the cache's batch size of these (record, radiance) pairs.  The pair count stays on the device, so the call
 * only enqueues work; it blocks only when loss (mean relative L2 loss before
 * the step) or n_samples (pairs used; 0 = the step saw no pairs) is
 * requested.  Both may be NULL. */
rt_status rt_render_cached(rt_scene *scene, const rt_render_opts *opts, rt_cache *cache, uint32_t cache_depth,
                           double *accum, uint32_t *nvalid, uint32_t flags, void *stream);

/* One frame of online radiance caching with its training inside the frame
 * (no separate training paths): rt_render_cached, except that one path in
 * 2^harvest_shift (chosen by a hash of its index) ignores the cache, goes
 * its full length like rt_render's paths (so it stays unbiased in the image)
 * and, at its first cache vertex, records that vertex and the radiance
 * gathered from there on, divided by the throughput on arrival: a training
 * sample with a full-path target.  After the frame, train_steps Adam steps
 * run, each on the next min(batch, samples) samples of the frame (wrapping
 * around).  plain = 1 renders every path without the cache (a warm-up
 * frame: plain path tracing whose paths all harvest).  Samples come from
 * the frame's last chunk, at most the wave capacity of them.
 *
 * Divergence: the call reads the last step's loss back (it synchronises the
 * stream).  A non-finite loss, or one above 100 times the moving average of
 * the steps before (after 16 of them), reinitialises the cache's parameters
 * and Adam state, and the next rt_render_cache_frame call renders plain
 * (stats->fallback = 1), so the plain path tracer stands in while the cache
 * starts over.  stats may be NULL. */
typedef struct {
    uint32_t cache_depth;    /* paths end at their first cache vertex at depth >= cache_depth */
    uint32_t harvest_shift;  /* one path in 2^harvest_shift trains the cache (>= 32: none) */
    uint32_t train_steps;    /* Adam steps after the frame (0: none) */
    uint32_t plain;          /* 1: no path ends in the cache */
} rt_cache_frame_opts;

typedef struct {
    uint32_t samples;  /* training samples harvested by the frame */
    uint32_t steps;    /* Adam steps taken */
    float loss;        /* mean relative L2 loss of the last step before its update (NaN: no step) */
    uint32_t fallback; /* 1: rendered plain because the cache was reset after diverging */
    uint32_t reset;    /* 1: this frame's training diverged and the cache was reinitialised */
} rt_cache_frame_stats;

rt_status rt_render_cache_frame(rt_scene *scene, const rt_render_opts *opts, rt_cache *cache,
                                const rt_cache_frame_opts *copts, double *accum, uint32_t *nvalid,
                                rt_cache_frame_stats *stats, uint32_t flags, void *stream);
rt_status rt_cache_train_scene(rt_cache *cache, rt_scene *scene, const rt_render_opts *opts,
                               uint32_t cache_depth, uint32_t n_paths, float *loss, uint32_t *n_samples,
                               void *stream);

/* ------------------------------------------------------------------ */
/* Synthetic workloads (host generators, deterministic)                 */
/* ------------------------------------------------------------------ */
/* Random triangle soup in the unit cube (centroids uniform, edge lengths
 * log-normal(ln mu, sigma), random orientation), PCG32-seeded.  Writes
 * 9*nt floats (v0 v1 v2 per triangle). */
rt_status rt_gen_soup(uint32_t nt, uint64_t seed, float edge_mu, float edge_sigma, float *tri9);
/* Rays: the first n_coherent are pinhole primary rays of a w x h camera
 * (row-major, pixel centres) at eye looking along +z with vertical field
 * of view vfov_deg; the rest are incoherent (origins uniform in the unit
 * cube, cosine-hemisphere directions about uniform random normals).
 * Output SoA arrays of n floats each; tmax = +inf. */
rt_status rt_gen_rays(uint32_t w, uint32_t h, const float eye[3], float vfov_deg,
                      uint32_t n_incoherent, uint64_t seed, float *ox, float *oy, float *oz,
                      float *dx, float *dy, float *dz);

#ifdef __cplusplus
}
#endif
#endif
