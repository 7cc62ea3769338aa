/*
 * oracle_shade.h -- oracle (test infrastructure only): scene description and
 * estimator entry points of shade.c.  The POD layouts match the library's
 * rt_material / rt_point_light / rt_camera (include/rt.h) field for field so
 * tests can pass the same arrays to both, but nothing here includes or
 * links the library.
 */
#ifndef RT_ORACLE_SHADE_H
#define RT_ORACLE_SHADE_H

#include <stdint.h>

#include "oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_MAT_LAMBERTIAN = 0,       /* materials/LambertianMaterial.cu */
    ORC_MAT_DIELECTRIC = 1,       /* materials/DielectricMaterial.cu (GlassMaterial: ior 1.5) */
    ORC_MAT_TROWBRIDGE = 2,       /* materials/TrowbridgeMaterial.cu (conductor Fresnel) */
    ORC_MAT_TROWBRIDGE_BASIC = 3, /* materials/TrowbridgeBasicMaterial.cu (F = 1) */
    ORC_MAT_LIGHT = 4,            /* materials/LightMaterial.cu (emissive, NullBSDF) */
    ORC_MAT_MIRROR = 5,           /* materials/MirrorMaterial.cu (PerfectReflectionBSDF) */
    ORC_MAT_PLAIN = 6             /* materials/PlainMaterial.cu (emits kd / ke, absorbs) */
};

typedef struct {
    uint32_t type;
    float kd[3];       /* albedo / colour */
    float ior;
    float roughness;   /* alpha = sqrt(roughness) (TrowbridgeBSDF.cu:56-59) */
    float eta[3], k[3];
    float intensity;   /* light: emitted radiance = kd * intensity */
    float ke[3];       /* plain: emitted towards rays that are not camera rays */
    uint32_t texture;  /* 0 none, k + 1: texture k (Lambertian kd, plain camera-ray kd) */
} orc_material;

typedef struct {
    float pos[3];
    float color[3];
    float intensity;
} orc_point_light;

typedef struct {
    float pos[3], lookat[3], up[3], cts[3];
    float frame_y;
    uint32_t width, height;
} orc_camera;

typedef struct orc_scene orc_scene;

/* Barycentrics b0..b2 = e_i / det of the watertight test (Triangle.cu:563-566). */
void orc_tri_barycentrics(const orc_ray *r, const float *p0, const float *p1, const float *p2, float b[3]);

/* Camera matrices exactly as the reference composes them. */
void orc_camera_setup(const orc_camera *c, float raster_to_camera[16], float camera_to_world[16]);
void orc_camera_ray(const float *r2c, const float *c2w, uint32_t w, uint32_t h, uint32_t px, uint32_t py,
                    float u0, float u1, float o[3], float d[3]);

float orc_fresnel_dielectric(float cosO, float nO, float nD);
uint64_t orc_path_initstate(uint32_t seed, uint32_t sample);
/* orc_path, also recording each segment's ray (origin, direction: 6 floats
 * per segment, at most max_depth segments) into seg and their count. */
int orc_path_segments(const orc_scene *s, uint32_t px, uint32_t py, uint32_t sample, uint32_t seed,
                      uint32_t max_depth, uint32_t rr_depth, uint64_t stream_base, float out[3], float *seg,
                      uint32_t *nseg);

/* spheres: 4 floats (cx, cy, cz, r) each; sph_light: 1 = also sampled as a
 * sphere area light.  Triangles are xyz[3*nv] indexed by idx[3*nt]. */
orc_scene *orc_scene_create(const float *xyz, const uint32_t *idx, uint32_t nt, const uint32_t *tri_mat,
                            const float *spheres, const uint32_t *sph_mat, const uint8_t *sph_light,
                            uint32_t nsph, const orc_material *mats, uint32_t nmat,
                            const orc_point_light *lights, uint32_t nlight, const orc_camera *cam);
/* Plane shapes (12 floats each: corner, axis_x, axis_y, normal), tested
 * after the spheres. */
void orc_scene_add_planes(orc_scene *s, const float *planes, const uint32_t *mats, uint32_t n);
/* A texture (3 bytes per texel, row y = pixelMatrix[y]); returns its index. */
uint32_t orc_scene_add_texture(orc_scene *s, const uint8_t *rgb, uint32_t width, uint32_t height);
/* Texture coordinates of every triangle (6 floats each, the scene's
 * triangle order); without them the default (0,0), (1,0), (1,1). */
void orc_scene_set_tri_uv(orc_scene *s, const float *uv);
/* Corner normals of every triangle (9 floats each, the scene's triangle
 * order; a NaN first component: none) of meshes that have vertex normals. */
void orc_scene_set_tri_normals(orc_scene *s, const float *nrm9);
void orc_scene_free(orc_scene *s);
void orc_scene_camera(const orc_scene *s, float r2c[16], float c2w[16]);

int orc_path(const orc_scene *s, uint32_t px, uint32_t py, uint32_t sample, uint32_t seed, uint32_t max_depth,
             uint32_t rr_depth, uint64_t stream_base, float out[3]);

/* Accumulates into accum (W*H*3 doubles) and nvalid (W*H) over the window
 * [x0,x1) x [y0,y1). */
void orc_render(const orc_scene *s, uint32_t spp, uint32_t seed, uint32_t max_depth, uint32_t rr_depth,
                uint32_t spp_offset, uint64_t stream_base, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1,
                double *accum, uint32_t *nvalid, int nthreads);

/* RENDER_MODE PHOTON_MAPPING (Scene.cu:912-967, pathTracing.cu:84-88): the
 * photon map (n photons: locations xyz, shape ids, the kd-tree of
 * orc_kd_build as order / axis, radiance rgb and directions dir) and the
 * search's k (<= 1024) and radius. */
typedef struct {
    const float *xyz;
    const int32_t *shape;
    const uint32_t *order;
    const uint8_t *axis;
    uint32_t n;
    const double *rgb;
    const float *dir;
    uint32_t k;
    float radius;
    uint32_t mode; /* 1: PHOTON_MAPPING, 2: HYBRID (Scene.cu:963-967) */
} orc_photon_render;

/* mode 1: orc_render without next-event estimation: a path's first
 * scattering hit adds beta times the photon map's estimate there (then its
 * emission) and the path ends.  mode 2: every scattering hit adds beta times
 * half the estimate, then half of its next-event terms, and the path goes
 * on. */
void orc_render_photon(const orc_scene *s, uint32_t spp, uint32_t seed, uint32_t max_depth, uint32_t rr_depth,
                       uint32_t spp_offset, uint64_t stream_base, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1,
                       double *accum, uint32_t *nvalid, int nthreads, const orc_photon_render *ph);

/* orc_render with paths that end at their first cacheable vertex (diffuse, or
 * GGX roughness >= 0.04) at depth >= cache_depth and add beta * cache_rgb
 * there: the library's cache termination with a cache that answers one
 * constant. */
void orc_render_const_cache(const orc_scene *s, uint32_t spp, uint32_t seed, uint32_t max_depth,
                            uint32_t rr_depth, uint32_t spp_offset, uint64_t stream_base, uint32_t x0, uint32_t y0,
                            uint32_t x1, uint32_t y1, double *accum, uint32_t *nvalid, int nthreads,
                            uint32_t cache_depth, const float cache_rgb[3]);

float orc_compute_rmse(const double *sum_a, double n_a, const double *sum_b, double n_b, uint32_t npix);

/* The shading functions one at a time (shading frame, z = normal), for the
 * tests that pin them:
 *   concentric disk warp (Sampler.cu:13-36); GGX D, Lambda and the visible-
 *   normal half-vector sampler with its uniformSampleDisk(2u - 1) mapping
 *   (TrowbridgeBSDF.cu:71-151); fresnelConductor (geometricMath.cu:242-271);
 *   every material's f * |cos wi|, pdf and sampleDirection with explicit
 *   draws; the dielectric branch over a PCG32 stream; Spectrum::brightness;
 *   Russian roulette (pathTracing.cu:91-102); the sphere light's cone
 *   sampling and pdf (the library's own light, no reference counterpart). */
void orc_concentric_disk(float u0, float u1, float out[2]);
void orc_ggx_sample_half(float alpha, const float wo[3], float u0, float u1, float h[3]);
void orc_ggx_D_n(float alpha, const float *h, uint32_t n, float *out);
void orc_ggx_lambda_n(float alpha, const float *w, uint32_t n, float *out);
float orc_fresnel_conductor(float cosi, float eta, float k);
int orc_bsdf_draws(const orc_material *m);
void orc_bsdf_eval_n(const orc_material *m, const float *wo, const float *wi, uint32_t n, float *f, float *pdf);
void orc_bsdf_sample_n(const orc_material *m, const float *wo, const float *u, uint32_t n, float *wi, float *pdf,
                       int32_t *delta, float *weight);
uint32_t orc_dielectric_refractions(uint32_t seed, uint32_t n);
float orc_luminance(const float rgb[3]);
int orc_russian_roulette(const float beta[3], float u, float out[3]);
int orc_sphere_light_sample(const float center[3], float radius, const float p[3], float u0, float u1, float wi[3],
                            float *dist, float *pdf);
float orc_sphere_light_pdf(const float center[3], float radius, const float p[3]);

/* orc_render with throughput, radiance and light terms in double, as the
 * reference's Spectrum (radiometry/Spectrum.cu:15): measures the float
 * estimator's distance from the reference's arithmetic (stream base 0). */
void orc_render_f64(const orc_scene *s, uint32_t spp, uint32_t seed, uint32_t max_depth, uint32_t rr_depth,
                    uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1, double *accum, uint32_t *nvalid, int nthreads);

/* Photon pre-pass (integration/photonDispersion.cu:26-192, Scene.cu:427-455,
 * 989-1021) with the library's draw order: photon path i draws from PCG32
 * stream i of the seed (light, two direction draws, then each bounce's
 * material draws and the roulette).  Stores at most cap photons in path /
 * bounce order (key = i << 5 | bounce); returns how many the pass made.
 * orc_photon_light_table: every point light's selection probability and
 * their running sum (nlight doubles each). */
void orc_photon_light_table(const orc_scene *s, double *p, double *cdf);
uint32_t orc_trace_photons(const orc_scene *s, uint32_t n, uint32_t seed, uint32_t max_depth, uint32_t rr_depth,
                           uint32_t cap, float *pos, float *dir, float *nrm, double *rgb, int32_t *shape,
                           uint32_t *key);
/* combinePhotonContributions for any material over given search results
 * (count[q], idx / dist rows of k): sum(rgb * lightReflected(view, -dir)) /
 * (PI dmax^2) in the shading frame of nrm[q]; Lambertian and Trowbridge. */
void orc_photon_gather_mat(const double *rgb, const float *dir, uint32_t m, const uint32_t *count, const uint32_t *idx,
                           const float *dist, uint32_t k, const float *nrm, const float *view,
                           const orc_material *mats, double *out);

#ifdef __cplusplus
}
#endif
#endif
