/*
 * oracle.h -- CPU restatement of the reference renderer's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code.  It is used by tests/ (as the checker), by
 * __graft_entry__.smoke() (as the checker) and by bench.py's cpu_baseline leg
 * (as the timed CPU baseline, "kind": "port").
 *
 * Every function restates the reference's algorithm in plain C11 with the
 * reference's floating-point operation order (compiled with
 * -ffp-contract=off, no fast-math), so FP32 results are bit-comparable with
 * the HIP kernels, which are compiled the same way.
 *
 * Pinning status (what checks each function against something other than
 * itself) is listed per function below and in tests/test_oracle_*.py.
 */
#ifndef RT_ORACLE_H
#define RT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Scalar helpers (cuda_libraries/math.cu:117-194)                      */
/* ------------------------------------------------------------------ */
float orc_error_bound(int n);            /* math.cu:117-120 */
float orc_next_float_up(float x);        /* math.cu:147-170 */
float orc_next_float_down(float x);      /* math.cu:172-194 */

/* A ray as the reference builds it: Ray(o, d, tMax) precomputes 1/d
 * (geometry/Ray.cu:32-39). */
typedef struct {
    float o[3], d[3], inv[3];
    float tmax;
} orc_ray;

void orc_ray_make(orc_ray *r, const float o[3], const float d[3], float tmax);

/* Per-candidate record of one Triangle::intersects call. */
typedef struct {
    int   reached;   /* passed every test up to shapes/Triangle.cu:569 */
    float t;         /* scaledOffset * invDet (the candidate distance) */
    int   masked;    /* failed the conservative deltaT test (Triangle.cu:592) */
    int   fp64;      /* took the FP64 edge-function fallback (Triangle.cu:507-522) */
} orc_tri_info;

/* Triangle::intersects (shapes/Triangle.cu:433-643) for one triangle.
 * hitOffset is in/out exactly as in the reference (lowered before the deltaT
 * rejection).  info may be NULL.  Returns 1 on an accepted hit. */
int orc_tri_intersects(const orc_ray *r, const float *p0, const float *p1,
                       const float *p2, float *hitOffset, orc_tri_info *info);
/* The same test with the deltaT rejection applied before hitOffset is
 * lowered (the library's order): a rejected candidate leaves hitOffset
 * alone, so a traversal returns the closest candidate that passes, in any
 * visit order. */
int orc_tri_intersects_guarded(const orc_ray *r, const float *p0, const float *p1,
                               const float *p2, float *hitOffset, orc_tri_info *info);

/* Bound3f::intersects (geometry/BoundingBox.cu:296-329), CUDA fminf/fmaxf
 * semantics for min/max (NaN operands are ignored). */
int orc_box_intersects(const float lo[3], const float hi[3], const orc_ray *r,
                       float *hitOffset);

/* ------------------------------------------------------------------ */
/* Reference BVH (cuda_libraries/BVH.cu, shapes/Triangle.cu:703-776)    */
/* ------------------------------------------------------------------ */
typedef struct orc_refbvh orc_refbvh;

/* Morton-sorted pairing tree, leaf size 1 (BVH.cu:153-341). */
orc_refbvh *orc_refbvh_build(const float *xyz, const uint32_t *idx, uint32_t nt);
void orc_refbvh_free(orc_refbvh *b);
uint32_t orc_refbvh_root(const orc_refbvh *b);
/* Leaf triangle ids in Morton order (for the sort pin). */
void orc_refbvh_sorted_ids(const orc_refbvh *b, uint32_t *out);
/* Morton code of triangle i exactly as Morton3DFunctor computes it. */
uint32_t orc_refbvh_morton(const orc_refbvh *b, uint32_t i);

/* BVH::intersectRay (BVH.cu:355-448). transparent may be NULL. */
int32_t orc_refbvh_intersect(const orc_refbvh *b, const orc_ray *r,
                             const uint8_t *transparent, int check_shadow,
                             float *t_out, uint64_t *n_box, uint64_t *n_tri);
/* The same walk with orc_tri_intersects_guarded (the renderer's queries). */
int32_t orc_refbvh_intersect_guarded(const orc_refbvh *b, const orc_ray *r, const uint8_t *transparent,
                                     int check_shadow, float *t_out);

/* Batch over SoA rays with nthreads pthreads (BVH.cu:586-635 analogue). */
void orc_refbvh_trace(const orc_refbvh *b, const float *ox, const float *oy,
                      const float *oz, const float *dx, const float *dy,
                      const float *dz, const float *tmax, uint32_t n,
                      int32_t *prim, float *t, int nthreads);

/* Brute force over all triangles in index order with the exact
 * Triangle::intersects arithmetic, plus the order-independence analysis:
 *   prim[i]  = closest accepted candidate (-1 = none)
 *   t[i]     = its distance
 *   flags[i] bit0 = amb_tie  (>=2 accepted candidates share the minimal t)
 *            bit1 = amb_mask (a deltaT-masked candidate has t <= minimal
 *                             accepted t, so the reference result depends
 *                             on visit order)
 *            bit2 = fp64 fallback taken by the winning candidate        */
void orc_brute_trace(const float *xyz, const uint32_t *idx, uint32_t nt,
                     const uint8_t *transparent, int check_shadow,
                     const float *ox, const float *oy, const float *oz,
                     const float *dx, const float *dy, const float *dz,
                     const float *tmax, uint32_t n, int32_t *prim, float *t,
                     uint8_t *flags, int nthreads);

/* Traversal of a two-wide BVH in the library's exported node layout (see
 * include/rt.h: rt_bvh_export), counting node visits and triangle tests.
 * Used to price the algorithmic bytes of the HIP traversal and as a
 * cross-check of the exported tree.  Returns closest-hit ids like the
 * kernel.  nodes: 16 floats per node; tris: 12 floats per triangle. */
void orc_bvh2_trace(const float *nodes, const float *tris, const float *ox,
                    const float *oy, const float *oz, const float *dx,
                    const float *dy, const float *dz, const float *tmax,
                    uint32_t n, int32_t *prim, float *t, uint64_t *n_node,
                    uint64_t *n_tri, int nthreads);

/* Traversal of the library's four-wide quantized layout (include/rt.h:
 * rt_bvh_export4), operation for operation as the kernel walks it: counts
 * node visits and triangle tests and reports the deepest stack used. */
/* The GPU-built BVH restated on the CPU (lbvh.c): nodes2 holds
 * max(1, nt - 1) x 16 floats, tris (nt + 1) x 12, nodes4 max(1, nt - 1) x 16
 * (written when wide); info = num_nodes2, num_leaves, max_depth,
 * num_nodes4, depth4, max_stack4.  Returns -1 on bad input. */
int orc_lbvh_build(const float *xyz, uint32_t nv, const uint32_t *idx, uint32_t nt, uint32_t max_leaf,
                   float trav_cost, int wide, float *nodes2, float *tris, float *nodes4, uint32_t *info);

void orc_bvh4_trace(const float *nodes4, const float *tris, const float *ox,
                    const float *oy, const float *oz, const float *dx,
                    const float *dy, const float *dz, const float *tmax,
                    uint32_t n, int32_t *prim, float *t, uint64_t *n_node,
                    uint64_t *n_tri, uint32_t *max_sp, int nthreads);

/* ------------------------------------------------------------------ */
/* Sphere (shapes/Sphere.cu)                                            */
/* ------------------------------------------------------------------ */
typedef struct {
    float o2w[16], w2o[16]; /* row-major 4x4 */
    float radius;
    float center[3];
} orc_sphere;

/* Sphere(center, radius): translate(c) * rotateZ(90) (Sphere.cu:233-245). */
void orc_sphere_init(orc_sphere *s, const float c[3], float radius);
/* Sphere::intersects (Sphere.cu:357-398) without the shadow/transparency
 * early-out (callers apply it).  Returns 1 and lowers *hitOffset on a hit. */
int orc_sphere_intersects(const orc_sphere *s, const orc_ray *r, float *hitOffset);
/* Sphere::intersect (Sphere.cu:277-355): also world hit point, outward
 * world normal and world pError. */
int orc_sphere_intersect(const orc_sphere *s, const orc_ray *r, float *hitOffset,
                         float p[3], float n[3], float perr[3]);

typedef struct {
    float o2w[16], w2o[16];
} orc_plane;

/* Plane(corner, axisX, axisY, normal) (Plane.cu:27-48). */
void orc_plane_init(orc_plane *p, const float corner[3], const float axis_x[3], const float axis_y[3],
                    const float normal[3]);
/* Plane::intersect (Plane.cu:58-125): 1 and lowers *hitOffset on a hit; p,
 * n, perr (may be NULL) get the world point, the normal facing the ray and
 * the point's error bound. */
int orc_plane_intersect(const orc_plane *pl, const orc_ray *r, float *hitOffset, float p[3], float n[3],
                        float perr[3]);

/* The same two hits, also giving SurfaceInteraction::surfacePoint: the
 * plane's (u, v), the sphere's (theta, phi) in degrees. */
int orc_plane_intersect_uv(const orc_plane *pl, const orc_ray *r, float *hitOffset, float p[3], float n[3],
                           float perr[3], float uv[2]);
int orc_sphere_intersect_uv(const orc_sphere *s, const orc_ray *r, float *hitOffset, float p[3], float n[3],
                            float perr[3], float uv[2]);

/* MESA cofactor inverse used by the reference (Matrix4x4.cu:303-440). */
void orc_mat4_inverse(const float *m, float *out);

#ifdef __cplusplus
}
#endif
#endif
