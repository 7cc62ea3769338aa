/*
 * oracle_photon.h -- CPU restatement of the reference's photon map search
 * (TEST INFRASTRUCTURE ONLY; see oracle.h): nn::KDTree's build and
 * nearest_neighbors (cuda_libraries/kdtree.h) as Scene::search_nearest calls
 * them (integration/Scene.cu:704-722), plus a brute-force k-nearest search
 * that states the same result by definition.
 *
 * Distances are float, sqrt(((0 + dx*dx) + dy*dy) + dz*dz), as the
 * reference's default norm.  A photon qualifies if its distance is strictly
 * below the radius and (for a query shape other than -1) its shape matches.
 *
 * Pinning: the brute force is the definition of the result set (the k
 * qualifying photons of smallest distance); tests/test_oracle_photon.py
 * checks the restated tree search against it on queries whose k-th and
 * (k+1)-th distances differ, and the restated build against the median and
 * widest-axis properties of every node.
 */
#ifndef RT_ORACLE_PHOTON_H
#define RT_ORACLE_PHOTON_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* kdtree.h build_tree: order[i] = input index at tree position i, axis[i] =
 * split axis of the range whose median is i.  Ties in a coordinate are put
 * in input order (nth_element leaves them unspecified).  Returns the depth. */
uint32_t orc_kd_build(const float *xyz, uint32_t n, uint32_t *order, uint8_t *axis);

/* kdtree.h nearest_neighbors_impl over a tree from orc_kd_build; writes up to
 * k input indices and distances in the reference's result order (insertion
 * order below k, then the layout of libstdc++'s make_heap / push_heap /
 * pop_heap), returns the count. */
uint32_t orc_kd_knn(const float *xyz, const int32_t *shape, const uint32_t *order, const uint8_t *axis, uint32_t n,
                    const float q[3], int32_t qshape, uint32_t k, float radius, uint32_t *idx, float *dist);

/* The bookkeeping of nearest_neighbors_impl for a sequence of n accepted
 * candidates (ids 0..n-1, distances dist): push_back below k, make_heap at
 * k, then push_heap + pop_heap + pop_back.  Writes the final ids and
 * distances in vector order and their count (min(n, k)). */
void orc_heap_sequence(const float *dist, uint32_t n, uint32_t k, uint32_t *ids, float *out_dist, uint32_t *count);

/* Brute force: the k qualifying photons of smallest (distance, index),
 * sorted that way. */
uint32_t orc_brute_knn(const float *xyz, const int32_t *shape, uint32_t n, const float q[3], int32_t qshape,
                       uint32_t k, float radius, uint32_t *idx, float *dist);

/* Many queries (m * 3 floats, qshape m ids or null), nthreads threads:
 * brute force (tree == 0) or tree search (tree != 0, over order / axis). */
void orc_knn_batch(const float *xyz, const int32_t *shape, const uint32_t *order, const uint8_t *axis, uint32_t n,
                   const float *q, const int32_t *qshape, uint32_t m, uint32_t k, float radius, int tree,
                   uint32_t *count, uint32_t *idx, float *dist, int nthreads);

/* combinePhotonContributions (integration/Scene.cu:607-634) on a Lambertian
 * surface for m queries, over the brute-force k nearest photons:
 * sum(rgb * (kd * INV_PI) * |cos(normal, -dir)|) * 1 / (PI * d_max^2). */
void orc_photon_gather_lambert(const float *xyz, const int32_t *shape, uint32_t n, const double *rgb, const float *dir,
                               const float *q, const float *nrm, const float *kd, const int32_t *qshape, uint32_t m,
                               uint32_t k, float radius, double *out);

/* Scene::combinePhotonsSubsurface (integration/Scene.cu:667-702) for m
 * queries over search_nearest(p, radius, unlimited, shape): every photon
 * strictly within radius on the query's shape, in the walk's insertion
 * order, summed as rgb * Sd * kd with Sd = INV_PI * Ft_i * Rd * Ft_o
 * (diffusionTermBSSRDF, geometricMath.cu:500-522), divided by
 * 2 PI dmax^2 (0 / 0 without photons).  params: kd[3], sigma_s[3],
 * sigma_a[3], g, eta.  orc_dipole_rd: impPaper_dipoleDiffusionAproximation
 * (geometricMath.cu:451-483) of one channel at distance r. */
void orc_photon_subsurface(const float *xyz, const int32_t *shape, const uint32_t *order, const uint8_t *axis,
                           uint32_t n, const double *rgb, const float *pdir, const float *pnrm, const float *q,
                           const float *qnrm, const float *qview, const int32_t *qshape, uint32_t m, float radius,
                           const float params[11], double *out);
double orc_dipole_rd(double sigma_s, double sigma_a, float g, float eta, float r);

#ifdef __cplusplus
}
#endif

#endif
