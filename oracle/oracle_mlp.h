/*
 * oracle_mlp.h -- oracle (test infrastructure only): FP64 model of the
 * radiance cache for gradient checks (mlp.c).
 */
#ifndef RT_ORACLE_MLP_H
#define RT_ORACLE_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_CACHE_PARAMS 17408u

typedef struct {
    uint32_t levels, table_size;
    uint32_t res[16], dense[16];
    float lo[3], inv_ext[3];
} orc_cache_cfg;

/* Mean relative L2 loss over n records and, if the outputs are non-NULL,
 * its gradients with respect to the MLP parameters (17408, layout of
 * rt_cache_get_params) and the grid features (levels * table_size * 2), and
 * the network outputs (n x 3).  den_fixed (n x 3, may be NULL) replaces
 * the stop-gradient denominators y^2 + 0.01 of the loss. */
double orc_cache_loss_grad(const orc_cache_cfg *c, const float *mlp, const float *grid, const float *in,
                           const float *target, uint32_t n, int emulate_bf16, double *mlp_grad,
                           double *grid_grad, double *y_out, const double *den_fixed);

#ifdef __cplusplus
}
#endif
#endif
