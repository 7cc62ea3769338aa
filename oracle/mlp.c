/*
 * mlp.c -- oracle (test infrastructure only): FP64 model of the radiance
 * cache (hash-grid encoding + bias-free 64-wide ReLU MLP + relative L2 loss)
 * with forward and backward passes, for checking the GPU's BF16/MFMA
 * training kernels.  With emulate_bf16 = 1 the values the GPU stores as BF16
 * (grid features, encodings, weights, activations, backpropagated
 * gradients) are rounded to BF16 at the same points, so the remaining
 * difference is FP32-versus-FP64 accumulation; with 0 it is the exact FP64
 * model.
 *
 * The encoding follows Mueller et al., "Instant Neural Graphics Primitives
 * with a Multiresolution Hash Encoding" (SIGGRAPH 2022); the loss follows
 * Mueller et al., "Real-time Neural Radiance Caching for Path Tracing"
 * (SIGGRAPH 2021).  The reference renderer has no such component.
 */
#include "oracle_mlp.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define W 64
#define LAYERS 5
#define OUTW 16
#define INSTRIDE 12

static double bf(double x, int emulate)
{
    if (!emulate) return x;
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

static uint32_t grid_index(const orc_cache_cfg *c, uint32_t l, uint32_t x, uint32_t y, uint32_t z)
{
    const uint32_t r1 = c->res[l] + 1;
    if (c->dense[l]) return x + r1 * (y + r1 * z);
    return (x * 1u ^ y * 2654435761u ^ z * 805459861u) & (c->table_size - 1u);
}

static void corners(const orc_cache_cfg *c, uint32_t l, const float p[3], uint32_t idx[8], double w[8])
{
    /* the same FP32 coordinate arithmetic as the kernel */
    float r = (float)c->res[l];
    float x = p[0] * r, y = p[1] * r, z = p[2] * r;
    float fx0 = floorf(x), fy0 = floorf(y), fz0 = floorf(z);
    uint32_t x0 = (uint32_t)fx0, y0 = (uint32_t)fy0, z0 = (uint32_t)fz0;
    float fx = x - fx0, fy = y - fy0, fz = z - fz0;
    for (int k = 0; k < 8; ++k) {
        uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
        uint32_t xi = x0 + dx, yi = y0 + dy, zi = z0 + dz;
        if (xi > c->res[l]) xi = c->res[l];
        if (yi > c->res[l]) yi = c->res[l];
        if (zi > c->res[l]) zi = c->res[l];
        idx[k] = grid_index(c, l, xi, yi, zi);
        float wf = (dx ? fx : 1 - fx) * (dy ? fy : 1 - fy) * (dz ? fz : 1 - fz);
        w[k] = (double)wf;
    }
}

static void norm_pos(const orc_cache_cfg *c, const float *q, float p[3])
{
    for (int k = 0; k < 3; ++k) {
        float v = (q[k] - c->lo[k]) * c->inv_ext[k];
        p[k] = fminf(fmaxf(v, 0.0f), 1.0f);
    }
}

static uint32_t layer_off(int l) { return (uint32_t)l * W * W; }
static int layer_out(int l) { return l == LAYERS - 1 ? OUTW : W; }

double orc_cache_loss_grad(const orc_cache_cfg *c, const float *mlp, const float *grid, const float *in,
                           const float *target, uint32_t n, int emulate, double *mlp_grad, double *grid_grad,
                           double *y_out, const double *den_fixed)
{
    const uint32_t L = c->levels;
    const size_t gp = (size_t)L * c->table_size * 2;
    if (mlp_grad) memset(mlp_grad, 0, sizeof(double) * ORC_CACHE_PARAMS);
    if (grid_grad) memset(grid_grad, 0, sizeof(double) * gp);
    double *wq = malloc(sizeof(double) * ORC_CACHE_PARAMS);
    for (uint32_t i = 0; i < ORC_CACHE_PARAMS; ++i) wq[i] = bf(mlp[i], emulate);
    double act[LAYERS][W], y[OUTW], g[OUTW], d[W], dn[W];
    double loss = 0;
    const double scale = 1.0 / (3.0 * (double)n);
    for (uint32_t s = 0; s < n; ++s) {
        const float *q = in + (size_t)s * INSTRIDE;
        float p[3];
        norm_pos(c, q, p);
        uint32_t cidx[16][8];
        double cw[16][8];
        for (uint32_t l = 0; l < L; ++l) {
            corners(c, l, p, cidx[l], cw[l]);
            double f0 = 0, f1 = 0;
            for (int k = 0; k < 8; ++k) {
                size_t e = ((size_t)l * c->table_size + cidx[l][k]) * 2;
                f0 += cw[l][k] * bf(grid[e], emulate);
                f1 += cw[l][k] * bf(grid[e + 1], emulate);
            }
            act[0][2 * l] = bf(f0, emulate);
            act[0][2 * l + 1] = bf(f1, emulate);
        }
        uint32_t o = 2 * L;
        for (int k = 3; k < 10; ++k) act[0][o++] = bf(q[k], emulate);
        act[0][o++] = 1.0;
        for (; o < W; ++o) act[0][o] = 0;
        /* forward */
        for (int l = 0; l < LAYERS; ++l) {
            const int no = layer_out(l);
            const double *Wl = wq + layer_off(l);
            for (int j = 0; j < no; ++j) {
                double acc = 0;
                for (int k = 0; k < W; ++k) acc += act[l][k] * Wl[k * no + j];
                if (l + 1 < LAYERS) {
                    act[l + 1][j] = bf(acc > 0 ? acc : 0, emulate);
                } else {
                    y[j] = acc;
                }
            }
        }
        if (y_out) for (int c3 = 0; c3 < 3; ++c3) y_out[(size_t)s * 3 + c3] = y[c3];
        /* relative L2 */
        for (int j = 0; j < OUTW; ++j) g[j] = 0;
        for (int c3 = 0; c3 < 3; ++c3) {
            double yy = y[c3], t = target[(size_t)s * 3 + c3];
            /* the denominator is a stop-gradient quantity; den_fixed lets a
             * finite-difference check hold it at its unperturbed value */
            double den = den_fixed ? den_fixed[(size_t)s * 3 + c3] : yy * yy + 0.01, dd = yy - t;
            loss += dd * dd / den;
            g[c3] = bf(2.0 * dd / den * scale, emulate);
        }
        if (!mlp_grad && !grid_grad) continue;
        /* backward: d holds dL/d(layer input pre-mask) of the current layer */
        double *cur = g;
        int cur_w = OUTW;
        for (int l = LAYERS - 1; l >= 0; --l) {
            const int no = layer_out(l);
            const double *Wl = wq + layer_off(l);
            if (mlp_grad)
                for (int k = 0; k < W; ++k)
                    for (int j = 0; j < no; ++j) mlp_grad[layer_off(l) + k * no + j] += act[l][k] * cur[j];
            for (int k = 0; k < W; ++k) {
                double acc = 0;
                for (int j = 0; j < cur_w; ++j) acc += cur[j] * Wl[k * no + j];
                if (l > 0) {
                    /* ReLU mask on the layer's input activation */
                    dn[k] = act[l][k] > 0 ? bf(acc, emulate) : 0.0;
                } else {
                    dn[k] = acc; /* encoding gradient stays FP32 on the GPU */
                }
            }
            memcpy(d, dn, sizeof d);
            cur = d;
            cur_w = W;
        }
        if (grid_grad)
            for (uint32_t l = 0; l < L; ++l)
                for (int k = 0; k < 8; ++k) {
                    size_t e = ((size_t)l * c->table_size + cidx[l][k]) * 2;
                    grid_grad[e] += cw[l][k] * d[2 * l];
                    grid_grad[e + 1] += cw[l][k] * d[2 * l + 1];
                }
    }
    free(wq);
    return loss / (3.0 * (double)n);
}
