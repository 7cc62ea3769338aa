// gen.cpp -- deterministic synthetic workloads (host).  All arithmetic is
// done in double and rounded once to float so the streams are independent of
// compiler contraction choices.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "rt_internal.h"
#include "rt_pcg.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

double gauss(rt::Pcg32 &g)
{
    // Box-Muller; u1 in (0, 1]
    double u1 = 1.0 - rt::pcg32_double(g);
    double u2 = rt::pcg32_double(g);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
}

// Uniform random rotation from three uniforms (Shoemake, "Uniform random
// rotations", Graphics Gems III, 1992): unit quaternion -> 3x3 matrix.
void random_rotation(rt::Pcg32 &g, double R[3][3])
{
    double u1 = rt::pcg32_double(g), u2 = rt::pcg32_double(g), u3 = rt::pcg32_double(g);
    double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
    double x = a * std::sin(2 * kPi * u2), y = a * std::cos(2 * kPi * u2);
    double z = b * std::sin(2 * kPi * u3), w = b * std::cos(2 * kPi * u3);
    R[0][0] = 1 - 2 * (y * y + z * z); R[0][1] = 2 * (x * y - z * w);     R[0][2] = 2 * (x * z + y * w);
    R[1][0] = 2 * (x * y + z * w);     R[1][1] = 1 - 2 * (x * x + z * z); R[1][2] = 2 * (y * z - x * w);
    R[2][0] = 2 * (x * z - y * w);     R[2][1] = 2 * (y * z + x * w);     R[2][2] = 1 - 2 * (x * x + y * y);
}

void uniform_sphere(rt::Pcg32 &g, double n[3])
{
    double z = 1.0 - 2.0 * rt::pcg32_double(g);
    double r = std::sqrt(std::max(0.0, 1.0 - z * z));
    double phi = 2.0 * kPi * rt::pcg32_double(g);
    n[0] = r * std::cos(phi); n[1] = r * std::sin(phi); n[2] = z;
}

// Orthonormal basis about n (Duff et al., "Building an Orthonormal Basis,
// Revisited", JCGT 2017).
void onb(const double n[3], double t[3], double b[3])
{
    double s = std::copysign(1.0, n[2]);
    double a = -1.0 / (s + n[2]);
    double c = n[0] * n[1] * a;
    t[0] = 1.0 + s * n[0] * n[0] * a; t[1] = s * c; t[2] = -s * n[0];
    b[0] = c; b[1] = s + n[1] * n[1] * a; b[2] = -n[1];
}

}  // namespace

extern "C" {

rt_status rt_gen_soup(uint32_t nt, uint64_t seed, float edge_mu, float edge_sigma, float *tri9)
{
    if (nt && !tri9) { rt::set_error("rt_gen_soup: null output"); return RT_ERR_INVALID; }
    rt::Pcg32 g = rt::pcg32_seed(seed, 0x5eed50u);
    const double lmu = std::log((double)edge_mu);
    for (uint32_t i = 0; i < nt; ++i) {
        double c[3] = {rt::pcg32_double(g), rt::pcg32_double(g), rt::pcg32_double(g)};
        double R[3][3];
        random_rotation(g, R);
        double phi0 = 2 * kPi * rt::pcg32_double(g);
        for (int k = 0; k < 3; ++k) {
            double len = std::exp(lmu + (double)edge_sigma * gauss(g));
            double phi = phi0 + 2 * kPi * k / 3.0 + (rt::pcg32_double(g) * 0.6 - 0.3);
            double l[3] = {len * std::cos(phi), len * std::sin(phi), 0.0};
            for (int a = 0; a < 3; ++a) {
                double v = c[a] + R[a][0] * l[0] + R[a][1] * l[1] + R[a][2] * l[2];
                tri9[9 * (size_t)i + 3 * k + a] = (float)v;
            }
        }
    }
    return RT_OK;
}

rt_status rt_gen_rays(uint32_t w, uint32_t h, const float eye[3], float vfov_deg, uint32_t n_incoherent,
                      uint64_t seed, float *ox, float *oy, float *oz, float *dx, float *dy, float *dz)
{
    const size_t nc = (size_t)w * h, n = nc + n_incoherent;
    if (n && (!ox || !oy || !oz || !dx || !dy || !dz || (nc && !eye))) {
        rt::set_error("rt_gen_rays: null argument");
        return RT_ERR_INVALID;
    }
    // Pinhole looking along +z, up +y, right +x, pixel centres, row 0 at top.
    const double th = std::tan(0.5 * (double)vfov_deg * kPi / 180.0);
    const double s = h ? 2.0 * th / (double)h : 0.0;
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            size_t i = (size_t)y * w + x;
            double d[3] = {((double)x + 0.5 - 0.5 * w) * s, (0.5 * h - (double)y - 0.5) * s, 1.0};
            double l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            ox[i] = eye[0]; oy[i] = eye[1]; oz[i] = eye[2];
            dx[i] = (float)(d[0] / l); dy[i] = (float)(d[1] / l); dz[i] = (float)(d[2] / l);
        }
    rt::Pcg32 g = rt::pcg32_seed(seed, 0x0ae1u);
    for (uint32_t j = 0; j < n_incoherent; ++j) {
        size_t i = nc + j;
        double o[3] = {rt::pcg32_double(g), rt::pcg32_double(g), rt::pcg32_double(g)};
        double nrm[3], t[3], b[3];
        uniform_sphere(g, nrm);
        onb(nrm, t, b);
        // cosine-weighted hemisphere about nrm (Malley's method)
        double u1 = rt::pcg32_double(g), u2 = rt::pcg32_double(g);
        double r = std::sqrt(u1), phi = 2 * kPi * u2;
        double lx = r * std::cos(phi), ly = r * std::sin(phi), lz = std::sqrt(std::max(0.0, 1.0 - u1));
        double d[3];
        for (int a = 0; a < 3; ++a) d[a] = lx * t[a] + ly * b[a] + lz * nrm[a];
        double l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        ox[i] = (float)o[0]; oy[i] = (float)o[1]; oz[i] = (float)o[2];
        dx[i] = (float)(d[0] / l); dy[i] = (float)(d[1] / l); dz[i] = (float)(d[2] / l);
    }
    return RT_OK;
}

}  // extern "C"
