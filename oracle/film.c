/*
 * film.c -- oracle restatement of the reference's film output (TEST
 * INFRASTRUCTURE ONLY; see oracle_film.h).
 *
 * Spectrum in the reference holds three doubles (radiometry/Spectrum.cu:15);
 * every operator below keeps its operation order and the float widening of
 * its constants: Spectrum::operator*(Float, Spectrum) takes a float
 * (cuda_libraries/types.h:22 defines Float as float), the float constants of
 * uncharted2_tonemap_partial are combined in float before they meet a
 * double, and brightness() weights are float literals stored in doubles.
 */
#include "oracle_film.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

/* Film.cu:13-21 operator ids must match the shipped library's (rt.h). */

void orc_film_pixel(const double *sums, const uint32_t *nvalid, uint64_t samples, size_t i, double rgb[3])
{
    if (!nvalid) {
        rgb[0] = sums[3 * i]; rgb[1] = sums[3 * i + 1]; rgb[2] = sums[3 * i + 2];
        return;
    }
    /* the renderer's pixel sum: the valid samples' sum plus their mean for
     * every invalid sample (integration/pathTracing.cu:474-486) */
    const double s = nvalid[i] ? (double)samples / (double)nvalid[i] : 0.0;
    rgb[0] = sums[3 * i] * s; rgb[1] = sums[3 * i + 1] * s; rgb[2] = sums[3 * i + 2] * s;
}

/* radiometry/Spectrum.cu:68-75 */
double orc_brightness(const double p[3])
{
    const double wr = 0.212671f, wg = 0.715160f, wb = 0.072169f;
    return wr * p[0] + wg * p[1] + wb * p[2];
}

/* Film.cu:120-154: maxima start at 0 and only a larger value replaces them */
void orc_film_maxima(const double *sums, const uint32_t *nvalid, size_t npix, uint64_t samples, double *max,
                     double *max_lum)
{
    double m = 0, ml = 0;
    for (size_t i = 0; i < npix; ++i) {
        double p[3];
        orc_film_pixel(sums, nvalid, samples, i, p);
        if (p[0] > m) m = p[0];
        if (p[1] > m) m = p[1];
        if (p[2] > m) m = p[2];
        const double l = orc_brightness(p);
        if (l > ml) ml = l;
    }
    const double n = (double)(long long)samples;
    *max = m / n;
    *max_lum = ml / n;
}

/* cuda_libraries/math.cu:44-47 and geometricMath.cu:127-130:
 * max(a, min(b, x)); a NaN x yields b */
static double clampd(double x, double a, double b)
{
    const double m = (x < b) ? x : b;
    return (a > m) ? a : m;
}

static void clamp3(double p[3])
{
    for (int c = 0; c < 3; ++c) p[c] = clampd(p[c], 0.0, 1.0);
}

/* Film.cu:202-211 */
static double uncharted2_partial(double x)
{
    const float A = 0.15, B = 0.50, C = 0.10, D = 0.20, E = 0.02, F = 0.30;
    const float cb = C * B, de = D * E, df = D * F, ef = E / F;
    return ((x * (x * (double)A + (double)cb) + (double)de) / (x * (x * (double)A + (double)B) + (double)df)) -
           (double)ef;
}

/* Film.cu:224-229 */
static double rtt_and_odt_fit(double v)
{
    const float k = 0.983729;
    const double a = v * (v + 0.0245786) - 0.000090537;
    const double b = v * (v * (double)k + 0.4329510) + 0.238081;
    return a / b;
}

/* Matrix3x3<double>::operator*(Spectrum) (cuda_libraries/Matrix3x3.cu:137-143) */
static void mat3_mul(const double m[9], const double v[3], double out[3])
{
    const double x = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
    const double y = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
    const double z = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
    out[0] = x; out[1] = y; out[2] = z;
}

/* toneMaxPixel (Film.cu:266-283) with the operators of Film.cu:156-264 */
void orc_tonemap_pixel(const double in[3], double max, double max_lum, uint64_t samples, int mode, double out[3])
{
    const double n = (double)(long long)samples;
    double p[3] = {in[0] / n, in[1] / n, in[2] / n};
    switch (mode) {
    case ORC_TONE_EQUALIZE: /* Film.cu:156-162 */
        for (int c = 0; c < 3; ++c) p[c] = p[c] / max;
        break;
    case ORC_TONE_REINHARD: { /* Film.cu:164-180 */
        if (orc_brightness(p) == 0) break;
        const double l_old = orc_brightness(p);
        const double numerator = l_old * (1.0 + (l_old / (max_lum * max_lum)));
        const double l_new = numerator / (1.0 + l_old);
        const double s = l_new / l_old;
        for (int c = 0; c < 3; ++c) p[c] = p[c] * s;
        clamp3(p);
        break;
    }
    case ORC_TONE_REINHARD_JODIE: { /* Film.cu:182-193: lerp(x, a, b) = (1 - x) * a + x * b */
        const double lum = orc_brightness(p);
        for (int c = 0; c < 3; ++c) {
            const double rp = p[c] / (1 + lum);
            const double x = p[c] / (p[c] + (double)1.0f);
            p[c] = (1.0 - x) * rp + x * rp;
        }
        clamp3(p);
        break;
    }
    case ORC_TONE_ACES: { /* Film.cu:231-247 */
        static const double in_m[9] = {0.59719, 0.35458, 0.04823, 0.07600, 0.90834, 0.01566, 0.02840, 0.13383, 0.83777};
        static const double out_m[9] = {1.60475, -0.53108, -0.07367, -0.10208, 1.10813, -0.00605,
                                        -0.00327, -0.07276, 1.07602};
        double cur[3];
        mat3_mul(in_m, p, cur);
        for (int c = 0; c < 3; ++c) cur[c] = rtt_and_odt_fit(cur[c]);
        mat3_mul(out_m, cur, p);
        clamp3(p);
        break;
    }
    case ORC_TONE_UNCHARTED2: { /* Film.cu:213-222, no clamp */
        const float exposure_bias = 2.0;
        const double white_scale = 1.0 / uncharted2_partial(11.2);
        for (int c = 0; c < 3; ++c) p[c] = uncharted2_partial(p[c] * (double)exposure_bias) * white_scale;
        break;
    }
    case ORC_TONE_GAMMA_CLAMP: /* Film.cu:255-264, no clamp */
        for (int c = 0; c < 3; ++c) p[c] = pow(p[c] / max, 1.0 / 2.2);
        break;
    default:
        break;
    }
    out[0] = p[0]; out[1] = p[1]; out[2] = p[2];
}

uint8_t orc_to_byte(double v)
{
    const double x = v * 255;
    /* cvttsd2si: out of range or NaN gives 0x80000000 */
    const int32_t i = (x > -2147483649.0 && x < 2147483648.0) ? (int32_t)x : INT32_MIN;
    return (uint8_t)(uint32_t)i;
}

static uint32_t bmp_row_bytes(uint32_t w) { return (3u * w + 3u) & ~3u; }

size_t orc_image_pixel_bytes(int format, uint32_t width, uint32_t height)
{
    return format == ORC_IMAGE_BMP ? (size_t)bmp_row_bytes(width) * height : (size_t)3 * width * height;
}

static void put_u16(uint8_t *p, uint16_t v) { p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); }
static void put_u32(uint8_t *p, uint32_t v)
{
    for (int k = 0; k < 4; ++k) p[k] = (uint8_t)(v >> (8 * k));
}

size_t orc_image_header(int format, uint32_t width, uint32_t height, uint8_t *out, size_t cap)
{
    uint8_t h[64];
    size_t len;
    if (format == ORC_IMAGE_BMP) {
        /* BMP_Format.cu: BitmapImage fields, little-endian */
        const uint32_t image_size = (width + 4u - (width % 4u)) * height;
        memset(h, 0, sizeof h);
        h[0] = 'B'; h[1] = 'M';
        put_u32(h + 2, 14u + 40u + image_size);
        put_u32(h + 10, 54u);
        put_u32(h + 14, 40u);
        put_u32(h + 18, width);
        put_u32(h + 22, height);
        put_u16(h + 26, 1);
        put_u16(h + 28, 24);
        put_u32(h + 30, 0);
        put_u32(h + 34, image_size);
        put_u32(h + 38, 3780u);
        put_u32(h + 42, 3780u);
        len = 54;
    } else {
        /* Film::writeToPPM (Film.cu:339-350) */
        len = (size_t)snprintf((char *)h, sizeof h, "P6\n%u %u\n255\n", width, height);
    }
    memcpy(out, h, len < cap ? len : cap);
    return len;
}

void orc_film_image(const double *sums, const uint32_t *nvalid, uint32_t width, uint32_t height, uint64_t samples,
                    int mode, int format, uint8_t *pixels, double *mapped)
{
    const size_t npix = (size_t)width * height;
    double max, max_lum;
    orc_film_maxima(sums, nvalid, npix, samples, &max, &max_lum);
    const uint32_t row = bmp_row_bytes(width);
    for (uint32_t y = 0; y < height; ++y) {
        for (uint32_t x = 0; x < width; ++x) {
            const size_t i = (size_t)y * width + x;
            double p[3], t[3];
            orc_film_pixel(sums, nvalid, samples, i, p);
            orc_tonemap_pixel(p, max, max_lum, samples, mode, t);
            if (mapped) { mapped[3 * i] = t[0]; mapped[3 * i + 1] = t[1]; mapped[3 * i + 2] = t[2]; }
            if (format == ORC_IMAGE_BMP) {
                /* toneMapBMP (Film.cu:307-336): array row j = film row height-1-j */
                uint8_t *o = pixels + (size_t)(height - 1 - y) * row + 3u * x;
                o[0] = orc_to_byte(t[2]); o[1] = orc_to_byte(t[1]); o[2] = orc_to_byte(t[0]);
            } else {
                uint8_t *o = pixels + 3 * i;
                o[0] = orc_to_byte(t[0]); o[1] = orc_to_byte(t[1]); o[2] = orc_to_byte(t[2]);
            }
        }
        if (format == ORC_IMAGE_BMP)
            for (uint32_t k = 3u * width; k < row; ++k) pixels[(size_t)(height - 1 - y) * row + k] = 0;
    }
}
