/*
 * oracle_film.h -- CPU restatement of the reference's film output: the
 * tone-mapping operators of camera/Film.cu:120-283 and the pixel arrays that
 * Film::toneMapBMP / toneMapPPM (Film.cu:285-337) hand to the BMP writer
 * (camera/BMP_Format.cu) and the PPM stream.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Plain C in double precision with
 * the reference's operation order, compiled without contraction, on x86-64:
 * the byte conversion (unsigned char)(v * 255) is restated with the x86
 * truncation the reference's host compiler emits (cvttsd2si to 32 bits, low
 * byte kept), so out-of-range and NaN values become 0 and values in
 * [256, 2^31) wrap modulo 256.
 *
 * Pinning: tests/test_oracle_film.py checks each operator against its
 * published closed form (Reinhard et al. 2002 eq. 4, Hable's filmic curve,
 * Hill's ACES fit with its two matrices, Jodie's Reinhard variant as the
 * reference composes it), special values that the mathematics fixes, the
 * BMP layout against an independent decoder, and the PPM stream against the
 * netpbm format definition.
 */
#ifndef RT_ORACLE_FILM_H
#define RT_ORACLE_FILM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ToneMappingType (camera/Film.cu:13-21). */
enum { ORC_TONE_EQUALIZE = 0, ORC_TONE_REINHARD = 1, ORC_TONE_REINHARD_JODIE = 2, ORC_TONE_ACES = 3,
       ORC_TONE_UNCHARTED2 = 4, ORC_TONE_GAMMA_CLAMP = 5 };
enum { ORC_IMAGE_BMP = 0, ORC_IMAGE_PPM = 1 };

/* Pixel sum i as the film holds it: sums[3i..3i+2], or, when nvalid is not
 * null, sums * (samples / nvalid[i]) (0 where nvalid is 0). */
void orc_film_pixel(const double *sums, const uint32_t *nvalid, uint64_t samples, size_t i, double rgb[3]);

/* getMaxChannel and getMaxLuminance (Film.cu:120-154): the largest channel
 * and the largest Spectrum::brightness over all pixel sums, each divided by
 * the sample count. */
void orc_film_maxima(const double *sums, const uint32_t *nvalid, size_t npix, uint64_t samples, double *max,
                     double *max_lum);

/* Spectrum::brightness (radiometry/Spectrum.cu:68-75). */
double orc_brightness(const double rgb[3]);

/* toneMaxPixel (Film.cu:266-283) on one pixel sum. */
void orc_tonemap_pixel(const double in[3], double max, double max_lum, uint64_t samples, int mode, double out[3]);

/* (unsigned char)(v * 255) as x86-64 code computes it. */
uint8_t orc_to_byte(double v);

/* Bytes of the pixel array: BMP rows of 3 * width bytes padded to a multiple
 * of 4 (BMP_Format.cu write_image), PPM 3 * width * height. */
size_t orc_image_pixel_bytes(int format, uint32_t width, uint32_t height);

/* File header; returns its length (at most cap bytes are written).  BMP: the
 * 54 bytes BitmapImage::write_image emits (with the reference's Image_Size of
 * (width + 4 - width % 4) * height); PPM: "P6\n<w> <h>\n255\n". */
size_t orc_image_header(int format, uint32_t width, uint32_t height, uint8_t *out, size_t cap);

/* The pixel array of writeToBMP / writeToPPM (Film.cu:285-362): BMP rows
 * bottom-up (array row j holds film row height-1-j) in B, G, R order with
 * zero padding; PPM rows top-down in R, G, B order.  If mapped is not null
 * it receives the tone-mapped values before the byte conversion, in film
 * order (width * height * 3 doubles, R, G, B). */
void orc_film_image(const double *sums, const uint32_t *nvalid, uint32_t width, uint32_t height, uint64_t samples,
                    int mode, int format, uint8_t *pixels, double *mapped);

#ifdef __cplusplus
}
#endif

#endif
