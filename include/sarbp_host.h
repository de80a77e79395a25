/*
 * sarbp_host — the host side of the `sarbp` driver: the PlatformData directory loader (npy files),
 * and the PNG writer for the u8 image. Plain C ABI over host memory; no GPU code.
 *
 * Mirrors the reference's platform_load (PlatformData.cpp:21-256, PlatformData.h:9-64) and the PNG
 * output of sarbp.cpp:312-327 (Halide::Tools::convert_and_save_image of a u8 {nu, nv} buffer).
 * Errors are returned as negative halide_error_code_t values with a message from
 * sarbp_host_last_error(); nothing throws across the ABI and nothing aborts.
 */
#ifndef SARBP_HOST_H
#define SARBP_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Platform data of one collection, as PlatformData.h:45-61 holds it. Optional scalars come with a
 * has_* flag, optional arrays are NULL when absent. Array layouts are the Halide buffers the
 * reference builds: pos f32 {3, npulses} (= [npulses][3]), phs f32 {2, nsamples, npulses}
 * (= complex64 [npulses][nsamples]). */
typedef struct sarbp_platform {
    int32_t nsamples, npulses;
    double delta_r;
    int32_t has_B, has_B_IF, has_delta_t, has_chirprate, has_f_0, has_vp;
    double B;
    float B_IF;
    double delta_t, chirprate, f_0, vp;
    const float *freq;   /* [nsamples] or NULL */
    const float *k_r;    /* [nsamples] */
    const double *k_y;   /* [npulses] or NULL */
    const float *n_hat;  /* [3] or NULL */
    const float *R_c;    /* [3] */
    const double *t;     /* [nsamples] or NULL */
    const float *pos;    /* [npulses][3] */
    const float *phs;    /* [npulses][nsamples][2] */
    int32_t reserved[8];
} sarbp_platform;

/* Replaces platform_load (PlatformData.cpp:21-256): reads DIR/{delta_r, nsamples, npulses, k_r, R_c,
 * pos, phs}.npy and the optional DIR/{B, B_IF, delta_t, chirprate, f_0, vp, freq, k_y, n_hat, t}.npy
 * with the reference's word-size checks and downcasts (f64 -> f32 for B_IF, freq, k_r, R_c, pos;
 * complex128 -> complex64 for phs; int64 -> int32 for the counts), and its shape checks. The arrays'
 * element kinds must match their word sizes (the reference reads by word size only). Little-endian,
 * C-ordered npy (format 1.0-3.0). On success *out owns every array; free it with
 * sarbp_platform_free. Messages use the reference's wording ("Bad shape: pos", "Bad word size: k_r",
 * "PlatformData: downcasting k_r from double to float" on stdout). */
int sarbp_platform_load(const char *dir, sarbp_platform **out);
void sarbp_platform_free(sarbp_platform *p);

/* 8-bit greyscale PNG of a width x height image (row y at pixels + y * row_stride), the format
 * convert_and_save_image writes for a u8 {width, height} buffer (sarbp.cpp:324). zlib deflate. */
int sarbp_png_write_gray8(const char *path, const uint8_t *pixels, int32_t width, int32_t height,
                          int64_t row_stride);

const char *sarbp_host_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SARBP_HOST_H */
