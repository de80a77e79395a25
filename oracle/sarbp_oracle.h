/*
 * CPU FP64 oracle for the SAR backprojection hot path — TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library, and
 * only as the checker (or as the timed CPU baseline). The product path never links or calls it.
 *
 * Every function restates one piece of the reference pipeline (ISI-apex/halide-sar-app) literally,
 * including its type-promotion behaviour; each definition cites the reference file:line it follows.
 */
#ifndef SARBP_ORACLE_H
#define SARBP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Policies for the f32 norm ||pos_p|| (backprojection.cpp:108, util.h:64-68). */
enum orc_norm_r0_policy {
    ORC_NORM_R0_F32_STRICT = 0, /* IEEE f32, no contraction: ((x*x + y*y) + z*z) */
    ORC_NORM_R0_F32_FMA = 1,    /* f32 with contraction: fma(z, z, fma(y, y, x*x)) */
    ORC_NORM_R0_F64 = 2,        /* what the comment at backprojection.cpp:107 claims */
};

/* Taylor window (signal.h:14-51). int_den selects integer (1) or f32 (0) division for the
 * (m+1)^2/(n+1)^2 ratio of F_m_den (signal.h:40). */
int orc_taylor(int num, double S_L, int int_den, double *out);
/* Clamped linspace (util.h:80-84) in f64. */
void orc_linspace(double start, double stop, int num, double *out);
/* ceil-mid roll (signal_complex.h:24-29): out[i] = in[(i + ceil(n/2)) % n]; returns the source index */
int orc_fftshift_src(int i, int n);
/* N_fft rule of sarbp.cpp:258 */
int orc_nfft(int nsamples, int upsample);

/* Forward unnormalised complex DFT, exp(-2*pi*i*k*n/N) (FFTW_FORWARD as used at dft.cpp:26,72).
 * Interleaved complex f64, in place. Radix-2 for powers of two, direct O(N^2) otherwise. */
int orc_fft_f64(double *data, int n);

/* f32 norm table with the given policy; out64 may be NULL. */
void orc_norm_r0(const float *pos, int np, int policy, float *out32, double *out64);

/* Range profiles Q {2, N, np} (interleaved complex f64, pulse-major) from phs {2, ns, np}:
 * window * |k_r| filter, centred pad, 2-axis fftshift, FFT, 2-axis fftshift
 * (backprojection.cpp:80-104). pulse_shift_quirk=0 drops the pulse-axis roll (RITSAR semantics). */
int orc_range_profiles(const float *phs, const float *k_r, int ns, int np, int taylor_s_l, int N,
                       int int_den, int pulse_shift_quirk, double *Q);

/* Per-pixel accumulation (backprojection.cpp:108-138) for the output region
 * x in [x0, x0+w), y in [y0, y0+h); out is {2, w, h} interleaved complex f64.
 * r is pixel_locs {r_extent, 3} with stride (1, r_extent); nr0 is the per-pulse norm as f64.
 * nthreads <= 0 means one thread. Loop nest follows the reference CPU schedule
 * (backprojection.cpp:302-314): rows serial, 4-pixel groups in parallel, norm_rr0[4 x np] staged. */
int orc_backproject(const double *Q, int N, int ns, int np, double delta_r, float k_c,
                    const float *pos, const double *nr0, const double *r, int64_t r_extent,
                    int nu, int nv, int x0, int y0, int w, int h, int nthreads, double *out);

/* Same, for an explicit list of output coordinates (xs[i], ys[i]); out is {2, count}. */
int orc_backproject_points(const double *Q, int N, int ns, int np, double delta_r, float k_c,
                           const float *pos, const double *nr0, const double *r, int64_t r_extent,
                           int nu, int nv, const int32_t *xs, const int32_t *ys, int count,
                           int nthreads, double *out);

/* The whole backprojection() pipeline (backprojection.cpp:67-139) over a {2, nu, nv} output. */
int orc_backprojection(const float *phs, const float *k_r, int ns, int np, int taylor_s_l, int N,
                       double delta_r, int nu, int nv, const float *pos, const double *r,
                       int norm_policy, int int_den, int pulse_shift_quirk, int nthreads,
                       double *out);

/* img_output_to_dB (img_output.cpp:10-27): out {nu, nv} f64 from img {2, nu, nv}. */
void orc_img_output_to_dB(const double *img, int64_t npix, double *out);
/* img_output_u8 (img_output.cpp:41-52, signal.h:55-57, util.h:13-15). */
void orc_img_output_u8(const double *dB, int64_t npix, double dB_min, double dB_max, uint8_t *out);

/* Image plane (ImgPlane.cpp:16-63, ip.h:5-17, img_plane.cpp:12-113).
 * dims_out = {nu, nv}; scal_out = {d_u, d_v}; u[nu], v[nv], k_u[nu], k_v[nv], u_hat[3], v_hat[3];
 * pixel_locs may be NULL, else {nu*nv, 3}. Call once with u == NULL to get dims only. */
int orc_img_plane(int nsamples, int npulses, double delta_r, const float *R_c, const float *n_hat,
                  double res_factor, double aspect, int upsample, int *dims_out, double *scal_out,
                  double *u, double *v, double *k_u, double *k_v, double *u_hat, double *v_hat,
                  double *pixel_locs);

#ifdef __cplusplus
}
#endif
#endif
