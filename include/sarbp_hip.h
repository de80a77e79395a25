/*
 * sarbp_hip — MI355X-native SAR backprojection behind the C ABI of the reference's
 * Halide-generated `backprojection` pipeline (ISI-apex/halide-sar-app).
 *
 * Part 1 declares the drop-in entry points, exactly the symbols and argument meaning a caller of the
 * reference binds (sarbp.cpp:183-229,277,299,316; dft.h), here with C linkage. The library also
 * exports them with the reference's own linkage: `backprojection`, `backprojection_cuda` and
 * `dft_init_fftw` / `dft_destroy_fftw` with C++ linkage (include/backprojection.h,
 * include/backprojection_cuda.h; the reference's dft.h), `img_output_to_dB` / `img_output_u8` with
 * C linkage (include/img_output_to_dB.h, include/img_output_u8.h), so a caller compiled against the
 * reference's headers links unchanged. A translation unit that includes the reference's dft.h must
 * not include this header (the two declare dft_init_fftw with different linkage); the pipeline
 * headers above may be included in either order with this one.
 * Part 2 declares extensions that have no counterpart in the reference: device-resident plans for
 * callers whose data already lives in HBM, and helpers used by the serving engine, the scene
 * simulator and the host driver.
 *
 * All entry points return 0 on success and a negative halide_error_code_t value on failure; none
 * aborts. Buffers are owned by the caller; inputs are read-only.
 */
#ifndef SARBP_HIP_H
#define SARBP_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "halide_buffer_compat.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ============================== Part 1: the reference's ABI ============================== */

/*
 * Replaces the AOT symbol generated from BackprojectionGenerator (backprojection.cpp:35-46,
 * registered at :354), called at sarbp.cpp:277 (bounds query at :269).
 *   phs        f32 {2, nsamples, npulses}          (PlatformData.cpp:238)
 *   k_r        f32 {nsamples}                      (PlatformData.cpp:137)
 *   taylor_s_l Taylor sidelobe level in dB         (backprojection.cpp:37)
 *   N_fft      FFT length (backprojection.cpp:38, sarbp.cpp:258): a power of two 2..131072
 *              (above 16384 a four-step FFT through HBM) or any other length 3..8192
 *              (Bluestein's chirp convolution); larger values return param_too_large
 *   delta_r    range resolution                    (backprojection.cpp:39)
 *   u, v       f64 {nu}, {nv}; only extents used   (backprojection.cpp:71-72)
 *   pos        f32 {3, npulses}                    (PlatformData.cpp:219)
 *   r          f64 {nu*nv, 3} pixel locations      (ImgPlane.cpp:103)
 *   output_img f64 {2, x-range, y-range}; any region of the nu x nv image (backprojection.cpp:46,138)
 * A buffer with host == NULL and device == 0 is a bounds query: required input regions are written
 * into its dims and 0 is returned (the protocol of dft.cpp:49-58).
 */
#ifndef SARBP_DECLARED_backprojection
#define SARBP_DECLARED_backprojection 1
int backprojection(halide_buffer_t *phs, halide_buffer_t *k_r, int32_t taylor_s_l, int32_t N_fft,
                   double delta_r, halide_buffer_t *u, halide_buffer_t *v, halide_buffer_t *pos,
                   halide_buffer_t *r, halide_buffer_t *output_img);
#endif

/* Replaces `backprojection_cuda` (backprojection.cpp:356, selected at sarbp.cpp:198-200): the same
 * pipeline; on this library both names run on the GPU. */
#ifndef SARBP_DECLARED_backprojection_cuda
#define SARBP_DECLARED_backprojection_cuda 1
int backprojection_cuda(halide_buffer_t *phs, halide_buffer_t *k_r, int32_t taylor_s_l,
                        int32_t N_fft, double delta_r, halide_buffer_t *u, halide_buffer_t *v,
                        halide_buffer_t *pos, halide_buffer_t *r, halide_buffer_t *output_img);
#endif

/* Replaces img_output_to_dB (img_output.cpp:10-35, registered :54), called at sarbp.cpp:299:
 * out(x, y) = 10 * log10(|img(x, y)| / max|img|); img f64 {2, nu, nv}, out f64 {nu, nv}. */
int img_output_to_dB(halide_buffer_t *img, halide_buffer_t *out);

/* Replaces img_output_u8 (img_output.cpp:41-52, registered :55), called at sarbp.cpp:316:
 * out = u8(clamp(255 * (clamp(dB, lo, hi) - lo) / (hi - lo), 0, 255)); dB f64 {nu, nv}, out u8. */
int img_output_u8(halide_buffer_t *dB, double dB_min, double dB_max, halide_buffer_t *out);

/* Replace dft_init_fftw / dft_destroy_fftw (dft.h, dft.cpp:20-35), called at sarbp.cpp:261,292.
 * Here they build / release the device twiddle table for N_fft; calling them is optional. */
void dft_init_fftw(size_t N_fft);
void dft_destroy_fftw(void);

/* ============================== Part 2: extensions ============================== */

/* Numerical policies for the two places where the reference's result depends on how Halide
 * compiled it (no counterpart in the reference; process-wide defaults for Part 1 calls). */
enum sarbp_norm_r0_policy {
    SARBP_NORM_R0_F32_STRICT = 0, /* ||pos|| in f32, IEEE, no contraction (default) */
    SARBP_NORM_R0_F32_FMA = 1,    /* f32 with fma(z, z, fma(y, y, x*x)) */
    SARBP_NORM_R0_F64 = 2,        /* f64 */
};
enum sarbp_taylor_den_policy {
    SARBP_TAYLOR_DEN_INT = 0, /* (m+1)^2/(n+1)^2 as Int(32) division (default) */
    SARBP_TAYLOR_DEN_F32 = 1, /* the same ratio in f32 */
};
/* In a serving job (sarbp_job), either policy may be SARBP_POLICY_DEFAULT: the process-wide default
 * at submission, which is what a plan or a Part 1 call made then would use. */
#define SARBP_POLICY_DEFAULT (-1)
enum sarbp_q_format { SARBP_Q_C32 = 0, SARBP_Q_C16 = 1 };
enum sarbp_range_mode {
    SARBP_RANGE_F64 = 0, /* each pixel's row coordinate in f64 (default) */
    SARBP_RANGE_DF32 = 1 /* grid mode, cubic form: each pixel's row coordinate as an exact integer
                            part per pulse and column plus an f32 cubic in the pixel's offset (a
                            double-float pair); tiles in other forms, and plans without a grid,
                            keep f64 */
};

int sarbp_set_default_policies(int norm_r0_policy, int taylor_den_policy);
int sarbp_get_default_policies(int *norm_r0_policy, int *taylor_den_policy);

/* Last error message of the calling thread ("" if none). */
const char *sarbp_last_error(void);

/* How the calling thread's last backprojection() / backprojection_cuda() call formed its image:
 * bands in all, and bands whose pixel locations were found on the image plane of the call's u, v
 * buffers (so that the accumulation generated them, in grid mode, instead of reading r). No
 * counterpart in the reference; for tests and for callers checking that their pixel_locs take the
 * faster path. */
int sarbp_abi_last_call_bands(int32_t *grid_bands, int32_t *bands);

/* Number of visible GPUs (0 when none; never aborts). */
int sarbp_device_count(void);

/* Host-side Taylor window (signal.h:14-51) with an explicit policy; out has num doubles. */
int sarbp_taylor_window(int32_t num, int32_t taylor_s_l, int32_t taylor_den_policy, double *out);

/* Host-side image plane (ImgPlane.cpp:16-63, img_plane.cpp:12-113, ip.h:5-17).
 * dims_out = {nu, nv}; scal_out = {d_u, d_v}. With u == NULL only dims/scal are written.
 * pixel_locs (may be NULL) is f64 {nu*nv, 3} with stride (1, nu*nv). */
int sarbp_img_plane_create(int32_t nsamples, int32_t npulses, double delta_r, const float *R_c,
                           const float *n_hat, double res_factor, double aspect, int32_t upsample,
                           int32_t *dims_out, double *scal_out, double *u, double *v, double *k_u,
                           double *k_v, double *u_hat, double *v_hat, double *pixel_locs);

/* Image-plane pixel locations generated on the device (img_plane.cpp:77-113):
 * pixel_locs(x, d) = u_hat[d] * u[x % nu] + v_hat[d] * v[x / nu], written as f64 {nu*nv, 3}
 * (stride (1, nu*nv)). u_dev/v_dev are device arrays; u_hat/v_hat are host arrays of 3. */
int sarbp_pixel_locs_device(const double *u_dev, int32_t nu, const double *v_dev, int32_t nv,
                            const double *u_hat, const double *v_hat, double *pixel_locs_dev,
                            void *stream);

/* Device-resident plan: everything the hot path needs besides the per-call inputs. */
typedef struct sarbp_plan sarbp_plan;

typedef struct sarbp_plan_desc {
    int32_t nsamples, npulses, n_fft, taylor_s_l;
    double delta_r;
    int32_t nu, nv;                  /* image-plane extents (only extents matter, like u/v) */
    int32_t x0, y0, width, height;   /* output region of the nu x nv image */
    int64_t r_min, r_extent;         /* pixel index range held by the r buffer */
    int32_t norm_r0_policy, taylor_den_policy;
    int32_t q_format;                /* sarbp_q_format */
    int32_t range_mode;              /* sarbp_range_mode */
    int32_t device;                  /* HIP device ordinal */
    int32_t tile_ppt;                /* pixels per thread of the accumulation tile (2, 3 or 4;
                                        0 picks the default: 4 for regions of at least 2048
                                        16 x 64 tiles, else 2) */
    int32_t fft_blocks;              /* range profiles: 0 = one 512-thread block per profile;
                                        > 0 = at most this many 256-thread blocks, each looping
                                        over profiles (N_fft <= 8192; for a device shared with
                                        other streams' accumulation launches). Same values. */
    int32_t row_windows;             /* grid mode, cubic form, complex-f32 rows, 16 x 48 and
                                        16 x 64 tiles: 0 = every pixel gathers its row elements
                                        from HBM, in blocks of 32 pulses (default); 1 = each tile
                                        stages the row elements its pixels read in LDS, per block
                                        of 8 pulses, where they fit (96 elements per pulse at
                                        16 x 64, 80 at 16 x 48). The same image bit for bit. */
    int32_t reserved[4];
} sarbp_plan_desc;

void sarbp_plan_desc_init(sarbp_plan_desc *d);
int sarbp_plan_create(const sarbp_plan_desc *d, sarbp_plan **out);
int sarbp_plan_destroy(sarbp_plan *p);

/* Run the whole pipeline on `stream` (hipStream_t, may be NULL for the null stream).
 *   phs_dev   complex f32 [npulses][nsamples]       k_r_dev  f32 [nsamples]
 *   pos_dev   f32 [npulses][3]                      r_dev    f64 [3][r_extent] (SoA)
 *   out_dev   f64 {2, width, height} dense (interleaved complex, row y-major)
 * Asynchronous: returns after enqueueing. */
int sarbp_plan_execute(sarbp_plan *p, const void *phs_dev, const void *k_r_dev,
                       const void *pos_dev, const void *r_dev, void *out_dev, void *stream);

/* The two stages separately (for profiling and tests). q_dev may be NULL to use the plan's own
 * range-profile buffer; otherwise it must hold sarbp_plan_q_bytes(p) bytes (the plan's stored
 * row format: npulses rows of edge-extended, phase-rotated interpolation elements). */
int sarbp_plan_range_profiles(sarbp_plan *p, const void *phs_dev, const void *k_r_dev,
                              void *q_dev, void *stream);
int sarbp_plan_accumulate(sarbp_plan *p, const void *q_dev, const void *k_r_dev,
                          const void *pos_dev, const void *r_dev, void *out_dev, void *stream);
/* Device pointer and size of the plan's range-profile buffer. */
void *sarbp_plan_q_buffer(sarbp_plan *p);
size_t sarbp_plan_q_bytes(sarbp_plan *p);
/* Pixels per thread of the plan's accumulation tiles (tile height 16 * this), the default resolved. */
int32_t sarbp_plan_tile_ppt(const sarbp_plan *p);

/* Which form of the range expansion the plan's grid-mode accumulation gives its expansion tiles
 * (decided from the plan's whole region, its tile height and the device's resident tiles; tiles
 * whose expansion bound fails take the direct form in every case). */
enum sarbp_grid_form {
    SARBP_GRID_FORM_NONE = 0,  /* no grid: pixel locations are read from r */
    SARBP_GRID_FORM_PQ = 1,    /* p, q form: 5.25 f64 operations per pixel-pulse */
    SARBP_GRID_FORM_CUBIC = 2  /* cubic in the row offset: 3 f64 FMAs per pixel-pulse */
};
int32_t sarbp_plan_grid_form(const sarbp_plan *p);

/* How many tiles of the plan's last accumulation (since its pulse table was built) took each loop:
 * counts[0] beyond the rows' cover (clamped, exact phase), [1] direct form, [2] expansion in its
 * p, q form, [3] cubic form gathering row elements from HBM, [4] cubic form reading row windows
 * staged in LDS, [5] cubic form with double-float row coordinates (counts has 6 entries).
 * Synchronises the device. */
int sarbp_plan_tile_paths(sarbp_plan *p, int64_t *counts);

/* dB and u8 output (img_output_to_dB, img_output_u8) of the image the plan's last accumulation
 * wrote (sarbp_plan_execute / sarbp_plan_accumulate), in one pass over it: that accumulation also
 * found max |z|. img_dev is that f64 {2, width, height} image; dB_dev (f64, width * height) and/or
 * u8_dev may be NULL. Asynchronous on `stream`; must follow the accumulation in stream order and
 * precede the plan's next one. The values equal sarbp_img_output_device's on the same image. */
int sarbp_plan_img_output(sarbp_plan *p, const void *img_dev, double *dB_dev, uint8_t *u8_dev,
                          double dB_min, double dB_max, void *stream);

/* Grid mode: the plan's pixel locations come from the image grid (img_plane.cpp:77-113:
 * r(x, y) = (0 + u_hat u[x]) + v_hat v[y], u[x] = (x + floor(-nu/2)) d_u, v likewise, nu and nv from
 * the plan) and are generated inside the accumulation kernel, which then needs no r buffer (r_dev
 * may be NULL in sarbp_plan_execute / sarbp_plan_accumulate). The kernel forms ranges from the
 * grid offsets of each pixel from its tile's centre (or, where that expansion's error bound fails,
 * derives the squared range of a thread's later pixels from its first one's), so the image agrees
 * with the r-buffer path to rounding (~1e-8 relative), not bit for bit. The plan's output
 * region must lie inside the image (constraint_violated otherwise). u_hat or v_hat NULL switches
 * back to r. */
int sarbp_plan_set_grid(sarbp_plan *p, double d_u, double d_v, const double *u_hat,
                        const double *v_hat);
/* Copy the plan's range profiles Q to host memory as complex f32 [npulses][n_fft] (undoing the
 * padded storage and, for f16, the per-profile scale); synchronises the device. */
int sarbp_plan_read_profiles(sarbp_plan *p, float *host_out);
/* Copy the first n entries of the plan's ||pos_p|| table (as computed by the last accumulate,
 * under the plan's norm_r0 policy) to host memory; synchronous. */
int sarbp_plan_read_pulse_norms(sarbp_plan *p, double *host_out, int32_t n);

/* dB and u8 image output on the device for an f64 complex image of npix pixels: max |z|, then
 * dB and u8 in a second pass. dB_dev and/or u8_dev may be NULL. Asynchronous on `stream`. The
 * library keeps one 8-byte max cell per (device, stream) this has run on, so calls on different
 * streams may overlap. */
int sarbp_img_output_device(const void *img_dev, int64_t npix, double *dB_dev, uint8_t *u8_dev,
                            double dB_min, double dB_max, void *stream);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops): sarbp_img_output_device's dB and u8 output without any
 * allocation. The max |z| is reduced through a static device array of the library (8200 bytes per
 * GPU, held until the process ends; it does not grow with npix, so no image is too large for it).
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU (on different streams
 * or host threads) race; calls ordered on one stream are safe. Asynchronous on `stream`; the same
 * values as sarbp_img_output_device. */
int sarbp_img_output_device_noalloc(const void *img_dev, int64_t npix, double *dB_dev,
                                    uint8_t *u8_dev, double dB_min, double dB_max, void *stream);

/* Synthetic phase history on the device: phs[p][k] = sum_s amp_s exp(j k_r64[k] * D[p][s]) with
 * D[p][s] = |pos_p| - |pos_p - tgt_s| in f64. pos f64 [np][3], tgt f64 [nt][3], amp complex f64
 * [nt], k_r64 f64 [ns]; out complex f32 [np][ns]. */
int sarbp_sim_phs_device(const double *pos_dev, int32_t npulses, const double *k_r64_dev,
                         int32_t nsamples, const double *tgt_dev, const double *amp_dev,
                         int32_t ntargets, void *phs_dev, void *stream);

/* ------------------------------------------------------------------ serving engine
 * Continuous batching of image-formation jobs on one GPU: jobs are admitted in arrival order while
 * their device memory fits the HBM budget (smaller jobs may pass a head that does not fit), run on
 * several streams so that transfers and kernels of different jobs overlap, and jobs at or below
 * `fuse_max_pixel_pulses` share one accumulation launch over a job-indexed tile list. A job's result
 * is the same whether it runs alone or fused (same per-pixel code). Host buffers of a job must stay
 * valid until sarbp_server_wait returns for it; page-locked buffers let the copies overlap.
 * The engine uses num_streams + 5 streams (uploads, small-job preparation, fused launches, two
 * download streams). HIP maps a process's streams onto GPU_MAX_HW_QUEUES hardware queues (4 unless
 * the environment sets it before the runtime starts), and packets sharing a queue run in order, so
 * with fewer queues than streams small jobs can wait behind large jobs' bands: results are the
 * same, latency is lower with num_streams + 6 queues (8 for the default, counting the null stream). */
typedef struct sarbp_server sarbp_server;

typedef struct sarbp_server_config {
    int32_t device;
    int32_t max_inflight;           /* admitted jobs at once (>= 1) */
    int64_t hbm_budget_bytes;       /* device memory the admitted jobs may hold together */
    int64_t fuse_max_pixel_pulses;  /* jobs with nu*nv*npulses <= this are fused (0: never) */
    int32_t num_streams;            /* large-job streams, 1..16 (default 2) */
    int32_t autotune;               /* 1: time the tile candidates on a shape class's first job */
    int64_t fuse_launch_pixel_pulses; /* pixel-pulses of one fused launch at most, which bounds
                                         how long other jobs' kernels wait behind one (0: the
                                         default, 8e9, ~4 ms on an idle device) */
    int32_t band_order;             /* larger jobs sharing a stream: 0 (default) the one with the
                                       least work left enqueues its bands, the others wait; 1 all
                                       of them enqueue, their bands interleaving */
    int32_t reserved[5];
} sarbp_server_config;

typedef struct sarbp_job {
    const void *phs;          /* host complex f32 [npulses][nsamples] */
    const float *k_r;         /* host f32 [nsamples] */
    const float *pos;         /* host f32 [npulses][3] */
    int32_t nsamples, npulses, n_fft, taylor_s_l;
    double delta_r;
    int32_t nu, nv;           /* image size; output rows flipped as in backprojection.cpp:130 */
    double d_u, d_v;          /* grid spacing: u = (x + floor(-nu/2)) d_u, v likewise */
    double u_hat[3], v_hat[3];
    const double *pixel_locs; /* optional host f64 {nu*nv, 3}; NULL: generated from the grid */
    /* Policies as for a plan, or SARBP_POLICY_DEFAULT for the process default (that of plans and
     * Part 1 calls: f32_strict and the Int(32) Taylor ratio unless changed). Under the Int(32)
     * reading a Taylor level of 31 dB or more divides by zero in the window's F_m denominator
     * (signal.h:40), so the window and the image are non-finite, as the reference's would be; a
     * job stream that wants images at those levels asks for SARBP_TAYLOR_DEN_F32. */
    int32_t norm_r0_policy, taylor_den_policy, q_format;
    int32_t range_mode;       /* sarbp_range_mode (0: f64) */
    void *out;                /* host complex f64 {2, nu, nv} */
    int32_t reserved[8];
} sarbp_job;

typedef struct sarbp_job_stats {
    int64_t t_submit_ns, t_admit_ns, t_end_ns; /* since server creation */
    int64_t device_bytes;                      /* HBM the job held */
    int32_t fused;                             /* accumulated in a shared launch */
    int32_t tile_ppt;                          /* tile shape used */
    int32_t status;                            /* 0 or a negative error code */
    int32_t host_enqueue_us;                   /* scheduler time spent enqueuing the job */
    int32_t device_us;                         /* device time from the job's first operation to
                                                  its output in host memory */
    int32_t row_windows;                       /* bands staged row windows in LDS */
    int32_t reserved[2];
    /* When the scheduler saw each step of the job end or began it, in microseconds after
     * t_admit_ns (-1: not reached): inputs uploaded, range profiles computed (a job accumulated
     * in bands: when its range profiles and bands were enqueued together), accumulation enqueued
     * (its fused launch or first band), accumulation complete. The scheduler polls every 20 us. */
    int32_t up_us, ready_us, launch_us, acc_us;
} sarbp_job_stats;

void sarbp_server_config_init(sarbp_server_config *cfg);
int sarbp_server_create(const sarbp_server_config *cfg, sarbp_server **out);
/* Queues a job (returns at once); fails with halide_error_code_buffer_allocation_too_large if the
 * job alone exceeds the HBM budget. */
int sarbp_server_submit(sarbp_server *s, const sarbp_job *job, uint64_t *job_id);
/* Waits for a job (timeout_ms < 0: no limit): 0 when done, 1 on timeout, < 0 if the job failed. */
int sarbp_server_wait(sarbp_server *s, uint64_t job_id, int32_t timeout_ms, sarbp_job_stats *stats);
/* Finishes every queued job, then releases the server. */
int sarbp_server_destroy(sarbp_server *s);

#ifdef __cplusplus
}
#endif
#endif
