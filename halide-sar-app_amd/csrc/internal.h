// Internal declarations shared by the HIP kernels, the plan layer and the C-ABI shim.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "sarbp_hip.h"

namespace sarbp {

// Per-pulse constants of the accumulation, computed on the device from pos and k_r.
// With d = |r - pos_p| (f64):   luts  = B - inv_lsr * d        (linspace index, f64)
//                               phase = fract(gamma * d + G)    (revolutions of -k_c * dr)
// and |r - pos_p|^2 = |r|^2 + C - 2 r.pos_p.
struct PulseConst {
    double px, py, pz;  // platform position, f32 widened
    double C;           // |pos_p|^2
    double B;           // (nr0 - lsa) * inv_lsr
    double G;           // -gamma * nr0
    double nr0;         // ||pos_p|| under the selected policy, widened
    double pad;
};

// Launch-wide constants written next to the pulse table.
struct BpHeader {
    double k_c;      // k_r[nsamples / 2] (f32 widened)
    double gamma;    // k_c / (2 pi)
    double lsa;      // floor(-ns * delta_r / 2)
    double inv_lsr;  // (N - 1) / (lsb - lsa)
};

// Range-axis constants shared by the host and the device.
struct RangeAxis {
    double lsa, lsb, lsr, inv_lsr;
};
RangeAxis range_axis(int ns, int n_fft, double delta_r);

// Thread-local last-error string.
void set_error(const char *fmt, ...);

// Host-side math (host_math.cpp).
int taylor_window(int num, int s_l, int den_policy, double *out);

// Kernel launchers (kernels.hip). All are asynchronous on `stream`; return hipError_t.
hipError_t launch_pulse_consts(const float *pos, const float *k_r, int ns, int np, int n_fft,
                               double delta_r, int norm_policy, PulseConst *pc, BpHeader *hdr,
                               hipStream_t stream);
hipError_t launch_range_profiles(const void *phs, const float *k_r, const double *ws,
                                 const double *wp, int ns, int np, int n_fft, const float2 *tw,
                                 void *q, int q_format, float *qscale, hipStream_t stream);
struct BpArgs {
    const void *q;
    int q_format;
    const float *qscale;  // per-profile inverse scale of f16 profiles (NULL for f32)
    int n_fft, np;
    const PulseConst *pc;
    const BpHeader *hdr;
    const double *r;
    int64_t r_min, r_extent;
    int nu, nv, x0, y0, width, height;
    double2 *out;
};
hipError_t launch_accumulate(const BpArgs &a, hipStream_t stream);
hipError_t launch_img_output(const double2 *img, int64_t npix, double *dB, uint8_t *u8,
                             double lo, double hi, unsigned long long *scratch_max,
                             hipStream_t stream);
hipError_t launch_u8_only(const double *dB, int64_t npix, uint8_t *u8, double lo, double hi,
                          hipStream_t stream);
hipError_t launch_sim_phs(const double *pos, int np, const double *k64, int ns,
                          const double *tgt, const double *amp, int nt, float2 *phs,
                          hipStream_t stream);

// Device twiddle tables exp(-2 pi i t / N), t in [0, N), cached per (device, N).
const float2 *twiddles(int device, int n_fft, hipStream_t stream);
void release_twiddles();

constexpr int kMaxFft = 16384;
// Range-profile rows are stored padded to N + kQPad elements (see the storage comment in kernels.hip).
constexpr int kQPad = 4;
// Bytes per stored range-profile element: (q, dq) in complex f32 or complex f16.
inline size_t q_elem_bytes(int q_format) { return q_format == SARBP_Q_C16 ? 8 : 16; }

}  // namespace sarbp
