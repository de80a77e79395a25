// HIP kernels of the backprojection hot path for gfx950 (MI355X), and their launchers.
//
//   pulse_const_kernel      ||pos_p|| table + per-pulse constants (backprojection.cpp:108,120-123)
//   range_profile_kernel    window * |k_r|, centred pad, fftshift, LDS-resident radix-4/2 Stockham
//                           FFT, fftshift -> range profiles Q (backprojection.cpp:80-104, dft.cpp)
//   range_profile_shared_kernel  the same with 256-thread blocks looping over pulses, for a device
//                           shared with accumulation launches (serving engine)
//   accumulate_kernel       per-pixel range, linear interpolation of Q, phase-corrected sum over
//                           pulses, mid-pulse phase and row flip (backprojection.cpp:52-65,108-138);
//                           pixel locations read from pixel_locs or generated from the image grid
//   accumulate_fused_kernel the same over a job-indexed tile list (several images in one launch)
//   pixel_locs_kernel       the image plane's pixel locations (img_plane.cpp:77-113)
//   absmax / db_u8 kernels  img_output_to_dB and img_output_u8 (img_output.cpp:10-52)
//   sim_phs_kernel          seeded point-scatterer phase histories for benchmarks and tests
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <mutex>
#include <set>
#include <type_traits>
#include <utility>
#include <vector>

#include "img_math.h"
#include "internal.h"
#include "tile_order.h"

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

// Structured buffer loads (the LLVM intrinsic; clang has no builtin for it): the address is
// base + vindex * stride + voffset, the multiply done by the load unit (buffer_load ... idxen)
// instead of a VALU shift per element.
__device__ u32x4 struct_load_b128(__amdgpu_buffer_rsrc_t rsrc, int vindex, int voffset,
                                  int soffset, int aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.v4i32");
__device__ u32x2 struct_load_b64(__amdgpu_buffer_rsrc_t rsrc, int vindex, int voffset,
                                 int soffset, int aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.v2i32");

namespace sarbp {

namespace {

// Complex product, each product and sum rounded once (no contraction), so that both range-profile
// kernels give the same bits whatever the compiler would fuse in either.
__device__ inline float2 cmul(float2 a, float2 b) {
#pragma clang fp contract(off)
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ inline int floordiv2(int a) { return (a >= 0) ? a / 2 : -((-a + 1) / 2); }

// Range-profile storage: rows of RowLayout::len elements (A_j, B_j), see RowLayout in internal.h.
struct QPairF32 {
    float qr, qi, dr, di;
};
struct QPairF16 {
    __half2 q, dq;
};

// f32 rows hold B scaled by 2^-32 (exact), so that the accumulation can weight it by fract(t) * 2^32,
// the integer its fixed-point row coordinate yields, without a multiply; f16 rows hold B itself.
constexpr float kF32RowBScale = 0x1p-32f;
__device__ inline void storeq(QPairF32 *p, float2 q, float2 dq) {
    QPairF32 v;
    v.qr = q.x;
    v.qi = q.y;
    v.dr = dq.x * kF32RowBScale;
    v.di = dq.y * kF32RowBScale;
    *p = v;
}
__device__ inline void storeq(QPairF16 *p, float2 q, float2 dq) {
    QPairF16 v;
    v.q = __float22half2_rn(q);
    v.dq = __float22half2_rn(dq);
    *p = v;
}
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));

// Buffer resource over one padded row of `len` elements of `stride` bytes (32-bit data format
// word as for gfx9 buffer loads). Every index the accumulation forms lies in [0, len): the fast path
// runs only for tiles whose pixels are all within the rows' cover radius (lanes outside the image
// take an in-image location) and the exact path clamps.
__device__ inline __amdgpu_buffer_rsrc_t row_rsrc(const void *row, int len, int stride) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(row), (short)stride, len,
                                             0x00020000);
}
__device__ inline float4 loadq(const QPairF32 *, __amdgpu_buffer_rsrc_t r, int idx) {
    const u32x4 v = struct_load_b128(r, idx, 0, 0, 0);
    const f32x4 f = __builtin_bit_cast(f32x4, v);
    return make_float4(f.x, f.y, f.z, f.w);
}
__device__ inline float4 loadq(const QPairF16 *, __amdgpu_buffer_rsrc_t r, int idx) {
    const u32x2 v = struct_load_b64(r, idx, 0, 0, 0);
    const f16x4 h = __builtin_bit_cast(f16x4, v);
    return make_float4((float)h.x, (float)h.y, (float)h.z, (float)h.w);
}

// Interpolation weights are carried in units of the rows' B scale: w * 2^32 for f32 rows, w for f16.
template <typename QT>
constexpr float weight_unit() {
    return sizeof(QT) == sizeof(QPairF16) ? 1.0f : 0x1p32f;
}

// trunc(t): v_cvt_i32_f64 saturates out-of-range values and gives 0 for NaN.
__device__ inline int trunc_i32(double t) {
    int i;
    asm("v_cvt_i32_f64 %0, %1" : "=v"(i) : "v"(t));
    return i;
}

// trunc(t) clamped to [0, hi]: v_cvt_i32_f64 saturates (NaN gives 0), v_med3_i32 clamps.
__device__ inline int trunc_clamp(double t, int hi) {
    int i;
    asm("v_cvt_i32_f64 %0, %1\n\tv_med3_i32 %0, %0, 0, %2" : "=&v"(i) : "v"(t), "s"(hi));
    return i;
}

// Row coordinate t = B - sqrt(s) for a squared range s in row units (sigma |r - pos|^2,
// sigma = inv_lsr^2, so sqrt(s) = d / lsr): the hardware reciprocal square root y (about 23
// correct bits) and one Newton-Raphson step in f64, a = s y, sqrt(s) = a (3 - a y) / 2, with the
// subtraction from B in the last fused multiply-add. Relative error of sqrt(s) ~1.5 (2^-23)^2.
__device__ inline double row_coord(double s, double B) {
    const double y = __builtin_amdgcn_rsq(s);
    const double a = s * y;
    const double h = a * y;
    const double c = __builtin_fma(-0.5, h, 1.5);
    return __builtin_fma(-a, c, B);
}

// One component of a grid pixel location, rounded as img_plane.cpp:77-113 computes it (product,
// sum with 0, product, sum; no contraction): bit-identical with pixel_locs_kernel and the host.
__device__ inline double grid_coord(double uh, double u, double vh, double v) {
#pragma clang fp contract(off)
    return (0.0 + uh * u) + vh * v;
}

// Raises *cell to the bit pattern of sqrt(m2) (m2 >= 0 orders like its bits), skipping the atomic
// when the cell already holds at least that much.
__device__ inline void raise_absmax(unsigned long long *cell, double m2) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(sqrt(m2));
    if (b > __hip_atomic_load(cell, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(cell, b);
}

}  // namespace

// ------------------------------------------------------------------ per-pulse constants

__global__ void pulse_const_kernel(const float *__restrict__ pos, const float *__restrict__ k_r,
                                   int ns, int np, double lsa, double lsr, double inv_lsr, int ext,
                                   int policy, double ux, double uy, double uz, double vx,
                                   double vy, double vz, PulseConst *__restrict__ pc,
                                   BpHeader *__restrict__ hdr,
                                   unsigned long long *__restrict__ absmax,
                                   unsigned *__restrict__ paths) {
#pragma clang fp contract(off)
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) {
        if (absmax) *absmax = 0ull;  // the image's max |z|, raised by the accumulation
        if (paths)
            for (int k = 0; k < kNumPaths; ++k) paths[k] = 0u;  // tiles per loop, counted likewise
        // k_c = k_r(nsamples / 2), f32 (backprojection.cpp:123); phase = -k_c dr in revolutions
        const double k_c = (double)k_r[ns / 2];
        const double gamma = k_c / (2.0 * M_PI);
        hdr->k_c = k_c;
        hdr->gamma = gamma;
        hdr->lsa = lsa;
        hdr->inv_lsr = inv_lsr;
        hdr->kappa = -gamma * lsr;
        hdr->g_ext = -gamma * (lsa - lsr * (double)ext);
    }
    if (p >= np) return;
    const float x = pos[3 * p], y = pos[3 * p + 1], z = pos[3 * p + 2];
    // Plain operators under `fp contract(off)` round each product and sum once; sqrtf and sqrt are
    // the IEEE correctly rounded forms (the __fsqrt_rn intrinsic maps to the 1-ulp native sqrt).
    double nr0;
    if (policy == SARBP_NORM_R0_F64) {
        const double dx = x, dy = y, dz = z;
        nr0 = sqrt((dx * dx + dy * dy) + dz * dz);
    } else if (policy == SARBP_NORM_R0_F32_FMA) {
        nr0 = (double)sqrtf(fmaf(z, z, fmaf(y, y, x * x)));
    } else {
        // norm3d of the f32 pos Func (util.h:64-68): f32 products and sums, in order
        const float xx = x * x, yy = y * y, zz = z * z;
        nr0 = (double)sqrtf((xx + yy) + zz);
    }
    PulseConst c;
    c.px = x;
    c.py = y;
    c.pz = z;
    // squared ranges in row units: sigma |r - pos|^2 = sigma |r|^2 + Cs - 2 sigma r.pos
    const double sigma = inv_lsr * inv_lsr;
    c.Cs = sigma * (c.px * c.px + c.py * c.py + c.pz * c.pz);
    // t = luts + ext with luts = (nr0 - d - lsa) / lsr (backprojection.cpp:53-54,117,120)
    c.B = (nr0 - lsa) * inv_lsr + (double)ext;
    c.nr0 = nr0;
    c.U = (-2.0 * sigma) * (ux * c.px + uy * c.py + uz * c.pz);
    c.V = (-2.0 * sigma) * (vx * c.px + vy * c.py + vz * c.pz);
    pc[p] = c;
}

// R_j = exp(2 pi i fract(g_ext + kappa j)) for the row elements, from k_c on the device.
__global__ void rotation_kernel(const float *__restrict__ k_r, int ns, double lsa, double lsr,
                                int ext, int len, float2 *__restrict__ rot) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= len) return;
    const double gamma = (double)k_r[ns / 2] / (2.0 * M_PI);
    const double g_ext = -gamma * (lsa - lsr * (double)ext);
    const double kappa = -gamma * lsr;
    const double ph = g_ext + kappa * (double)j;
    const double f = ph - floor(ph);
    double sn, cs;
    sincospi(2.0 * f, &sn, &cs);
    rot[j] = make_float2((float)cs, (float)sn);
}

// ------------------------------------------------------------------ range profiles + FFT

constexpr int kFftThreads = 512;

template <int R>
__device__ inline void dft_r(float2 *v);

template <>
__device__ inline void dft_r<2>(float2 *v) {
    const float2 a = v[0], b = v[1];
    v[0] = make_float2(a.x + b.x, a.y + b.y);
    v[1] = make_float2(a.x - b.x, a.y - b.y);
}

// forward 4-point DFT, exp(-2 pi i r k / 4)
template <>
__device__ inline void dft_r<4>(float2 *v) {
    const float2 a0 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
    const float2 a1 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
    const float2 a2 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
    const float2 d = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
    const float2 a3 = make_float2(d.y, -d.x);  // -i * (v1 - v3)
    v[0] = make_float2(a0.x + a2.x, a0.y + a2.y);
    v[2] = make_float2(a0.x - a2.x, a0.y - a2.y);
    v[1] = make_float2(a1.x + a3.x, a1.y + a3.y);
    v[3] = make_float2(a1.x - a3.x, a1.y - a3.y);
}

// forward 16-point DFT in registers as 4 x 4: with n = c + 4 l and k = q + 4 p,
// X[q + 4 p] = sum_c W_4^(c p) W_16^(c q) sum_l x[c + 4 l] W_4^(l q): four 4-point DFTs over l, the
// nine non-trivial factors W_16^(c q), four 4-point DFTs over c.
__device__ inline float2 mul_w16(float2 v, int e) {  // v * W_16^e, e in {1, 2, 3, 6, 9}
    constexpr float c8 = 0.92387953251128674f, s8 = 0.38268343236508978f, r2 = 0.70710678118654752f;
    switch (e) {
        case 1: return cmul(v, make_float2(c8, -s8));
        case 2: return cmul(v, make_float2(r2, -r2));
        case 3: return cmul(v, make_float2(s8, -c8));
        case 6: return cmul(v, make_float2(-r2, -r2));
        default: return cmul(v, make_float2(-c8, s8));  // e = 9
    }
}
// The 16-point DFT of v, written to out[base + k Ns] (k = 0..15); v is overwritten. The first
// stage works in place (y[c][q] at v[c + 4 q]), the second writes its outputs straight out.
__device__ inline void dft16_store(float2 *v, float2 *out, int base, int Ns) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float2 t[4] = {v[c], v[c + 4], v[c + 8], v[c + 12]};
        dft_r<4>(t);
#pragma unroll
        for (int q = 0; q < 4; ++q) v[c + 4 * q] = t[q];
    }
#pragma unroll
    for (int c = 1; c < 4; ++c)
#pragma unroll
        for (int q = 1; q < 4; ++q)
            v[c + 4 * q] = c * q == 4 ? make_float2(v[c + 4 * q].y, -v[c + 4 * q].x)  // W_16^4 = -i
                                      : mul_w16(v[c + 4 * q], c * q);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float2 s[4] = {v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]};
        dft_r<4>(s);
#pragma unroll
        for (int p = 0; p < 4; ++p) out[base + (q + 4 * p) * Ns] = s[p];
    }
}

// Pass schedule of the one-row FFTs of N = 2^logN points (fft_pass / fft_pass_oop): a radix-2
// pass first where logN is odd, then radix 4 while the span Ns is below 32 (at smaller spans a
// radix-16 pass's strided LDS writes would meet in few banks), then radix 16 while four or more
// bits are left, then radix 4. `left` = log2(N / Ns). N = 8192: 2, 4, 4, 16, 16 (five LDS passes
// where radix 4 takes seven).
__host__ __device__ constexpr int fft_next_radix(int Ns, int left) {
    return (left & 1) ? 2 : (Ns < 32 || left < 4) ? 4 : 16;
}
// Span of the first radix-16 pass (the first span >= 32 of the schedule).
__host__ __device__ constexpr int fft_first_span16(int logN) { return (logN & 1) ? 32 : 64; }

// Twiddles of a radix-16 pass of span Ns: W_N^(r k N / (16 Ns)) for r = 1..15 and k in [0, Ns), at
// stage16[(r - 1) Ns + k]; the pass of span Ns at tw + 2N + Ns - fft_first_span16(logN) (after the
// N-point table and the radix-4 stage tables).
__device__ inline const float2 *fft_stage16_twiddles(const float2 *tw, int N, int logN, int Ns) {
    return tw + 2 * N + Ns - fft_first_span16(logN);
}

// Twiddles of one radix-4 pass of span Ns over an N-point row: W_N^(r k N / (4 Ns)) for r = 1..3
// and k in [0, Ns), stored at stage[(r - 1) Ns + k] so that the lanes of a wave (consecutive k)
// read consecutive entries. The N-point table tw (W_N^t, t in [0, N)) is followed by these stage
// tables of every radix-4 pass, the pass of span Ns at tw + N + Ns - Ns0 (Ns0 = 2 when log2 N is
// odd, the radix-2 pass coming first, else 1); their values are entries of tw, so a butterfly
// rounds alike with either table.
__device__ inline const float2 *fft_stage_twiddles(const float2 *tw, int N, int logN, int Ns) {
    return tw + N + Ns - ((logN & 1) ? 2 : 1);
}

// One Stockham autosort pass of radix R over the LDS-resident row, in place: every thread loads
// its butterflies' inputs and (radix 4) their twiddles, the block synchronises, then results are
// written to their sorted slots. The twiddle loads do not depend on the row, so they are all issued
// before the barrier, where their latency overlaps the other waves' reads. stage: this pass's
// twiddles (fft_stage_twiddles); radix-2 passes run only at Ns = 1, where every twiddle is 1 and
// none is read. A butterfly with k = 0 takes no twiddle (its factors are 1).
template <int R, int NB>
__device__ inline void fft_pass(float2 *buf, int N, int Ns, const float2 *__restrict__ stage,
                                int tid) {
    const int nbf = N / R;
    float2 v[NB][R];
    float2 w[NB][R - 1];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = tid + b * kFftThreads;
        if (j < nbf) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = buf[j + r * nbf];
            if constexpr (R == 4) {
                const int k = j & (Ns - 1);
#pragma unroll
                for (int r = 1; r < R; ++r) w[b][r - 1] = stage[(r - 1) * Ns + k];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = tid + b * kFftThreads;
        if (j < nbf) {
            const int k = j & (Ns - 1);
            if constexpr (R == 4) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const float2 t = cmul(v[b][r], w[b][r - 1]);
                    v[b][r] = k != 0 ? t : v[b][r];
                }
            }
            dft_r<R>(v[b]);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[base + r * Ns] = v[b][r];
        }
    }
    __syncthreads();
}

// A radix-16 pass of the same Stockham form: the 16 inputs and 15 twiddles of a butterfly are read
// before the barrier (lanes past the last butterfly read the last one's); then the 16-point DFT
// writes its outputs. NB butterflies per
// thread (N / (16 kFftThreads), rounded up). k = 0 takes no twiddle, as in fft_pass.
template <int NB>
__device__ inline void fft_pass16(float2 *buf, int N, int Ns, const float2 *__restrict__ stage,
                                  int tid) {
    const int nbf = N / 16;
    float2 v[NB][16], w[NB][15];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = min(tid + b * kFftThreads, nbf - 1), k = j & (Ns - 1);
#pragma unroll
        for (int r = 0; r < 16; ++r) v[b][r] = buf[j + r * nbf];
#pragma unroll
        for (int r = 1; r < 16; ++r) w[b][r - 1] = stage[(r - 1) * Ns + k];
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = tid + b * kFftThreads;
        if (j < nbf) {  // whole waves: nbf is a multiple of 64 wherever a radix-16 pass runs
            const int k = j & (Ns - 1);
            if (k != 0) {
#pragma unroll
                for (int r = 1; r < 16; ++r) v[b][r] = cmul(v[b][r], w[b][r - 1]);
            }
            dft16_store(v[b], buf, (j - k) * 16 + k, Ns);
        }
    }
    __syncthreads();
}

// Source pulse of profile p: s = (p + 2 ceil(np/2)) mod np; the pulse axis is rolled by fftshift
// before and after the FFT (backprojection.cpp:98,104).
__device__ inline int profile_source(int p, int np) {
    const int pmid = np / 2 + np % 2;
    return (int)(((long long)p + 2LL * pmid) % np);
}

// Sample x of the FFT input of a pulse whose phase-history row is `row` and whose pulse window is
// wps: the padded, windowed, fftshifted sample.
// The filtered, windowed value of phase-history sample a with k_r value kr and range window wsi
// (pulse window wps).
__device__ inline float2 windowed_sample(float2 a, float kr, double wsi, double wps) {
    const float f = fabsf(kr);          // filt = |k_r| (backprojection.cpp:85)
    const double w = wsi * wps;         // win (backprojection.cpp:82)
    // phs * filt rounds in f32, then the f64 window applies (backprojection.cpp:90)
    return make_float2((float)((double)(a.x * f) * w), (float)((double)(a.y * f) * w));
}

__device__ inline float2 profile_sample(const float2 *__restrict__ row,
                                        const float *__restrict__ k_r,
                                        const double *__restrict__ ws, double wps, int ns, int N,
                                        int xpad, int x) {
    const int xs = (x + N / 2) & (N - 1);  // fftshift (signal_complex.h:24-29)
    const int i = xs - xpad;
    if (i >= 0 && i < ns) return windowed_sample(row[i], k_r[i], ws[i], wps);
    return make_float2(0.f, 0.f);
}

// Fills buf[0, N) with the padded, windowed, fftshifted samples of pulse s: sample i of the pulse
// lands at xs = i + xpad of the padded row (signal_complex.h:10-21), which fftshift moves to
// x = (xs + N/2) mod N (signal_complex.h:24-29); every other slot is zero. The slots that hold no
// sample are cleared first (no loads); then each thread places kRowUnroll samples per trip with
// all their loads issued before any is used. A lane past the last sample redoes the last one,
// writing the same value to the same slot as its owner.
constexpr int kRowUnroll = 4;

template <int T>
__device__ inline void load_profile_row(float2 *buf, const float2 *__restrict__ phs,
                                        const float *__restrict__ k_r,
                                        const double *__restrict__ ws,
                                        const double *__restrict__ wp, int ns, int N, int s,
                                        int tid) {
    const int xpad = floordiv2(N - ns);     // pad's leading offset (signal_complex.h:14)
    const double wps = wp[s];
    const float2 *row = phs + (size_t)s * ns;
    const int lo = max(xpad, 0), hi = min(xpad + ns, N);  // padded positions holding samples
    for (int x = tid; x < N; x += T) {
        const int xs = (x + N / 2) & (N - 1);
        if (xs < lo || xs >= hi) buf[x] = make_float2(0.f, 0.f);
    }
    const int i_lo = lo - xpad, i_hi = hi - xpad;  // samples that land in the row
    for (int i0 = i_lo + tid; i0 < i_hi; i0 += kRowUnroll * T) {
        float2 a[kRowUnroll];
        float kr[kRowUnroll];
        double w[kRowUnroll];
        int x[kRowUnroll];
#pragma unroll
        for (int u = 0; u < kRowUnroll; ++u) {
            const int i = min(i0 + u * T, i_hi - 1);
            x[u] = (i + xpad + N / 2) & (N - 1);
            a[u] = row[i];
            kr[u] = k_r[i];
            w[u] = ws[i];
        }
#pragma unroll
        for (int u = 0; u < kRowUnroll; ++u) buf[x[u]] = windowed_sample(a[u], kr[u], w[u], wps);
    }
}

// Writes the transformed row buf[0, N) as profile row p: Q(k) = dft((k + mid) mod N) with
// mid = N/2 + N%2 (fftshift, signal_complex.h:24-29; backprojection.cpp:104) as rotated (A, B)
// elements; with qscale (f16 storage) the row is first scaled by a power of two. `red` holds T / 64
// floats. Ends with a barrier.
template <typename QT, int T>
__device__ inline void store_profile_row(const float2 *buf, float *red, int N,
                                         const float2 *__restrict__ rot, int ext, int len,
                                         QT *__restrict__ q, float *__restrict__ qscale, int p,
                                         int tid) {
#pragma clang fp contract(off)
    const int mid = N / 2 + N % 2;
    QT *dst = q + (size_t)p * len;
    float scale = 1.0f;
    if (qscale) {
        // f16 storage: scale the profile by 2^e so that its largest component is about 2^14
        // (exact to undo; 11-bit significands keep ~5e-4 relative rounding per value)
        float m = 0.0f;
        for (int k = tid; k < N; k += T) m = fmaxf(m, fmaxf(fabsf(buf[k].x), fabsf(buf[k].y)));
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
        if ((tid & 63) == 0) red[tid >> 6] = m;
        __syncthreads();
        m = red[0];
        for (int w = 1; w < T / 64; ++w) m = fmaxf(m, red[w]);
        int e = 0;
        if (m > 0.0f && m <= 3.0e38f) (void)frexpf(m, &e);
        e = max(-100, min(100, 14 - e));
        scale = ldexpf(1.0f, e);
        if (tid == 0) qscale[p] = ldexpf(1.0f, -e);
    }
    // kStoreUnroll elements per thread and trip: every load of the trip (the rotation and the two
    // row values) is issued before the first store; stores share the memory counter with loads, so
    // a load waited for after a store would wait for that store too. A lane past the row's end redoes
    // its last element, storing the same value to the same place as the element's owner.
    constexpr int kStoreUnroll = 8;
    for (int j0 = tid; j0 < len; j0 += kStoreUnroll * T) {
        float2 a[kStoreUnroll], b[kStoreUnroll], r[kStoreUnroll];
        int jj[kStoreUnroll];
#pragma unroll
        for (int u = 0; u < kStoreUnroll; ++u) {
            const int j = min(j0 + u * T, len - 1);
            const int k0 = min(max(j - ext, 0), N - 1), k1 = min(max(j - ext + 1, 0), N - 1);
            jj[u] = j;
            const int s0 = k0 + mid, s1 = k1 + mid;  // < 2N
            a[u] = buf[s0 >= N ? s0 - N : s0];
            b[u] = buf[s1 >= N ? s1 - N : s1];
            r[u] = rot[j];
        }
#pragma unroll
        for (int u = 0; u < kStoreUnroll; ++u) {
            const float2 q0 = make_float2(a[u].x * scale, a[u].y * scale);
            const float2 dq = make_float2(b[u].x * scale - q0.x, b[u].y * scale - q0.y);
            storeq(dst + jj[u], cmul(q0, r[u]), cmul(dq, r[u]));
        }
    }
    __syncthreads();
}

// Range profile p in one block of 512 threads, the FFT in place in one LDS row.
template <typename QT, int NB4>
__device__ inline void range_profile_one(
    float2 *buf, float *red, int p, const float2 *__restrict__ phs, const float *__restrict__ k_r,
    const double *__restrict__ ws, const double *__restrict__ wp, int ns, int np, int N, int logN,
    const float2 *__restrict__ tw, const float2 *__restrict__ rot, int ext, int len,
    QT *__restrict__ q, float *__restrict__ qscale, int tid) {
    load_profile_row<kFftThreads>(buf, phs, k_r, ws, wp, ns, N, profile_source(p, np), tid);
    __syncthreads();
    for (int Ns = 1, left = logN; left > 0;) {
        const int R = fft_next_radix(Ns, left);
        if (R == 2) {
            fft_pass<2, 2 * NB4>(buf, N, 1, tw, tid);
        } else if (R == 4) {
            fft_pass<4, NB4>(buf, N, Ns, fft_stage_twiddles(tw, N, logN, Ns), tid);
        } else {
            fft_pass16<(NB4 + 3) / 4>(buf, N, Ns, fft_stage16_twiddles(tw, N, logN, Ns), tid);
        }
        Ns *= R;
        left -= R == 2 ? 1 : R == 4 ? 2 : 4;
    }
    store_profile_row<QT, kFftThreads>(buf, red, N, rot, ext, len, q, qscale, p, tid);
}

// Range profiles, one block of 512 threads per profile. NB4: radix-4 butterflies per thread and
// pass, N / (4 * kFftThreads) rounded up (4 up to N = 8192, 8 at 16384), so that the registers of
// smaller rows are sized for them.
//   NB4 = 4: two blocks fit a CU's LDS (4 waves per SIMD, at most 128 VGPRs); one profile per
//            block (blockIdx.x), the grid holds every profile.
//   NB4 = 8: one block per CU (128 KiB of LDS); blocks loop over profiles p = blockIdx.x +
//            k gridDim.x on a grid of kRowBlocksPerCu blocks per CU, so that a block moves on to
//            its next profile while the previous one's row stores drain (C3: 1.29 -> 1.23 ms;
//            at N = 8192 the loop made the kernel slower: 0.118 -> 0.123 ms).
constexpr int kRowBlocksPerCu = 4;

template <typename QT, int NB4>
__global__ __launch_bounds__(kFftThreads) __attribute__((amdgpu_waves_per_eu(NB4 == 4 ? 4 : 2)))
void range_profile_kernel(
    const float2 *__restrict__ phs, const float *__restrict__ k_r, const double *__restrict__ ws,
    const double *__restrict__ wp, int ns, int np, int N, int logN,
    const float2 *__restrict__ tw, const float2 *__restrict__ rot, int ext, int len,
    QT *__restrict__ q, float *__restrict__ qscale) {
    extern __shared__ float2 buf[];
    __shared__ float red[kFftThreads / 64];
    const int tid = threadIdx.x;
    if constexpr (NB4 == 4) {
        range_profile_one<QT, NB4>(buf, red, blockIdx.x, phs, k_r, ws, wp, ns, np, N, logN, tw,
                                   rot, ext, len, q, qscale, tid);
    } else {
        for (int p = blockIdx.x; p < np; p += gridDim.x)
            range_profile_one<QT, NB4>(buf, red, p, phs, k_r, ws, wp, ns, np, N, logN, tw, rot,
                                       ext, len, q, qscale, tid);
    }
}

// Range profiles of an N_fft that is not a power of two (N <= kMaxFftBluestein), by Bluestein's
// chirp-z identity nk = (n^2 + k^2 - (k - n)^2) / 2:
//   X_k = w_k sum_n (x_n w_n) conj(w_(k-n)),   w_n = exp(-i pi n^2 / N)
// a circular convolution of length M = 2^ceil(log2(2N - 1)) computed with two M-point FFTs in LDS:
//   buf = FFT_M(a), a_n = x_n w_n (n < N, else 0)
//   buf = FFT_M(conj(buf * Bs)),  Bs = FFT_M(b) / M,  b_m = conj(w_|m|) for |m| < N (mod M)
//   X_k = w_k conj(buf_k)                     (the inverse FFT as conj(FFT(conj(.))) / M)
// then the row store of the power-of-two kernels. x_n is the padded, windowed, fftshifted sample
// (fftshift by mid = N/2 + N%2, signal_complex.h:24-29). Tables (bluestein_tables, after the
// twiddles of M in the same allocation): Bs at tw + 2M, w at tw + 3M.
template <typename QT, int NB4>
__global__ __launch_bounds__(kFftThreads) void range_profile_bluestein_kernel(
    const float2 *__restrict__ phs, const float *__restrict__ k_r, const double *__restrict__ ws,
    const double *__restrict__ wp, int ns, int np, int N, int logM,
    const float2 *__restrict__ tw, const float2 *__restrict__ rot, int ext, int len,
    QT *__restrict__ q, float *__restrict__ qscale) {
    extern __shared__ float2 buf[];
    __shared__ float red[kFftThreads / 64];
    const int M = 1 << logM;
    const float2 *Bs = tw + 2 * (size_t)M;
    const float2 *w = tw + 3 * (size_t)M;
    const int p = blockIdx.x;
    const int tid = threadIdx.x;
    const int s = profile_source(p, np);
    const int xpad = floordiv2(N - ns);
    const double wps = wp[s];
    const float2 *row = phs + (size_t)s * ns;
    const int mid = N / 2 + N % 2;
    for (int n = tid; n < M; n += kFftThreads) {
        float2 v = make_float2(0.f, 0.f);
        if (n < N) {
            const int xs = n + mid >= N ? n + mid - N : n + mid;
            const int i = xs - xpad;
            if (i >= 0 && i < ns) v = cmul(windowed_sample(row[i], k_r[i], ws[i], wps), w[n]);
        }
        buf[n] = v;
    }
    __syncthreads();
    for (int pass = 0; pass < 2; ++pass) {
        int Ns = 1;
        if (logM & 1) {
            fft_pass<2, 2 * NB4>(buf, M, 1, tw, tid);
            Ns = 2;
        }
        for (; Ns < M; Ns *= 4)
            fft_pass<4, NB4>(buf, M, Ns, fft_stage_twiddles(tw, M, logM, Ns), tid);
        for (int k = tid; k < M; k += kFftThreads) {
            const float2 y = buf[k];
            if (pass == 0) {
                const float2 t = cmul(y, Bs[k]);
                buf[k] = make_float2(t.x, -t.y);
            } else if (k < N) {
                buf[k] = cmul(w[k], make_float2(y.x, -y.y));
            }
        }
        __syncthreads();
    }
    store_profile_row<QT, kFftThreads>(buf, red, N, rot, ext, len, q, qscale, p, tid);
}

// Variant for a device shared with long-running accumulation workgroups: blocks of 256 threads and
// few registers, the shape of an accumulation workgroup, so that each one that retires leaves room
// for one of these; each block transforms the profiles p = blockIdx.x + k gridDim.x. The FFT runs
// out of place between two LDS rows (2 N float2, N <= kMaxFftShared), one barrier per pass; the
// arithmetic of every butterfly is that of fft_pass.
constexpr int kFftSharedThreads = 256;

template <int R>
__device__ inline void fft_pass_oop(const float2 *src, float2 *dst, int N, int Ns,
                                    const float2 *__restrict__ stage, int tid) {
    const int nbf = N / R;
    // full trips of kOopGroup butterflies per thread load every input and twiddle of the trip
    // before the first butterfly; the rest go one at a time
    constexpr int kOopGroup = 4;
    int j0 = tid;
    for (; j0 + (kOopGroup - 1) * kFftSharedThreads < nbf; j0 += kOopGroup * kFftSharedThreads) {
        float2 v[kOopGroup][R], w[kOopGroup][R - 1];
#pragma unroll
        for (int g = 0; g < kOopGroup; ++g) {
            const int j = j0 + g * kFftSharedThreads;
#pragma unroll
            for (int r = 0; r < R; ++r) v[g][r] = src[j + r * nbf];
            if constexpr (R == 4) {
                const int k = j & (Ns - 1);
#pragma unroll
                for (int r = 1; r < R; ++r) w[g][r - 1] = stage[(r - 1) * Ns + k];
            }
        }
#pragma unroll
        for (int g = 0; g < kOopGroup; ++g) {
            const int j = j0 + g * kFftSharedThreads;
            const int k = j & (Ns - 1);
            if constexpr (R == 4) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const float2 t = cmul(v[g][r], w[g][r - 1]);
                    v[g][r] = k != 0 ? t : v[g][r];
                }
            }
            dft_r<R>(v[g]);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[base + r * Ns] = v[g][r];
        }
    }
    for (int j = j0; j < nbf; j += kFftSharedThreads) {
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = src[j + r * nbf];
        const int k = j & (Ns - 1);
        if (k != 0) {
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul(v[r], stage[(r - 1) * Ns + k]);
        }
        dft_r<R>(v);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[base + r * Ns] = v[r];
    }
    __syncthreads();
}

// The radix-16 pass of fft_pass16, out of place, one butterfly at a time (the same arithmetic).
__device__ inline void fft_pass16_oop(const float2 *src, float2 *dst, int N, int Ns,
                                      const float2 *__restrict__ stage, int tid) {
    const int nbf = N / 16;
#pragma nounroll
    for (int j = tid; j < nbf; j += kFftSharedThreads) {
        float2 v[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = src[j + r * nbf];
        const int k = j & (Ns - 1);
        if (k != 0) {
#pragma unroll
            for (int r = 1; r < 16; ++r) v[r] = cmul(v[r], stage[(r - 1) * Ns + k]);
        }
        dft16_store(v, dst, (j - k) * 16 + k, Ns);
    }
    __syncthreads();
}

template <typename QT>
__global__ __launch_bounds__(kFftSharedThreads) void range_profile_shared_kernel(
    const float2 *__restrict__ phs, const float *__restrict__ k_r, const double *__restrict__ ws,
    const double *__restrict__ wp, int ns, int np, int N, int logN,
    const float2 *__restrict__ tw, const float2 *__restrict__ rot, int ext, int len,
    QT *__restrict__ q, float *__restrict__ qscale) {
    extern __shared__ float2 buf[];
    __shared__ float red[kFftSharedThreads / 64];
    const int tid = threadIdx.x;
    for (int p = blockIdx.x; p < np; p += gridDim.x) {
        float2 *a = buf, *b = buf + N;
        load_profile_row<kFftSharedThreads>(a, phs, k_r, ws, wp, ns, N, profile_source(p, np),
                                            tid);
        __syncthreads();
        for (int Ns = 1, left = logN; left > 0;) {
            const int R = fft_next_radix(Ns, left);
            if (R == 2)
                fft_pass_oop<2>(a, b, N, 1, tw, tid);
            else if (R == 4)
                fft_pass_oop<4>(a, b, N, Ns, fft_stage_twiddles(tw, N, logN, Ns), tid);
            else
                fft_pass16_oop(a, b, N, Ns, fft_stage16_twiddles(tw, N, logN, Ns), tid);
            Ns *= R;
            left -= R == 2 ? 1 : R == 4 ? 2 : 4;
            float2 *t = a;
            a = b;
            b = t;
        }
        store_profile_row<QT, kFftSharedThreads>(a, red, N, rot, ext, len, q, qscale, p, tid);
    }
}

// Range profiles longer than one LDS row (kMaxFft < N <= kMaxFftTwoPass): a four-step FFT of
// N = N1 N2 (N1 = 2^ceil(m/2), N2 = 2^floor(m/2)) in three launches over a batch of profiles, with
// n = n2 + N2 n1 and k = k1 + N1 k2:
//   fft_cols_kernel  per profile and block of C = kMaxFft / N1 consecutive n2: the length-N1 DFTs
//                    over n1 of x[n2 + N2 n1] in LDS, times W_N^(n2 k1), into Y[n2][k1]
//   fft_rows_kernel  per profile and block of D = kMaxFft / N2 consecutive k1: the length-N2 DFTs
//                    over n2 of Y[n2][k1] in LDS, into X[k1 + N1 k2] (natural order)
//   rows_store_kernel per profile: the rotated (A, B) row elements from X, as store_profile_row
// Sequences sit in LDS at a stride of L + 1 elements, so the transposed accesses of the loads and
// stores hit distinct banks while global accesses stay contiguous. The butterflies are those of
// fft_pass; the twiddles come from the N-point table.
constexpr int kTwoPassThreads = kFftThreads;

// One Stockham pass of radix R over kMaxFft / L sequences of length L = 2^logL at stride L + 1:
// the butterflies of fft_pass, applied to each sequence (twiddle W_L^t = tw[t * N / L]).
template <int R>
__device__ inline void fft_pass_multi(float2 *buf, int logL, int Ns, int tw_scale,
                                      const float2 *__restrict__ tw, int tid) {
    constexpr int NB = kMaxFft / R / kTwoPassThreads;
    const int L = 1 << logL, nbf = L / R, lp = L + 1;
    int lognbf = 0;
    while ((1 << lognbf) < nbf) ++lognbf;
    const int stride = L / (Ns * R);
    float2 v[NB][R];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = tid + b * kTwoPassThreads;
        const int c = j >> lognbf, jj = j & (nbf - 1);
        const float2 *src = buf + c * lp + jj;
#pragma unroll
        for (int r = 0; r < R; ++r) v[b][r] = src[r * nbf];
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = tid + b * kTwoPassThreads;
        const int c = j >> lognbf, jj = j & (nbf - 1);
        const int k = jj & (Ns - 1);
        if (k != 0) {
#pragma unroll
            for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], tw[r * k * stride * tw_scale]);
        }
        dft_r<R>(v[b]);
        float2 *dst = buf + c * lp + (jj - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[r * Ns] = v[b][r];
    }
    __syncthreads();
}

// The length-2^logL DFTs of every sequence in LDS, in place (radix 4, one radix-2 pass first for
// an odd logL).
__device__ inline void fft_multi(float2 *buf, int logL, int logN, const float2 *__restrict__ tw,
                                 int tid) {
    const int tw_scale = 1 << (logN - logL);
    int Ns = 1;
    if (logL & 1) {
        fft_pass_multi<2>(buf, logL, 1, tw_scale, tw, tid);
        Ns = 2;
    }
    for (; Ns < (1 << logL); Ns *= 4) fft_pass_multi<4>(buf, logL, Ns, tw_scale, tw, tid);
}

__global__ __launch_bounds__(kTwoPassThreads) void fft_cols_kernel(
    const float2 *__restrict__ phs, const float *__restrict__ k_r, const double *__restrict__ ws,
    const double *__restrict__ wp, int ns, int np, int logN, int logN1, int p0,
    const float2 *__restrict__ tw, float2 *__restrict__ Y) {
    extern __shared__ float2 buf[];
    const int N = 1 << logN, N1 = 1 << logN1, N2 = N >> logN1;
    const int logC = 14 - logN1, C = 1 << logC;  // kMaxFft = 2^14 elements per block
    const int blocks_per_profile = N2 >> logC;
    const int pl = blockIdx.x / blocks_per_profile, c0 = (blockIdx.x % blocks_per_profile) << logC;
    const int p = p0 + pl, s = profile_source(p, np);
    const int xpad = floordiv2(N - ns);
    const double wps = wp[s];
    const float2 *row = phs + (size_t)s * ns;
    const int tid = threadIdx.x;
    // x[n2 + N2 n1], consecutive threads along n2 (contiguous input)
    for (int e = tid; e < kMaxFft; e += kTwoPassThreads) {
        const int c = e & (C - 1), n1 = e >> logC;
        buf[c * (N1 + 1) + n1] = profile_sample(row, k_r, ws, wps, ns, N, xpad, c0 + c + N2 * n1);
    }
    __syncthreads();
    fft_multi(buf, logN1, logN, tw, tid);
    // Y[n2][k1] = DFT * W_N^(n2 k1), consecutive threads along k1
    float2 *out = Y + (size_t)pl * N;
    for (int e = tid; e < kMaxFft; e += kTwoPassThreads) {
        const int k1 = e & (N1 - 1), c = e >> logN1, n2 = c0 + c;
        out[(size_t)n2 * N1 + k1] = cmul(buf[c * (N1 + 1) + k1], tw[(n2 * k1) & (N - 1)]);
    }
}

__global__ __launch_bounds__(kTwoPassThreads) void fft_rows_kernel(
    const float2 *__restrict__ Y, int logN, int logN1, const float2 *__restrict__ tw,
    float2 *__restrict__ X) {
    extern __shared__ float2 buf[];
    const int N = 1 << logN, N1 = 1 << logN1, logN2 = logN - logN1, N2 = 1 << logN2;
    const int logD = 14 - logN2, D = 1 << logD;
    const int blocks_per_profile = N1 >> logD;
    const int pl = blockIdx.x / blocks_per_profile, d0 = (blockIdx.x % blocks_per_profile) << logD;
    const float2 *in = Y + (size_t)pl * N;
    const int tid = threadIdx.x;
    // Y[n2][d0 + d], consecutive threads along d
    for (int e = tid; e < kMaxFft; e += kTwoPassThreads) {
        const int d = e & (D - 1), n2 = e >> logD;
        buf[d * (N2 + 1) + n2] = in[(size_t)n2 * N1 + d0 + d];
    }
    __syncthreads();
    fft_multi(buf, logN2, logN, tw, tid);
    // X[d0 + d + N1 k2], consecutive threads along d
    float2 *out = X + (size_t)pl * N;
    for (int e = tid; e < kMaxFft; e += kTwoPassThreads) {
        const int d = e & (D - 1), k2 = e >> logD;
        out[(size_t)k2 * N1 + d0 + d] = buf[d * (N2 + 1) + k2];
    }
}

template <typename QT>
__global__ __launch_bounds__(kTwoPassThreads) void rows_store_kernel(
    const float2 *__restrict__ X, int N, int p0, const float2 *__restrict__ rot, int ext, int len,
    QT *__restrict__ q, float *__restrict__ qscale) {
    __shared__ float red[kTwoPassThreads / 64];
    const int pl = blockIdx.x;
    store_profile_row<QT, kTwoPassThreads>(X + (size_t)pl * N, red, N, rot, ext, len, q, qscale,
                                           p0 + pl, threadIdx.x);
}

// ------------------------------------------------------------------ accumulation

constexpr int kBpThreads = 256;
// Pulses per f32 partial sum (flushed to the f64 accumulator after each block) and the unroll
// factor of the pulse loop; compile-time knobs for tuning sweeps.
#ifndef SARBP_BP_PULSE_BLOCK
#define SARBP_BP_PULSE_BLOCK 64
#endif
#ifndef SARBP_BP_UNROLL
#define SARBP_BP_UNROLL 1
#endif
constexpr int kBpPulseBlock = SARBP_BP_PULSE_BLOCK;
constexpr int kDefaultPpt = 2;

// Tile of 16 x (16 * PPT) output pixels per block; thread (tx, ty) owns pixels (tx, ty + 16 j).
// Per pulse p the pixel's row coordinate is t = B_p - |r - pos_p| / lsr. Most tiles form it from
// the expansion about the tile's centre (accumulate_pulses_expanded, below); the direct form, used
// where the expansion's error bound fails, keeps squared ranges in row units,
// s = sigma |r - pos_p|^2 with sigma = inv_lsr^2, so that sqrt(s) is d / lsr. Per pulse p
// (constants from the pulse table):
//   s  = sigma |r|^2 + Cs_p - 2 sigma r.pos_p          f64, from per-pixel terms (PixTerms)
//   t  = B_p - sqrt(s)                                 f64: v_rsq_f64 + one Newton step (row_coord)
//   j  = floor(t), w = t - j                           int, f64 -> f32
//   term = (A_j + w B_j) * exp(2 pi i kappa w)         one 16-byte gather; f32
//   acc += term                                        f32, flushed to f64 every 64 pulses
// This fast path needs every t of the block inside the row; a block whose pixels reach beyond the
// row's cover radius takes the exact path instead: j clamped to the row, the phase
// kappa (t - j) reduced in f64 (the clamped edge elements have B = 0, so the weight is moot).
//
// Per-pixel terms, in registers for the whole pulse loop:
//   r mode:    s_j = fma(a_j, px, fma(b_j, py, fma(c_j, pz, r2_j + Cs)))  with (a, b, c) = -2 sigma r
//   grid mode: r(u, v) = u u_hat + v v_hat, so -2 sigma r.pos = u U_p + v V_p (pulse table) and
//              s_0 = fma(u, U, fma(v, V, r2_0 + Cs))  with a_0 = u, b_0 = v of the thread's pixel 0;
//              pixel j lies 16 j output rows further: s_j = s_0 + fma(c_j, V, dr2_j) with
//              c_j = v_j - v_0 and dr2_j = r2_j - r2_0
// (r2 = sigma |r|^2). Lanes whose pixel is outside the region take pixel 0's location (grid) or the
// origin, so that their loads stay inside the row.
template <int PPT>
struct PixTerms {
    double a[PPT], b[PPT], c[PPT], r2[PPT], dr2[PPT];
};

// One pulse's gather for the thread's pixels: the loaded row elements (in flight until consumed),
// interpolation weights, phases (revolutions; on the fast path formed when consumed, kappa w) and,
// for f16 rows, the row's scale.
template <int PPT>
struct Gather {
    float4 e[PPT];
    float ww[PPT], xx[PPT];
    float qs;
};

template <typename QT, int PPT, bool FAST, bool GRID>
__device__ inline void gather_pulse(const BpArgs &a, const PixTerms<PPT> &P, int p,
                                    Gather<PPT> &g) {
    constexpr bool kScaled = sizeof(QT) == sizeof(QPairF16);
    const QT *qbase = reinterpret_cast<const QT *>(a.q);
    const PulseConst c = a.pc[p];
    const __amdgpu_buffer_rsrc_t row = row_rsrc(qbase + (size_t)p * a.len, a.len, (int)sizeof(QT));
    g.qs = kScaled ? a.qscale[p] : 1.0f;
    double sq[PPT];
    if (GRID) {
        sq[0] = __builtin_fma(P.a[0], c.U, __builtin_fma(P.b[0], c.V, P.r2[0] + c.Cs));
#pragma unroll
        for (int j = 1; j < PPT; ++j) sq[j] = sq[0] + __builtin_fma(P.c[j], c.V, P.dr2[j]);
    } else {
#pragma unroll
        for (int j = 0; j < PPT; ++j)
            sq[j] = __builtin_fma(P.a[j], c.px,
                                  __builtin_fma(P.b[j], c.py, __builtin_fma(P.c[j], c.pz, P.r2[j] + c.Cs)));
    }
    int ii[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const double t = row_coord(sq[j], c.B);
        if (FAST) {
            ii[j] = trunc_i32(t);
            g.ww[j] = (float)__builtin_amdgcn_fract(t) * weight_unit<QT>();
        } else {
            ii[j] = trunc_clamp(t, a.len - 2);
            const double w = t - (double)ii[j];
            g.ww[j] = (float)w * weight_unit<QT>();
            g.xx[j] = (float)__builtin_amdgcn_fract(a.hdr->kappa * w);
        }
    }
#pragma unroll
    for (int j = 0; j < PPT; ++j) g.e[j] = loadq(qbase, row, ii[j]);
}

// term = (A + w B) exp(2 pi i x) added to the f32 partial sums; ww and the stored B carry inverse
// scales (weight_unit), and on the fast path x = kappa_f ww with kappa_f = kappa / weight_unit.
template <typename QT, int PPT, bool FAST>
__device__ inline void add_terms(const Gather<PPT> &g, float kappa_f, float (&sr)[PPT],
                                 float (&si)[PPT]) {
    constexpr bool kScaled = sizeof(QT) == sizeof(QPairF16);
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const float x = FAST ? kappa_f * g.ww[j] : g.xx[j];
        const float sn = __builtin_amdgcn_sinf(x);
        const float cs = __builtin_amdgcn_cosf(x);
        float qr = __builtin_fmaf(g.ww[j], g.e[j].z, g.e[j].x);
        float qi = __builtin_fmaf(g.ww[j], g.e[j].w, g.e[j].y);
        if (kScaled) {
            qr *= g.qs;
            qi *= g.qs;
        }
        sr[j] = __builtin_fmaf(qr, cs, sr[j]);
        sr[j] = __builtin_fmaf(-qi, sn, sr[j]);
        si[j] = __builtin_fmaf(qr, sn, si[j]);
        si[j] = __builtin_fmaf(qi, cs, si[j]);
    }
}

// The pulse loop, software-pipelined by one pulse: the gather of pulse p + 1 is issued before the
// terms of pulse p are formed, so a wave's load latency overlaps its own arithmetic. Two gather
// sets alternate (two pulses per trip), so nothing is copied between them, and every gather in the
// loop is issued unconditionally (the last trip's second one is clamped to pulse np - 1 and not
// used): a load skipped on some path would make the compiler wait for all of them. The f32 partial
// sums are flushed to f64 after every kBpPulseBlock pulses, as in a plain loop.
static_assert((kBpPulseBlock & (kBpPulseBlock - 1)) == 0 && kBpPulseBlock % 2 == 0,
              "the flush test needs an even power of two");
template <typename QT, int PPT, bool FAST, bool GRID>
__device__ inline void accumulate_pulses(const BpArgs &a, const PixTerms<PPT> &P,
                                         double (&accr)[PPT], double (&acci)[PPT]) {
    const int np = a.np;
    const float kappa_f = (float)a.hdr->kappa / weight_unit<QT>();
    float sr[PPT], si[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) sr[j] = si[j] = 0.f;
    auto flush = [&] {
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            accr[j] += (double)sr[j];
            acci[j] += (double)si[j];
            sr[j] = si[j] = 0.f;
        }
    };
    Gather<PPT> g0, g1;
    gather_pulse<QT, PPT, FAST, GRID>(a, P, 0, g0);
    for (int p = 0; p + 1 < np; p += 2) {
        gather_pulse<QT, PPT, FAST, GRID>(a, P, p + 1, g1);
        add_terms<QT, PPT, FAST>(g0, kappa_f, sr, si);
        gather_pulse<QT, PPT, FAST, GRID>(a, P, min(p + 2, np - 1), g0);
        add_terms<QT, PPT, FAST>(g1, kappa_f, sr, si);
        if (((p + 2) & (kBpPulseBlock - 1)) == 0) flush();
    }
    if (np & 1) add_terms<QT, PPT, FAST>(g0, kappa_f, sr, si);  // pulse np - 1
    flush();
}

// Tile-centred range expansion. With R = (r_c - pos_p) / lsr the vector from the platform to the
// tile's centre in row units, d_c = |R|, n = R / d_c and delta = (r - r_c) / lsr the pixel's offset,
//   |R + delta| = d_c + p + q h - 2 h^2 p q + E,   p = n.delta,  q = |delta|^2 - p^2,  h = 1 / (2 d_c)
// (the Taylor series of d_c sqrt(1 + 2 p / d_c + |delta|^2 / d_c^2) to third order in
// eps = |delta| / d_c), with |E| <= d_c eps^4 / 8 for eps <= 0.1 (checked against 40-digit
// arithmetic over random directions). So t = T - p - q (h + g p) with T = B_p - d_c and g = -2 h^2:
// per pixel-pulse 7 f64 operations with pixel_locs, and 3 in grid mode, which evaluates the same
// polynomial in another form (below); no square root, against 8 (6.5 in grid mode) plus v_rsq_f64
// for the direct form. A tile takes this path when
// |delta|_max^4 / (4 d_min^3) <= kExpandTol, twice the bound on E, over all of its pixels and pulses
// (d_min: the least d_c over the pulses); otherwise the direct form. On the C1-C3 scenes the bound
// is below 2e-10 row elements (3e-9 radians of phase).
//
// With pixel_locs, the per-pulse constants {T, n, h, g} of the tile are built kExpBlock pulses at a
// time into one of two LDS buffers by the first kExpBlock threads, one pulse each, while the
// workgroup consumes the other buffer; one barrier per block.
constexpr int kExpBlock = 128;
constexpr double kExpandTol = 1e-8;
constexpr double kRowMagic = 1572864.0;  // 1.5 * 2^20
static_assert(kMaxFftTwoPass + kMaxFftTwoPass / 2 + 2 < (1 << 18), "rows fit the fixed-point index");
// Row coordinates in fixed point: the expansion's T carries 1.5 * 2^20, so that the f64 result
// y = t + 1.5 * 2^20 (t in [-2^19, 2^19)) has exponent 20 and its mantissa is (t + 2^19) * 2^32: the
// high word's low 19 bits hold floor(t) for t >= 0 (a t in [-2^18, 0) gives an index of at least
// 2^18, beyond every row, which the buffer's bounds check reads as zeros) and the low word holds
// fract(t) * 2^32. One AND and one conversion instead of three quarter-rate conversions (cvt_i32,
// fract, cvt_f32); t keeps 2^-32 of a row element (2e-10; the phase 1e-9 rad).
static_assert(kExpBlock % kBpPulseBlock == 0 && kExpBlock <= kBpThreads,
              "a block of pulses holds whole flush periods and is filled in one pass");

// Grid mode evaluates the same expansion as a cubic in the pixel's offsets (du, dv) from the centre
// (metres along u_hat and v_hat): with p = a du + b dv (a = n.u_hat / lsr, b = n.v_hat / lsr) and
// q = Quu du^2 + 2 Quv du dv + Qvv dv^2 (Quu = sigma |u_hat|^2 - a^2, Quv = sigma u_hat.v_hat - a b,
// Qvv = sigma |v_hat|^2 - b^2; sigma = inv_lsr^2),
//   t = T - p - h q - g p q = sum c_ij du^i dv^j (i + j <= 3).
// A thread's pixels share one column (one du), so per pulse and column the cubic in dv has
// coefficients C0..C3 (C_k a polynomial of degree 3 - k in du; C3 = c_03 is the same for every
// column), and each pixel takes three f64 FMAs (Horner in dv) where the form above takes five.
// Per block of kGridBlock pulses the coefficients are built in two stages that run one block apart:
// c_ij per pulse (one wave's lanes, one pulse each; the wave rotates with the block), then C0..C2
// per pulse and column (all threads). One barrier per block. A block of 32 pulses takes 30 KiB of
// LDS per workgroup (4 or 5 workgroups per CU at 4 or 3 pixels per thread, within 160 KiB); at 2
// pixels per thread (8 per CU) blocks of 16 take 15 KiB. Halving the block measured 1-2 % slower.
template <int PPT>
constexpr int kGridBlock = PPT <= 2 ? 16 : 32;
static_assert(kBpPulseBlock % kGridBlock<2> == 0 && kBpPulseBlock % kGridBlock<4> == 0 &&
                  kGridBlock<4> <= 64,
              "flush periods hold whole blocks; one wave builds a block's c_ij");

template <bool GRID, int PPT>
struct ExpShared {
    double k[2][kExpBlock][6];  // per pulse: T, n, h, g
};
template <int B>
struct GridCubicB {
    double c[2][B][10];       // per pulse: c_00 (+ kRowMagic), c_10, c_20, c_30, c_01, c_11, c_21,
                              // c_02, c_12, c_03
    double2 c01[2][B][16];    // per pulse and column: (C0, C1)
    double c2[2][B][16];      // C2
    double c3[2][B];          // C3
};
template <int PPT>
using GridCubic = GridCubicB<kGridBlock<PPT>>;
// The double-float form of the cubic (range_mode SARBP_RANGE_DF32): per pulse and column the
// constant term as its exact integer part H = floor(C0 - kRowMagic) and the f32 remainder
// L = C0 - kRowMagic - H, and C1..C3 in f32, so that a pixel's row coordinate is H plus the f32
// cubic L + dv (C1 + dv (C2 + dv C3)).
template <int B>
struct GridCubicDF {
    double c[2][B][10];  // per pulse: the c_ij of the f64 form (c_00 + kRowMagic)
    int h[2][B][16];
    float l[2][B][16], c1[2][B][16], c2[2][B][16];
    float c3[2][B];
};
template <int PPT>
struct ExpShared<true, PPT> {
    union {
        double k[2][kExpBlock][6];  // per pulse: T, n.u_hat / lsr, n.v_hat / lsr, -, h, g
        GridCubic<PPT> cub;
        GridCubicDF<kGridBlock<PPT>> cdf;
    };
};

// Row windows in LDS (grid mode, cubic form, f32 rows). Per block of kWinBlock pulses and per
// pulse, the W consecutive row elements [j0_p, j0_p + W) that hold every row coordinate the tile's
// pixels take at that pulse are copied into LDS (global_load_lds: the texture path moves W elements
// per pulse instead of one gathered element per pixel, ~1/14 of the bytes on C2), and the pixels
// read them there. Windows are double-buffered: block b + 1's are loaded while block b's are read.
// W is per tile: 2 H + 2 with H a bound on |t - t_c| over the tile's pixels and all pulses, from
// the cubic's coefficients (triangle inequality over its terms); a tile whose W exceeds the slot
// keeps the gathers. The window offset enters the cubic's constant term: c_00 - j0_p + S s, with
// s the pulse's slot in the two buffers and S the slot stride, so the pixel's fixed-point
// coordinate y is its LDS element index plus kRowMagic, and both c_00 and y stay in [2^20, 2^21),
// where subtracting an integer is exact: every weight and element is the one the gathers read, and
// the image the same bits.
//
// Per-block costs are kept off the VALU: the c_ij of kWinGroup pulses are built at once by one
// wave (one pulse per lane) into a two-group ring, and each slot's window is loaded by
// ceil(W / 64) wave instructions whose source is a scalar base plus the lane's offset.
#ifndef SARBP_WIN_BLOCK
#define SARBP_WIN_BLOCK 8
#endif
constexpr int kWinBlock = SARBP_WIN_BLOCK;
constexpr int kWinGroup = 32;
static_assert(kBpPulseBlock % kWinBlock == 0 && kWinGroup % kWinBlock == 0 && kWinGroup <= 64,
              "flush periods hold whole window blocks; one wave builds a group's c_ij");
// Slot stride S in elements (the largest W) per tile height; 0: no window path. LDS per workgroup:
// ~11.4 KiB of coefficients + 2 kWinBlock S 16 B of windows, within 160 KiB over the workgroups a
// CU holds (4 at 4 pixels per thread: S 96, 35.9 KiB; 5 at 3: S 80, 31.4 KiB).
template <int PPT>
constexpr int kWinSlot = PPT == 4 ? 96 : PPT == 3 ? 80 : 0;
struct WinCubic {
    double c[2][kWinGroup][10];     // c_ij ring, two groups (c_00 + kRowMagic - j0 + S s)
    int j0[2][kWinGroup];           // window start of each pulse
    double2 c01[2][kWinBlock][16];  // per pulse and column: (C0, C1)
    double c2[2][kWinBlock][16];    // C2
    double c3[2][kWinBlock];        // C3
};
template <int PPT>
struct WinShared {
    union {
        double k[2][kExpBlock][6];
        GridCubicB<kGridBlock<PPT>> cub;  // tiles whose window does not fit
        GridCubicDF<kGridBlock<PPT>> cdf;  // double-float form
        struct {
            WinCubic wc;
            QPairF32 e[2 * kWinBlock * (kWinSlot<PPT> > 0 ? kWinSlot<PPT> : 1)];
        } w;
    };
};

template <bool GRID, int PPT, bool WIN = false>
struct alignas(16) TileShared {
    typename std::conditional<WIN, WinShared<PPT>, ExpShared<GRID, PPT>>::type x;
    double rc[3];  // centre of the expansion (metres)
    double dd_max[kBpThreads / 64], d2_min[kBpThreads / 64], h_max[kBpThreads / 64];
    int far[kBpThreads / 64];
};

// Per-pixel offsets from the tile's centre. Grid mode (used for the tile's error bound only):
// e0[0] = du (metres along u_hat, shared by the thread's pixels, which lie in one column),
// e1[j] = dv (metres along v_hat). pixel_locs mode: e0, e1, e2 = delta in row units.
// dd = |delta|^2 in row units.
template <int PPT>
struct ExpTerms {
    double e0[PPT], e1[PPT], e2[PPT], dd[PPT];
};

template <bool GRID, int PPT, class SH>
__device__ inline void fill_expansion(const BpArgs &a, SH &sh, int blk, int buf) {
    const int i = threadIdx.x, p = blk * kExpBlock + i;
    if (i >= kExpBlock || p >= a.np) return;
    const double inv_lsr = a.hdr->inv_lsr;
    const PulseConst c = a.pc[p];
    const double Rx = (sh.rc[0] - c.px) * inv_lsr, Ry = (sh.rc[1] - c.py) * inv_lsr,
                 Rz = (sh.rc[2] - c.pz) * inv_lsr;
    const double d = sqrt(Rx * Rx + Ry * Ry + Rz * Rz);
    const double id = 1.0 / d;
    const double h = 0.5 * id;
    double *k = sh.x.k[buf][i];
    k[0] = (c.B - d) + kRowMagic;
    if (GRID) {
        const double s = id * inv_lsr;
        k[1] = (Rx * a.g.uh[0] + Ry * a.g.uh[1] + Rz * a.g.uh[2]) * s;
        k[2] = (Rx * a.g.vh[0] + Ry * a.g.vh[1] + Rz * a.g.vh[2]) * s;
        k[3] = 0.0;
    } else {
        k[1] = Rx * id;
        k[2] = Ry * id;
        k[3] = Rz * id;
    }
    k[4] = h;
    k[5] = -2.0 * h * h;
}

// The cubic's coefficients c_ij for pulse p about the tile centre rc (metres): k[0] = c_00 +
// kRowMagic, then c_10, c_20, c_30, c_01, c_11, c_21, c_02, c_12, c_03 (row units per metre^(i+j)).
__device__ inline void cubic_coeffs(const BpArgs &a, const double *rc, int p, double (&k)[10]) {
    const double inv_lsr = a.hdr->inv_lsr, sigma = inv_lsr * inv_lsr;
    const PulseConst c = a.pc[p];
    const double Rx = (rc[0] - c.px) * inv_lsr, Ry = (rc[1] - c.py) * inv_lsr,
                 Rz = (rc[2] - c.pz) * inv_lsr;
    const double d = sqrt(Rx * Rx + Ry * Ry + Rz * Rz);
    const double id = 1.0 / d;
    const double h = 0.5 * id, g = -2.0 * h * h;
    const double *uh = a.g.uh, *vh = a.g.vh;
    const double s = id * inv_lsr;
    const double na = (Rx * uh[0] + Ry * uh[1] + Rz * uh[2]) * s;
    const double nb = (Rx * vh[0] + Ry * vh[1] + Rz * vh[2]) * s;
    const double quu = sigma * (uh[0] * uh[0] + uh[1] * uh[1] + uh[2] * uh[2]) - na * na;
    const double quv = sigma * (uh[0] * vh[0] + uh[1] * vh[1] + uh[2] * vh[2]) - na * nb;
    const double qvv = sigma * (vh[0] * vh[0] + vh[1] * vh[1] + vh[2] * vh[2]) - nb * nb;
    k[0] = (c.B - d) + kRowMagic;
    k[1] = -na;
    k[2] = -h * quu;
    k[3] = -g * na * quu;
    k[4] = -nb;
    k[5] = -2.0 * h * quv;
    k[6] = -g * (nb * quu + 2.0 * na * quv);
    k[7] = -h * qvv;
    k[8] = -g * (2.0 * nb * quv + na * qvv);
    k[9] = -g * nb * qvv;
}

// Bound on |t - t_c| over a tile whose offsets satisfy |du| <= U, |dv| <= V, from the cubic's
// coefficients (each term at its largest), plus a margin for the rounding of y.
__device__ inline double cubic_half_span(const double (&k)[10], double U, double V) {
    const double u2 = U * U, v2 = V * V;
    return fabs(k[1]) * U + fabs(k[2]) * u2 + fabs(k[3]) * u2 * U + fabs(k[4]) * V +
           fabs(k[5]) * U * V + fabs(k[6]) * u2 * V + fabs(k[7]) * v2 + fabs(k[8]) * U * v2 +
           fabs(k[9]) * v2 * V + 1e-3;
}

// Window parameters of a tile on the window path: W elements per pulse, the largest offsets U, V
// and the row length.
struct WinTile {
    int W;
    double U, V;
    int len;
};

// Grid mode, first stage: c_ij of block blk's pulses, one lane per pulse, in wave blk mod 4.
static_assert(kBpThreads == 256, "fill_grid_cubic rotates its writer wave over four waves");
template <int B, class CUB>
__device__ inline void fill_grid_cubic(const BpArgs &a, const double *rc, CUB &cub, int blk) {
    const int i = (int)threadIdx.x - 64 * (blk & 3), p = blk * B + i;
    if (i < 0 || i >= B || p >= a.np) return;
    double k[10];
    cubic_coeffs(a, rc, p, k);
    double *dst = cub.c[blk & 1][i];
#pragma unroll
    for (int m = 0; m < 10; ++m) dst[m] = k[m];
}

// Window path, first stage: c_ij of group g's kWinGroup pulses (one lane per pulse, wave g mod 4),
// each with its window start j0 (clamped to the row) and c_00 - j0 + S s in place of c_00, s the
// pulse's slot: block b = 4 g + i / 8 uses buffer b mod 2.
template <int S>
__device__ inline void fill_win_group(const BpArgs &a, const double *rc, WinCubic &wc, int g,
                                      const WinTile &win) {
    const int i = (int)threadIdx.x - 64 * (g & 3), p = g * kWinGroup + i;
    if (i < 0 || i >= kWinGroup || p >= a.np) return;
    double k[10];
    cubic_coeffs(a, rc, p, k);
    const double H = cubic_half_span(k, win.U, win.V);
    const double t_c = k[0] - kRowMagic;
    const int j = min(max((int)floor(t_c - H), 0), win.len - win.W);
    const int blk = p / kWinBlock, slot = (blk & 1) * kWinBlock + p % kWinBlock;
    wc.j0[g & 1][i] = j;
    k[0] -= (double)(j - S * slot);
    double *dst = wc.c[g & 1][i];
#pragma unroll
    for (int m = 0; m < 10; ++m) dst[m] = k[m];
}

// Grid mode, second stage: C0..C3 of block blk's pulses for the tile's 16 columns
// (du = (x - 7.5) d_u), from the c_ij of the first stage.
template <int B>
__device__ inline void fill_grid_columns(const BpArgs &a, GridCubicB<B> &cub, int blk) {
    const int n = min(B, a.np - blk * B), buf = blk & 1;
    for (int job = threadIdx.x; job < 16 * n; job += kBpThreads) {
        const int i = job >> 4, x = job & 15;
        const double *k = cub.c[buf][i];
        const double du = ((double)x - 7.5) * a.g.d_u;
        const double c0 = __builtin_fma(__builtin_fma(__builtin_fma(k[3], du, k[2]), du, k[1]), du, k[0]);
        const double c1 = __builtin_fma(__builtin_fma(k[6], du, k[5]), du, k[4]);
        cub.c01[buf][i][x] = make_double2(c0, c1);
        cub.c2[buf][i][x] = __builtin_fma(k[8], du, k[7]);
        if (x == 0) cub.c3[buf][i] = k[9];
    }
}

// Double-float form, second stage: the f64 C0..C2 of fill_grid_columns, split and rounded into
// the GridCubicDF layout (C0 - kRowMagic is exact: both lie in [2^20, 2^21)).
template <int B>
__device__ inline void fill_grid_columns_df(const BpArgs &a, GridCubicDF<B> &cub, int blk) {
    const int n = min(B, a.np - blk * B), buf = blk & 1;
    for (int job = threadIdx.x; job < 16 * n; job += kBpThreads) {
        const int i = job >> 4, x = job & 15;
        const double *k = cub.c[buf][i];
        const double du = ((double)x - 7.5) * a.g.d_u;
        const double c0 = __builtin_fma(__builtin_fma(__builtin_fma(k[3], du, k[2]), du, k[1]), du, k[0]);
        const double c1 = __builtin_fma(__builtin_fma(k[6], du, k[5]), du, k[4]);
        const double t0 = c0 - kRowMagic, hf = floor(t0);
        cub.h[buf][i][x] = (int)hf;
        cub.l[buf][i][x] = (float)(t0 - hf);
        cub.c1[buf][i][x] = (float)c1;
        cub.c2[buf][i][x] = (float)__builtin_fma(k[8], du, k[7]);
        if (x == 0) cub.c3[buf][i] = (float)k[9];
    }
}

// Window path, second stage: C0..C3 of block blk's pulses for the tile's 16 columns, from the
// group ring.
__device__ inline void fill_win_columns(const BpArgs &a, WinCubic &wc, int blk) {
    const int n = min(kWinBlock, a.np - blk * kWinBlock), buf = blk & 1;
    const int g = blk / (kWinGroup / kWinBlock), i0 = (blk % (kWinGroup / kWinBlock)) * kWinBlock;
    for (int job = threadIdx.x; job < 16 * n; job += kBpThreads) {
        const int i = job >> 4, x = job & 15;
        const double *k = wc.c[g & 1][i0 + i];
        const double du = ((double)x - 7.5) * a.g.d_u;
        const double c0 = __builtin_fma(__builtin_fma(__builtin_fma(k[3], du, k[2]), du, k[1]), du, k[0]);
        const double c1 = __builtin_fma(__builtin_fma(k[6], du, k[5]), du, k[4]);
        wc.c01[buf][i][x] = make_double2(c0, c1);
        wc.c2[buf][i][x] = __builtin_fma(k[8], du, k[7]);
        if (x == 0) wc.c3[buf][i] = k[9];
    }
}

// Window path: copies block blk's windows into its LDS buffer with global_load_lds: slot s's
// window is elements [S s, S s + W) of the two buffers, loaded by ceil(W / 64) wave instructions
// (one per 64 elements; each lane writes the element at its lane index past the instruction's
// LDS address). The source of an instruction is the row element j0 + 64 c (a wave-uniform address)
// plus the lane. Asynchronous: the caller waits for the wave's loads (vmcnt) before the barrier
// that publishes them.
typedef __attribute__((address_space(3))) void lds_void;
template <int S>
__device__ inline void load_windows(const BpArgs &a, const WinTile &win, const WinCubic &wc,
                                    QPairF32 *buf, int blk) {
    // every value but the lane's is uniform: loop and addresses on the scalar unit
    const int W = __builtin_amdgcn_readfirstlane(win.W), bk = __builtin_amdgcn_readfirstlane(blk);
    const int len = __builtin_amdgcn_readfirstlane(a.len);
    const int n = min(kWinBlock, a.np - bk * kWinBlock);
    const int chunks = (W + 63) >> 6;  // 1 or 2 (W <= S <= 128)
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = bk / (kWinGroup / kWinBlock), i0 = (bk % (kWinGroup / kWinBlock)) * kWinBlock;
    const QPairF32 *q = reinterpret_cast<const QPairF32 *>(a.q);
    // lanes that load in a slot's first and second chunk (the second holds elements 64 .. W - 1)
    const bool in0 = lane < W, in1 = lane + 64 < W;
    for (int m = wave; m < n * chunks; m += kBpThreads / 64) {
        const int sl = chunks == 2 ? m >> 1 : m, c = m - sl * chunks;
        const int j0 = __builtin_amdgcn_readfirstlane(wc.j0[g & 1][i0 + sl]);
        const QPairF32 *row = q + (size_t)(bk * kWinBlock + sl) * len + (j0 + 64 * c);
        QPairF32 *dst = buf + ((bk & 1) * kWinBlock + sl) * S + 64 * c;
        if (c == 0 ? in0 : in1) __builtin_amdgcn_global_load_lds(row + lane, (lds_void *)dst, 16, 0, 0);
    }
}

// One pulse's expansion constants from the LDS block (the address is the same in every lane, so
// each read is one broadcast). Issuing these reads earlier in the pipelined loop measured no faster.
struct ExpConst {
    double T, n0, n1, n2, h, g;
};
__device__ inline ExpConst read_expansion(const double *kp) {
    ExpConst k;
    k.T = kp[0];
    k.n0 = kp[1];
    k.n1 = kp[2];
    k.n2 = kp[3];
    k.h = kp[4];
    k.g = kp[5];
    return k;
}

// Index and weight of the thread's pixels from y = t + kRowMagic, then their loads.
template <typename QT, int PPT>
__device__ inline void gather_fixed(const BpArgs &a, const double (&y)[PPT], int p, Gather<PPT> &g) {
    constexpr bool kScaled = sizeof(QT) == sizeof(QPairF16);
    const QT *qbase = reinterpret_cast<const QT *>(a.q);
    const __amdgpu_buffer_rsrc_t row = row_rsrc(qbase + (size_t)p * a.len, a.len, (int)sizeof(QT));
    g.qs = kScaled ? a.qscale[p] : 1.0f;
    int ii[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const u32x2 b = __builtin_bit_cast(u32x2, y[j]);
        ii[j] = (int)(b.y & 0x7FFFFu);                        // floor(t)
        g.ww[j] = (float)b.x * (0x1p-32f * weight_unit<QT>());  // fract(t) in weight units
    }
#pragma unroll
    for (int j = 0; j < PPT; ++j) g.e[j] = loadq(qbase, row, ii[j]);
}

template <typename QT, int PPT, bool GRID>
__device__ inline void gather_expanded(const BpArgs &a, const ExpTerms<PPT> &E, const ExpConst &k,
                                       int p, Gather<PPT> &g) {
    double y[PPT];
    const double au = GRID ? k.n0 * E.e0[0] : 0.0;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const double pr =
            GRID ? __builtin_fma(k.n1, E.e1[j], au)
                 : __builtin_fma(k.n0, E.e0[j], __builtin_fma(k.n1, E.e1[j], k.n2 * E.e2[j]));
        const double q = __builtin_fma(-pr, pr, E.dd[j]);
        const double m = __builtin_fma(k.g, pr, k.h);
        y[j] = __builtin_fma(-q, m, k.T - pr);  // t + kRowMagic
    }
    gather_fixed<QT, PPT>(a, y, p, g);
}

// Grid mode: y = C0 + dv (C1 + dv (C2 + dv C3)) for the thread's pixels, with the pulse's C0..C3 of
// the thread's column.
template <typename QT, int PPT>
__device__ inline void gather_grid(const BpArgs &a, const double (&dv)[PPT], const double2 &c01,
                                   double c2, double c3, int p, Gather<PPT> &g) {
    double y[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j)
        y[j] = __builtin_fma(__builtin_fma(__builtin_fma(c3, dv[j], c2), dv[j], c01.y), dv[j], c01.x);
    gather_fixed<QT, PPT>(a, y, p, g);
}

// Double-float form: s = L + dv (C1 + dv (C2 + dv C3)) in f32 is the row coordinate minus the
// column's integer part H; index H + floor(s), weight fract(s) in the rows' weight unit.
template <typename QT, int PPT>
__device__ inline void gather_grid_df(const BpArgs &a, const float (&dv)[PPT], int h, float l,
                                      float c1, float c2, float c3, int p, Gather<PPT> &g) {
    constexpr bool kScaled = sizeof(QT) == sizeof(QPairF16);
    const QT *qbase = reinterpret_cast<const QT *>(a.q);
    const __amdgpu_buffer_rsrc_t row = row_rsrc(qbase + (size_t)p * a.len, a.len, (int)sizeof(QT));
    g.qs = kScaled ? a.qscale[p] : 1.0f;
    int ii[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const float sv =
            __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(c3, dv[j], c2), dv[j], c1), dv[j], l);
        int fl;
        asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(fl) : "v"(sv));  // floor(sv)
        ii[j] = h + fl;  // a negative index reads zeros (buffer bounds check), as in gather_fixed
        g.ww[j] = __builtin_amdgcn_fractf(sv) * weight_unit<QT>();
    }
#pragma unroll
    for (int j = 0; j < PPT; ++j) g.e[j] = loadq(qbase, row, ii[j]);
}

// The same with the pulse's row window in LDS: y = kRowMagic + (the pixel's element index in the
// window buffers) + fract, so the high word of y is 0x41380000 + that index, and (hi << 4) + K with
// K = (LDS address of the buffers) - 0x13800000 (mod 2^32) is the element's LDS address: one
// shift-add per pixel in place of the index mask.
typedef __attribute__((address_space(3))) const f32x4 lds_f32x4;
template <int PPT>
__device__ inline void gather_grid_win(const double (&dv)[PPT], const double2 &c01, double c2,
                                       double c3, unsigned K, Gather<PPT> &g) {
    double y[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j)
        y[j] = __builtin_fma(__builtin_fma(__builtin_fma(c3, dv[j], c2), dv[j], c01.y), dv[j], c01.x);
    unsigned addr[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const u32x2 b = __builtin_bit_cast(u32x2, y[j]);
        addr[j] = (b.y << 4) + K;
        g.ww[j] = (float)b.x;  // fract(t) * 2^32: the weight unit of f32 rows
    }
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const f32x4 v = *(lds_f32x4 *)(size_t)addr[j];
        g.e[j] = make_float4(v.x, v.y, v.z, v.w);
    }
}

// The fast pulse loop over the cubic (grid mode), row elements gathered from HBM: blocks of B
// pulses; in the period of block b the threads build block b + 2's c_ij and block b + 1's C0..C3,
// then consume block b's (pipelined as in accumulate_pulses; flush after every kBpPulseBlock pulses
// of the whole sequence). Each write goes to the buffer of the block two back, whose last reader
// ran in an earlier period: c_ij of block b + 2 replace those of b (read by the second stage in
// period b - 1), C of b + 1 those of b - 1 (consumed in period b - 1).
template <typename QT, int PPT, int B>
__device__ inline void accumulate_pulses_grid(const BpArgs &a, const double (&dv)[PPT], int col,
                                              const double *rc, GridCubicB<B> &cub,
                                              double (&accr)[PPT], double (&acci)[PPT]) {
    const int np = a.np;
    const float kappa_f = (float)a.hdr->kappa / weight_unit<QT>();
    float sr[PPT], si[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) sr[j] = si[j] = 0.f;
    auto flush = [&] {
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            accr[j] += (double)sr[j];
            acci[j] += (double)si[j];
            sr[j] = si[j] = 0.f;
        }
    };
    const int nblk = (np + B - 1) / B;
    fill_grid_cubic<B>(a, rc, cub, 0);
    if (nblk > 1) fill_grid_cubic<B>(a, rc, cub, 1);
    __syncthreads();
    fill_grid_columns<B>(a, cub, 0);
    for (int b = 0; b < nblk; ++b) {
        __syncthreads();
        if (b + 2 < nblk) fill_grid_cubic<B>(a, rc, cub, b + 2);
        if (b + 1 < nblk) fill_grid_columns<B>(a, cub, b + 1);
        const int buf = b & 1, p0 = b * B, p1 = min(np, p0 + B);
        const double2 *c01 = &cub.c01[buf][0][col];
        const double *c2 = &cub.c2[buf][0][col], *c3 = cub.c3[buf];
        Gather<PPT> g0, g1;
        gather_grid<QT, PPT>(a, dv, c01[0], c2[0], c3[0], p0, g0);
        for (int p = p0; p + 1 < p1; p += 2) {
            const int i1 = p + 1 - p0;
            gather_grid<QT, PPT>(a, dv, c01[16 * i1], c2[16 * i1], c3[i1], p + 1, g1);
            add_terms<QT, PPT, true>(g0, kappa_f, sr, si);
            const int pn = min(p + 2, p1 - 1), i2 = pn - p0;
            gather_grid<QT, PPT>(a, dv, c01[16 * i2], c2[16 * i2], c3[i2], pn, g0);
            add_terms<QT, PPT, true>(g1, kappa_f, sr, si);
            if (((p + 2) & (kBpPulseBlock - 1)) == 0) flush();
        }
        if ((p1 - p0) & 1) add_terms<QT, PPT, true>(g0, kappa_f, sr, si);  // pulse p1 - 1
    }
    flush();
}

// The loop of accumulate_pulses_grid in the double-float form (same blocks and pipelining).
template <typename QT, int PPT, int B>
__device__ inline void accumulate_pulses_grid_df(const BpArgs &a, const float (&dv)[PPT], int col,
                                                 const double *rc, GridCubicDF<B> &cub,
                                                 double (&accr)[PPT], double (&acci)[PPT]) {
    const int np = a.np;
    const float kappa_f = (float)a.hdr->kappa / weight_unit<QT>();
    float sr[PPT], si[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) sr[j] = si[j] = 0.f;
    auto flush = [&] {
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            accr[j] += (double)sr[j];
            acci[j] += (double)si[j];
            sr[j] = si[j] = 0.f;
        }
    };
    const int nblk = (np + B - 1) / B;
    fill_grid_cubic<B>(a, rc, cub, 0);
    if (nblk > 1) fill_grid_cubic<B>(a, rc, cub, 1);
    __syncthreads();
    fill_grid_columns_df<B>(a, cub, 0);
    for (int b = 0; b < nblk; ++b) {
        __syncthreads();
        if (b + 2 < nblk) fill_grid_cubic<B>(a, rc, cub, b + 2);
        if (b + 1 < nblk) fill_grid_columns_df<B>(a, cub, b + 1);
        const int buf = b & 1, p0 = b * B, p1 = min(np, p0 + B);
        const int *h = &cub.h[buf][0][col];
        const float *l = &cub.l[buf][0][col], *c1 = &cub.c1[buf][0][col];
        const float *c2 = &cub.c2[buf][0][col], *c3 = cub.c3[buf];
        Gather<PPT> g0, g1;
        gather_grid_df<QT, PPT>(a, dv, h[0], l[0], c1[0], c2[0], c3[0], p0, g0);
        for (int p = p0; p + 1 < p1; p += 2) {
            const int i1 = 16 * (p + 1 - p0);
            gather_grid_df<QT, PPT>(a, dv, h[i1], l[i1], c1[i1], c2[i1], c3[p + 1 - p0], p + 1, g1);
            add_terms<QT, PPT, true>(g0, kappa_f, sr, si);
            const int pn = min(p + 2, p1 - 1), i2 = 16 * (pn - p0);
            gather_grid_df<QT, PPT>(a, dv, h[i2], l[i2], c1[i2], c2[i2], c3[pn - p0], pn, g0);
            add_terms<QT, PPT, true>(g1, kappa_f, sr, si);
            if (((p + 2) & (kBpPulseBlock - 1)) == 0) flush();
        }
        if ((p1 - p0) & 1) add_terms<QT, PPT, true>(g0, kappa_f, sr, si);  // pulse p1 - 1
    }
    flush();
}

// The same loop reading row windows from LDS (f32 rows), in blocks of kWinBlock pulses. In the
// period of block b the threads build the c_ij of group (b + 2) / 4 when b = 2 mod 4 (one wave),
// block b + 1's C0..C3 and block b + 1's windows, then consume block b. Group g's ring slot was last
// read in period 4 g - 6 (its columns and windows of blocks up to 4 g - 5), and C and windows of
// b + 1 replace those of b - 1, consumed in period b - 1. Window loads are waited for (vmcnt)
// before the barrier that opens the period reading them.
template <int PPT, int S>
__device__ inline void accumulate_pulses_win(const BpArgs &a, const double (&dv)[PPT], int col,
                                             const double *rc, WinCubic &wc, QPairF32 *wbuf,
                                             const WinTile &win, double (&accr)[PPT],
                                             double (&acci)[PPT]) {
    constexpr int B = kWinBlock, BPG = kWinGroup / kWinBlock;
    const int np = a.np;
    const float kappa_f = (float)a.hdr->kappa / weight_unit<QPairF32>();
    float sr[PPT], si[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) sr[j] = si[j] = 0.f;
    auto flush = [&] {
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            accr[j] += (double)sr[j];
            acci[j] += (double)si[j];
            sr[j] = si[j] = 0.f;
        }
    };
    const unsigned K = (unsigned)(size_t)(lds_void *)wbuf - 0x13800000u;
    const int nblk = (np + B - 1) / B;
    fill_win_group<S>(a, rc, wc, 0, win);
    __syncthreads();
    fill_win_columns(a, wc, 0);
    load_windows<S>(a, win, wc, wbuf, 0);
    for (int b = 0; b < nblk; ++b) {
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's window loads have landed
        __syncthreads();
        if (b % BPG == BPG - 2 && (b + 2) * B < np) fill_win_group<S>(a, rc, wc, (b + 2) / BPG, win);
        if (b + 1 < nblk) {
            fill_win_columns(a, wc, b + 1);
            load_windows<S>(a, win, wc, wbuf, b + 1);
        }
        const int buf = b & 1, p0 = b * B, p1 = min(np, p0 + B);
        const double2 *c01 = &wc.c01[buf][0][col];
        const double *c2 = &wc.c2[buf][0][col], *c3 = wc.c3[buf];
        Gather<PPT> g0, g1;
        gather_grid_win<PPT>(dv, c01[0], c2[0], c3[0], K, g0);
        for (int p = p0; p + 1 < p1; p += 2) {
            const int i1 = p + 1 - p0;
            gather_grid_win<PPT>(dv, c01[16 * i1], c2[16 * i1], c3[i1], K, g1);
            add_terms<QPairF32, PPT, true>(g0, kappa_f, sr, si);
            const int i2 = min(p + 2, p1 - 1) - p0;
            gather_grid_win<PPT>(dv, c01[16 * i2], c2[16 * i2], c3[i2], K, g0);
            add_terms<QPairF32, PPT, true>(g1, kappa_f, sr, si);
            if (((p + 2) & (kBpPulseBlock - 1)) == 0) flush();
        }
        if ((p1 - p0) & 1) add_terms<QPairF32, PPT, true>(g0, kappa_f, sr, si);  // pulse p1 - 1
    }
    flush();
}

// The fast pulse loop over the expansion in the form above: per block of kExpBlock pulses a barrier,
// the next block's constants, then the block's pulses pipelined as in accumulate_pulses (flush after
// every kBpPulseBlock pulses of the whole sequence).
template <typename QT, int PPT, bool GRID, class SH>
__device__ inline void accumulate_pulses_expanded(const BpArgs &a, const ExpTerms<PPT> &E, SH &sh,
                                                  double (&accr)[PPT], double (&acci)[PPT]) {
    const int np = a.np;
    const float kappa_f = (float)a.hdr->kappa / weight_unit<QT>();
    float sr[PPT], si[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) sr[j] = si[j] = 0.f;
    auto flush = [&] {
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            accr[j] += (double)sr[j];
            acci[j] += (double)si[j];
            sr[j] = si[j] = 0.f;
        }
    };
    const int nblk = (np + kExpBlock - 1) / kExpBlock;
    fill_expansion<GRID, PPT>(a, sh, 0, 0);
    for (int b = 0; b < nblk; ++b) {
        __syncthreads();  // block b's constants written; block b - 1's buffer no longer read
        if (b + 1 < nblk) fill_expansion<GRID, PPT>(a, sh, b + 1, (b + 1) & 1);
        const double *kb = &sh.x.k[b & 1][0][0];
        const int p0 = b * kExpBlock, p1 = min(np, p0 + kExpBlock);
        Gather<PPT> g0, g1;
        gather_expanded<QT, PPT, GRID>(a, E, read_expansion(kb), p0, g0);
        for (int p = p0; p + 1 < p1; p += 2) {
            gather_expanded<QT, PPT, GRID>(a, E, read_expansion(kb + 6 * (p + 1 - p0)), p + 1, g1);
            add_terms<QT, PPT, true>(g0, kappa_f, sr, si);
            const int pn = min(p + 2, p1 - 1);
            gather_expanded<QT, PPT, GRID>(a, E, read_expansion(kb + 6 * (pn - p0)), pn, g0);
            add_terms<QT, PPT, true>(g1, kappa_f, sr, si);
            if (((p + 2) & (kBpPulseBlock - 1)) == 0) flush();
        }
        if ((p1 - p0) & 1) add_terms<QT, PPT, true>(g0, kappa_f, sr, si);  // pulse p1 - 1
    }
    flush();
}

// One 16 x (16 * PPT) tile whose top-left output pixel is (bx, by) relative to the region.
// Location of output pixel (lx, ly) of the region: from the grid, or read from r (pixel index
// xy_to_pixel = nu * (nv - y - 1) + x, backprojection.cpp:130; r is constant_exterior 0).
template <bool GRID>
__device__ inline void pixel_loc(const BpArgs &a, int lx, int ly, bool ok, double &rx, double &ry,
                                 double &rz, double &u, double &v) {
    const int x = a.x0 + lx, y = a.y0 + ly;
    u = v = 0.0;
    if (GRID) {
        // grid regions lie inside the image (sarbp_plan_set_grid checks); lanes past the region's
        // edge take location 0, as with r
        rx = ry = rz = 0.0;
        if (!ok) return;
        u = (double)(a.g.u0 + x) * a.g.d_u;
        v = (double)(a.g.v0 + (a.nv - y - 1)) * a.g.d_v;
        rx = grid_coord(a.g.uh[0], u, a.g.vh[0], v);
        ry = grid_coord(a.g.uh[1], u, a.g.vh[1], v);
        rz = grid_coord(a.g.uh[2], u, a.g.vh[2], v);
        return;
    }
    const int64_t px = (int64_t)a.nu * (int64_t)(a.nv - y - 1) + x;
    const int64_t idx = px - a.r_min;
    rx = ry = rz = 0.0;
    if (ok && idx >= 0 && idx < a.r_extent) {
        rx = a.r[idx];
        ry = a.r[a.r_extent + idx];
        rz = a.r[2 * a.r_extent + idx];
    }
}

// A copy of x that the compiler cannot relate to x, so that values computed from it are not merged
// with (and kept live from) earlier computations from x.
__device__ inline int opaque(int x) {
    asm volatile("" : "+v"(x));
    return x;
}

// Per-pixel terms of the direct form (PixTerms) for thread (tx, ty) of the tile at (bx, by).
template <bool GRID, int PPT>
__device__ inline void direct_terms(const BpArgs &a, int bx, int by, int tx, int ty,
                                    PixTerms<PPT> &P) {
    const double sigma = a.hdr->inv_lsr * a.hdr->inv_lsr;
    bool ok0 = true;
    double v0 = 0.0;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const int lx = bx + tx, ly = by + ty + 16 * j;
        const bool ok = lx < a.width && ly < a.height;
        if (j == 0) ok0 = ok;
        double rx, ry, rz, u, v;
        pixel_loc<GRID>(a, lx, ly, ok, rx, ry, rz, u, v);
        const double r2s = sigma * (rx * rx + ry * ry + rz * rz);
        if (GRID) {
            if (j == 0) {
                P.a[0] = u;
                P.b[0] = v;
                P.r2[0] = r2s;
                v0 = v;
            }
            // a pixel outside the region computes pixel 0's range
            P.c[j] = ok && ok0 ? v - v0 : 0.0;
            P.dr2[j] = ok && ok0 ? r2s - P.r2[0] : 0.0;
        } else {
            P.a[j] = (-2.0 * sigma) * rx;
            P.b[j] = (-2.0 * sigma) * ry;
            P.c[j] = (-2.0 * sigma) * rz;
            P.r2[j] = r2s;
        }
    }
}

// Per-pixel terms of the expansion about rc (ExpTerms). A pixel outside the region takes pixel 0's
// offsets (grid mode), or those of the tile's first pixel (bx, by), which is inside it.
template <bool GRID, int PPT>
__device__ inline void expansion_terms(const BpArgs &a, int bx, int by, int tx, int ty,
                                       const double (&rc)[3], ExpTerms<PPT> &E) {
    const double inv_lsr = a.hdr->inv_lsr;
    const bool ok0 = bx + tx < a.width && by + ty < a.height;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const int lx = bx + tx, ly = by + ty + 16 * j;
        const bool ok = lx < a.width && ly < a.height;
        if (GRID) {
            const double ox = ok0 ? (double)tx - 7.5 : -7.5;
            const double oy = ok0 ? (double)(ty + 16 * (ok ? j : 0)) : 0.0;
            const double du = ox * a.g.d_u, dv = ((8.0 * PPT - 0.5) - oy) * a.g.d_v;
            if (j == 0) E.e0[0] = du;
            E.e1[j] = dv;
            double s2 = 0.0;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const double c = a.g.uh[i] * du + a.g.vh[i] * dv;
                s2 += c * c;
            }
            E.dd[j] = inv_lsr * inv_lsr * s2;
        } else {
            double rx, ry, rz, u, v;
            if (ok)
                pixel_loc<GRID>(a, lx, ly, true, rx, ry, rz, u, v);
            else
                pixel_loc<GRID>(a, bx, by, true, rx, ry, rz, u, v);
            E.e0[j] = (rx - rc[0]) * inv_lsr;
            E.e1[j] = (ry - rc[1]) * inv_lsr;
            E.e2[j] = (rz - rc[2]) * inv_lsr;
            E.dd[j] = E.e0[j] * E.e0[j] + E.e1[j] * E.e1[j] + E.e2[j] * E.e2[j];
        }
    }
}

// Grid mode: the thread's pixel offsets dv from the centre (metres along v_hat) and its column of
// the cubic's coefficients. As in expansion_terms, a pixel outside the region takes pixel 0's row,
// and a thread whose pixel 0 is outside it takes the tile's first pixel (column 0, row 0).
template <int PPT>
__device__ inline void grid_offsets(const BpArgs &a, int bx, int by, int tx, int ty, double (&dv)[PPT],
                                    int &col) {
    const bool ok0 = bx + tx < a.width && by + ty < a.height;
    col = ok0 ? tx : 0;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const bool ok = by + ty + 16 * j < a.height;
        const double oy = ok0 ? (double)(ty + 16 * (ok ? j : 0)) : 0.0;
        dv[j] = ((8.0 * PPT - 0.5) - oy) * a.g.d_v;
    }
}

template <typename QT, int PPT, bool GRID, bool WIN = false>
__device__ inline void accumulate_tile_impl(const BpArgs &a, int bx, int by,
                                            TileShared<GRID, PPT, WIN> &sh, bool cubic) {
    static_assert(!WIN || (GRID && sizeof(QT) == sizeof(QPairF32) && kWinSlot<PPT> > 0),
                  "row windows: grid mode, f32 rows, tile heights with a window slot");
    const double inv_lsr = a.hdr->inv_lsr;
    // Wave w holds rows 4w..4w+3 (+16 j) of the tile's 16 columns. The memory pipeline serves a
    // gather four lanes at a time and its cost grows with the cache lines each quad of lanes
    // touches: a quad of 4 columns spans 4 g_x range bins (g_x: bins per pixel along x). Where
    // g_x >= 2.4 and the range changes less along y, a quad takes 4 rows of one column instead
    // (lane = 4 x + row). Below that, 4 columns of one row are cheaper: there the 4 lanes of a
    // column quad would mostly read one element. Measured on the C2 geometry at 1/4 to 2 times its
    // pixel spacing (0.95 to 7.6 bins per pixel): columns are faster up to 1.9 bins per pixel, rows
    // from 2.85. g_x and g_y are taken at the tile's first pixel for the mid pulse.
    bool y_quads;
    {
        const PulseConst cm = a.pc[a.np / 2];
        const int x1 = bx + 1 < a.width ? bx + 1 : bx - 1, y1 = by + 1 < a.height ? by + 1 : by - 1;
        double r[3][3], u, v;
        pixel_loc<GRID>(a, bx, by, true, r[0][0], r[0][1], r[0][2], u, v);
        pixel_loc<GRID>(a, x1 >= 0 ? x1 : bx, by, true, r[1][0], r[1][1], r[1][2], u, v);
        pixel_loc<GRID>(a, bx, y1 >= 0 ? y1 : by, true, r[2][0], r[2][1], r[2][2], u, v);
        double d[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double dx = r[i][0] - cm.px, dy = r[i][1] - cm.py, dz = r[i][2] - cm.pz;
            d[i] = sqrt(dx * dx + dy * dy + dz * dz);
        }
        const double gx = fabs(d[1] - d[0]) * inv_lsr, gy = fabs(d[2] - d[0]) * inv_lsr;
        y_quads = gx >= 2.4 && gx > gy;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tx = y_quads ? lane >> 2 : lane & 15;
    const int ty = 4 * wave + (y_quads ? lane & 3 : lane >> 4);

    // Centre of the expansion: the middle of the tile's full 16 x (16 PPT) footprint (grid mode; it
    // may lie past the region's edge), or the location of the tile pixel nearest to it (pixel_locs).
    double rc[3];
    if (GRID) {
        const double uc = ((double)(a.g.u0 + a.x0 + bx) + 7.5) * a.g.d_u;
        const double vc = ((double)(a.g.v0 + a.nv - (a.y0 + by) - 1) - (8.0 * PPT - 0.5)) * a.g.d_v;
#pragma unroll
        for (int i = 0; i < 3; ++i) rc[i] = a.g.uh[i] * uc + a.g.vh[i] * vc;
    } else {
        double u, v;
        pixel_loc<GRID>(a, min(bx + 8, a.width - 1), min(by + 8 * PPT, a.height - 1), true, rc[0],
                        rc[1], rc[2], u, v);
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < 3; ++i) sh.rc[i] = rc[i];
    }

    // Which loop: pixels beyond the rows' cover radius need the exact one; otherwise the expansion
    // where its error bound holds, else the direct fast loop.
    bool far = false;
    double dd_max = 0.0;
    {
        ExpTerms<PPT> E;
        expansion_terms<GRID>(a, bx, by, tx, ty, rc, E);
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            const int lx = bx + tx, ly = by + ty + 16 * j;
            double rx, ry, rz, u, v;
            pixel_loc<GRID>(a, lx, ly, lx < a.width && ly < a.height, rx, ry, rz, u, v);
            far = far || !(rx * rx + ry * ry + rz * rz <= a.r_cover2);
            dd_max = fmax(dd_max, E.dd[j]);
        }
    }
    for (int off = 32; off > 0; off >>= 1) dd_max = fmax(dd_max, __shfl_xor(dd_max, off, 64));
    const bool wave_far = __ballot(far) != 0;
    if (lane == 0) {
        sh.far[wave] = wave_far;
        sh.dd_max[wave] = dd_max;
    }
    __syncthreads();
    bool any_far = false;
#pragma unroll
    for (int w = 0; w < kBpThreads / 64; ++w) {
        any_far = any_far || sh.far[w] != 0;
        dd_max = fmax(dd_max, sh.dd_max[w]);
    }
    // the bound needs the least distance from the centre to the platform; a cubic tile of a window
    // kernel also needs the largest half-span of its row coordinates over the pulses
    bool expand = false;
    WinTile wt{};
    const bool want_win = WIN && cubic && !a.range_df;
    if (WIN) {
        wt.U = 7.5 * fabs(a.g.d_u);
        wt.V = (8.0 * PPT - 0.5) * fabs(a.g.d_v);
        wt.len = a.len;
    }
    if (!any_far) {
        double d2 = INFINITY, hm = 0.0;
        for (int p = threadIdx.x; p < a.np; p += kBpThreads) {
            const double dx = rc[0] - a.pc[p].px, dy = rc[1] - a.pc[p].py, dz = rc[2] - a.pc[p].pz;
            d2 = fmin(d2, dx * dx + dy * dy + dz * dz);
            if (want_win) {
                double k[10];
                cubic_coeffs(a, rc, p, k);
                hm = fmax(hm, cubic_half_span(k, wt.U, wt.V));
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            d2 = fmin(d2, __shfl_xor(d2, off, 64));
            hm = fmax(hm, __shfl_xor(hm, off, 64));
        }
        if (lane == 0) {
            sh.d2_min[wave] = d2;
            sh.h_max[wave] = hm;
        }
        __syncthreads();
#pragma unroll
        for (int w = 0; w < kBpThreads / 64; ++w) {
            d2 = fmin(d2, sh.d2_min[w]);
            hm = fmax(hm, sh.h_max[w]);
        }
        const double dm = sqrt(inv_lsr * inv_lsr * d2);  // row units
        expand = dm > 0.0 && dd_max <= 1e-4 * dm * dm &&
                 dd_max * dd_max <= 4.0 * kExpandTol * dm * dm * dm;
        if (want_win && hm < 0.5 * kWinSlot<PPT>) {
            // every floor(t) of a pulse lies in [floor(t_c - H), floor(t_c + H)]: 2 H + 2 elements
            const int W = min((int)floor(2.0 * hm) + 2, a.len);
            wt.W = __builtin_amdgcn_readfirstlane(W <= kWinSlot<PPT> ? W : 0);
        }
    }

    // The decision is the same in every lane (it comes from LDS); readfirstlane tells the compiler,
    // so that the loops become plain branches. Each branch rebuilds its per-pixel terms from an
    // opaque copy of the lane's coordinates, so that nothing computed above stays live through a
    // pulse loop that does not use it.
    const int path = __builtin_amdgcn_readfirstlane(any_far ? 0 : expand ? 1 : 2);
    if (a.paths && threadIdx.x == 0) {
        const int k = path == 0   ? kPathExact
                      : path == 2 ? kPathDirect
                      : !(GRID && cubic) ? kPathPQ
                      : a.range_df ? kPathCubicDF
                      : wt.W > 0  ? kPathCubicWindow
                                  : kPathCubicGather;
        atomicAdd(&a.paths[k], 1u);
    }
    double accr[PPT], acci[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) accr[j] = acci[j] = 0.0;
    if (path == 1 && GRID && cubic) {
        if constexpr (GRID) {
            double dv[PPT];
            int col;
            grid_offsets<PPT>(a, bx, by, opaque(tx), opaque(ty), dv, col);
            if (a.range_df) {
                float dvf[PPT];
#pragma unroll
                for (int j = 0; j < PPT; ++j) dvf[j] = (float)dv[j];
                accumulate_pulses_grid_df<QT, PPT, kGridBlock<PPT>>(a, dvf, col, sh.rc, sh.x.cdf,
                                                                    accr, acci);
            } else if constexpr (WIN) {
                if (wt.W > 0)
                    accumulate_pulses_win<PPT, kWinSlot<PPT>>(a, dv, col, sh.rc, sh.x.w.wc,
                                                              sh.x.w.e, wt, accr, acci);
                else
                    accumulate_pulses_grid<QT, PPT, kGridBlock<PPT>>(a, dv, col, sh.rc, sh.x.cub,
                                                                     accr, acci);
            } else {
                accumulate_pulses_grid<QT, PPT, kGridBlock<PPT>>(a, dv, col, sh.rc, sh.x.cub,
                                                                 accr, acci);
            }
        }
    } else if (path == 1) {
        ExpTerms<PPT> E;
        const double rcs[3] = {sh.rc[0], sh.rc[1], sh.rc[2]};
        expansion_terms<GRID>(a, bx, by, opaque(tx), opaque(ty), rcs, E);
        accumulate_pulses_expanded<QT, PPT, GRID>(a, E, sh, accr, acci);
    } else {
        PixTerms<PPT> P;
        direct_terms<GRID>(a, bx, by, opaque(tx), opaque(ty), P);
        if (path == 0)
            accumulate_pulses<QT, PPT, false, GRID>(a, P, accr, acci);
        else
            accumulate_pulses<QT, PPT, true, GRID>(a, P, accr, acci);
    }

    // fimg = img * expj(k_c * dr_i(pixel, npulses / 2)) (backprojection.cpp:135)
    const PulseConst cm = a.pc[a.np / 2];
    const double k_c = a.hdr->k_c;
    const int txo = opaque(tx), tyo = opaque(ty);
    double zmax2 = 0.0;  // largest |z|^2 the thread writes
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const int lx = bx + txo, ly = by + tyo + 16 * j;
        if (lx >= a.width || ly >= a.height) continue;
        double rx, ry, rz, u, v;
        pixel_loc<GRID>(a, lx, ly, true, rx, ry, rz, u, v);
        const double dx = rx - cm.px, dy = ry - cm.py, dz = rz - cm.pz;
        const double dm = sqrt(dx * dx + dy * dy + dz * dz);
        const double ang = k_c * (cm.nr0 - dm);
        double sn, cs;
        sincos(ang, &sn, &cs);
        const double2 z = make_double2(accr[j] * cs - acci[j] * sn, accr[j] * sn + acci[j] * cs);
        a.out[(size_t)ly * a.width + lx] = z;
        const double m2 = abs2(z);
        zmax2 = m2 > zmax2 ? m2 : zmax2;
    }
    if (a.absmax) {  // the same for every lane of the launch
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(zmax2, off, 64);
            zmax2 = o > zmax2 ? o : zmax2;
        }
        if ((threadIdx.x & 63) == 0) raise_absmax(a.absmax, zmax2);
    }
}

// Waves per SIMD the f32-row kernels are built for (at most 64, 96 and 128 VGPRs): the register
// budget of a whole kernel is set by its largest loop, and the exact loop for tiles beyond the rows'
// cover, which almost never runs, must not take occupancy from the others (left to itself it took
// 160 VGPRs at 4 pixels per thread). Its spills stay in that loop; the expansion and direct fast
// loops have none. The f16-row kernels keep the compiler's own allocation (1: no minimum), except
// the grid kernel at 4 pixels per thread, which fits 128 VGPRs without spills when asked to (left
// to itself it took 129, and one wave less per SIMD).
template <typename QT, bool GRID>
constexpr int waves_per_simd(int ppt) {
    return sizeof(QT) == sizeof(QPairF16) ? (GRID && ppt == 4 ? 4 : 1)
                                          : ppt <= 2 ? 8 : ppt == 3 ? 5 : 4;
}

// Grid mode is a kernel template parameter (not a branch in one kernel): each variant gets its
// own register allocation, so the grid kernels keep the occupancy their smaller state allows.
// One workgroup per tile (a 1-D grid of all the tiles), each XCD taking a band of columns or of rows
// (xcd_band_tile): columns when the range changes more across the region's width than across its
// height. Grid mode measures that with the mid pulse's U and V (the gradient of the squared range
// along u and v at the scene origin, where image planes are centred); pixel_locs launches use
// columns, the orientation of the reference's image plane (u_hat along range). Every workgroup of
// a launch takes the same decision, so the order stays a bijection. a.cubic (grid mode): expansion
// tiles take the cubic form (accumulate_pulses_grid) rather than the p, q form.
// WIN: cubic tiles read their rows from windows in LDS where the window fits (see WinShared).
#ifndef SARBP_WIN_WAVES
#define SARBP_WIN_WAVES 0
#endif
template <typename QT, int PPT, bool GRID, bool WIN = false>
__global__ __launch_bounds__(kBpThreads) __attribute__((amdgpu_waves_per_eu(
    WIN && SARBP_WIN_WAVES > 0 ? SARBP_WIN_WAVES : waves_per_simd<QT, GRID>(PPT))))
void accumulate_kernel(BpArgs a) {
    __shared__ TileShared<GRID, PPT, WIN> sh;
    constexpr int th = 16 * PPT;
    const int tiles_x = (a.width + 15) / 16, tiles_y = (a.height + th - 1) / th;
    bool cols = true;
    if (GRID) {
        const PulseConst cm = a.pc[a.np / 2];
        cols = fabs(cm.U) * a.g.d_u * (double)a.width >= fabs(cm.V) * a.g.d_v * (double)a.height;
    }
    int tx, ty;
    xcd_band_tile((int)blockIdx.x, tiles_x, tiles_y, cols, tx, ty);
    accumulate_tile_impl<QT, PPT, GRID, WIN>(a, tx * 16, ty * th, sh, a.cubic != 0);
}

// Several jobs in one launch over a tile list (job by job, row-major): block b computes tile
// tiles[xcd_rank(b, ntiles)], so each XCD takes a contiguous run of the list and reads the profile
// rows of few jobs. All the jobs of a launch share the grid mode.
template <typename QT, int PPT, bool GRID>
__global__ __launch_bounds__(kBpThreads) __attribute__((amdgpu_waves_per_eu(waves_per_simd<QT, GRID>(PPT))))
void accumulate_fused_kernel(
    const BpArgs *__restrict__ jobs, const FusedTile *__restrict__ tiles, int ntiles) {
    __shared__ TileShared<GRID, PPT> sh;
    const FusedTile t = tiles[xcd_rank((int)blockIdx.x, ntiles)];
    const BpArgs a = jobs[t.job];
    accumulate_tile_impl<QT, PPT, GRID>(a, t.bx, t.by, sh, a.cubic != 0);
}

// ------------------------------------------------------------------ image output

constexpr int kImgThreads = 256;
constexpr int kAbsmaxBlocks = 1024;

__global__ void zero_u64_kernel(unsigned long long *p) { *p = 0ull; }

// max |z| over the image (img_output.cpp:24); |z| >= 0, so the f64 bit pattern orders like the value.
// The correctly rounded sqrt is monotone, so each block takes the largest z.x^2 + z.y^2 and one sqrt
// of it: the same bits as the largest sqrt. Atomics on one address are served one at a time (one
// per block of 4096 held the kernel at 1.3 TB/s), so the launch has at most kAbsmaxBlocks blocks and
// a block issues its atomicMax only when its value exceeds the maximum it reads first.
__global__ __launch_bounds__(kImgThreads) void absmax_kernel(const double2 *__restrict__ img,
                                                             int64_t n,
                                                             unsigned long long *__restrict__ out) {
    __shared__ double red[kImgThreads / 64];
    double m = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {  // four loads in flight per thread
        double2 z[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) z[k] = img[i + k * stride];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double a2 = abs2(z[k]);
            m = a2 > m ? a2 : m;
        }
    }
    for (; i < n; i += stride) {
        const double a2 = abs2(img[i]);
        m = a2 > m ? a2 : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(m, off, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kImgThreads / 64; ++w) m = red[w] > m ? red[w] : m;
        // most blocks' maxima are below the running one by now: their atomics are skipped
        raise_absmax(out, m);
    }
}

// Each thread converts 4 consecutive pixels: four 16-byte loads, two 16-byte dB stores and one
// 4-byte u8 store (single-byte stores would take a store instruction per pixel); a tail of
// n mod 4 pixels is taken one at a time by the first threads. VEC false (dB not 16-byte or u8 not
// 4-byte aligned): every pixel on its own.
template <bool VEC>
__global__ __launch_bounds__(kImgThreads) void db_u8_kernel(
    const double2 *__restrict__ img, int64_t n, const unsigned long long *__restrict__ maxbits,
    double db_per_neper, double *__restrict__ dB, uint8_t *__restrict__ u8, double lo, double hi) {
    const double m = __longlong_as_double((long long)*maxbits);
    if (!VEC) {
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * blockDim.x) {
            const double v = to_db(img[i], m, db_per_neper);
            if (dB) dB[i] = v;
            if (u8) u8[i] = to_u8(v, lo, hi);
        }
        return;
    }
    const int64_t n4 = n / 4;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = t0; q < n4; q += stride) {
        double2 z[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) z[k] = img[4 * q + k];
        double v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = to_db(z[k], m, db_per_neper);
        if (dB) {
            double2 *d2 = reinterpret_cast<double2 *>(dB + 4 * q);
            d2[0] = make_double2(v[0], v[1]);
            d2[1] = make_double2(v[2], v[3]);
        }
        if (u8) {
            const unsigned w = (unsigned)to_u8(v[0], lo, hi) | ((unsigned)to_u8(v[1], lo, hi) << 8) |
                               ((unsigned)to_u8(v[2], lo, hi) << 16) |
                               ((unsigned)to_u8(v[3], lo, hi) << 24);
            *reinterpret_cast<unsigned *>(u8 + 4 * q) = w;
        }
    }
    const int64_t i = 4 * n4 + t0;
    if (i < n) {
        const double v = to_db(img[i], m, db_per_neper);
        if (dB) dB[i] = v;
        if (u8) u8[i] = to_u8(v, lo, hi);
    }
}

__global__ __launch_bounds__(kImgThreads) void u8_kernel(const double *__restrict__ dB, int64_t n,
                                                         uint8_t *__restrict__ u8, double lo,
                                                         double hi) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        u8[i] = to_u8(dB[i], lo, hi);
}

// ------------------------------------------------------------------ image plane on the device

// pixel_locs(x, d) = sum_r A(r, d) b(x, r) = (0 + u_hat[d] u[x % nu]) + v_hat[d] v[x / nu]
// (img_plane.cpp:77-113), f64, each product and sum rounded once as on the host.
__global__ void pixel_locs_kernel(const double *__restrict__ u, int nu, const double *__restrict__ v,
                                  int nv, double uh0, double uh1, double uh2, double vh0,
                                  double vh1, double vh2, double *__restrict__ out) {
#pragma clang fp contract(off)
    const int64_t npix = (int64_t)nu * nv;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npix;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double uu = u[i % nu], vv = v[i / nu];
        out[i] = (0.0 + uh0 * uu) + vh0 * vv;
        out[npix + i] = (0.0 + uh1 * uu) + vh1 * vv;
        out[2 * npix + i] = (0.0 + uh2 * uu) + vh2 * vv;
    }
}

hipError_t launch_pixel_locs(const double *u, int nu, const double *v, int nv,
                             const double *u_hat, const double *v_hat, double *out,
                             hipStream_t stream) {
    const int64_t npix = (int64_t)nu * nv;
    if (npix <= 0) return hipSuccess;
    const int64_t want = (npix + 255) / 256;
    const int blocks = (int)(want < 65536 ? want : 65536);
    hipLaunchKernelGGL(pixel_locs_kernel, dim3(blocks), dim3(256), 0, stream, u, nu, v, nv,
                       u_hat[0], u_hat[1], u_hat[2], v_hat[0], v_hat[1], v_hat[2], out);
    return hipGetLastError();
}

// ------------------------------------------------------------------ scene simulator

constexpr int kSimThreads = 256;

__global__ __launch_bounds__(kSimThreads) void sim_phs_kernel(
    const double *__restrict__ pos, int np, const double *__restrict__ k64, int ns,
    const double *__restrict__ tgt, const double *__restrict__ amp, int nt,
    float2 *__restrict__ phs) {
    __shared__ double D[kSimThreads];
    __shared__ float2 A[kSimThreads];
    const int p = blockIdx.y;
    const int k = blockIdx.x * kSimThreads + threadIdx.x;
    const double px = pos[3 * p], py = pos[3 * p + 1], pz = pos[3 * p + 2];
    const double r0 = sqrt(px * px + py * py + pz * pz);
    const double kr = k < ns ? k64[k] * (1.0 / (2.0 * M_PI)) : 0.0;
    float accr = 0.f, acci = 0.f;
    for (int t0 = 0; t0 < nt; t0 += kSimThreads) {
        const int t = t0 + threadIdx.x;
        if (t < nt) {
            const double dx = px - tgt[3 * t], dy = py - tgt[3 * t + 1], dz = pz - tgt[3 * t + 2];
            D[threadIdx.x] = r0 - sqrt(dx * dx + dy * dy + dz * dz);
            A[threadIdx.x] = make_float2((float)amp[2 * t], (float)amp[2 * t + 1]);
        }
        __syncthreads();
        const int tn = min(kSimThreads, nt - t0);
        for (int u = 0; u < tn; ++u) {
            const float f = (float)__builtin_amdgcn_fract(kr * D[u]);
            const float sn = __builtin_amdgcn_sinf(f), cs = __builtin_amdgcn_cosf(f);
            const float2 am = A[u];
            accr += am.x * cs - am.y * sn;
            acci += am.x * sn + am.y * cs;
        }
        __syncthreads();
    }
    if (k < ns) phs[(size_t)p * ns + k] = make_float2(accr, acci);
}

// ------------------------------------------------------------------ launchers

hipError_t launch_pulse_consts(const float *pos, const float *k_r, int ns, int np,
                               const RowLayout &lay, int norm_policy, const PixelGrid *grid,
                               PulseConst *pc, BpHeader *hdr, hipStream_t stream,
                               unsigned long long *absmax, unsigned *paths) {
    const int threads = 256;
    const int blocks = (np + threads - 1) / threads;
    const bool g = grid && grid->on;
    hipLaunchKernelGGL(pulse_const_kernel, dim3(blocks > 0 ? blocks : 1), dim3(threads), 0, stream,
                       pos, k_r, ns, np, lay.lsa, lay.lsr, lay.inv_lsr, lay.ext, norm_policy,
                       g ? grid->uh[0] : 0.0, g ? grid->uh[1] : 0.0, g ? grid->uh[2] : 0.0,
                       g ? grid->vh[0] : 0.0, g ? grid->vh[1] : 0.0, g ? grid->vh[2] : 0.0, pc,
                       hdr, absmax, paths);
    return hipGetLastError();
}

// Allows a kernel the largest dynamic LDS the CU offers next to its static LDS (160 KiB total on
// gfx950). Failures are cleared so that they do not surface from a later hipGetLastError.
static void allow_max_dynamic_lds(const void *fn) {
    hipFuncAttributes attr;
    size_t stat = 0;
    if (hipFuncGetAttributes(&attr, fn) == hipSuccess) stat = attr.sharedSizeBytes;
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(163840 - stat));
    (void)hipGetLastError();
}

// Four-step range profiles (kMaxFft < N): batches of profiles through the column, row and store
// kernels, with Y and X of a batch in `scratch` (2 batch N elements).
template <typename QT>
static hipError_t launch_two_pass(const float2 *phs, const float *k_r, const double *ws,
                                  const double *wp, int ns, int np, int logN, const RowLayout &lay,
                                  const float2 *tw, const float2 *rot, QT *q, float *qscale,
                                  float2 *scratch, size_t scratch_elems, hipStream_t stream) {
    const int N = 1 << logN, logN1 = (logN + 1) / 2, N2 = N >> logN1;
    const int batch = (int)std::min<size_t>((size_t)np, scratch_elems / (2 * (size_t)N));
    if (batch < 1) return hipErrorInvalidValue;
    const size_t lds = sizeof(float2) * (size_t)(kMaxFft + kMaxFft / N2);  // sequences at L + 1
    static std::once_flag once;
    std::call_once(once, [] {
        allow_max_dynamic_lds((const void *)&fft_cols_kernel);
        allow_max_dynamic_lds((const void *)&fft_rows_kernel);
    });
    float2 *Y = scratch, *X = scratch + (size_t)batch * N;
    const int blocks = N / kMaxFft;  // per profile, in both passes
    for (int p0 = 0; p0 < np; p0 += batch) {
        const int nb = std::min(batch, np - p0);
        hipLaunchKernelGGL(fft_cols_kernel, dim3(nb * blocks), dim3(kTwoPassThreads), lds, stream,
                           phs, k_r, ws, wp, ns, np, logN, logN1, p0, tw, Y);
        hipLaunchKernelGGL(fft_rows_kernel, dim3(nb * blocks), dim3(kTwoPassThreads), lds, stream,
                           (const float2 *)Y, logN, logN1, tw, X);
        hipLaunchKernelGGL(rows_store_kernel<QT>, dim3(nb), dim3(kTwoPassThreads), 0, stream,
                           (const float2 *)X, N, p0, rot, lay.ext, lay.len, q, qscale);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

size_t two_pass_scratch_elems(int n_fft, int np) {
    if (n_fft <= kMaxFft) return 0;
    // batches of up to 2^27 elements (1 GiB) of Y and X together, at least one profile
    const size_t per = 2 * (size_t)n_fft;
    const size_t batch = std::max<size_t>(1, std::min<size_t>((size_t)np, ((size_t)1 << 27) / per));
    return batch * per;
}

hipError_t launch_range_profiles(const void *phs, const float *k_r, const double *ws,
                                 const double *wp, int ns, int np, const RowLayout &lay,
                                 const float2 *tw, float2 *rot, void *q, int q_format,
                                 float *qscale, hipStream_t stream, int shared_blocks,
                                 float2 *scratch, size_t scratch_elems) {
    const int n_fft = lay.n_fft;
    int logN = 0;
    while ((1 << logN) < n_fft) ++logN;
    hipLaunchKernelGGL(rotation_kernel, dim3((lay.len + 255) / 256), dim3(256), 0, stream, k_r, ns,
                       lay.lsa, lay.lsr, lay.ext, lay.len, rot);
    if ((n_fft & (n_fft - 1)) != 0) {
        if (n_fft > kMaxFftBluestein) return hipErrorInvalidValue;
        const int logM = bluestein_log_m(n_fft);
        const size_t lds = sizeof(float2) * ((size_t)1 << logM);
        const bool big = logM > 13;  // NB4 = 8 for M = 16384
        const void *fn = q_format == SARBP_Q_C16
                             ? (big ? (const void *)&range_profile_bluestein_kernel<QPairF16, 8>
                                    : (const void *)&range_profile_bluestein_kernel<QPairF16, 4>)
                             : (big ? (const void *)&range_profile_bluestein_kernel<QPairF32, 8>
                                    : (const void *)&range_profile_bluestein_kernel<QPairF32, 4>);
        {
            static std::mutex m;
            static std::set<const void *> done;
            std::lock_guard<std::mutex> lk(m);
            if (done.insert(fn).second) allow_max_dynamic_lds(fn);
        }
        const float2 *rot_c = rot;
        if (q_format == SARBP_Q_C16) {
            if (big)
                hipLaunchKernelGGL((range_profile_bluestein_kernel<QPairF16, 8>), dim3(np),
                                   dim3(kFftThreads), lds, stream, (const float2 *)phs, k_r, ws, wp,
                                   ns, np, n_fft, logM, tw, rot_c, lay.ext, lay.len,
                                   (QPairF16 *)q, qscale);
            else
                hipLaunchKernelGGL((range_profile_bluestein_kernel<QPairF16, 4>), dim3(np),
                                   dim3(kFftThreads), lds, stream, (const float2 *)phs, k_r, ws, wp,
                                   ns, np, n_fft, logM, tw, rot_c, lay.ext, lay.len,
                                   (QPairF16 *)q, qscale);
        } else {
            if (big)
                hipLaunchKernelGGL((range_profile_bluestein_kernel<QPairF32, 8>), dim3(np),
                                   dim3(kFftThreads), lds, stream, (const float2 *)phs, k_r, ws, wp,
                                   ns, np, n_fft, logM, tw, rot_c, lay.ext, lay.len,
                                   (QPairF32 *)q, (float *)nullptr);
            else
                hipLaunchKernelGGL((range_profile_bluestein_kernel<QPairF32, 4>), dim3(np),
                                   dim3(kFftThreads), lds, stream, (const float2 *)phs, k_r, ws, wp,
                                   ns, np, n_fft, logM, tw, rot_c, lay.ext, lay.len,
                                   (QPairF32 *)q, (float *)nullptr);
        }
        return hipGetLastError();
    }
    if (n_fft > kMaxFft) {
        if (n_fft > kMaxFftTwoPass || !scratch) return hipErrorInvalidValue;
        if (q_format == SARBP_Q_C16)
            return launch_two_pass<QPairF16>((const float2 *)phs, k_r, ws, wp, ns, np, logN, lay,
                                             tw, rot, (QPairF16 *)q, qscale, scratch,
                                             scratch_elems, stream);
        return launch_two_pass<QPairF32>((const float2 *)phs, k_r, ws, wp, ns, np, logN, lay, tw,
                                         rot, (QPairF32 *)q, nullptr, scratch, scratch_elems,
                                         stream);
    }
    if (shared_blocks > 0 && n_fft <= kMaxFftShared) {
        const size_t lds = 2 * sizeof(float2) * (size_t)n_fft;
        const dim3 grid((unsigned)std::min(np, shared_blocks));
        if (q_format == SARBP_Q_C16) {
            static std::once_flag once;
            std::call_once(once, [] { allow_max_dynamic_lds((const void *)&range_profile_shared_kernel<QPairF16>); });
            hipLaunchKernelGGL(range_profile_shared_kernel<QPairF16>, grid, dim3(kFftSharedThreads),
                               lds, stream, (const float2 *)phs, k_r, ws, wp, ns, np, n_fft, logN,
                               tw, (const float2 *)rot, lay.ext, lay.len, (QPairF16 *)q, qscale);
        } else {
            static std::once_flag once;
            std::call_once(once, [] { allow_max_dynamic_lds((const void *)&range_profile_shared_kernel<QPairF32>); });
            hipLaunchKernelGGL(range_profile_shared_kernel<QPairF32>, grid, dim3(kFftSharedThreads),
                               lds, stream, (const float2 *)phs, k_r, ws, wp, ns, np, n_fft, logN,
                               tw, (const float2 *)rot, lay.ext, lay.len, (QPairF32 *)q,
                               (float *)nullptr);
        }
        return hipGetLastError();
    }
    const size_t lds = sizeof(float2) * (size_t)n_fft;
    static_assert(kMaxFft == 8 * 4 * kFftThreads, "NB4 = 8 covers the largest one-row FFT");
    const bool big = n_fft > 4 * 4 * kFftThreads;  // NB4 = 8, else 4
    unsigned rp_grid = (unsigned)np;  // NB4 = 4: one block per profile
    if (big) {
        int dev = 0, cus = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        rp_grid = (unsigned)std::min<int64_t>(np, (int64_t)kRowBlocksPerCu * std::max(cus, 1));
    }
    const void *fn = q_format == SARBP_Q_C16
                         ? (big ? (const void *)&range_profile_kernel<QPairF16, 8>
                                : (const void *)&range_profile_kernel<QPairF16, 4>)
                         : (big ? (const void *)&range_profile_kernel<QPairF32, 8>
                                : (const void *)&range_profile_kernel<QPairF32, 4>);
    {
        static std::mutex m;
        static std::set<const void *> done;
        std::lock_guard<std::mutex> lk(m);
        if (done.insert(fn).second) allow_max_dynamic_lds(fn);
    }
    const float2 *rot_c = rot;
    if (q_format == SARBP_Q_C16) {
        QPairF16 *qq = (QPairF16 *)q;
        if (big)
            hipLaunchKernelGGL((range_profile_kernel<QPairF16, 8>), dim3(rp_grid), dim3(kFftThreads), lds,
                               stream, (const float2 *)phs, k_r, ws, wp, ns, np, n_fft, logN, tw,
                               rot_c, lay.ext, lay.len, qq, qscale);
        else
            hipLaunchKernelGGL((range_profile_kernel<QPairF16, 4>), dim3(rp_grid), dim3(kFftThreads), lds,
                               stream, (const float2 *)phs, k_r, ws, wp, ns, np, n_fft, logN, tw,
                               rot_c, lay.ext, lay.len, qq, qscale);
    } else {
        QPairF32 *qq = (QPairF32 *)q;
        if (big)
            hipLaunchKernelGGL((range_profile_kernel<QPairF32, 8>), dim3(rp_grid), dim3(kFftThreads), lds,
                               stream, (const float2 *)phs, k_r, ws, wp, ns, np, n_fft, logN, tw,
                               rot_c, lay.ext, lay.len, qq, (float *)nullptr);
        else
            hipLaunchKernelGGL((range_profile_kernel<QPairF32, 4>), dim3(rp_grid), dim3(kFftThreads), lds,
                               stream, (const float2 *)phs, k_r, ws, wp, ns, np, n_fft, logN, tw,
                               rot_c, lay.ext, lay.len, qq, (float *)nullptr);
    }
    return hipGetLastError();
}

// The cubic form builds its coefficients every kGridBlock pulses behind a barrier, a latency that
// other workgroups on the CU hide once the region holds at least two device-fulls of tiles; smaller
// regions (C1: 512 tiles of 16 x 32, 2 per CU) ran 19 % slower with it than with the p, q form,
// whose blocks are 128 pulses.
int grid_cubic(const BpArgs &a) {
    if (!a.g.on) return 0;
    const int th = tile_height(a.ppt), ppt = th / 16;
    const int64_t tiles = (int64_t)((a.full_w + 15) / 16) * ((a.full_h + th - 1) / th);
    const int q_format = a.q_format;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return 0;
    static std::mutex m;
    static std::map<std::pair<int, int>, int> resident;  // (device, 8 q_format + ppt) -> tiles
    int full;
    {
        std::lock_guard<std::mutex> lk(m);
        const auto key = std::make_pair(device, 8 * q_format + ppt);
        auto it = resident.find(key);
        if (it == resident.end())
            it = resident.emplace(key, accumulate_resident_tiles(device, q_format, ppt, true)).first;
        full = it->second;
    }
    return full > 0 && tiles >= 2 * (int64_t)full;
}

template <int PPT, bool GRID>
static hipError_t launch_accumulate_ppt(const BpArgs &a, hipStream_t stream) {
    const dim3 grid((unsigned)(((a.width + 15) / 16) * ((a.height + 16 * PPT - 1) / (16 * PPT))));
    BpArgs b = a;
    b.cubic = grid_cubic(a);
    if constexpr (GRID && kWinSlot<PPT> > 0) {
        // the window kernel gives the same bits as the gather kernel; it differs in LDS use only
        if (b.cubic && a.q_format == SARBP_Q_C32 && a.row_windows) {
            hipLaunchKernelGGL((accumulate_kernel<QPairF32, PPT, true, true>), grid,
                               dim3(kBpThreads), 0, stream, b);
            return hipGetLastError();
        }
    }
    if (a.q_format == SARBP_Q_C16)
        hipLaunchKernelGGL((accumulate_kernel<QPairF16, PPT, GRID>), grid, dim3(kBpThreads), 0,
                           stream, b);
    else
        hipLaunchKernelGGL((accumulate_kernel<QPairF32, PPT, GRID>), grid, dim3(kBpThreads), 0,
                           stream, b);
    return hipGetLastError();
}

template <int PPT, bool GRID>
static hipError_t launch_fused_ppt(const BpArgs *jobs, const FusedTile *tiles, int ntiles,
                                   int q_format, hipStream_t stream) {
    if (q_format == SARBP_Q_C16)
        hipLaunchKernelGGL((accumulate_fused_kernel<QPairF16, PPT, GRID>), dim3(ntiles),
                           dim3(kBpThreads), 0, stream, jobs, tiles, ntiles);
    else
        hipLaunchKernelGGL((accumulate_fused_kernel<QPairF32, PPT, GRID>), dim3(ntiles),
                           dim3(kBpThreads), 0, stream, jobs, tiles, ntiles);
    return hipGetLastError();
}

template <bool GRID>
static hipError_t launch_fused_mode(const BpArgs *jobs_dev, const FusedTile *tiles_dev, int ntiles,
                                    int q_format, int ppt, hipStream_t stream) {
    switch (ppt) {
    case 3: return launch_fused_ppt<3, GRID>(jobs_dev, tiles_dev, ntiles, q_format, stream);
    case 4: return launch_fused_ppt<4, GRID>(jobs_dev, tiles_dev, ntiles, q_format, stream);
    default: return launch_fused_ppt<2, GRID>(jobs_dev, tiles_dev, ntiles, q_format, stream);
    }
}

hipError_t launch_accumulate_fused(const BpArgs *jobs_dev, const FusedTile *tiles_dev, int ntiles,
                                   int q_format, int ppt, bool grid, hipStream_t stream) {
    if (ntiles <= 0) return hipSuccess;
    return grid ? launch_fused_mode<true>(jobs_dev, tiles_dev, ntiles, q_format, ppt, stream)
                : launch_fused_mode<false>(jobs_dev, tiles_dev, ntiles, q_format, ppt, stream);
}

int tile_height(int ppt) { return 16 * (ppt >= 2 && ppt <= 4 ? ppt : kDefaultPpt); }

// With the range expansion, 16 x 64 tiles (4 pixels per thread, 4 waves per SIMD) run 2.5 % (C2) to
// 5 % (C3) faster than 16 x 32 (8 waves per SIMD): the per-pulse constants and the gather setup are
// shared by more pixels. Below ~8 tiles per CU the taller tile leaves CUs idle at the end.
int default_tile_ppt(int width, int height) {
    const int64_t tiles4 = (int64_t)((width + 15) / 16) * ((height + 63) / 64);
    return tiles4 >= 2048 ? 4 : kDefaultPpt;
}

template <typename QT, int PPT, bool GRID>
static int resident_blocks() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n, reinterpret_cast<const void *>(&accumulate_kernel<QT, PPT, GRID>), kBpThreads, 0) !=
        hipSuccess)
        return 0;
    return n;
}

template <int PPT, bool GRID>
static int resident_blocks_fmt(int q_format) {
    return q_format == SARBP_Q_C16 ? resident_blocks<QPairF16, PPT, GRID>()
                                   : resident_blocks<QPairF32, PPT, GRID>();
}

int accumulate_resident_tiles(int device, int q_format, int ppt, bool grid) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return 0;
    int per_cu = 0;
    switch (ppt) {
    case 3: per_cu = grid ? resident_blocks_fmt<3, true>(q_format) : resident_blocks_fmt<3, false>(q_format); break;
    case 4: per_cu = grid ? resident_blocks_fmt<4, true>(q_format) : resident_blocks_fmt<4, false>(q_format); break;
    default: per_cu = grid ? resident_blocks_fmt<2, true>(q_format) : resident_blocks_fmt<2, false>(q_format); break;
    }
    return cus * per_cu;
}

hipError_t launch_accumulate(const BpArgs &a, hipStream_t stream) {
    if (a.width <= 0 || a.height <= 0) return hipSuccess;
    if (a.g.on) {
        switch (a.ppt) {
        case 2: return launch_accumulate_ppt<2, true>(a, stream);
        case 3: return launch_accumulate_ppt<3, true>(a, stream);
        case 4: return launch_accumulate_ppt<4, true>(a, stream);
        default: return launch_accumulate_ppt<kDefaultPpt, true>(a, stream);
        }
    }
    switch (a.ppt) {
    case 2: return launch_accumulate_ppt<2, false>(a, stream);
    case 3: return launch_accumulate_ppt<3, false>(a, stream);
    case 4: return launch_accumulate_ppt<4, false>(a, stream);
    default: return launch_accumulate_ppt<kDefaultPpt, false>(a, stream);
    }
}

hipError_t launch_zero_u64(unsigned long long *p, hipStream_t stream) {
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(1), 0, stream, p);
    return hipGetLastError();
}

hipError_t launch_db_u8(const double2 *img, int64_t npix, const unsigned long long *maxbits,
                        double *dB, uint8_t *u8, double lo, double hi, hipStream_t stream) {
    if (npix <= 0 || (!dB && !u8)) return hipSuccess;
    // 4 pixels per thread; the tail (npix mod 4 < 4 pixels) is taken by the first threads
    const int64_t want = (npix / 4 + kImgThreads - 1) / kImgThreads;
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(want, 4096));
    // 10 / log(10) with the f32 log(10) of the reference (util.h:56-58), rounded once
    const double db_per_neper = 10.0 / (double)std::log(10.0f);
    const bool vec = (uintptr_t)dB % 16 == 0 && (uintptr_t)u8 % 4 == 0;
    if (vec)
        hipLaunchKernelGGL(db_u8_kernel<true>, dim3(blocks), dim3(kImgThreads), 0, stream, img,
                           npix, maxbits, db_per_neper, dB, u8, lo, hi);
    else
        hipLaunchKernelGGL(db_u8_kernel<false>, dim3(blocks), dim3(kImgThreads), 0, stream, img,
                           npix, maxbits, db_per_neper, dB, u8, lo, hi);
    return hipGetLastError();
}

hipError_t launch_img_output(const double2 *img, int64_t npix, double *dB, uint8_t *u8,
                             double lo, double hi, unsigned long long *scratch_max,
                             hipStream_t stream) {
    if (npix <= 0) return hipSuccess;
    const int64_t want = (npix + kImgThreads - 1) / kImgThreads;
    const int blocks = (int)(want < kAbsmaxBlocks ? want : kAbsmaxBlocks);
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(1), 0, stream, scratch_max);
    hipLaunchKernelGGL(absmax_kernel, dim3(blocks), dim3(kImgThreads), 0, stream, img, npix,
                       scratch_max);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_db_u8(img, npix, scratch_max, dB, u8, lo, hi, stream);
}

hipError_t launch_u8_only(const double *dB, int64_t npix, uint8_t *u8, double lo, double hi,
                          hipStream_t stream) {
    if (npix <= 0) return hipSuccess;
    const int64_t want = (npix + kImgThreads - 1) / kImgThreads;
    const int blocks = (int)(want < 4096 ? want : 4096);
    hipLaunchKernelGGL(u8_kernel, dim3(blocks), dim3(kImgThreads), 0, stream, dB, npix, u8, lo,
                       hi);
    return hipGetLastError();
}

hipError_t launch_sim_phs(const double *pos, int np, const double *k64, int ns,
                          const double *tgt, const double *amp, int nt, float2 *phs,
                          hipStream_t stream) {
    if (np <= 0 || ns <= 0) return hipSuccess;
    const dim3 grid((ns + kSimThreads - 1) / kSimThreads, np);
    hipLaunchKernelGGL(sim_phs_kernel, grid, dim3(kSimThreads), 0, stream, pos, np, k64, ns, tgt,
                       amp, nt, phs);
    return hipGetLastError();
}

// ------------------------------------------------------------------ twiddle cache

namespace {
std::mutex g_tw_mu;
std::map<std::pair<int, int>, float2 *> g_tw;
}  // namespace

namespace {

// Stage tables of the radix-4 passes of an n-point FFT (fft_stage_twiddles), written to
// h[n, 2n) from the n-point table in h[0, n).
void fill_stage_twiddles(std::vector<float2> &h, size_t base, int n) {
    int logN = 0;
    while ((1 << logN) < n) ++logN;
    const int ns0 = (logN & 1) ? 2 : 1;
    for (int Ns = ns0; Ns < n; Ns *= 4) {
        const size_t stride = (size_t)n / (4 * (size_t)Ns);
        for (int r = 1; r < 4; ++r)
            for (int k = 0; k < Ns; ++k)
                h[base + n + (size_t)(Ns - ns0) + (size_t)(r - 1) * Ns + k] =
                    h[base + (size_t)r * k * stride];
    }
}

// Stage tables of the radix-16 passes of an n-point FFT (fft_stage16_twiddles), written to
// h[2n, 3n) from the n-point table in h[0, n): the passes of the schedule fft_next_radix.
void fill_stage16_twiddles(std::vector<float2> &h, int n) {
    int logN = 0;
    while ((1 << logN) < n) ++logN;
    const int ns16 = fft_first_span16(logN);
    for (int Ns = 1, left = logN; left > 0;) {
        const int R = fft_next_radix(Ns, left);
        if (R == 16) {
            const size_t stride = (size_t)n / (16 * (size_t)Ns);
            for (int r = 1; r < 16; ++r)
                for (int k = 0; k < Ns; ++k)
                    h[2 * (size_t)n + (size_t)(Ns - ns16) + (size_t)(r - 1) * Ns + k] =
                        h[(size_t)r * k * stride];
        }
        Ns *= R;
        left -= R == 2 ? 1 : R == 4 ? 2 : 4;
    }
}

// In-place forward FFT of a power-of-two length in f64 (iterative radix 2, twiddles evaluated
// directly), for the host-side Bluestein table.
void fft_f64_host(std::vector<std::complex<double>> &a) {
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const size_t half = len / 2;
        for (size_t k = 0; k < half; ++k) {
            const double ang = -2.0 * M_PI * (double)k / (double)len;
            const std::complex<double> wk(std::cos(ang), std::sin(ang));
            for (size_t i = 0; i < n; i += len) {
                const std::complex<double> u = a[i + k], v = a[i + k + half] * wk;
                a[i + k] = u + v;
                a[i + k + half] = u - v;
            }
        }
    }
}

}  // namespace

int bluestein_log_m(int n_fft) {
    int logM = 0;
    while ((1 << logM) < 2 * n_fft - 1) ++logM;
    return logM;
}

const float2 *twiddles(int device, int n_fft, hipStream_t stream) {
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto key = std::make_pair(device, n_fft);
    auto it = g_tw.find(key);
    if (it != g_tw.end()) return it->second;
    std::vector<float2> h;
    if ((n_fft & (n_fft - 1)) == 0) {
        // [0, N): W_N^t; [N, 2N): the stage tables of the radix-4 passes (fft_stage_twiddles);
        // [2N, 3N): those of the radix-16 passes (fft_stage16_twiddles); copies of entries of the
        // first part
        const size_t n = (size_t)n_fft;
        h.assign(3 * n, make_float2(0.f, 0.f));
        for (int t = 0; t < n_fft; ++t) {
            const double a = -2.0 * M_PI * (double)t / (double)n_fft;
            h[t] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
        fill_stage_twiddles(h, 0, n_fft);
        fill_stage16_twiddles(h, n_fft);
    } else {
        // Bluestein tables of range_profile_bluestein_kernel, M = 2^ceil(log2(2N - 1)):
        // [0, M) W_M^t, [M, 2M) the stage tables of M, [2M, 3M) FFT_M(b) / M, [3M, 3M + N) the chirp
        // w_n = exp(-i pi n^2 / N) (n^2 reduced mod 2N in integers), b_m = conj(w_|m|) for |m| < N
        const int M = 1 << bluestein_log_m(n_fft);
        h.assign(3 * (size_t)M + (size_t)n_fft, make_float2(0.f, 0.f));
        for (int t = 0; t < M; ++t) {
            const double a = -2.0 * M_PI * (double)t / (double)M;
            h[t] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
        fill_stage_twiddles(h, 0, M);
        std::vector<std::complex<double>> w((size_t)n_fft), b((size_t)M);
        for (int m = 0; m < n_fft; ++m) {
            const long long r = ((long long)m * m) % (2LL * n_fft);
            const double a = -M_PI * (double)r / (double)n_fft;
            w[m] = std::complex<double>(std::cos(a), std::sin(a));
        }
        for (int m = 0; m < n_fft; ++m) {
            b[m] = std::conj(w[m]);
            if (m > 0) b[(size_t)M - m] = std::conj(w[m]);
        }
        fft_f64_host(b);
        for (int k = 0; k < M; ++k)
            h[2 * (size_t)M + k] = make_float2((float)(b[k].real() / M), (float)(b[k].imag() / M));
        for (int m = 0; m < n_fft; ++m)
            h[3 * (size_t)M + m] = make_float2((float)w[m].real(), (float)w[m].imag());
    }
    float2 *d = nullptr;
    if (hipMalloc(&d, sizeof(float2) * h.size()) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    (void)stream;
    g_tw[key] = d;
    return d;
}

void release_twiddles() {
    std::lock_guard<std::mutex> lk(g_tw_mu);
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (auto &kv : g_tw) {
        (void)hipSetDevice(kv.first.first);
        (void)hipFree(kv.second);
    }
    (void)hipSetDevice(cur);
    g_tw.clear();
}

}  // namespace sarbp
