// Internal declarations shared by the HIP kernels, the plan layer and the C-ABI shim.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "sarbp_hip.h"

namespace sarbp {

// Per-pulse constants of the accumulation, computed on the device from pos and k_r.
// With d = |r - pos_p| (f64) the row coordinate of the interpolation is t = B - inv_lsr * d, and
// |r - pos_p|^2 = |r|^2 + C - 2 r.pos_p.
// Squared ranges are taken in row units: sigma |r - pos_p|^2 with sigma = inv_lsr^2.
struct PulseConst {
    double px, py, pz;  // platform position, f32 widened
    double Cs;          // sigma |pos_p|^2
    double B;           // (nr0 - lsa) * inv_lsr + ext
    double nr0;         // ||pos_p|| under the selected policy, widened
    double U, V;        // grid mode: -2 sigma u_hat.pos_p, -2 sigma v_hat.pos_p, so that for
                        // r = u u_hat + v v_hat: sigma |r - pos_p|^2 = sigma |r|^2 + Cs + u U + v V
                        // (0 without a grid)
};

// A pixel grid generated in the accumulation kernel instead of read from pixel_locs
// (img_plane.cpp:77-113): r(x, vy) = (0 + u_hat (u0 + x) d_u) + v_hat (v0 + vy) d_v, rounded in that
// order, for image column x and grid row vy = nv - 1 - y of output row y.
struct PixelGrid {
    int on;
    int u0, v0;  // floor(-nu / 2), floor(-nv / 2)
    double d_u, d_v;
    double uh[3], vh[3];
};
// The grid of an image nu x nv with spacing d_u, d_v and unit vectors u_hat, v_hat.
PixelGrid make_grid(int nu, int nv, double d_u, double d_v, const double *u_hat,
                    const double *v_hat);

// Launch-wide constants written next to the pulse table.
struct BpHeader {
    double k_c;       // k_r[nsamples / 2] (f32 widened)
    double gamma;     // k_c / (2 pi): revolutions of phase per metre of dr
    double lsa;       // floor(-ns * delta_r / 2)
    double inv_lsr;   // (N - 1) / (lsb - lsa)
    double kappa;     // -gamma * lsr: revolutions of phase per unit of t
    double g_ext;     // -gamma * (lsa - lsr * ext): phase at t = 0, revolutions
};

// Range profiles are stored per pulse as rows of `len` elements. Element j is the interpolation
// interval starting at t = j, where t = luts + ext and luts is the reference's index on the
// linspace axis (backprojection.cpp:53-57):
//   A_j = Q[clamp(j - ext)] * R_j,  B_j = (Q[clamp(j - ext + 1)] - Q[clamp(j - ext)]) * R_j,
//   R_j = exp(2 pi i fract(g_ext + kappa j)),  clamp to [0, N - 1].
// The pixel's term q(t) * expj(-k_c dr) is then (A_j + w B_j) * exp(2 pi i kappa w) with
// j = floor(t), w = t - j: the phase of the reference depends on dr only, and so on t only.
// `ext` edge-replicated elements on each side make the reference's clamped interpolation
// (backprojection.cpp:59-62) exact for |dr| up to `r_cover`.
struct RowLayout {
    int n_fft, ext, len;
    double lsa, lsb, lsr, inv_lsr;
    double r_cover;  // pixels with |r| <= r_cover have every t in [0, len - 1)
};
RowLayout row_layout(int ns, int n_fft, double delta_r);

// Thread-local last-error string.
void set_error(const char *fmt, ...);

// Host-side math (host_math.cpp).
int taylor_window(int num, int s_l, int den_policy, double *out);

// Kernel launchers (kernels.hip). All are asynchronous on `stream`; return hipError_t.
// grid: the pixel grid of grid mode (host memory) or NULL.
// absmax, paths (may be NULL): cleared by the same launch, for an accumulation that raises them.
hipError_t launch_pulse_consts(const float *pos, const float *k_r, int ns, int np,
                               const RowLayout &lay, int norm_policy, const PixelGrid *grid,
                               PulseConst *pc, BpHeader *hdr, hipStream_t stream,
                               unsigned long long *absmax = nullptr, unsigned *paths = nullptr);
// N_fft above kMaxFft (up to kMaxFftTwoPass) takes a four-step FFT through `scratch`
// (two_pass_scratch_elems(N_fft, np) complex f32 elements; ignored for smaller N_fft).
hipError_t launch_range_profiles(const void *phs, const float *k_r, const double *ws,
                                 const double *wp, int ns, int np, const RowLayout &lay,
                                 const float2 *tw, float2 *rot, void *q, int q_format,
                                 float *qscale, hipStream_t stream, int shared_blocks = 0,
                                 float2 *scratch = nullptr, size_t scratch_elems = 0);
// shared_blocks > 0 (and N_fft <= kMaxFftShared): the profiles are computed by at most that many
// 256-thread blocks, which suits a device busy with accumulation launches of other streams.
constexpr int kMaxFftShared = 8192;
// Scratch the four-step FFT of N_fft > kMaxFft needs (complex f32 elements; 0 up to kMaxFft).
size_t two_pass_scratch_elems(int n_fft, int np);
struct BpArgs {
    const void *q;
    int q_format;
    const float *qscale;  // per-profile inverse scale of f16 profiles (NULL for f32)
    int n_fft, np;
    const PulseConst *pc;
    const BpHeader *hdr;
    int ext, len;
    double r_cover2;
    const double *r;
    int ppt;  // pixels per thread (tile height 16 * ppt)
    int64_t r_min, r_extent;
    int nu, nv, x0, y0, width, height;
    int full_w, full_h;  // the whole region of the plan or job (a band launch covers part of it)
    int cubic;           // grid mode: expansion tiles take the cubic form (set by the launchers)
    double2 *out;
    PixelGrid g;  // g.on: locations from the grid, r unused
    // When not NULL: the f64 bit pattern of max |z| over the pixels written is raised into *absmax
    // (atomicMax; |z| >= 0 orders like its bits), so that the dB output needs no pass of its own.
    // The caller clears it before the first launch of an image.
    unsigned long long *absmax;
    // Grid mode, cubic form, f32 rows: tiles may read their rows from windows staged in LDS
    // (same bits as the gathers; 0 keeps the gathers).
    int row_windows;
    // When not NULL: per tile one count added to paths[k], k = the tile's loop (TilePath).
    unsigned *paths;
    // Grid mode, cubic form: the row coordinate of each pixel in double-float form (an exact
    // integer part per pulse and column, an f32 cubic per pixel) instead of f64.
    int range_df;
};
// The loop a tile takes (counted in BpArgs::paths).
enum TilePath {
    kPathExact = 0,        // pixels beyond the rows' cover: clamped index, f64 phase fraction
    kPathDirect = 1,       // direct form (expansion bound fails)
    kPathPQ = 2,           // range expansion, p, q form
    kPathCubicGather = 3,  // grid mode, cubic form, row elements gathered from HBM
    kPathCubicWindow = 4,  // grid mode, cubic form, row elements from windows in LDS
    kPathCubicDF = 5,      // grid mode, cubic form, double-float row coordinates
    kNumPaths = 6
};
hipError_t launch_accumulate(const BpArgs &a, hipStream_t stream);
// Grid mode: whether a plan's or job's expansion tiles take the cubic form, from its whole region
// (full_w x full_h at a.ppt), so that every launch of one job, banded, fused or alone, takes the
// same form. 0 without a grid.
int grid_cubic(const BpArgs &a);
// A tile of a fused launch: job index and the tile's top-left output pixel within that job's region.
struct FusedTile {
    int job, bx, by, pad;
};
// All jobs of one fused launch share q_format, ppt and the grid mode (BpArgs::g.on == grid); each
// job's BpArgs::cubic is read as the caller set it (grid_cubic).
hipError_t launch_accumulate_fused(const BpArgs *jobs_dev, const FusedTile *tiles_dev, int ntiles,
                                   int q_format, int ppt, bool grid, hipStream_t stream);
// Output rows covered by one accumulation tile for a pixels-per-thread setting (16 columns wide).
int tile_height(int ppt);
// Pixels per thread for a region when the caller leaves the choice to the library (tile_ppt 0).
int default_tile_ppt(int width, int height);
// Accumulation workgroups (tiles) resident on the whole device at once for this instantiation
// (occupancy per CU x CUs); 0 if the runtime cannot tell.
int accumulate_resident_tiles(int device, int q_format, int ppt, bool grid);
hipError_t launch_img_output(const double2 *img, int64_t npix, double *dB, uint8_t *u8,
                             double lo, double hi, unsigned long long *scratch_max,
                             hipStream_t stream);
// dB / u8 from an image whose max |z| bits are already in *maxbits (e.g. from the accumulation).
hipError_t launch_db_u8(const double2 *img, int64_t npix, const unsigned long long *maxbits,
                        double *dB, uint8_t *u8, double lo, double hi, hipStream_t stream);
// *p = 0 on `stream` (a kernel, not a memset, so that the step can sit in a captured graph).
hipError_t launch_zero_u64(unsigned long long *p, hipStream_t stream);
hipError_t launch_u8_only(const double *dB, int64_t npix, uint8_t *u8, double lo, double hi,
                          hipStream_t stream);
hipError_t launch_pixel_locs(const double *u, int nu, const double *v, int nv,
                             const double *u_hat, const double *v_hat, double *out,
                             hipStream_t stream);
hipError_t launch_sim_phs(const double *pos, int np, const double *k64, int ns,
                          const double *tgt, const double *amp, int nt, float2 *phs,
                          hipStream_t stream);

// Device twiddle tables, cached per (device, N). A power of two N: exp(-2 pi i t / N) for t in
// [0, N), followed by N entries of per-pass copies laid out for the radix-4 passes of the one-row
// FFT. Any other N (<= kMaxFftBluestein): the tables of range_profile_bluestein_kernel, whose FFT
// length is M = 2^bluestein_log_m(N) >= 2N - 1.
const float2 *twiddles(int device, int n_fft, hipStream_t stream);
int bluestein_log_m(int n_fft);
// 0 if the pipeline takes this N_fft (a power of two in [2, kMaxFftTwoPass], or any other value
// in [3, kMaxFftBluestein]); else sets the error message and returns param_too_small/large.
int check_n_fft(int n_fft);
void release_twiddles();

// Longest FFT done in one LDS row (one workgroup per profile), and the longest done in two passes
// through HBM: rows of 1.5 N_fft + 2 elements must index in 18 bits (the fixed-point row
// coordinate of the accumulation).
constexpr int kMaxFft = 16384;
constexpr int kMaxFftTwoPass = 131072;
// Largest N_fft that is not a power of two: Bluestein's convolution length 2^ceil(log2(2N - 1))
// must fit one LDS row (kMaxFft).
constexpr int kMaxFftBluestein = kMaxFft / 2;

// Plan steps for callers that split the accumulation into row bands (the Halide-buffer entry point
// overlaps transfers with computation this way). plan_prepare enqueues the range profiles and the
// pulse table; plan_accumulate_rows enqueues rows [row0, row0 + nrows) of the plan's output
// region into out_rows (dense complex f64, width x nrows), reading r over [r_min, r_min + r_extent).
// Bands whose row0 is a multiple of plan_tile_rows give the image of one whole-region launch bit
// for bit.
int plan_prepare(sarbp_plan *p, const void *phs_dev, const void *k_r_dev, const void *pos_dev,
                 hipStream_t stream);
// use_grid = false runs the rows from r even when the plan has a grid (sarbp_plan_set_grid): the
// pulse table holds the terms of both forms, so the bands of one call may mix them.
int plan_accumulate_rows(sarbp_plan *p, const double *r_dev, int64_t r_min, int64_t r_extent,
                         int row0, int nrows, void *out_rows, hipStream_t stream,
                         bool use_grid = true);
int plan_tile_rows(const sarbp_plan *p);
// Tiles of the plan's accumulation kernel resident on its device at once (0 if unknown).
int plan_resident_tiles(const sarbp_plan *p);
const sarbp_plan_desc &plan_desc(const sarbp_plan *p);
// Bytes per stored range-profile element: (q, dq) in complex f32 or complex f16.
inline size_t q_elem_bytes(int q_format) { return q_format == SARBP_Q_C16 ? 8 : 16; }

}  // namespace sarbp
