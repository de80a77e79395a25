// Device-resident plans: the windows, range-profile buffer and pulse table of one problem shape,
// reused across calls. The Halide-buffer entry points and the serving engine both run on plans.
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "internal.h"

using namespace sarbp;

struct sarbp_plan {
    sarbp_plan_desc d;
    double *ws = nullptr;       // Taylor window over samples, f64 [ns]
    double *wp = nullptr;       // Taylor window over pulses, f64 [np]
    void *q = nullptr;          // range profiles [np][N] in q_format
    float *qscale = nullptr;    // per-profile inverse scale of f16 profiles [np]
    float2 *rot = nullptr;      // row rotation R_j [len]
    RowLayout lay;              // row layout and interpolation axis
    PulseConst *pc = nullptr;   // per-pulse constants [np]
    BpHeader *hdr = nullptr;    // launch-wide constants
    const float2 *tw = nullptr; // twiddles exp(-2 pi i t / N)
    PixelGrid grid{};           // grid.on: pixel locations generated in the accumulation kernel
    unsigned long long *absmax = nullptr;  // max |z| (f64 bits) of the last accumulated image
    unsigned *paths = nullptr;             // tiles per loop (TilePath) of the last accumulation
    float2 *fft_scratch = nullptr;         // four-step FFT work space (N_fft > kMaxFft)
    size_t fft_scratch_elems = 0;
};

namespace {

std::mutex g_policy_mu;
int g_norm_policy = SARBP_NORM_R0_F32_STRICT;
int g_den_policy = SARBP_TAYLOR_DEN_INT;


struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

int check_desc(const sarbp_plan_desc *d) {
    if (!d) {
        set_error("plan descriptor is NULL");
        return halide_error_code_buffer_argument_is_null;
    }
    if (d->nsamples < 2 || d->npulses < 2) {
        set_error("nsamples (%d) and npulses (%d) must be >= 2: the Taylor window's linspace "
                  "divides by num - 1 (util.h:80-84)",
                  d->nsamples, d->npulses);
        return halide_error_code_param_too_small;
    }
    if (int rv = check_n_fft(d->n_fft)) return rv;
    if (d->nu < 1 || d->nv < 1 || d->width < 0 || d->height < 0 || d->r_extent < 0) {
        set_error("image extents must be positive and region extents non-negative");
        return halide_error_code_buffer_extents_negative;
    }
    if ((int64_t)d->nu * d->nv > INT32_MAX) {
        set_error("nu * nv exceeds the 32-bit extent of the pixel_locs buffer");
        return halide_error_code_buffer_extents_too_large;
    }
    if (d->norm_r0_policy < 0 || d->norm_r0_policy > 2 || d->taylor_den_policy < 0 ||
        d->taylor_den_policy > 1 || d->q_format < 0 || d->q_format > 1) {
        set_error("invalid policy or format in plan descriptor");
        return halide_error_code_constraint_violated;
    }
    if (d->tile_ppt < 0 || d->tile_ppt > 4 || d->tile_ppt == 1) {
        set_error("tile_ppt %d is not one of 0 (default), 2, 3, 4", d->tile_ppt);
        return halide_error_code_constraint_violated;
    }
    if (d->row_windows < 0 || d->row_windows > 1) {
        set_error("row_windows %d is not 0 (gathers) or 1 (windows where they fit)",
                  d->row_windows);
        return halide_error_code_constraint_violated;
    }
    if (d->fft_blocks < 0) {
        set_error("fft_blocks %d is negative", d->fft_blocks);
        return halide_error_code_constraint_violated;
    }
    if (d->range_mode != SARBP_RANGE_F64 && d->range_mode != SARBP_RANGE_DF32) {
        set_error("range_mode %d is not SARBP_RANGE_F64 or SARBP_RANGE_DF32", d->range_mode);
        return halide_error_code_constraint_violated;
    }
    return 0;
}

}  // namespace

extern "C" int sarbp_set_default_policies(int norm_r0_policy, int taylor_den_policy) {
    if (norm_r0_policy < 0 || norm_r0_policy > 2 || taylor_den_policy < 0 || taylor_den_policy > 1) {
        set_error("invalid policy (%d, %d)", norm_r0_policy, taylor_den_policy);
        return halide_error_code_constraint_violated;
    }
    std::lock_guard<std::mutex> lk(g_policy_mu);
    g_norm_policy = norm_r0_policy;
    g_den_policy = taylor_den_policy;
    return 0;
}

extern "C" int sarbp_get_default_policies(int *norm_r0_policy, int *taylor_den_policy) {
    std::lock_guard<std::mutex> lk(g_policy_mu);
    if (norm_r0_policy) *norm_r0_policy = g_norm_policy;
    if (taylor_den_policy) *taylor_den_policy = g_den_policy;
    return 0;
}

extern "C" int sarbp_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" void sarbp_plan_desc_init(sarbp_plan_desc *d) {
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    int np_, dp_;
    sarbp_get_default_policies(&np_, &dp_);
    d->norm_r0_policy = np_;
    d->taylor_den_policy = dp_;
    d->q_format = SARBP_Q_C32;
    d->range_mode = SARBP_RANGE_F64;
}

extern "C" int sarbp_plan_destroy(sarbp_plan *p) {
    if (!p) return 0;
    DeviceGuard g(p->d.device);
    (void)hipFree(p->ws);
    (void)hipFree(p->wp);
    (void)hipFree(p->q);
    (void)hipFree(p->qscale);
    (void)hipFree(p->rot);
    (void)hipFree(p->pc);
    (void)hipFree(p->hdr);
    (void)hipFree(p->absmax);
    (void)hipFree(p->paths);
    (void)hipFree(p->fft_scratch);
    delete p;
    return 0;
}

extern "C" int sarbp_plan_create(const sarbp_plan_desc *d, sarbp_plan **out) {
    if (!out) {
        set_error("sarbp_plan_create: out is NULL");
        return halide_error_code_buffer_argument_is_null;
    }
    *out = nullptr;
    int rv = check_desc(d);
    if (rv) return rv;
    int ndev = sarbp_device_count();
    if (ndev <= 0) {
        set_error("no HIP device is visible");
        return halide_error_code_no_device_interface;
    }
    if (d->device < 0 || d->device >= ndev) {
        set_error("device %d out of range (%d visible)", d->device, ndev);
        return halide_error_code_no_device_interface;
    }
    DeviceGuard g(d->device);
    sarbp_plan *p = new sarbp_plan();
    p->d = *d;
    if (p->d.tile_ppt == 0) p->d.tile_ppt = default_tile_ppt(d->width, d->height);
    const int ns = d->nsamples, np = d->npulses, N = d->n_fft;
    std::vector<double> ws((size_t)ns), wp((size_t)np);
    taylor_window(ns, d->taylor_s_l, d->taylor_den_policy, ws.data());
    taylor_window(np, d->taylor_s_l, d->taylor_den_policy, wp.data());
    const size_t qelem = q_elem_bytes(d->q_format);
    p->lay = row_layout(ns, N, d->delta_r);
    const int len = p->lay.len;
    bool ok = hipMalloc(&p->ws, sizeof(double) * ns) == hipSuccess &&
              hipMalloc(&p->wp, sizeof(double) * np) == hipSuccess &&
              hipMalloc(&p->q, qelem * (size_t)np * (size_t)len) == hipSuccess &&
              hipMalloc(&p->rot, sizeof(float2) * (size_t)len) == hipSuccess &&
              hipMalloc(&p->pc, sizeof(PulseConst) * (size_t)np) == hipSuccess &&
              hipMalloc(&p->hdr, sizeof(BpHeader)) == hipSuccess &&
              hipMalloc(&p->absmax, sizeof(unsigned long long)) == hipSuccess &&
              hipMalloc(&p->paths, sizeof(unsigned) * kNumPaths) == hipSuccess &&
              ((p->fft_scratch_elems = two_pass_scratch_elems(N, np)) == 0 ||
               hipMalloc(&p->fft_scratch, sizeof(float2) * p->fft_scratch_elems) == hipSuccess) &&
              (d->q_format != SARBP_Q_C16 ||
               hipMalloc(&p->qscale, sizeof(float) * (size_t)np) == hipSuccess);
    if (!ok) {
        sarbp_plan_destroy(p);
        set_error("device allocation failed for plan (np=%d, N=%d)", np, N);
        return halide_error_code_device_malloc_failed;
    }
    if (hipMemcpy(p->ws, ws.data(), sizeof(double) * ns, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->wp, wp.data(), sizeof(double) * np, hipMemcpyHostToDevice) != hipSuccess) {
        sarbp_plan_destroy(p);
        set_error("copy of Taylor windows to device failed");
        return halide_error_code_copy_to_device_failed;
    }
    p->tw = twiddles(d->device, N, nullptr);
    if (!p->tw) {
        sarbp_plan_destroy(p);
        set_error("twiddle table allocation failed");
        return halide_error_code_device_malloc_failed;
    }
    *out = p;
    return 0;
}

extern "C" void *sarbp_plan_q_buffer(sarbp_plan *p) { return p ? p->q : nullptr; }

extern "C" size_t sarbp_plan_q_bytes(sarbp_plan *p) {
    return p ? q_elem_bytes(p->d.q_format) * (size_t)p->d.npulses * (size_t)p->lay.len : 0;
}

extern "C" int32_t sarbp_plan_tile_ppt(const sarbp_plan *p) { return p ? p->d.tile_ppt : 0; }

extern "C" int32_t sarbp_plan_grid_form(const sarbp_plan *p) {
    if (!p || !p->grid.on) return SARBP_GRID_FORM_NONE;
    DeviceGuard g(p->d.device);
    BpArgs a{};
    a.q_format = p->d.q_format;
    a.ppt = p->d.tile_ppt;
    a.full_w = p->d.width;
    a.full_h = p->d.height;
    a.g = p->grid;
    return grid_cubic(a) ? SARBP_GRID_FORM_CUBIC : SARBP_GRID_FORM_PQ;
}

extern "C" int sarbp_plan_tile_paths(sarbp_plan *p, int64_t *counts) {
    if (!p || !counts) {
        set_error("sarbp_plan_tile_paths: NULL argument");
        return halide_error_code_buffer_argument_is_null;
    }
    DeviceGuard g(p->d.device);
    unsigned h[kNumPaths];
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(h, p->paths, sizeof h, hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("sarbp_plan_tile_paths: copy failed");
        return halide_error_code_copy_to_host_failed;
    }
    for (int k = 0; k < kNumPaths; ++k) counts[k] = h[k];
    return 0;
}

extern "C" int sarbp_plan_img_output(sarbp_plan *p, const void *img_dev, double *dB_dev,
                                     uint8_t *u8_dev, double dB_min, double dB_max, void *stream) {
    if (!p || !img_dev) {
        set_error("sarbp_plan_img_output: NULL argument");
        return halide_error_code_buffer_argument_is_null;
    }
    DeviceGuard g(p->d.device);
    const int64_t npix = (int64_t)p->d.width * p->d.height;
    hipError_t e = launch_db_u8((const double2 *)img_dev, npix, p->absmax, dB_dev, u8_dev, dB_min,
                                dB_max, (hipStream_t)stream);
    if (e != hipSuccess) {
        set_error("image-output launch failed: %s", hipGetErrorString(e));
        return halide_error_code_device_run_failed;
    }
    return 0;
}

namespace {
float half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h >> 15) << 31;
    int e = (h >> 10) & 0x1f;
    uint32_t m = h & 0x3ff;
    uint32_t bits;
    if (e == 0x1f) {
        bits = sign | 0x7f800000u | (m << 13);
    } else if (e == 0) {
        if (m == 0) {
            bits = sign;
        } else {  // subnormal: renormalise
            int sh = 0;
            while (!(m & 0x400)) {
                m <<= 1;
                ++sh;
            }
            m &= 0x3ff;
            bits = sign | ((uint32_t)(127 - 15 + 1 - sh) << 23) | (m << 13);
        }
    } else {
        bits = sign | ((uint32_t)(e - 15 + 127) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}
}  // namespace

extern "C" int sarbp_plan_read_profiles(sarbp_plan *p, float *host_out) {
    if (!p || !host_out) {
        set_error("sarbp_plan_read_profiles: NULL argument");
        return halide_error_code_buffer_argument_is_null;
    }
    DeviceGuard g(p->d.device);
    const int np = p->d.npulses, N = p->d.n_fft, len = p->lay.len, ext = p->lay.ext;
    const size_t eb = q_elem_bytes(p->d.q_format);
    const size_t row = (size_t)len * eb;
    std::vector<uint8_t> h(row * (size_t)np);
    std::vector<float> sc((size_t)np, 1.0f);
    std::vector<float2> rot((size_t)len);
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(h.data(), p->q, h.size(), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(rot.data(), p->rot, sizeof(float2) * len, hipMemcpyDeviceToHost) != hipSuccess ||
        (p->qscale && hipMemcpy(sc.data(), p->qscale, sizeof(float) * np,
                                hipMemcpyDeviceToHost) != hipSuccess)) {
        set_error("sarbp_plan_read_profiles: copy failed");
        return halide_error_code_copy_to_host_failed;
    }
    for (int pp = 0; pp < np; ++pp) {
        for (int k = 0; k < N; ++k) {
            const int j = k + ext;  // element j holds A_j = Q[j - ext] * R_j
            const uint8_t *e = h.data() + (size_t)pp * row + (size_t)j * eb;
            float re, im;
            if (p->d.q_format == SARBP_Q_C16) {
                uint16_t a, b;
                std::memcpy(&a, e, 2);
                std::memcpy(&b, e + 2, 2);
                re = half_to_float(a);
                im = half_to_float(b);
            } else {
                std::memcpy(&re, e, 4);
                std::memcpy(&im, e + 4, 4);
            }
            const double cr = rot[j].x, ci = rot[j].y;  // multiply by conj(R_j), |R_j| = 1
            host_out[2 * ((size_t)pp * N + k)] = (float)((re * cr + im * ci) * sc[pp]);
            host_out[2 * ((size_t)pp * N + k) + 1] = (float)((im * cr - re * ci) * sc[pp]);
        }
    }
    return 0;
}

extern "C" int sarbp_plan_read_pulse_norms(sarbp_plan *p, double *host_out, int32_t n) {
    if (!p || !host_out) {
        set_error("sarbp_plan_read_pulse_norms: NULL argument");
        return halide_error_code_buffer_argument_is_null;
    }
    if (n < 0 || n > p->d.npulses) {
        set_error("sarbp_plan_read_pulse_norms: n (%d) outside [0, npulses]", n);
        return halide_error_code_param_too_large;
    }
    DeviceGuard g(p->d.device);
    std::vector<PulseConst> h((size_t)n);
    if (n > 0 && hipMemcpy(h.data(), p->pc, sizeof(PulseConst) * (size_t)n,
                           hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("sarbp_plan_read_pulse_norms: copy failed");
        return halide_error_code_copy_to_host_failed;
    }
    for (int i = 0; i < n; ++i) host_out[i] = h[i].nr0;
    return 0;
}

extern "C" int sarbp_plan_range_profiles(sarbp_plan *p, const void *phs_dev, const void *k_r_dev,
                                         void *q_dev, void *stream) {
    if (!p || !phs_dev || !k_r_dev) {
        set_error("sarbp_plan_range_profiles: NULL argument");
        return halide_error_code_buffer_argument_is_null;
    }
    DeviceGuard g(p->d.device);
    hipError_t e = launch_range_profiles(phs_dev, (const float *)k_r_dev, p->ws, p->wp,
                                         p->d.nsamples, p->d.npulses, p->lay, p->tw, p->rot,
                                         q_dev ? q_dev : p->q, p->d.q_format, p->qscale,
                                         (hipStream_t)stream, p->d.fft_blocks, p->fft_scratch,
                                         p->fft_scratch_elems);
    if (e != hipSuccess) {
        set_error("range-profile launch failed: %s", hipGetErrorString(e));
        return halide_error_code_device_run_failed;
    }
    return 0;
}

extern "C" int sarbp_plan_accumulate(sarbp_plan *p, const void *q_dev, const void *k_r_dev,
                                     const void *pos_dev, const void *r_dev, void *out_dev,
                                     void *stream) {
    if (!p || !k_r_dev || !pos_dev || !out_dev ||
        (!r_dev && p->d.r_extent > 0 && !p->grid.on)) {
        set_error("sarbp_plan_accumulate: NULL argument");
        return halide_error_code_buffer_argument_is_null;
    }
    DeviceGuard g(p->d.device);
    const sarbp_plan_desc &d = p->d;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = launch_pulse_consts((const float *)pos_dev, (const float *)k_r_dev, d.nsamples,
                                       d.npulses, p->lay, d.norm_r0_policy,
                                       p->grid.on ? &p->grid : nullptr, p->pc, p->hdr, s,
                                       p->absmax, p->paths);
    if (e == hipSuccess) {
        BpArgs a{};
        a.q = q_dev ? q_dev : p->q;
        a.q_format = d.q_format;
        a.qscale = d.q_format == SARBP_Q_C16 ? p->qscale : nullptr;
        a.n_fft = d.n_fft;
        a.np = d.npulses;
        a.pc = p->pc;
        a.hdr = p->hdr;
        a.ext = p->lay.ext;
        a.len = p->lay.len;
        a.r_cover2 = p->lay.r_cover * p->lay.r_cover;
        a.r = (const double *)r_dev;
        a.ppt = d.tile_ppt;
        a.r_min = d.r_min;
        a.r_extent = d.r_extent;
        a.nu = d.nu;
        a.nv = d.nv;
        a.x0 = d.x0;
        a.y0 = d.y0;
        a.width = d.width;
        a.height = d.height;
        a.full_w = d.width;
        a.full_h = d.height;
        a.cubic = 0;
        a.out = (double2 *)out_dev;
        a.g = p->grid;
        a.absmax = p->absmax;
        a.paths = p->paths;
        a.range_df = d.range_mode == SARBP_RANGE_DF32;
        a.row_windows = d.row_windows;
        e = launch_accumulate(a, s);
    }
    if (e != hipSuccess) {
        set_error("accumulation launch failed: %s", hipGetErrorString(e));
        return halide_error_code_device_run_failed;
    }
    return 0;
}

extern "C" int sarbp_plan_set_grid(sarbp_plan *p, double d_u, double d_v, const double *u_hat,
                                   const double *v_hat) {
    if (!p) {
        set_error("sarbp_plan_set_grid: NULL plan");
        return halide_error_code_buffer_argument_is_null;
    }
    if (!u_hat || !v_hat) {  // back to pixel locations read from r
        p->grid = PixelGrid{};
        return 0;
    }
    const sarbp_plan_desc &d = p->d;
    if (d.x0 < 0 || d.y0 < 0 || (int64_t)d.x0 + d.width > d.nu || (int64_t)d.y0 + d.height > d.nv) {
        set_error("sarbp_plan_set_grid: the plan's output region must lie inside the nu x nv image");
        return halide_error_code_constraint_violated;
    }
    p->grid = make_grid(d.nu, d.nv, d_u, d_v, u_hat, v_hat);
    return 0;
}

extern "C" int sarbp_plan_execute(sarbp_plan *p, const void *phs_dev, const void *k_r_dev,
                                  const void *pos_dev, const void *r_dev, void *out_dev,
                                  void *stream) {
    int rv = sarbp_plan_range_profiles(p, phs_dev, k_r_dev, nullptr, stream);
    if (rv) return rv;
    return sarbp_plan_accumulate(p, nullptr, k_r_dev, pos_dev, r_dev, out_dev, stream);
}

namespace sarbp {

int plan_prepare(sarbp_plan *p, const void *phs_dev, const void *k_r_dev, const void *pos_dev,
                 hipStream_t stream) {
    int rv = sarbp_plan_range_profiles(p, phs_dev, k_r_dev, nullptr, stream);
    if (rv) return rv;
    DeviceGuard g(p->d.device);
    const sarbp_plan_desc &d = p->d;
    hipError_t e = launch_pulse_consts((const float *)pos_dev, (const float *)k_r_dev, d.nsamples,
                                       d.npulses, p->lay, d.norm_r0_policy,
                                       p->grid.on ? &p->grid : nullptr, p->pc, p->hdr, stream,
                                       p->absmax, p->paths);
    if (e != hipSuccess) {
        set_error("pulse-table launch failed: %s", hipGetErrorString(e));
        return halide_error_code_device_run_failed;
    }
    return 0;
}

int plan_accumulate_rows(sarbp_plan *p, const double *r_dev, int64_t r_min, int64_t r_extent,
                         int row0, int nrows, void *out_rows, hipStream_t stream, bool use_grid) {
    DeviceGuard g(p->d.device);
    const sarbp_plan_desc &d = p->d;
    BpArgs a{};
    a.q = p->q;
    a.q_format = d.q_format;
    a.qscale = d.q_format == SARBP_Q_C16 ? p->qscale : nullptr;
    a.n_fft = d.n_fft;
    a.np = d.npulses;
    a.pc = p->pc;
    a.hdr = p->hdr;
    a.ext = p->lay.ext;
    a.len = p->lay.len;
    a.r_cover2 = p->lay.r_cover * p->lay.r_cover;
    a.r = r_dev;
    a.ppt = d.tile_ppt;
    a.r_min = r_min;
    a.r_extent = r_extent;
    a.nu = d.nu;
    a.nv = d.nv;
    a.x0 = d.x0;
    a.y0 = d.y0 + row0;
    a.width = d.width;
    a.height = nrows;
    a.full_w = d.width;
    a.full_h = d.height;
    a.cubic = 0;
    a.out = (double2 *)out_rows;
    a.g = use_grid ? p->grid : PixelGrid{};
    a.absmax = p->absmax;
    a.paths = p->paths;
    a.range_df = d.range_mode == SARBP_RANGE_DF32;
    a.row_windows = d.row_windows;
    hipError_t e = launch_accumulate(a, stream);
    if (e != hipSuccess) {
        set_error("accumulation launch failed: %s", hipGetErrorString(e));
        return halide_error_code_device_run_failed;
    }
    return 0;
}

int plan_tile_rows(const sarbp_plan *p) { return tile_height(p->d.tile_ppt); }

int plan_resident_tiles(const sarbp_plan *p) {
    return accumulate_resident_tiles(p->d.device, p->d.q_format, p->d.tile_ppt, p->grid.on != 0);
}

const sarbp_plan_desc &plan_desc(const sarbp_plan *p) { return p->d; }

}  // namespace sarbp

namespace {
// One 8-byte max cell per (device, stream) that image output has run on: calls on one stream are
// ordered by the stream, and calls on different streams never share a cell. Held until exit.
std::mutex g_scratch_mu;
std::map<std::pair<int, hipStream_t>, unsigned long long *> g_scratch;
}  // namespace

extern "C" int sarbp_img_output_device(const void *img_dev, int64_t npix, double *dB_dev,
                                       uint8_t *u8_dev, double dB_min, double dB_max,
                                       void *stream) {
    if (!img_dev && npix > 0) {
        set_error("sarbp_img_output_device: img is NULL");
        return halide_error_code_buffer_argument_is_null;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("no HIP device is visible");
        return halide_error_code_no_device_interface;
    }
    unsigned long long *cell = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        unsigned long long *&slot = g_scratch[std::make_pair(dev, (hipStream_t)stream)];
        if (!slot && hipMalloc(&slot, sizeof(unsigned long long)) != hipSuccess) {
            slot = nullptr;
            set_error("scratch allocation failed");
            return halide_error_code_device_malloc_failed;
        }
        cell = slot;
    }
    hipError_t e = launch_img_output((const double2 *)img_dev, npix, dB_dev, u8_dev, dB_min, dB_max,
                                     cell, (hipStream_t)stream);
    if (e != hipSuccess) {
        set_error("image-output launch failed: %s", hipGetErrorString(e));
        return halide_error_code_device_run_failed;
    }
    return 0;
}

extern "C" int sarbp_sim_phs_device(const double *pos_dev, int32_t npulses,
                                    const double *k_r64_dev, int32_t nsamples,
                                    const double *tgt_dev, const double *amp_dev,
                                    int32_t ntargets, void *phs_dev, void *stream) {
    if (!pos_dev || !k_r64_dev || !phs_dev || (ntargets > 0 && (!tgt_dev || !amp_dev))) {
        set_error("sarbp_sim_phs_device: NULL argument");
        return halide_error_code_buffer_argument_is_null;
    }
    hipError_t e = launch_sim_phs(pos_dev, npulses, k_r64_dev, nsamples, tgt_dev, amp_dev,
                                  ntargets, (float2 *)phs_dev, (hipStream_t)stream);
    if (e != hipSuccess) {
        set_error("simulator launch failed: %s", hipGetErrorString(e));
        return halide_error_code_device_run_failed;
    }
    return 0;
}

extern "C" int sarbp_pixel_locs_device(const double *u_dev, int32_t nu, const double *v_dev,
                                       int32_t nv, const double *u_hat, const double *v_hat,
                                       double *pixel_locs_dev, void *stream) {
    if (!u_dev || !v_dev || !u_hat || !v_hat || !pixel_locs_dev) {
        set_error("sarbp_pixel_locs_device: NULL argument");
        return halide_error_code_buffer_argument_is_null;
    }
    if (nu < 0 || nv < 0) {
        set_error("sarbp_pixel_locs_device: negative extent");
        return halide_error_code_buffer_extents_negative;
    }
    hipError_t e = launch_pixel_locs(u_dev, nu, v_dev, nv, u_hat, v_hat, pixel_locs_dev,
                                     (hipStream_t)stream);
    if (e != hipSuccess) {
        set_error("pixel_locs launch failed: %s", hipGetErrorString(e));
        return halide_error_code_device_run_failed;
    }
    return 0;
}
