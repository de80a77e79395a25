// Host-side pieces of the pipeline that are O(n) or smaller: the Taylor windows, the range-axis
// constants and the image plane. Built with -ffp-contract=off so each operation rounds once.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "internal.h"

namespace sarbp {

namespace {
thread_local std::string g_last_error;

// Halide integer division rounds toward negative infinity.
int floordiv(int a, int b) {
    int q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
    return q;
}
}  // namespace

void set_error(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    std::fprintf(stderr, "sarbp_hip: %s\n", buf);
}

const char *last_error() { return g_last_error.c_str(); }

// dft.cpp:20-30 plans FFTW for any N_fft; here powers of two run the LDS (N <= kMaxFft) or
// four-step (N <= kMaxFftTwoPass) FFT and other lengths Bluestein's convolution (N <=
// kMaxFftBluestein).
int check_n_fft(int n_fft) {
    const bool pow2 = n_fft > 0 && (n_fft & (n_fft - 1)) == 0;
    if (n_fft < 2) {
        set_error("N_fft (%d) must be at least 2", n_fft);
        return halide_error_code_param_too_small;
    }
    if (pow2 ? n_fft > kMaxFftTwoPass : n_fft > kMaxFftBluestein) {
        set_error("N_fft (%d) must be a power of two up to %d or another length up to %d", n_fft,
                  kMaxFftTwoPass, kMaxFftBluestein);
        return halide_error_code_param_too_large;
    }
    return 0;
}

// backprojection.cpp:120 and :53 — the interpolation axis is linspace(floor(-ns dr / 2),
// floor(ns dr / 2), N); rows extend it by N / 4 elements on each side.
RowLayout row_layout(int ns, int n_fft, double delta_r) {
    RowLayout a;
    a.n_fft = n_fft;
    a.lsa = std::floor((double)(-ns) * delta_r / 2.0);
    a.lsb = std::floor((double)ns * delta_r / 2.0);
    a.lsr = (a.lsb - a.lsa) / (double)(n_fft - 1);
    a.inv_lsr = 1.0 / a.lsr;
    a.ext = n_fft / 4 > 2 ? n_fft / 4 : 2;
    a.len = n_fft + 2 * a.ext + 2;
    // |dr| = | ||pos|| - |r - pos| | <= |r| (+ the f32 rounding of ||pos||, < 1 % of a metre here);
    // keep 2 elements of margin at each end of the row.
    const double lo = -(a.lsa + (2 - a.ext) * a.lsr);               // dr >= lsa + (2 - ext) lsr
    const double hi = a.lsa + (double)(n_fft - 1 + a.ext - 2) * a.lsr;  // t <= len - 3
    a.r_cover = (lo < hi ? lo : hi) - 0.01;
    if (a.r_cover < 0.0) a.r_cover = 0.0;
    return a;
}

PixelGrid make_grid(int nu, int nv, double d_u, double d_v, const double *u_hat,
                    const double *v_hat) {
    PixelGrid g{};
    g.on = 1;
    g.u0 = floordiv(-nu, 2);  // img_plane.cpp:12-14
    g.v0 = floordiv(-nv, 2);
    g.d_u = d_u;
    g.d_v = d_v;
    for (int c = 0; c < 3; ++c) {
        g.uh[c] = u_hat[c];
        g.vh[c] = v_hat[c];
    }
    return g;
}

// Taylor window (signal.h:14-51). The arithmetic follows the Halide expression types:
// A's pow/acosh in f32, n_bar and sigma_p in f64, the F_m products in f64, and the
// (m+1)^2/(n+1)^2 ratio of F_m_den as Int(32) division (den_policy 0) or in f32 (1).
int taylor_window(int num, int s_l, int den_policy, double *out) {
    if (num < 2) return -1;
    std::vector<double> xi((size_t)num);
    const double step = (0.5 - (-0.5)) / (double)(num - 1);  // util.h:80-84
    for (int x = 0; x < num; ++x) {
        double v = -0.5 + (double)x * step;
        v = v < 0.5 ? v : 0.5;
        v = v > -0.5 ? v : -0.5;
        xi[x] = v;
    }
    const float expo = (float)((double)s_l * (1.0 / 20));
    const double A = (1.0 / M_PI) * (double)acoshf(powf(10.0f, expo));
    const int n_bar = (int)(2.0 * (A * A) + 0.5) + 1;
    const double nbm = (double)n_bar - 0.5;
    const double sigma_p = (double)n_bar / std::sqrt(A * A + nbm * nbm);
    const int nr = n_bar - 1;
    std::vector<double> F((size_t)(nr > 0 ? nr : 1));
    for (int m = 0; m < nr; ++m) {
        const int m1sq = (m + 1) * (m + 1);
        const float sgn = powf(-1.0f, (float)(m + 2));
        double fnum = 1.0;
        for (int r = 0; r < nr; ++r) {
            const double rh = (double)r + 0.5;
            const double t = (double)m1sq / (sigma_p * sigma_p) / (A * A + rh * rh);
            fnum = fnum * ((double)sgn * (1.0 - t));
        }
        double fden = 1.0;
        for (int r = 0; r < nr; ++r) {
            if (r == m) continue;
            const int r1sq = (r + 1) * (r + 1);
            const double qv = den_policy == SARBP_TAYLOR_DEN_F32
                                  ? (double)((float)m1sq / (float)r1sq)
                                  : (double)(m1sq / r1sq);
            fden = fden * (1.0 - qv);
        }
        F[m] = fnum / fden;
    }
    double wmax = -INFINITY;
    for (int x = 0; x < num; ++x) {
        double w = 1.0;
        for (int r = 0; r < nr; ++r) w = w + F[r] * std::cos((2.0 * M_PI) * (double)(r + 1) * xi[x]);
        out[x] = w;
        wmax = w > wmax ? w : wmax;
    }
    for (int x = 0; x < num; ++x) out[x] = out[x] / wmax;
    return 0;
}

}  // namespace sarbp

using namespace sarbp;

extern "C" const char *sarbp_last_error(void) { return last_error(); }

extern "C" int sarbp_taylor_window(int32_t num, int32_t taylor_s_l, int32_t taylor_den_policy,
                                   double *out) {
    if (!out) {
        set_error("sarbp_taylor_window: out is NULL");
        return halide_error_code_buffer_argument_is_null;
    }
    if (num < 2) {
        set_error("sarbp_taylor_window: num must be >= 2 (got %d)", num);
        return halide_error_code_param_too_small;
    }
    return taylor_window(num, taylor_s_l, taylor_den_policy, out) == 0
               ? 0
               : halide_error_code_generic_error;
}

// img_plane_create (ImgPlane.cpp:16-63) with the generators of img_plane.cpp:12-113 and ip.h.
extern "C" int sarbp_img_plane_create(int32_t nsamples, int32_t npulses, double delta_r,
                                      const float *R_c, const float *n_hat, double res_factor,
                                      double aspect, int32_t upsample, int32_t *dims_out,
                                      double *scal_out, double *u, double *v, double *k_u,
                                      double *k_v, double *u_hat, double *v_hat,
                                      double *pixel_locs) {
    if (nsamples < 1 || npulses < 1) {
        set_error("sarbp_img_plane_create: nsamples and npulses must be >= 1");
        return halide_error_code_param_too_small;
    }
    auto up = [](int n) {  // ip.h:5-9: next power of two >= n
        const double l = std::log2((double)n);
        const int r = (l - (int)l) > 0;
        return (int)std::pow(2.0, (double)(int)(l + r));
    };
    const int nu = upsample ? up(nsamples) : nsamples;
    const int nv = upsample ? up(npulses) : npulses;
    const double d_u = delta_r * res_factor * (double)nsamples / (double)nu;  // ip.h:11-13
    const double d_v = aspect * d_u;                                           // ip.h:15-17
    if (dims_out) {
        dims_out[0] = nu;
        dims_out[1] = nv;
    }
    if (scal_out) {
        scal_out[0] = d_u;
        scal_out[1] = d_v;
    }
    if (!u) return 0;
    if (!v || !k_u || !k_v || !u_hat || !v_hat || !R_c || !n_hat) {
        set_error("sarbp_img_plane_create: NULL output or input array");
        return halide_error_code_buffer_argument_is_null;
    }
    for (int x = 0; x < nu; ++x) u[x] = (double)(floordiv(-nu, 2) + x) * d_u;  // img_plane.cpp:12-14
    for (int x = 0; x < nv; ++x) v[x] = (double)(floordiv(-nv, 2) + x) * d_v;
    auto kaxis = [](int n, double d, double *k) {  // img_plane.cpp:17-19
        const double start = -1.0 / (2.0 * d), stop = 1.0 / (2.0 * d);
        const double step = (stop - start) / (double)(n - 1);
        for (int x = 0; x < n; ++x) {
            double t = start + (double)x * step;
            t = t < stop ? t : stop;
            t = t > start ? t : start;
            k[x] = (2.0 * M_PI) * t;
        }
    };
    kaxis(nu, d_u, k_u);
    kaxis(nv, d_v, k_v);
    // v_hat = f64(cross(n_hat, R_c) / |.|) computed on the f32 inputs (img_plane.cpp:22-27,29-39)
    float cr[3];
    for (int x = 0; x < 3; ++x)
        cr[x] = n_hat[(x + 1) % 3] * R_c[(x + 2) % 3] - n_hat[(x + 2) % 3] * R_c[(x + 1) % 3];
    float n2 = 0.0f;
    for (int x = 0; x < 3; ++x) n2 = n2 + cr[x] * cr[x];
    const float nn = std::sqrt(n2);
    for (int x = 0; x < 3; ++x) v_hat[x] = (double)(cr[x] / nn);
    // u_hat = cross(v_hat, n_hat) / |.| in f64 (img_plane.cpp:41-47)
    double cu[3];
    for (int x = 0; x < 3; ++x)
        cu[x] = v_hat[(x + 1) % 3] * (double)n_hat[(x + 2) % 3] -
                v_hat[(x + 2) % 3] * (double)n_hat[(x + 1) % 3];
    double m2 = 0.0;
    for (int x = 0; x < 3; ++x) m2 = m2 + cu[x] * cu[x];
    const double mn = std::sqrt(m2);
    for (int x = 0; x < 3; ++x) u_hat[x] = cu[x] / mn;
    if (pixel_locs) {  // img_plane.cpp:77-113
        const int64_t npix = (int64_t)nu * nv;
        for (int d = 0; d < 3; ++d) {
            double *dst = pixel_locs + (size_t)d * (size_t)npix;
            for (int64_t y = 0; y < nv; ++y) {
                const double b = v_hat[d] * v[y];
                for (int64_t x = 0; x < nu; ++x) dst[y * nu + x] = (0.0 + u_hat[d] * u[x]) + b;
            }
        }
    }
    return 0;
}
