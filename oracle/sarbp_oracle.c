/*
 * CPU FP64 oracle for the SAR backprojection hot path — TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C, the algorithm of the reference's Halide `backprojection`
 * generator and its helpers (ISI-apex/halide-sar-app). It is the checker for the HIP product
 * path and the CPU baseline timed beside it; nothing in the product links it. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Type promotion follows Halide's rules as the reference expressions trigger them:
 *   - int * f64 Expr -> f64; f32 * f64 -> f64; pow(int_expr, 2) -> int_expr*int_expr (Int(32));
 *   - pow(10, e) with a non-constant exponent -> powf on f32 operands;
 *   - log(Expr(10)) -> logf(10.0f) (int argument cast to Float(32)).
 * Compiled with -ffp-contract=off so every f32/f64 operation rounds once, in program order.
 *
 * Parity notes (what each function is pinned by) are listed in tests/test_oracle.py.
 */
#define _GNU_SOURCE
#include "sarbp_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---------------------------------------------------------------- helpers (util.h) */

/* util.h:80-84: step = (stop - start) / (num - 1); clamp(start + x*step, start, stop) */
void orc_linspace(double start, double stop, int num, double *out) {
    double step = (stop - start) / (double)(num - 1);
    for (int x = 0; x < num; x++) {
        double v = start + (double)x * step;
        /* Halide clamp(a, lo, hi) = max(min(a, hi), lo) */
        v = v < stop ? v : stop;
        v = v > start ? v : start;
        out[x] = v;
    }
}

/* signal_complex.h:24-29: in((x + x_mid) % x_extent) with x_mid = x_extent/2 + x_extent%2 */
int orc_fftshift_src(int i, int n) {
    int mid = n / 2 + n % 2;
    return (i + mid) % n;
}

/* sarbp.cpp:258: N_fft = pow(2, int(log2(nsamples * upsample)) + 1) */
int orc_nfft(int nsamples, int upsample) {
    return (int)pow(2.0, (double)((int)log2((double)(nsamples * upsample)) + 1));
}

/* ---------------------------------------------------------------- Taylor window (signal.h:14-51) */

int orc_taylor(int num, double S_L, int int_den, double *out) {
    if (num < 2) return -1;
    double *xi = (double *)malloc(sizeof(double) * (size_t)num);
    if (!xi) return -1;
    /* signal.h:27 */
    orc_linspace(-0.5, 0.5, num, xi);
    /* signal.h:28: Expr(1/pi) * acosh(pow(10, S_L * Expr(1.0/20))); pow(int, non-const) is f32 */
    float expo = (float)(S_L * (1.0 / 20));
    float pw = powf(10.0f, expo);
    double A = (1.0 / (double)M_PI) * (double)acoshf(pw);
    /* signal.h:29: i32(2 * pow(A, 2) + 0.5) + 1 */
    int n_bar = (int)(2.0 * (A * A) + 0.5) + 1;
    /* signal.h:30 */
    double nbm = (double)n_bar - 0.5;
    double sigma_p = (double)n_bar / sqrt(A * A + nbm * nbm);
    int nr = n_bar - 1; /* RDom r(0, n_bar - 1), signal.h:36 */
    double *F = (double *)malloc(sizeof(double) * (size_t)(nr > 0 ? nr : 1));
    if (!F) {
        free(xi);
        return -1;
    }
    for (int m = 0; m < nr; m++) {
        /* signal.h:37,39: F_m_num(x) *= pow(-1, x+2) * (1 - (x+1)^2 / sigma^2 / (A^2 + (r+0.5)^2)) */
        double fnum = 1.0;
        float sgn = powf(-1.0f, (float)(m + 2));
        int m1sq = (m + 1) * (m + 1);
        for (int r = 0; r < nr; r++) {
            double rh = (double)r + 0.5;
            double t = (double)m1sq / (sigma_p * sigma_p) / (A * A + rh * rh);
            fnum = fnum * ((double)sgn * (1.0 - t));
        }
        /* signal.h:38,40: F_m_den(x) = select(x == r, F, F * (1 - (x+1)^2 / (r+1)^2)) */
        double fden = 1.0;
        for (int r = 0; r < nr; r++) {
            if (m == r) continue;
            int r1sq = (r + 1) * (r + 1);
            double q;
            if (int_den) {
                q = (double)(m1sq / r1sq); /* Int(32) / Int(32): both positive, floor == trunc */
            } else {
                q = (double)((float)m1sq / (float)r1sq);
            }
            fden = fden * (1.0 - q);
        }
        F[m] = fnum / fden; /* signal.h:41 */
    }
    /* signal.h:44-45: w(x) = 1 + sum_r F_m(r) * cos(Expr(2 pi) * (r + 1) * xi) */
    double wmax = -INFINITY;
    for (int x = 0; x < num; x++) {
        double w = 1.0;
        for (int r = 0; r < nr; r++) {
            double arg = (2.0 * (double)M_PI) * (double)(r + 1) * xi[x];
            w = w + F[r] * cos(arg);
        }
        out[x] = w;
        if (w > wmax) wmax = w;
    }
    /* signal.h:48 */
    for (int x = 0; x < num; x++) out[x] = out[x] / wmax;
    free(F);
    free(xi);
    return 0;
}

/* ---------------------------------------------------------------- FFT (dft.cpp:40-79 via FFTW_FORWARD) */

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* Halide integer division rounds toward negative infinity */
static int floordiv_i(int a, int b) {
    int q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q--;
    return q;
}

int orc_fft_f64(double *data, int n) {
    if (n <= 0) return -1;
    if (!is_pow2(n)) {
        /* direct definition X_k = sum_j x_j exp(-2 pi i j k / n), the exponent reduced to
           (j k) mod n and its cos / sin evaluated once per residue */
        double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)n);
        double *cs = (double *)malloc(sizeof(double) * 2 * (size_t)n);
        if (!tmp || !cs) {
            free(tmp);
            free(cs);
            return -1;
        }
        for (int m = 0; m < n; m++) {
            double a = -2.0 * M_PI * (double)m / (double)n;
            cs[2 * m] = cos(a);
            cs[2 * m + 1] = sin(a);
        }
        for (int k = 0; k < n; k++) {
            double sr = 0.0, si = 0.0;
            long long jk = 0; /* (j k) mod n, advanced by k per step */
            for (int j = 0; j < n; j++) {
                double c = cs[2 * jk], s = cs[2 * jk + 1];
                sr += data[2 * j] * c - data[2 * j + 1] * s;
                si += data[2 * j] * s + data[2 * j + 1] * c;
                jk += k;
                if (jk >= n) jk -= n;
            }
            tmp[2 * k] = sr;
            tmp[2 * k + 1] = si;
        }
        memcpy(data, tmp, sizeof(double) * 2 * (size_t)n);
        free(tmp);
        free(cs);
        return 0;
    }
    /* iterative radix-2 decimation in time with bit-reversed input order */
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double tr = data[2 * i], ti = data[2 * i + 1];
            data[2 * i] = data[2 * j];
            data[2 * i + 1] = data[2 * j + 1];
            data[2 * j] = tr;
            data[2 * j + 1] = ti;
        }
    }
    /* twiddles w_k = exp(-2 pi i k / n), each evaluated directly (no recurrence drift) */
    double *tw = (double *)malloc(sizeof(double) * (size_t)(n > 1 ? n : 2));
    if (!tw) return -1;
    for (int k = 0; k < n / 2; k++) {
        double a = -2.0 * M_PI * (double)k / (double)n;
        tw[2 * k] = cos(a);
        tw[2 * k + 1] = sin(a);
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, tstep = n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < half; k++) {
                double wr = tw[2 * k * tstep], wi = tw[2 * k * tstep + 1];
                int a = i + k, b = i + k + half;
                double br = data[2 * b] * wr - data[2 * b + 1] * wi;
                double bi = data[2 * b] * wi + data[2 * b + 1] * wr;
                double ar = data[2 * a], ai = data[2 * a + 1];
                data[2 * a] = ar + br;
                data[2 * a + 1] = ai + bi;
                data[2 * b] = ar - br;
                data[2 * b + 1] = ai - bi;
            }
        }
    }
    free(tw);
    return 0;
}

/* ---------------------------------------------------------------- norm_r0 (backprojection.cpp:108) */

void orc_norm_r0(const float *pos, int np, int policy, float *out32, double *out64) {
    for (int p = 0; p < np; p++) {
        float x = pos[3 * p], y = pos[3 * p + 1], z = pos[3 * p + 2];
        double v;
        if (policy == ORC_NORM_R0_F64) {
            double dx = x, dy = y, dz = z;
            v = sqrt(dx * dx + dy * dy + dz * dz);
            if (out32) out32[p] = (float)v;
        } else if (policy == ORC_NORM_R0_F32_FMA) {
            float s = fmaf(z, z, fmaf(y, y, x * x));
            float n = sqrtf(s);
            v = n;
            if (out32) out32[p] = n;
        } else {
            float xx = x * x, yy = y * y, zz = z * z;
            float s = (xx + yy) + zz;
            float n = sqrtf(s);
            v = n;
            if (out32) out32[p] = n;
        }
        if (out64) out64[p] = v;
    }
}

/* ---------------------------------------------------------------- range profiles (backprojection.cpp:80-104) */

int orc_range_profiles(const float *phs, const float *k_r, int ns, int np, int taylor_s_l, int N,
                       int int_den, int pulse_shift_quirk, double *Q) {
    if (ns < 2 || np < 2 || N < 1) return -1;
    double *ws = (double *)malloc(sizeof(double) * (size_t)ns);
    double *wp = (double *)malloc(sizeof(double) * (size_t)np);
    double *row = (double *)malloc(sizeof(double) * 2 * (size_t)N);
    if (!ws || !wp || !row) {
        free(ws);
        free(wp);
        free(row);
        return -1;
    }
    orc_taylor(ns, (double)taylor_s_l, int_den, ws); /* backprojection.cpp:80 */
    orc_taylor(np, (double)taylor_s_l, int_den, wp); /* backprojection.cpp:81 */
    int x_pad = floordiv_i(N - ns, 2);                /* signal_complex.h:14 (Halide floor division) */
    int pmid = np / 2 + np % 2;
    for (int p = 0; p < np; p++) {
        /* Q(k, p) = dft(shift_N(k), shift_np(p)); dft row q transforms fftsh(., q) which reads
         * phs_pad(shift_N(x), shift_np(q)) — the pulse axis is rolled twice. */
        int q = pulse_shift_quirk ? (p + pmid) % np : p;
        int s = pulse_shift_quirk ? (q + pmid) % np : p;
        for (int x = 0; x < N; x++) {
            int xs = orc_fftshift_src(x, N); /* fftsh, backprojection.cpp:98 */
            double re = 0.0, im = 0.0;
            if (xs >= x_pad && xs < ns + x_pad) { /* pad, signal_complex.h:17-20 */
                int i = xs - x_pad;
                float filt = fabsf(k_r[i]);                         /* backprojection.cpp:85 */
                float fr = phs[2 * ((size_t)s * ns + i)] * filt;    /* f32 * f32 -> f32 */
                float fi = phs[2 * ((size_t)s * ns + i) + 1] * filt;
                double win = ws[i] * wp[s];                         /* backprojection.cpp:82 */
                re = (double)fr * win;                              /* backprojection.cpp:90 */
                im = (double)fi * win;
            }
            row[2 * x] = re;
            row[2 * x + 1] = im;
        }
        orc_fft_f64(row, N); /* backprojection.cpp:101 */
        double *dst = Q + 2 * (size_t)p * N;
        for (int k = 0; k < N; k++) { /* backprojection.cpp:104 */
            int ks = orc_fftshift_src(k, N);
            dst[2 * k] = row[2 * ks];
            dst[2 * k + 1] = row[2 * ks + 1];
        }
    }
    free(ws);
    free(wp);
    free(row);
    return 0;
}

/* ---------------------------------------------------------------- accumulation (backprojection.cpp:52-138) */

typedef struct {
    const double *Q;
    int N, ns, np;
    double lsa, lsr;
    float k_c;
    const float *pos;
    const double *nr0;
    const double *r;
    int64_t r_extent;
    int nu, nv;
    /* region mode */
    int x0, y0, w, h;
    /* point mode */
    const int32_t *xs, *ys;
    int count;
    double *out;
    long long ntasks;
    long long next; /* atomic task counter */
} bp_job;

/* r = constant_exterior(r_in, 0.0) (backprojection.cpp:77) */
static inline double r_at(const bp_job *J, long long px, int d) {
    if (px < 0 || px >= J->r_extent) return 0.0;
    return J->r[(size_t)d * (size_t)J->r_extent + (size_t)px];
}

/* dr_i(pixel, pulse) = norm_r0(pulse) - ||r(pixel) - pos(pulse)|| (backprojection.cpp:111-117) */
static inline double dr_at(const bp_job *J, double rx, double ry, double rz, int p) {
    double dx = rx - (double)J->pos[3 * p];
    double dy = ry - (double)J->pos[3 * p + 1];
    double dz = rz - (double)J->pos[3 * p + 2];
    double n = sqrt(dx * dx + dy * dy + dz * dz);
    return J->nr0[p] - n;
}

/* Q_hat: linear interpolation on the implicit linspace axis (backprojection.cpp:52-65,120) */
static inline void qhat(const bp_job *J, int p, double dr, double *qr, double *qi) {
    double luts = (dr - J->lsa) / J->lsr;
    double fl = floor(luts);
    int lutl;
    if (fl < -2147483648.0) lutl = INT32_MIN;
    else if (fl > 2147483647.0) lutl = INT32_MAX;
    else lutl = (int)fl;
    double luto = luts - (double)lutl;
    int cll = lutl < 0 ? 0 : (lutl > J->N - 1 ? J->N - 1 : lutl);
    int lutu = lutl == INT32_MAX ? lutl : lutl + 1;
    int clu = lutu < 0 ? 0 : (lutu > J->N - 1 ? J->N - 1 : lutu);
    double wgt = luto < 1.0 ? luto : 1.0;
    wgt = wgt > 0.0 ? wgt : 0.0;
    const double *row = J->Q + 2 * (size_t)p * J->N;
    /* Halide lerp on floats: zero_val * (1 - w) + one_val * w */
    double omw = 1.0 - wgt;
    *qr = row[2 * cll] * omw + row[2 * clu] * wgt;
    *qi = row[2 * cll + 1] * omw + row[2 * clu + 1] * wgt;
}

/* One output pixel; stage holds dr for every pulse (norm_rr0 staging of the CPU schedule). */
static void bp_pixel(const bp_job *J, int x, int y, double *stage, double *dst) {
    long long px = (long long)J->nu * (long long)(J->nv - y - 1) + x; /* backprojection.cpp:130 */
    double rx = r_at(J, px, 0), ry = r_at(J, px, 1), rz = r_at(J, px, 2);
    for (int p = 0; p < J->np; p++) stage[p] = dr_at(J, rx, ry, rz, p);
    double ar = 0.0, ai = 0.0; /* img(pixel) = 0 (backprojection.cpp:126) */
    double nk = -(double)J->k_c;
    for (int p = 0; p < J->np; p++) { /* backprojection.cpp:127 */
        double dr = stage[p];
        double qr, qi;
        qhat(J, p, dr, &qr, &qi);
        double ph = nk * dr;
        double c = cos(ph), s = sin(ph);
        ar = ar + (qr * c - qi * s);
        ai = ai + (qr * s + qi * c);
    }
    /* fimg = img * expj(k_c * dr_i(pixel, npulses/2)) (backprojection.cpp:135) */
    double ph = (double)J->k_c * dr_at(J, rx, ry, rz, J->np / 2);
    double c = cos(ph), s = sin(ph);
    dst[0] = ar * c - ai * s;
    dst[1] = ar * s + ai * c;
}

static void *bp_worker(void *arg) {
    bp_job *J = (bp_job *)arg;
    double *stage = (double *)malloc(sizeof(double) * (size_t)J->np);
    if (!stage) return NULL;
    for (;;) {
        long long t = __atomic_fetch_add(&J->next, 1, __ATOMIC_RELAXED);
        if (t >= J->ntasks) break;
        if (J->xs) {
            /* point mode: tasks of 4 listed points */
            for (int j = 0; j < 4; j++) {
                long long i = 4 * t + j;
                if (i >= J->count) break;
                bp_pixel(J, J->xs[i], J->ys[i], stage, J->out + 2 * i);
            }
        } else {
            /* region mode: task = (row, 4-pixel group), rows outermost (backprojection.cpp:311-314) */
            long long groups = (J->w + 3) / 4;
            int yy = (int)(t / groups);
            int g = (int)(t % groups);
            for (int j = 0; j < 4; j++) {
                int xx = 4 * g + j;
                if (xx >= J->w) break;
                double *dst = J->out + 2 * ((size_t)yy * J->w + xx);
                bp_pixel(J, J->x0 + xx, J->y0 + yy, stage, dst);
            }
        }
    }
    free(stage);
    return NULL;
}

static int run_job(bp_job *J, int nthreads) {
    if (nthreads <= 1) {
        bp_worker(J);
        return 0;
    }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    if (!th) return -1;
    int started = 0;
    for (int i = 0; i < nthreads; i++) {
        if (pthread_create(&th[i], NULL, bp_worker, J) != 0) break;
        started++;
    }
    if (started == 0) bp_worker(J);
    for (int i = 0; i < started; i++) pthread_join(th[i], NULL);
    free(th);
    return 0;
}

static void job_init(bp_job *J, const double *Q, int N, int ns, int np, double delta_r, float k_c,
                     const float *pos, const double *nr0, const double *r, int64_t r_extent, int nu,
                     int nv) {
    memset(J, 0, sizeof(*J));
    J->Q = Q;
    J->N = N;
    J->ns = ns;
    J->np = np;
    /* backprojection.cpp:120: lsa = floor(-ns*dr/2), lsb = floor(ns*dr/2), lsn = N;
     * backprojection.cpp:53: lsr = (lsb - lsa) / (lsn - 1) */
    double lsa = floor((double)(-ns) * delta_r / 2.0);
    double lsb = floor((double)ns * delta_r / 2.0);
    J->lsa = lsa;
    J->lsr = (lsb - lsa) / (double)(N - 1);
    J->k_c = k_c;
    J->pos = pos;
    J->nr0 = nr0;
    J->r = r;
    J->r_extent = r_extent;
    J->nu = nu;
    J->nv = nv;
}

int orc_backproject(const double *Q, int N, int ns, int np, double delta_r, float k_c,
                    const float *pos, const double *nr0, const double *r, int64_t r_extent,
                    int nu, int nv, int x0, int y0, int w, int h, int nthreads, double *out) {
    if (N < 2 || np < 1 || w < 0 || h < 0) return -1;
    bp_job J;
    job_init(&J, Q, N, ns, np, delta_r, k_c, pos, nr0, r, r_extent, nu, nv);
    J.x0 = x0;
    J.y0 = y0;
    J.w = w;
    J.h = h;
    J.out = out;
    J.ntasks = (long long)h * ((w + 3) / 4);
    return run_job(&J, nthreads);
}

int orc_backproject_points(const double *Q, int N, int ns, int np, double delta_r, float k_c,
                           const float *pos, const double *nr0, const double *r, int64_t r_extent,
                           int nu, int nv, const int32_t *xs, const int32_t *ys, int count,
                           int nthreads, double *out) {
    if (N < 2 || np < 1 || count < 0) return -1;
    bp_job J;
    job_init(&J, Q, N, ns, np, delta_r, k_c, pos, nr0, r, r_extent, nu, nv);
    J.xs = xs;
    J.ys = ys;
    J.count = count;
    J.out = out;
    J.ntasks = (count + 3) / 4;
    return run_job(&J, nthreads);
}

int orc_backprojection(const float *phs, const float *k_r, int ns, int np, int taylor_s_l, int N,
                       double delta_r, int nu, int nv, const float *pos, const double *r,
                       int norm_policy, int int_den, int pulse_shift_quirk, int nthreads,
                       double *out) {
    double *Q = (double *)malloc(sizeof(double) * 2 * (size_t)N * (size_t)np);
    double *nr0 = (double *)malloc(sizeof(double) * (size_t)np);
    if (!Q || !nr0) {
        free(Q);
        free(nr0);
        return -1;
    }
    int rv = orc_range_profiles(phs, k_r, ns, np, taylor_s_l, N, int_den, pulse_shift_quirk, Q);
    if (rv == 0) {
        orc_norm_r0(pos, np, norm_policy, NULL, nr0);
        float k_c = k_r[ns / 2]; /* backprojection.cpp:123 */
        rv = orc_backproject(Q, N, ns, np, delta_r, k_c, pos, nr0, r, (int64_t)nu * nv, nu, nv, 0,
                             0, nu, nv, nthreads, out);
    }
    free(Q);
    free(nr0);
    return rv;
}

/* ---------------------------------------------------------------- image output (img_output.cpp) */

void orc_img_output_to_dB(const double *img, int64_t npix, double *out) {
    /* img_output.cpp:24: m() = maximum(abs(cimg)); abs = sqrt(re^2 + im^2) (halide_complexfunc.h:284) */
    double m = -INFINITY;
    for (int64_t i = 0; i < npix; i++) {
        double a = sqrt(img[2 * i] * img[2 * i] + img[2 * i + 1] * img[2 * i + 1]);
        if (a > m) m = a;
    }
    /* img_output.cpp:26: 10 * log10(|z| / m); util.h:56-58 log10(x) = log(x) / log(Expr(10)),
     * and log of the Int(32) constant 10 is evaluated in f32 */
    double ln10 = (double)logf(10.0f);
    for (int64_t i = 0; i < npix; i++) {
        double a = sqrt(img[2 * i] * img[2 * i] + img[2 * i + 1] * img[2 * i + 1]);
        out[i] = 10.0 * (log(a / m) / ln10);
    }
}

void orc_img_output_u8(const double *dB, int64_t npix, double dB_min, double dB_max, uint8_t *out) {
    for (int64_t i = 0; i < npix; i++) {
        double v = dB[i];
        /* util.h:13-15: normalize = (clamp(v, lo, hi) - lo) / (hi - lo) */
        double c = v < dB_max ? v : dB_max;
        c = c > dB_min ? c : dB_min;
        double nrm = (c - dB_min) / (dB_max - dB_min);
        /* signal.h:55-57: clamp(normalize * 255, 0, 255); img_output.cpp:50: u8 truncating cast */
        double s = nrm * 255.0;
        s = s < 255.0 ? s : 255.0;
        s = s > 0.0 ? s : 0.0;
        out[i] = (s == s) ? (uint8_t)s : 0; /* NaN (all-zero image) maps to 0 */
    }
}

/* ---------------------------------------------------------------- image plane (ImgPlane.cpp, img_plane.cpp, ip.h) */

/* ip.h:5-9 */
static int ip_upsample(int n) {
    double l = log2((double)n);
    int r = (l - (int)l) > 0;
    return (int)pow(2.0, (double)(int)(l + r));
}

int orc_img_plane(int nsamples, int npulses, double delta_r, const float *R_c, const float *n_hat,
                  double res_factor, double aspect, int upsample, int *dims_out, double *scal_out,
                  double *u, double *v, double *k_u, double *k_v, double *u_hat, double *v_hat,
                  double *pixel_locs) {
    /* ImgPlane.cpp:20-26 */
    int nu = upsample ? ip_upsample(nsamples) : nsamples;
    int nv = upsample ? ip_upsample(npulses) : npulses;
    /* ip.h:11-17 */
    double d_u = delta_r * res_factor * (double)nsamples / (double)nu;
    double d_v = aspect * d_u;
    if (dims_out) {
        dims_out[0] = nu;
        dims_out[1] = nv;
    }
    if (scal_out) {
        scal_out[0] = d_u;
        scal_out[1] = d_v;
    }
    if (!u) return 0;
    /* img_plane.cpp:12-14: arange(-n/2, 1, x) * d, with Halide floor division */
    for (int x = 0; x < nu; x++) u[x] = (double)(floordiv_i(-nu, 2) + x) * d_u;
    for (int x = 0; x < nv; x++) v[x] = (double)(floordiv_i(-nv, 2) + x) * d_v;
    /* img_plane.cpp:17-19: 2 pi * linspace(-1/(2d), 1/(2d), n) */
    double *tmp = (double *)malloc(sizeof(double) * (size_t)(nu > nv ? nu : nv));
    if (!tmp) return -1;
    orc_linspace(-1.0 / (2.0 * d_u), 1.0 / (2.0 * d_u), nu, tmp);
    for (int x = 0; x < nu; x++) k_u[x] = (2.0 * M_PI) * tmp[x];
    orc_linspace(-1.0 / (2.0 * d_v), 1.0 / (2.0 * d_v), nv, tmp);
    for (int x = 0; x < nv; x++) k_v[x] = (2.0 * M_PI) * tmp[x];
    free(tmp);
    /* img_plane.cpp:22-27 + util.h:18-26: v_hat = f64(cross(n_hat, R_c) / norm) in f32 */
    float cr[3];
    for (int x = 0; x < 3; x++)
        cr[x] = n_hat[(x + 1) % 3] * R_c[(x + 2) % 3] - n_hat[(x + 2) % 3] * R_c[(x + 1) % 3];
    float n2 = 0.0f;
    for (int x = 0; x < 3; x++) n2 = n2 + cr[x] * cr[x];
    float nn = sqrtf(n2);
    for (int x = 0; x < 3; x++) v_hat[x] = (double)(cr[x] / nn);
    /* u_hat = cross(v_hat, n_hat) / norm in f64 */
    double cu[3];
    for (int x = 0; x < 3; x++)
        cu[x] = v_hat[(x + 1) % 3] * (double)n_hat[(x + 2) % 3] -
                v_hat[(x + 2) % 3] * (double)n_hat[(x + 1) % 3];
    double m2 = 0.0;
    for (int x = 0; x < 3; x++) m2 = m2 + cu[x] * cu[x];
    double mn = sqrt(m2);
    for (int x = 0; x < 3; x++) u_hat[x] = cu[x] / mn;
    if (pixel_locs) {
        /* img_plane.cpp:77-113: pixel_locs(x, d) = 0 + u_hat[d]*u[x % nu] + v_hat[d]*v[x / nu] */
        int64_t npix = (int64_t)nu * nv;
        for (int d = 0; d < 3; d++) {
            for (int64_t x = 0; x < npix; x++) {
                double a = u_hat[d] * u[x % nu];
                double b = v_hat[d] * v[x / nu];
                pixel_locs[(size_t)d * (size_t)npix + (size_t)x] = (0.0 + a) + b;
            }
        }
    }
    return 0;
}
