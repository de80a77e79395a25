// The reference's Halide-buffer entry points (backprojection, backprojection_cuda,
// img_output_to_dB, img_output_u8, dft_init_fftw, dft_destroy_fftw) over device-resident plans.
//
// Each call validates the buffers the way Halide-generated code would (types, dimensions, regions),
// answers bounds queries, packs the host inputs into pinned staging memory, runs the pipeline on the
// GPU and scatters the result into the caller's (possibly strided, possibly partial) output.
// backprojection() splits the output into row bands so that host packing, uploads of each band's
// pixel locations, the accumulation (which writes each band into page-locked host memory) and the
// scatter of different bands overlap; bands alternate between two compute streams so that one
// band's last wave overlaps the next band.
#include <hip/hip_runtime.h>
#include <sched.h>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "abi_entry.h"
#include "internal.h"

using namespace sarbp;

namespace {

bool is_bq(const halide_buffer_t *b) { return b->host == nullptr && b->device == 0; }

bool has_type(const halide_buffer_t *b, uint8_t code, uint8_t bits) {
    return b->type.code == code && b->type.bits == bits && b->type.lanes == 1;
}

// Byte offset of coordinate c[] within a buffer (Halide: sum (c_i - min_i) * stride_i elements).
inline int64_t offset_of(const halide_buffer_t *b, const int64_t *c) {
    int64_t off = 0;
    for (int i = 0; i < b->dimensions; ++i)
        off += (c[i] - b->dim[i].min) * (int64_t)b->dim[i].stride;
    return off;
}

inline bool in_dim(const halide_buffer_t *b, int i, int64_t c) {
    return c >= b->dim[i].min && c < (int64_t)b->dim[i].min + b->dim[i].extent;
}

bool covers(const halide_buffer_t *b, int i, int64_t lo, int64_t hi) {
    return b->dim[i].min <= lo && (int64_t)b->dim[i].min + b->dim[i].extent >= hi;
}

void set_dense(halide_buffer_t *b, std::initializer_list<std::pair<int, int>> mins_exts) {
    int i = 0;
    int64_t stride = 1;
    for (auto me : mins_exts) {
        b->dim[i].min = me.first;
        b->dim[i].extent = me.second;
        b->dim[i].stride = (int32_t)stride;
        stride *= me.second;
        ++i;
    }
}

// Growable device / pinned-host buffers of one device. Pinned buffers take hipHostMalloc flags.
#ifndef SARBP_OUT_HOST_FLAGS
#define SARBP_OUT_HOST_FLAGS hipHostMallocCoherent
#endif
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    bool pinned = false;
    unsigned host_flags = hipHostMallocDefault;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (p) {
            if (pinned) (void)hipHostFree(p);
            else (void)hipFree(p);
            p = nullptr;
            cap = 0;
        }
        size_t want = bytes + bytes / 8;
        hipError_t e = pinned ? hipHostMalloc(&p, want, host_flags) : hipMalloc(&p, want);
        if (e == hipSuccess) cap = want;
        else p = nullptr;
        return e;
    }
};

struct Context {
    std::mutex mu;
    int device = 0;
    hipStream_t stream = nullptr;   // computation (range profiles, even bands)
    hipStream_t stream3 = nullptr;  // computation (odd bands)
    hipStream_t stream2 = nullptr;  // uploads of the bands' pixel locations
    std::vector<hipEvent_t> events;
    hipError_t ensure_events(size_t n) {
        while (events.size() < n) {
            hipEvent_t e;
            hipError_t rv = hipEventCreateWithFlags(&e, hipEventDisableTiming);
            if (rv != hipSuccess) return rv;
            events.push_back(e);
        }
        return hipSuccess;
    }
    DevBuf d_aux, d_out;  // device staging (d_out: img_output_*)
    // Inputs are packed into h_in and uploaded. The accumulation writes each band straight into
    // h_out (page-locked, mapped into the device's address space; fine-grained, so the writes
    // leave as the kernel makes them): no download, and no copy kernel competing with the next
    // band for compute units.
    DevBuf h_in{nullptr, 0, true}, h_out{nullptr, 0, true, SARBP_OUT_HOST_FLAGS};
    std::map<std::tuple<int, int, int, int, int, int, int, int, int, int, int, int, long long,
                        long long, double>,
             sarbp_plan *>
        plans;
    ~Context() {
        for (auto &kv : plans) sarbp_plan_destroy(kv.second);
    }
};

std::mutex g_ctx_mu;
std::map<int, std::unique_ptr<Context>> g_ctx;

int get_context(Context **out) {
    int ndev = sarbp_device_count();
    if (ndev <= 0) {
        set_error("no HIP device is visible; the HIP backprojection path cannot run");
        return halide_error_code_no_device_interface;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    auto &slot = g_ctx[dev];
    if (!slot) {
        auto c = std::make_unique<Context>();
        c->device = dev;
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&c->stream3, hipStreamNonBlocking) != hipSuccess) {
            set_error("stream creation failed");
            return halide_error_code_device_run_failed;
        }
        slot = std::move(c);
    }
    *out = slot.get();
    return 0;
}

// Host work of a call spread over up to 16 threads (at most the usable cores; one per 2 MB of
// work): the copies between the caller's buffers and
// the pinned staging (pageable <-> page-locked copies are bound by one core's copy rate otherwise:
// ~28 GB/s on one thread, ~100 GB/s on eight) and the check of pixel locations against the image
// plane. The helper threads live for the process (a C2 call makes ~35 such copies; a fresh set of
// threads per copy cost thread start-up each time). Each batch waits only for its own pieces, so
// callers on several host threads may share the pool.
class HostPool {
  public:
    static constexpr size_t kMaxThreads = 16;

    static HostPool &get() {
        static HostPool pool;
        return pool;
    }

    // Runs fn(0) .. fn(parts - 1), part 0 on the calling thread and the rest on the helpers.
    template <typename F>
    void run(size_t parts, const F &fn) {
        Batch batch;
        Thunk thunk{&fn, [](const void *f, size_t i) { (*static_cast<const F *>(f))(i); }};
        {
            std::lock_guard<std::mutex> lk(mu_);
            start_workers(parts - 1);
            for (size_t i = 1; i < parts; ++i) {
                tasks_.push_back({thunk, i, &batch});
                ++batch.left;
            }
        }
        work_cv_.notify_all();
        fn(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return batch.left == 0; });
    }

    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        work_cv_.notify_all();
        for (auto &t : workers_) t.join();
    }

  private:
    struct Batch {
        size_t left = 0;  // pieces not yet done (guarded by mu_)
    };
    struct Thunk {
        const void *fn;
        void (*call)(const void *, size_t);
    };
    struct Task {
        Thunk thunk;
        size_t index;
        Batch *batch;
    };

    void start_workers(size_t want) {  // mu_ held
        while (workers_.size() < std::min(want, kMaxThreads - 1))
            workers_.emplace_back([this] { work(); });
    }

    void work() {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            work_cv_.wait(lk, [&] { return stop_ || !tasks_.empty(); });
            if (tasks_.empty()) return;  // stop_ with nothing left
            const Task t = tasks_.back();
            tasks_.pop_back();
            lk.unlock();
            t.thunk.call(t.thunk.fn, t.index);
            lk.lock();
            if (--t.batch->left == 0) done_cv_.notify_all();
        }
    }

    std::mutex mu_;
    std::condition_variable work_cv_, done_cv_;
    std::vector<Task> tasks_;
    std::vector<std::thread> workers_;
    bool stop_ = false;
};

// Pieces for n bytes of host work: one per 2 MB, at most the pool's threads and the host's cores.
// CPUs this process may run on (its affinity mask; the machine's count where that is unknown).
size_t usable_cpus() {
    static const size_t n = [] {
        cpu_set_t set;
        if (sched_getaffinity(0, sizeof set, &set) == 0 && CPU_COUNT(&set) > 0)
            return (size_t)CPU_COUNT(&set);
        return (size_t)std::max(1u, std::thread::hardware_concurrency());
    }();
    return n;
}

size_t host_parts(size_t n) {
    constexpr size_t kMin = (size_t)2 << 20;
    return std::max<size_t>(1, std::min<size_t>({HostPool::kMaxThreads, usable_cpus(), n / kMin}));
}

void par_copy(void *dst, const void *src, size_t n) {
    const size_t parts = host_parts(n);
    if (parts <= 1) {
        std::memcpy(dst, src, n);
        return;
    }
    const size_t chunk = (n / parts + 63) & ~(size_t)63;
    HostPool::get().run(parts, [&](size_t i) {
        const size_t lo = i * chunk;
        if (lo < n) std::memcpy((char *)dst + lo, (const char *)src + lo, std::min(chunk, n - lo));
    });
}

// The image plane of img_plane_create (img_plane.cpp:12-14,77-113), which pixel_locs usually
// holds: u[x] = (floor(-nu/2) + x) d_u, v[y] likewise, and for grid row y and column x (pixel
// x + nu y) r = (0 + u_hat u[x]) + v_hat v[y], rounded in that order. The reference reads only
// the extents of u and v (backprojection.cpp:71-72), but callers pass the plane's own arrays. When
// they are that linspace, u_hat and v_hat follow from the two pixels where v or u is 0, whose
// locations are single products. A band of the call whose every pixel location lies within tol of
// the plane then runs in grid mode: the accumulation generates the locations (sarbp_plan_set_grid)
// and takes the cubic form of the range expansion, and the band's pixel locations are neither
// copied nor uploaded. Any other band reads r as before. tol = 2^-40 of the plane's extent (3e-10 m
// for a 300 m scene: a range error of ~1e-7 rad of phase at X band).
struct PlaneGuess {
    bool on = false;
    double d_u = 0.0, d_v = 0.0, uh[3] = {}, vh[3] = {}, tol = 0.0;
    std::vector<double> u, v;
    std::vector<double> uu[3];  // u_hat[c] u[x], the plane's first product
    const double *rc[3] = {};  // pixel i's component c at rc[c][i - rmin]
    int64_t rmin = 0;
    int nu = 0;
};

bool linspace_of(const halide_buffer_t *b, int n, std::vector<double> &out, double &step) {
    if (n < 2 || !b->host) return false;
    out.resize((size_t)n);
    const double *h = (const double *)b->host;
    for (int x = 0; x < n; ++x) {
        const int64_t co[1] = {b->dim[0].min + x};
        out[x] = h[offset_of(b, co)];
    }
    const int k0 = -((n + 1) / 2);                   // floor(-n / 2) (img_plane.cpp:12-14)
    const int xa = -k0 + 1 < n ? -k0 + 1 : -k0 - 1;  // the sample at +-1 step
    step = out[xa] * (double)(k0 + xa);              // k0 + xa = +-1
    if (!(std::fabs(step) > 0.0) || !std::isfinite(step)) return false;
    for (int x = 0; x < n; ++x)
        if (out[x] != (double)(k0 + x) * step) return false;
    return true;
}

void guess_plane(const halide_buffer_t *u, const halide_buffer_t *v, const halide_buffer_t *r,
                 int nu, int nv, int64_t rmin, int64_t rext, PlaneGuess &g) {
    g.on = false;
    if (!r->host || r->dim[0].stride != 1) return;
    for (int c = 0; c < 3; ++c)
        if (!in_dim(r, 1, c)) return;
    if (!linspace_of(u, nu, g.u, g.d_u) || !linspace_of(v, nv, g.v, g.d_v)) return;
    const int x0 = (nu + 1) / 2, y0 = (nv + 1) / 2;  // u[x0] = 0, v[y0] = 0
    const int xa = x0 + 1 < nu ? x0 + 1 : x0 - 1, ya = y0 + 1 < nv ? y0 + 1 : y0 - 1;
    const int64_t ia = (int64_t)y0 * nu + xa, ib = (int64_t)ya * nu + x0;
    if (ia < rmin || ia >= rmin + rext || ib < rmin || ib >= rmin + rext) return;
    double umax = 0.0, vmax = 0.0;
    for (double a : g.u) umax = std::max(umax, std::fabs(a));
    for (double a : g.v) vmax = std::max(vmax, std::fabs(a));
    const double *rs = (const double *)r->host;
    for (int c = 0; c < 3; ++c) {
        const int64_t co[2] = {rmin, c};
        g.rc[c] = rs + offset_of(r, co);
        g.uh[c] = g.rc[c][ia - rmin] / g.u[xa];  // r(xa, y0) = u_hat u[xa]
        g.vh[c] = g.rc[c][ib - rmin] / g.v[ya];  // r(x0, ya) = v_hat v[ya]
        if (!std::isfinite(g.uh[c]) || !std::isfinite(g.vh[c])) return;
    }
    for (int c = 0; c < 3; ++c) {
        g.uu[c].resize((size_t)nu);
        for (int x = 0; x < nu; ++x) g.uu[c][x] = g.uh[c] * g.u[x];
    }
    g.tol = std::ldexp(umax + vmax, -40);
    g.rmin = rmin;
    g.nu = nu;
    g.on = true;
}

// Whether pixels [lo, hi) of r lie within g.tol of the plane (checked on the host pool), one row
// segment at a time: m = (0 + u_hat u[x]) + v_hat v[y] with the row's product taken once and the
// first from g.uu, the same roundings as per pixel; a NaN location fails.
bool band_on_plane(const PlaneGuess &g, int64_t lo, int64_t hi) {
    if (hi <= lo) return true;
    const size_t parts = host_parts((size_t)(hi - lo) * 24);
    const int64_t chunk = (hi - lo + (int64_t)parts - 1) / (int64_t)parts;
    std::vector<char> ok(parts, 1);
    HostPool::get().run(parts, [&](size_t k) {
        const int64_t a = lo + (int64_t)k * chunk, b = std::min(hi, a + chunk);
        for (int64_t i = a; i < b;) {
            const int64_t y = i / g.nu, x0 = i - y * g.nu, n = std::min(b - i, g.nu - x0);
            int64_t bad = 0;
            for (int c = 0; c < 3; ++c) {
                const double vv = g.vh[c] * g.v[y], tol = g.tol;
                const double *uu = g.uu[c].data() + x0, *rr = g.rc[c] + (i - g.rmin);
                for (int64_t t = 0; t < n; ++t)
                    bad += !(std::fabs(rr[t] - ((0.0 + uu[t]) + vv)) <= tol);
            }
            if (bad) {
                ok[k] = 0;
                return;
            }
            i += n;
        }
    });
    for (char o : ok)
        if (!o) return false;
    return true;
}

thread_local int g_last_grid_bands = 0, g_last_bands = 0;

int hip_fail(hipError_t e, int code, const char *what) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return code;
}

int backprojection_impl(halide_buffer_t *phs, halide_buffer_t *k_r, int32_t taylor_s_l,
                        int32_t N_fft, double delta_r, halide_buffer_t *u, halide_buffer_t *v,
                        halide_buffer_t *pos, halide_buffer_t *r, halide_buffer_t *out) {
    halide_buffer_t *all[] = {phs, k_r, u, v, pos, r, out};
    const char *names[] = {"phs", "k_r", "u", "v", "pos", "r", "output_img"};
    const int dims[] = {3, 1, 1, 1, 2, 2, 3};
    const uint8_t codes[] = {halide_type_float, halide_type_float, halide_type_float,
                             halide_type_float, halide_type_float, halide_type_float,
                             halide_type_float};
    const uint8_t bits[] = {32, 32, 64, 64, 32, 64, 64};
    for (int i = 0; i < 7; ++i) {
        if (!all[i]) {
            set_error("buffer argument %s is NULL", names[i]);
            return halide_error_code_buffer_argument_is_null;
        }
        if (all[i]->dimensions != dims[i] || !all[i]->dim) {
            set_error("%s has %d dimensions, expected %d", names[i], all[i]->dimensions, dims[i]);
            return halide_error_code_bad_dimensions;
        }
        if (!has_type(all[i], codes[i], bits[i])) {
            set_error("%s has type code %d bits %d lanes %d, expected float%d", names[i],
                      all[i]->type.code, all[i]->type.bits, all[i]->type.lanes, bits[i]);
            return halide_error_code_bad_type;
        }
        for (int dd = 0; dd < dims[i]; ++dd) {
            if (all[i]->dim[dd].extent < 0) {
                set_error("%s has a negative extent in dimension %d", names[i], dd);
                return halide_error_code_buffer_extents_negative;
            }
        }
    }
    // backprojection.cpp:69-72: extents define the problem
    const int ns = phs->dim[1].extent, np = phs->dim[2].extent;
    const int nu = u->dim[0].extent, nv = v->dim[0].extent;
    const int64_t npix = (int64_t)nu * nv;

    // Bounds query (dft.cpp:49-58 protocol): fill the required regions and return.
    bool bq = false;
    if (is_bq(k_r)) { set_dense(k_r, {{0, ns}}); bq = true; }
    if (is_bq(pos)) { set_dense(pos, {{0, 3}, {0, np}}); bq = true; }
    if (is_bq(r)) {
        if (npix > INT32_MAX) {
            set_error("nu * nv overflows the pixel_locs extent");
            return halide_error_code_buffer_extents_too_large;
        }
        set_dense(r, {{0, (int)npix}, {0, 3}});
        bq = true;
    }
    if (is_bq(phs)) { set_dense(phs, {{0, 2}, {0, ns}, {0, np}}); bq = true; }
    if (is_bq(u) || is_bq(v) || is_bq(out)) bq = true;
    if (bq) return 0;

    // Required regions (Halide's access checks)
    if (!covers(phs, 0, 0, 2) || phs->dim[1].min != 0 || phs->dim[2].min != 0) {
        set_error("phs must cover {[0,2), [0,nsamples), [0,npulses)}");
        return halide_error_code_access_out_of_bounds;
    }
    if (!covers(k_r, 0, 0, ns) || !covers(k_r, 0, ns / 2, ns / 2 + 1)) {
        set_error("k_r must cover [0, nsamples)");
        return halide_error_code_access_out_of_bounds;
    }
    if (out->dim[0].min < 0 || out->dim[0].min + out->dim[0].extent > 2) {
        set_error("output_img dimension 0 must lie within [0, 2) (real, imaginary)");
        return halide_error_code_access_out_of_bounds;
    }
    if (int rv = check_n_fft(N_fft)) return rv;
    const int ox0 = out->dim[1].min, oy0 = out->dim[2].min;
    const int ow = out->dim[1].extent, oh = out->dim[2].extent;
    if (ow == 0 || oh == 0 || out->dim[0].extent == 0) return 0;

    Context *ctx = nullptr;
    int rv = get_context(&ctx);
    if (rv) return rv;
    std::lock_guard<std::mutex> lk(ctx->mu);

    int norm_policy, den_policy;
    sarbp_get_default_policies(&norm_policy, &den_policy);
    sarbp_plan_desc d;
    sarbp_plan_desc_init(&d);
    d.nsamples = ns;
    d.npulses = np;
    d.n_fft = N_fft;
    d.taylor_s_l = taylor_s_l;
    d.delta_r = delta_r;
    d.nu = nu;
    d.nv = nv;
    d.x0 = ox0;
    d.y0 = oy0;
    d.width = ow;
    d.height = oh;
    d.r_min = r->dim[0].min;
    d.r_extent = r->dim[0].extent;
    d.norm_r0_policy = norm_policy;
    d.taylor_den_policy = den_policy;
    d.device = ctx->device;
    auto key = std::make_tuple(ns, np, N_fft, taylor_s_l, nu, nv, ox0, oy0, ow, oh, norm_policy,
                               den_policy, (long long)d.r_min, (long long)d.r_extent, delta_r);
    sarbp_plan *plan = nullptr;
    auto it = ctx->plans.find(key);
    if (it != ctx->plans.end()) {
        plan = it->second;
    } else {
        if (ctx->plans.size() >= 8) {
            (void)hipStreamSynchronize(ctx->stream);
            for (auto &kv : ctx->plans) sarbp_plan_destroy(kv.second);
            ctx->plans.clear();
        }
        rv = sarbp_plan_create(&d, &plan);
        if (rv) return rv;
        ctx->plans[key] = plan;
    }

    // Pinned staging and device layout: phs | k_r | pos | r (each 256-byte aligned; r as three
    // component planes over the caller's region [rmin, rmin + rext)).
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t b_phs = align(sizeof(float) * 2 * (size_t)ns * np);
    const size_t b_kr = align(sizeof(float) * (size_t)ns);
    const size_t b_pos = align(sizeof(float) * 3 * (size_t)np);
    const int64_t rext = r->dim[0].extent;
    const int64_t rmin = r->dim[0].min;
    const size_t b_r = align(sizeof(double) * 3 * (size_t)rext);
    const size_t b_out = sizeof(double) * 2 * (size_t)ow * oh;
    hipError_t e = ctx->h_in.ensure(b_phs + b_kr + b_pos + b_r);
    if (e == hipSuccess) e = ctx->h_out.ensure(b_out);
    if (e != hipSuccess) return hip_fail(e, halide_error_code_out_of_memory, "pinned staging");
    if ((e = ctx->d_aux.ensure(b_phs + b_kr + b_pos + b_r)) != hipSuccess)
        return hip_fail(e, halide_error_code_device_malloc_failed, "device staging");
    uint8_t *h = (uint8_t *)ctx->h_in.p;
    float *hphs = (float *)h;
    float *hkr = (float *)(h + b_phs);
    float *hpos = (float *)(h + b_phs + b_kr);
    double *hr = (double *)(h + b_phs + b_kr + b_pos);
    uint8_t *dbase = (uint8_t *)ctx->d_aux.p;
    double *dr = (double *)(dbase + b_phs + b_kr + b_pos);
    hipStream_t s1 = ctx->stream, s2 = ctx->stream2, s3 = ctx->stream3;
    auto drain = [&](int code, const char *what) {
        (void)hipStreamSynchronize(s1);
        (void)hipStreamSynchronize(s2);
        (void)hipStreamSynchronize(s3);
        return hip_fail(e, code, what);
    };
    size_t phs_sent = 0;  // bytes of phs already enqueued for upload
    {
        const float *src = (const float *)phs->host;
        const bool dense = phs->dim[0].stride == 1 && phs->dim[1].stride == 2 &&
                           phs->dim[2].stride == 2 * ns && phs->dim[0].min == 0;
        if (dense) {
            // packed and uploaded in pulse chunks, so that each chunk's DMA overlaps the packing
            // of the next
            const size_t row = sizeof(float) * 2 * (size_t)ns;
            const int64_t chunk = std::max<int64_t>(1, (int64_t)((8u << 20) / row));
            for (int64_t p0 = 0; p0 < np; p0 += chunk) {
                const size_t off = row * (size_t)p0, n = row * (size_t)std::min(chunk, np - p0);
                par_copy((uint8_t *)hphs + off, (const uint8_t *)src + off, n);
                if ((e = hipMemcpyAsync(dbase + off, h + off, n, hipMemcpyHostToDevice, s1)) !=
                    hipSuccess)
                    return drain(halide_error_code_copy_to_device_failed, "input upload");
            }
            phs_sent = row * (size_t)np;
        } else {
            for (int64_t p = 0; p < np; ++p)
                for (int64_t s = 0; s < ns; ++s)
                    for (int64_t c = 0; c < 2; ++c) {
                        const int64_t co[3] = {c, s, p};
                        hphs[(p * ns + s) * 2 + c] = src[offset_of(phs, co)];
                    }
        }
        const float *ks = (const float *)k_r->host;
        for (int64_t s = 0; s < ns; ++s) {
            const int64_t co[1] = {s};
            hkr[s] = ks[offset_of(k_r, co)];
        }
        // pos = constant_exterior(pos_in, 0.0f) (backprojection.cpp:76)
        const float *ps = (const float *)pos->host;
        for (int64_t p = 0; p < np; ++p)
            for (int64_t c = 0; c < 3; ++c) {
                const int64_t co[2] = {c, p};
                hpos[p * 3 + c] =
                    (in_dim(pos, 0, c) && in_dim(pos, 1, p)) ? ps[offset_of(pos, co)] : 0.0f;
            }
    }
    if ((e = hipMemcpyAsync(dbase + phs_sent, h + phs_sent, b_phs + b_kr + b_pos - phs_sent,
                            hipMemcpyHostToDevice, s1)) != hipSuccess)
        return drain(halide_error_code_copy_to_device_failed, "input upload");
    PlaneGuess plane;
    guess_plane(u, v, r, nu, nv, rmin, rext, plane);
    if (plane.on && sarbp_plan_set_grid(plan, plane.d_u, plane.d_v, plane.uh, plane.vh) != 0)
        plane.on = false;
    if (!plane.on) (void)sarbp_plan_set_grid(plan, 0.0, 0.0, nullptr, nullptr);
    rv = plan_prepare(plan, dbase, dbase + b_phs, dbase + b_phs + b_kr, s1);
    if (rv) {
        (void)hipStreamSynchronize(s1);
        return rv;
    }

    // Row bands of half as many tiles as the device holds at once (two bands in flight fill it;
    // 1024 tiles at 16 x 32, 512 at 16 x 64), tile-aligned, at most 8, alternating between the two
    // compute streams. Against bands of a whole device wave the first band waits for half as many
    // pixel locations and the last one runs alone for half as long (C2, 16 x 32 tiles: 8.6-9.1 ms per
    // call instead of 9.0-9.4). Bands over three streams in turn measured slower (C2: 6.46 against
    // 6.02 ms per call).
    const int th = plan_tile_rows(plan);
    const int64_t tile_rows = (oh + th - 1) / th, tile_cols = (ow + 15) / 16;
    const int64_t band_tiles = std::max(256, plan_resident_tiles(plan) / 2);
    const int nb = (int)std::max<int64_t>(
        1, std::min<int64_t>({8, tile_rows, tile_rows * tile_cols / band_tiles}));
    const int band = (int)((tile_rows + nb - 1) / nb) * th;
    const int nbands = (oh + band - 1) / band;
    if ((e = ctx->ensure_events(2 * (size_t)nbands + 1)) != hipSuccess)
        return drain(halide_error_code_device_run_failed, "event creation");
    hipEvent_t *ev_r = ctx->events.data(), *ev_acc = ev_r + nbands;
    hipEvent_t ev_prep = ev_acc[nbands];
    if ((e = hipEventRecord(ev_prep, s1)) != hipSuccess ||
        (e = hipStreamWaitEvent(s3, ev_prep, 0)) != hipSuccess)
        return drain(halide_error_code_device_run_failed, "band ordering");

    // r = constant_exterior(r_in, 0.0) (backprojection.cpp:77): a band reads r only at its own pixels,
    // px = nu * (nv - y - 1) + x (backprojection.cpp:130), a contiguous index range per band.
    const double *rs = (const double *)r->host;
    // band b's pixel range [plo, phi), and within r's extent [lo, hi)
    auto band_pixels = [&](int b, int64_t &lo, int64_t &hi, int64_t *plo = nullptr,
                           int64_t *phi = nullptr) {
        const int row0 = b * band, nrows = std::min(band, oh - row0);
        const int64_t ylo = oy0 + row0, yhi = ylo + nrows - 1;
        const int64_t pl = (int64_t)nu * (nv - yhi - 1) + ox0, ph = (int64_t)nu * (nv - ylo - 1) + ox0 + ow;
        lo = std::max<int64_t>(rmin, pl);
        hi = std::min<int64_t>(rmin + rext, ph);
        if (plo) *plo = pl;
        if (phi) *phi = ph;
    };
    // Whether band b is on the image plane (checked once, in band order). A band with pixels
    // outside r's extent reads r (0 there); once one band is off the plane, the rest read r
    // unchecked.
    std::vector<signed char> grid_band((size_t)nbands, -1);
    auto band_grid = [&](int b) {
        if (grid_band[b] < 0) {
            int64_t lo, hi, plo, phi;
            band_pixels(b, lo, hi, &plo, &phi);
            plane.on = plane.on && plo >= rmin && phi <= rmin + rext && band_on_plane(plane, lo, hi);
            grid_band[b] = plane.on ? 1 : 0;
        }
        return grid_band[b] == 1;
    };
    g_last_grid_bands = 0;
    g_last_bands = nbands;
    for (int b = 0; b < nbands; ++b) {
        const int row0 = b * band, nrows = std::min(band, oh - row0);
        int64_t lo, hi;
        band_pixels(b, lo, hi);
        hipStream_t sc = (b & 1) ? s3 : s1;
        // The first two bands are both checked before the first is launched, so that they run
        // side by side from the start (a band is half a device of tiles) and the pairs after them
        // stay aligned to the last band.
        if (b == 0 && nbands > 1) (void)band_grid(0), (void)band_grid(1);
        if (band_grid(b)) {
            ++g_last_grid_bands;
            rv = plan_accumulate_rows(plan, dr, rmin, rext, row0, nrows,
                                      (double *)ctx->h_out.p + (size_t)2 * ow * row0, sc, true);
            if (rv) {
                (void)hipStreamSynchronize(s1);
                (void)hipStreamSynchronize(s2);
                (void)hipStreamSynchronize(s3);
                            return rv;
            }
            if ((e = hipEventRecord(ev_acc[b], sc)) != hipSuccess)
                return drain(halide_error_code_device_run_failed, "band ordering");
            continue;
        }
        if (hi > lo) {
            const int64_t n = hi - lo;
            for (int64_t c = 0; c < 3; ++c) {
                double *dst = hr + c * rext + (lo - rmin);
                if (!in_dim(r, 1, c)) {
                    std::memset(dst, 0, sizeof(double) * (size_t)n);
                } else if (r->dim[0].stride == 1) {
                    const int64_t co[2] = {lo, c};
                    par_copy(dst, rs + offset_of(r, co), sizeof(double) * (size_t)n);
                } else {
                    for (int64_t i = 0; i < n; ++i) {
                        const int64_t co[2] = {lo + i, c};
                        dst[i] = rs[offset_of(r, co)];
                    }
                }
                e = hipMemcpyAsync(dr + c * rext + (lo - rmin), dst, sizeof(double) * (size_t)n,
                                   hipMemcpyHostToDevice, s2);
                if (e != hipSuccess) return drain(halide_error_code_copy_to_device_failed, "r upload");
            }
        }
        if ((e = hipEventRecord(ev_r[b], s2)) != hipSuccess ||
            (e = hipStreamWaitEvent(sc, ev_r[b], 0)) != hipSuccess)
            return drain(halide_error_code_device_run_failed, "band ordering");
        rv = plan_accumulate_rows(plan, dr, rmin, rext, row0, nrows,
                                  (double *)ctx->h_out.p + (size_t)2 * ow * row0, sc, false);
        if (rv) {
            (void)hipStreamSynchronize(s1);
            (void)hipStreamSynchronize(s2);
            (void)hipStreamSynchronize(s3);
                return rv;
        }
        if ((e = hipEventRecord(ev_acc[b], sc)) != hipSuccess)
            return drain(halide_error_code_device_run_failed, "band ordering");
    }
    // Scatter each band {2, ow, nrows} into the caller's output region as its accumulation
    // completes (the kernel wrote it into h_out).
    const double *res = (const double *)ctx->h_out.p;
    double *dst = (double *)out->host;
    const int c0 = out->dim[0].min, c1 = out->dim[0].min + out->dim[0].extent;
    const bool dense = c0 == 0 && c1 == 2 && out->dim[0].stride == 1 && out->dim[1].stride == 2 &&
                       out->dim[2].stride == 2 * ow;
    for (int b = 0; b < nbands; ++b) {
        if ((e = hipEventSynchronize(ev_acc[b])) != hipSuccess)
            return drain(halide_error_code_device_run_failed, "pipeline execution");
        const int row0 = b * band, nrows = std::min(band, oh - row0);
        if (dense) {
            const size_t off = (size_t)2 * ow * row0;
            par_copy(dst + off, res + off, sizeof(double) * 2 * (size_t)ow * nrows);
        } else {
            for (int64_t y = row0; y < row0 + nrows; ++y)
                for (int64_t x = 0; x < ow; ++x)
                    for (int64_t c = c0; c < c1; ++c) {
                        const int64_t co[3] = {c, ox0 + x, oy0 + y};
                        dst[offset_of(out, co)] = res[(y * ow + x) * 2 + c];
                    }
        }
    }
    if ((e = hipStreamSynchronize(s1)) != hipSuccess || (e = hipStreamSynchronize(s3)) != hipSuccess)
        return drain(halide_error_code_device_run_failed, "pipeline execution");
    return 0;
}

}  // namespace

namespace sarbp {
int abi_backprojection(halide_buffer_t *phs, halide_buffer_t *k_r, int32_t taylor_s_l,
                       int32_t N_fft, double delta_r, halide_buffer_t *u, halide_buffer_t *v,
                       halide_buffer_t *pos, halide_buffer_t *r, halide_buffer_t *output_img) {
    return backprojection_impl(phs, k_r, taylor_s_l, N_fft, delta_r, u, v, pos, r, output_img);
}
}  // namespace sarbp

extern "C" int backprojection(halide_buffer_t *phs, halide_buffer_t *k_r, int32_t taylor_s_l,
                              int32_t N_fft, double delta_r, halide_buffer_t *u, halide_buffer_t *v,
                              halide_buffer_t *pos, halide_buffer_t *r,
                              halide_buffer_t *output_img) {
    return backprojection_impl(phs, k_r, taylor_s_l, N_fft, delta_r, u, v, pos, r, output_img);
}

extern "C" int backprojection_cuda(halide_buffer_t *phs, halide_buffer_t *k_r, int32_t taylor_s_l,
                                   int32_t N_fft, double delta_r, halide_buffer_t *u,
                                   halide_buffer_t *v, halide_buffer_t *pos, halide_buffer_t *r,
                                   halide_buffer_t *output_img) {
    return backprojection_impl(phs, k_r, taylor_s_l, N_fft, delta_r, u, v, pos, r, output_img);
}

extern "C" int img_output_to_dB(halide_buffer_t *img, halide_buffer_t *out) {
    if (!img || !out) {
        set_error("img_output_to_dB: NULL buffer");
        return halide_error_code_buffer_argument_is_null;
    }
    if (img->dimensions != 3 || out->dimensions != 2 || !img->dim || !out->dim) {
        set_error("img_output_to_dB: expected img {2, nu, nv} and out {nu, nv}");
        return halide_error_code_bad_dimensions;
    }
    if (!has_type(img, halide_type_float, 64) || !has_type(out, halide_type_float, 64)) {
        set_error("img_output_to_dB: img and out must be float64");
        return halide_error_code_bad_type;
    }
    // img_output.cpp:23: RDom r(0, img.dim(1).extent(), 0, img.dim(2).extent())
    const int W = img->dim[1].extent, H = img->dim[2].extent;
    if (is_bq(img)) {
        set_dense(img, {{0, 2}, {0, W}, {0, H}});
        return 0;
    }
    if (is_bq(out)) return 0;
    if (W < 0 || H < 0 || out->dim[0].extent < 0 || out->dim[1].extent < 0) {
        set_error("img_output_to_dB: negative extent");
        return halide_error_code_buffer_extents_negative;
    }
    if (!covers(img, 0, 0, 2) || img->dim[1].min != 0 || img->dim[2].min != 0 ||
        !covers(img, 1, out->dim[0].min, (int64_t)out->dim[0].min + out->dim[0].extent) ||
        !covers(img, 2, out->dim[1].min, (int64_t)out->dim[1].min + out->dim[1].extent)) {
        set_error("img_output_to_dB: img must cover [0,2) x [0,W) x [0,H) and the output region");
        return halide_error_code_access_out_of_bounds;
    }
    const int64_t ix0 = img->dim[1].min, iy0 = img->dim[2].min;
    const int64_t IW = img->dim[1].extent, IH = img->dim[2].extent;
    const int64_t npix = IW * IH;
    if (npix == 0) return 0;
    Context *ctx = nullptr;
    int rv = get_context(&ctx);
    if (rv) return rv;
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipError_t e;
    // Upload the whole img buffer region dense: the max runs over [0, W) x [0, H), which is the
    // region itself when mins are 0 (required above unless the region is larger).
    if ((e = ctx->h_in.ensure(sizeof(double) * 2 * npix)) != hipSuccess ||
        (e = ctx->h_out.ensure(sizeof(double) * npix)) != hipSuccess)
        return hip_fail(e, halide_error_code_out_of_memory, "pinned staging");
    if ((e = ctx->d_aux.ensure(sizeof(double) * 2 * npix)) != hipSuccess ||
        (e = ctx->d_out.ensure(sizeof(double) * npix)) != hipSuccess)
        return hip_fail(e, halide_error_code_device_malloc_failed, "device staging");
    double *h = (double *)ctx->h_in.p;
    const double *src = (const double *)img->host;
    // Only the RDom's [0, W) x [0, H) contributes to the max: pixels of a larger region outside it
    // are packed as zeros, which cannot raise a max of magnitudes.
    for (int64_t y = 0; y < IH; ++y)
        for (int64_t x = 0; x < IW; ++x) {
            const int64_t gx = ix0 + x, gy = iy0 + y;
            const bool in_rdom = gx >= 0 && gx < W && gy >= 0 && gy < H;
            for (int64_t c = 0; c < 2; ++c) {
                const int64_t co[3] = {c, gx, gy};
                h[(y * IW + x) * 2 + c] = in_rdom ? src[offset_of(img, co)] : 0.0;
            }
        }
    if ((e = hipMemcpyAsync(ctx->d_aux.p, h, sizeof(double) * 2 * npix, hipMemcpyHostToDevice,
                            ctx->stream)) != hipSuccess)
        return hip_fail(e, halide_error_code_copy_to_device_failed, "img upload");
    rv = sarbp_img_output_device(ctx->d_aux.p, npix, (double *)ctx->d_out.p, nullptr, 0.0, 1.0,
                                 ctx->stream);
    if (rv) return rv;
    if ((e = hipMemcpyAsync(ctx->h_out.p, ctx->d_out.p, sizeof(double) * npix,
                            hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess)
        return hip_fail(e, halide_error_code_copy_to_host_failed, "dB download");
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
        return hip_fail(e, halide_error_code_device_run_failed, "dB execution");
    // Values for pixels outside the RDom but inside img's region are recomputed with the true
    // magnitude: dB depends on the pixel and the global max only.
    const double *res = (const double *)ctx->h_out.p;
    double *dst = (double *)out->host;
    for (int64_t y = out->dim[1].min; y < (int64_t)out->dim[1].min + out->dim[1].extent; ++y)
        for (int64_t x = out->dim[0].min; x < (int64_t)out->dim[0].min + out->dim[0].extent; ++x) {
            const int64_t co[2] = {x, y};
            dst[offset_of(out, co)] = res[(y - iy0) * IW + (x - ix0)];
        }
    return 0;
}

extern "C" int img_output_u8(halide_buffer_t *dB, double dB_min, double dB_max,
                             halide_buffer_t *out) {
    if (!dB || !out) {
        set_error("img_output_u8: NULL buffer");
        return halide_error_code_buffer_argument_is_null;
    }
    if (dB->dimensions != 2 || out->dimensions != 2 || !dB->dim || !out->dim) {
        set_error("img_output_u8: expected 2-D buffers");
        return halide_error_code_bad_dimensions;
    }
    if (!has_type(dB, halide_type_float, 64) || !has_type(out, halide_type_uint, 8)) {
        set_error("img_output_u8: dB must be float64 and out uint8");
        return halide_error_code_bad_type;
    }
    if (is_bq(dB)) {
        set_dense(dB, {{out->dim[0].min, out->dim[0].extent}, {out->dim[1].min, out->dim[1].extent}});
        return 0;
    }
    if (is_bq(out)) return 0;
    const int64_t x0 = out->dim[0].min, y0 = out->dim[1].min;
    const int64_t W = out->dim[0].extent, H = out->dim[1].extent;
    if (W < 0 || H < 0) {
        set_error("img_output_u8: negative extent");
        return halide_error_code_buffer_extents_negative;
    }
    if (!covers(dB, 0, x0, x0 + W) || !covers(dB, 1, y0, y0 + H)) {
        set_error("img_output_u8: dB must cover the output region");
        return halide_error_code_access_out_of_bounds;
    }
    const int64_t npix = W * H;
    if (npix == 0) return 0;
    Context *ctx = nullptr;
    int rv = get_context(&ctx);
    if (rv) return rv;
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipError_t e;
    if ((e = ctx->h_in.ensure(sizeof(double) * npix)) != hipSuccess ||
        (e = ctx->h_out.ensure((size_t)npix)) != hipSuccess)
        return hip_fail(e, halide_error_code_out_of_memory, "pinned staging");
    if ((e = ctx->d_aux.ensure(sizeof(double) * npix)) != hipSuccess ||
        (e = ctx->d_out.ensure((size_t)npix)) != hipSuccess)
        return hip_fail(e, halide_error_code_device_malloc_failed, "device staging");
    double *h = (double *)ctx->h_in.p;
    const double *src = (const double *)dB->host;
    for (int64_t y = 0; y < H; ++y)
        for (int64_t x = 0; x < W; ++x) {
            const int64_t co[2] = {x0 + x, y0 + y};
            h[y * W + x] = src[offset_of(dB, co)];
        }
    if ((e = hipMemcpyAsync(ctx->d_aux.p, h, sizeof(double) * npix, hipMemcpyHostToDevice,
                            ctx->stream)) != hipSuccess)
        return hip_fail(e, halide_error_code_copy_to_device_failed, "dB upload");
    if ((e = launch_u8_only((const double *)ctx->d_aux.p, npix, (uint8_t *)ctx->d_out.p, dB_min,
                            dB_max, ctx->stream)) != hipSuccess)
        return hip_fail(e, halide_error_code_device_run_failed, "u8 launch");
    if ((e = hipMemcpyAsync(ctx->h_out.p, ctx->d_out.p, (size_t)npix, hipMemcpyDeviceToHost,
                            ctx->stream)) != hipSuccess)
        return hip_fail(e, halide_error_code_copy_to_host_failed, "u8 download");
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
        return hip_fail(e, halide_error_code_device_run_failed, "u8 execution");
    const uint8_t *res = (const uint8_t *)ctx->h_out.p;
    uint8_t *dst = out->host;
    for (int64_t y = 0; y < H; ++y)
        for (int64_t x = 0; x < W; ++x) {
            const int64_t co[2] = {x0 + x, y0 + y};
            dst[offset_of(out, co)] = res[y * W + x];
        }
    return 0;
}

namespace sarbp {
// dft.cpp:20-30 builds the process-wide FFTW plan; here it builds the device twiddle table.
void abi_dft_init(size_t N_fft) {
    if (N_fft > (size_t)kMaxFftTwoPass || check_n_fft((int)N_fft) != 0) return;
    if (sarbp_device_count() <= 0) return;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    (void)twiddles(dev, (int)N_fft, nullptr);
}

// dft.cpp:32-35 destroys the plan. Twiddle tables are shared by every plan of the process and are
// kept until it exits, so that calls in flight on other threads are unaffected.
void abi_dft_destroy() {}
}  // namespace sarbp

extern "C" int sarbp_abi_last_call_bands(int32_t *grid_bands, int32_t *bands) {
    if (grid_bands) *grid_bands = g_last_grid_bands;
    if (bands) *bands = g_last_bands;
    return 0;
}

extern "C" void dft_init_fftw(size_t N_fft) { abi_dft_init(N_fft); }

extern "C" void dft_destroy_fftw(void) { abi_dft_destroy(); }
