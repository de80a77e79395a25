// PlatformData directory loader (PlatformData.cpp:21-256) behind the C ABI of sarbp_host.h.
#include <complex>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "halide_buffer_compat.h"
#include "npy.hpp"
#include "sarbp_host.h"

using sarbp_host::file_exists;
using sarbp_host::npy_load;
using sarbp_host::NpyArray;

namespace {
thread_local std::string g_error;
}  // namespace

void sarbp_host::set_error(const std::string &msg) { g_error = msg; }

namespace {

// The C view first, so a sarbp_platform * is also a Holder *.
struct Holder {
    sarbp_platform view{};
    std::vector<float> freq, k_r, n_hat, R_c, pos, phs;
    std::vector<double> k_y, t;
};

[[noreturn]] void bad(const char *what, const char *name) {
    throw std::runtime_error(std::string(what) + ": " + name);
}

void need_kind(const NpyArray &a, const char *kinds, const char *name) {
    if (std::strchr(kinds, a.kind) == nullptr) bad("Bad element type", name);
}

// f64 scalar with word size 8 (PlatformData.cpp:25-30 and the like)
double scalar_f64(const NpyArray &a, const char *name) {
    if (a.word_size != sizeof(double)) bad("Bad word size", name);
    need_kind(a, "f", name);
    if (a.num_vals() < 1) bad("Bad shape", name);
    return *a.data<double>();
}

// f32 scalar, downcast from f64 (PlatformData.cpp:33-44)
float scalar_f32(const NpyArray &a, const char *name) {
    need_kind(a, "f", name);
    if (a.num_vals() < 1) bad("Bad shape", name);
    if (a.word_size == sizeof(float)) return *a.data<float>();
    if (a.word_size == sizeof(double)) {
        std::cout << "PlatformData: downcasting " << name << " from double to float" << std::endl;
        return (float)*a.data<double>();
    }
    bad("Bad word size", name);
}

// int32 count, int64 narrowed (PlatformData.cpp:79-99)
int count_i32(const NpyArray &a, const char *name) {
    need_kind(a, "iu", name);
    if (a.num_vals() < 1) bad("Bad shape", name);
    if (a.word_size == sizeof(int32_t)) return *a.data<int32_t>();
    if (a.word_size == sizeof(int64_t)) return (int)*a.data<int64_t>();
    bad("Bad word size", name);
}

// f32 array, downcast from f64 (PlatformData.cpp:111-150, 180-197, 215-232)
void array_f32(const NpyArray &a, const char *name, std::vector<float> &out) {
    need_kind(a, "f", name);
    out.resize(a.num_vals());
    if (a.word_size == sizeof(float)) {
        std::memcpy(out.data(), a.bytes.data(), a.bytes.size());
    } else if (a.word_size == sizeof(double)) {
        std::cout << "PlatformData: downcasting " << name << " from double to float" << std::endl;
        const double *src = a.data<double>();
        for (size_t i = 0; i < out.size(); ++i) out[i] = (float)src[i];
    } else {
        bad("Bad word size", name);
    }
}

// f64 array, word size 8 only (PlatformData.cpp:152-164, 199-211)
void array_f64(const NpyArray &a, const char *name, std::vector<double> &out) {
    if (a.word_size != sizeof(double)) bad("Bad word size", name);
    need_kind(a, "f", name);
    out.resize(a.num_vals());
    std::memcpy(out.data(), a.bytes.data(), a.bytes.size());
}

bool shape_is(const NpyArray &a, std::initializer_list<size_t> s) {
    if (a.shape.size() != s.size()) return false;
    size_t i = 0;
    for (size_t v : s)
        if (a.shape[i++] != v) return false;
    return true;
}

std::unique_ptr<Holder> load(const std::string &dir) {
    auto h = std::make_unique<Holder>();
    sarbp_platform &p = h->view;
    auto path = [&](const char *name) { return dir + "/" + name + ".npy"; };

    if (file_exists(path("B"))) {
        p.B = scalar_f64(npy_load(path("B")), "B");
        p.has_B = 1;
    }
    if (file_exists(path("B_IF"))) {
        p.B_IF = scalar_f32(npy_load(path("B_IF")), "B_IF");
        p.has_B_IF = 1;
    }
    p.delta_r = scalar_f64(npy_load(path("delta_r")), "delta_r");
    if (file_exists(path("delta_t"))) {
        p.delta_t = scalar_f64(npy_load(path("delta_t")), "delta_t");
        p.has_delta_t = 1;
    }
    if (file_exists(path("chirprate"))) {
        p.chirprate = scalar_f64(npy_load(path("chirprate")), "chirprate");
        p.has_chirprate = 1;
    }
    if (file_exists(path("f_0"))) {
        p.f_0 = scalar_f64(npy_load(path("f_0")), "f_0");
        p.has_f_0 = 1;
    }
    p.nsamples = count_i32(npy_load(path("nsamples")), "nsamples");
    p.npulses = count_i32(npy_load(path("npulses")), "npulses");
    if (p.nsamples < 1 || p.npulses < 1) throw std::runtime_error("Bad value: nsamples/npulses");
    const size_t ns = (size_t)p.nsamples, np = (size_t)p.npulses;
    if (file_exists(path("vp"))) {
        p.vp = scalar_f64(npy_load(path("vp")), "vp");
        p.has_vp = 1;
    }
    if (file_exists(path("freq"))) {
        NpyArray a = npy_load(path("freq"));
        if (!shape_is(a, {ns})) bad("Bad shape", "freq");
        array_f32(a, "freq", h->freq);
        p.freq = h->freq.data();
    }
    {
        NpyArray a = npy_load(path("k_r"));
        if (!shape_is(a, {ns})) bad("Bad shape", "k_r");
        array_f32(a, "k_r", h->k_r);
        p.k_r = h->k_r.data();
    }
    if (file_exists(path("k_y"))) {
        NpyArray a = npy_load(path("k_y"));
        if (!shape_is(a, {np})) bad("Bad shape", "k_y");
        array_f64(a, "k_y", h->k_y);
        p.k_y = h->k_y.data();
    }
    if (file_exists(path("n_hat"))) {
        NpyArray a = npy_load(path("n_hat"));
        if (!shape_is(a, {3})) bad("Bad shape", "n_hat");
        if (a.word_size != sizeof(float)) bad("Bad word size", "n_hat");
        array_f32(a, "n_hat", h->n_hat);
        p.n_hat = h->n_hat.data();
    }
    {
        NpyArray a = npy_load(path("R_c"));
        if (!shape_is(a, {3})) bad("Bad shape", "R_c");
        array_f32(a, "R_c", h->R_c);
        p.R_c = h->R_c.data();
    }
    if (file_exists(path("t"))) {
        NpyArray a = npy_load(path("t"));
        if (!shape_is(a, {ns})) bad("Bad shape", "t");
        array_f64(a, "t", h->t);
        p.t = h->t.data();
    }
    {
        NpyArray a = npy_load(path("pos"));
        if (!shape_is(a, {np, 3})) bad("Bad shape", "pos");
        array_f32(a, "pos", h->pos);
        p.pos = h->pos.data();
    }
    {
        NpyArray a = npy_load(path("phs"));
        if (!shape_is(a, {np, ns})) bad("Bad shape", "phs");
        need_kind(a, "c", "phs");
        h->phs.resize(2 * a.num_vals());
        if (a.word_size == sizeof(std::complex<float>)) {
            std::memcpy(h->phs.data(), a.bytes.data(), a.bytes.size());
        } else if (a.word_size == sizeof(std::complex<double>)) {
            std::cout << "PlatformData: downcasting phs from complex<double> to complex<float>"
                      << std::endl;
            const double *src = a.data<double>();
            for (size_t i = 0; i < h->phs.size(); ++i) h->phs[i] = (float)src[i];
        } else {
            bad("Bad word size", "phs");
        }
        p.phs = h->phs.data();
    }
    return h;
}

}  // namespace

extern "C" const char *sarbp_host_last_error(void) { return g_error.c_str(); }

extern "C" int sarbp_platform_load(const char *dir, sarbp_platform **out) {
    if (!dir || !out) {
        sarbp_host::set_error("sarbp_platform_load: NULL argument");
        return halide_error_code_buffer_argument_is_null;
    }
    *out = nullptr;
    try {
        std::unique_ptr<Holder> h = load(dir);
        *out = &h.release()->view;
        return 0;
    } catch (const std::exception &e) {
        sarbp_host::set_error(e.what());
        return halide_error_code_generic_error;
    }
}

extern "C" void sarbp_platform_free(sarbp_platform *p) {
    delete reinterpret_cast<Holder *>(p);
}
