// Per-pixel arithmetic of the dB / u8 image output (img_output.cpp:10-52), shared by every kernel
// that forms |z|, max |z|, dB or u8, so that they all round alike.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "log_f64.h"

namespace sarbp {

// |z|^2 rounded as the reference rounds it (abs of a ComplexFunc: re * re + im * im, each product
// and the sum rounded once; img_output.cpp:24-26). Every max and every dB value of this library
// takes |z| as the correctly rounded sqrt of this, so a maximum found anywhere (the accumulation's
// epilogue, the max kernels) equals the |z| the dB kernels form for the same pixel.
__device__ inline double abs2(double2 z) {
#pragma clang fp contract(off)
    return z.x * z.x + z.y * z.y;
}

// normalize (util.h:13-15), dB_scale (signal.h:55-57), u8 truncation (img_output.cpp:50)
__device__ inline uint8_t to_u8(double dB, double lo, double hi) {
    double c = dB < hi ? dB : hi;
    c = c > lo ? c : lo;
    double s = (c - lo) / (hi - lo) * 255.0;
    s = s < 255.0 ? s : 255.0;
    s = s > 0.0 ? s : 0.0;
    return (s == s) ? (uint8_t)s : (uint8_t)0;
}

// 10 * log10(|z| / m), log10(x) = log(x) / log(Expr(10)) with the f32 log(10) (util.h:56-58),
// formed as log(|z| / m) * db_per_neper with db_per_neper = 10 / (f64) logf(10), rounded once on the
// host (launch_db_u8): within ~2 ulp of the reference's two roundings, one multiply instead of a
// division per pixel. |z| / m is the correctly rounded sqrt and division, as the reference forms
// it; the log is log_f64 (within 1 ulp, a fraction of the device library's instructions), so the
// brightest pixel is exactly 0 dB.
__device__ inline double to_db(double2 z, double m, double db_per_neper) {
    return log_f64(sqrt(abs2(z)) / m) * db_per_neper;
}

}  // namespace sarbp
