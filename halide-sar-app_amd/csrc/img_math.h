// Per-pixel arithmetic of the dB / u8 image output (img_output.cpp:10-52), shared by every kernel
// that forms |z|, max |z|, dB or u8, so that they all round alike.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

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

// 10 * log10(|z| / m), log10(x) = log(x) / log(Expr(10)) with the f32 log(10) (util.h:56-58)
__device__ inline double to_db(double2 z, double m, double ln10f) {
    return 10.0 * (log(sqrt(abs2(z)) / m) / ln10f);
}

}  // namespace sarbp
