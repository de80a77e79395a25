// The entry points behind the exported names, for the translation unit that exports them with C++
// linkage (abi_cxx.cpp). That unit cannot see include/sarbp_hip.h, which declares the same names
// with C linkage, so these are declared here on their own.
#pragma once

#include <stddef.h>
#include <stdint.h>

struct halide_buffer_t;

namespace sarbp {

int abi_backprojection(halide_buffer_t *phs, halide_buffer_t *k_r, int32_t taylor_s_l,
                       int32_t N_fft, double delta_r, halide_buffer_t *u, halide_buffer_t *v,
                       halide_buffer_t *pos, halide_buffer_t *r, halide_buffer_t *output_img);
void abi_dft_init(size_t N_fft);
void abi_dft_destroy();

}  // namespace sarbp
