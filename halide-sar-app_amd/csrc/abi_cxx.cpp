// The reference's pipeline names with the linkage its build gives them:
//   backprojection, backprojection_cuda   C++ linkage (c_plus_plus_name_mangling,
//                                          CMakeLists.txt:73-84), plus their _argv forms
//   dft_init_fftw(size_t), dft_destroy_fftw()  C++ linkage (dft.h:5-6)
// so that the reference's sarbp.cpp, compiled against its own headers, links against this library.
// Each forwards to the code behind the C-linkage names of include/sarbp_hip.h.
#include "abi_entry.h"
#include "backprojection.h"
#include "backprojection_cuda.h"

// The reference's dft.h declares these two with C++ linkage.
void dft_init_fftw(size_t N_fft);
void dft_destroy_fftw(void);

int backprojection(halide_buffer_t *phs, halide_buffer_t *k_r, int32_t taylor_s_l, int32_t N_fft,
                   double delta_r, halide_buffer_t *u, halide_buffer_t *v, halide_buffer_t *pos,
                   halide_buffer_t *r, halide_buffer_t *output_img) {
    return sarbp::abi_backprojection(phs, k_r, taylor_s_l, N_fft, delta_r, u, v, pos, r,
                                     output_img);
}

int backprojection_cuda(halide_buffer_t *phs, halide_buffer_t *k_r, int32_t taylor_s_l,
                        int32_t N_fft, double delta_r, halide_buffer_t *u, halide_buffer_t *v,
                        halide_buffer_t *pos, halide_buffer_t *r, halide_buffer_t *output_img) {
    return sarbp::abi_backprojection(phs, k_r, taylor_s_l, N_fft, delta_r, u, v, pos, r,
                                     output_img);
}

namespace {
// Halide's argv convention: buffers by pointer, scalars by pointer to their value. A NULL buffer
// is reported by the entry point itself; a NULL scalar pointer cannot be read.
int from_argv(void **args) {
    constexpr int kNull = -12;  // halide_error_code_buffer_argument_is_null
    if (!args || !args[2] || !args[3] || !args[4]) return kNull;
    return sarbp::abi_backprojection(
        (halide_buffer_t *)args[0], (halide_buffer_t *)args[1], *(const int32_t *)args[2],
        *(const int32_t *)args[3], *(const double *)args[4], (halide_buffer_t *)args[5],
        (halide_buffer_t *)args[6], (halide_buffer_t *)args[7], (halide_buffer_t *)args[8],
        (halide_buffer_t *)args[9]);
}
}  // namespace

int backprojection_argv(void **args) { return from_argv(args); }

int backprojection_cuda_argv(void **args) { return from_argv(args); }

void dft_init_fftw(size_t N_fft) { sarbp::abi_dft_init(N_fft); }

void dft_destroy_fftw(void) { sarbp::abi_dft_destroy(); }
