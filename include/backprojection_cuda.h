/*
 * `backprojection_cuda` with C++ linkage, as the reference's Halide AOT build declares its GPU
 * variant (`add_halide_library(backprojection_cuda ... FEATURES c_plus_plus_name_mangling ... cuda)`,
 * CMakeLists.txt:77-84), selected by the reference's driver at sarbp.cpp:24,198-200. Here it is the
 * same function as `backprojection` (include/backprojection.h): both run on the GPU. Symbol:
 * _Z19backprojection_cudaP15halide_buffer_tS0_iidS0_S0_S0_S0_S0_.
 */
#ifndef SARBP_BACKPROJECTION_CUDA_H
#define SARBP_BACKPROJECTION_CUDA_H

#include <stdint.h>

struct halide_buffer_t;

#if defined(__cplusplus) && !defined(SARBP_DECLARED_backprojection_cuda)
#define SARBP_DECLARED_backprojection_cuda 1

int backprojection_cuda(struct halide_buffer_t *phs, struct halide_buffer_t *k_r,
                        int32_t taylor_s_l, int32_t N_fft, double delta_r,
                        struct halide_buffer_t *u, struct halide_buffer_t *v,
                        struct halide_buffer_t *pos, struct halide_buffer_t *r,
                        struct halide_buffer_t *output_img);

/* argv form: see backprojection_argv in include/backprojection.h. */
int backprojection_cuda_argv(void **args);

#endif
#endif
