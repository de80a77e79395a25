/*
 * The backprojection pipeline with C++ linkage, declared as the reference's Halide AOT build
 * declares it: `add_halide_library(backprojection ... FEATURES c_plus_plus_name_mangling)`
 * (CMakeLists.txt:73-76) over the generator inputs of backprojection.cpp:35-46, bound by the
 * reference's driver at sarbp.cpp:23,184,277. A caller built against the reference's generated
 * header links against libsarbp_hip.so with this one unchanged: the symbol is
 * _Z14backprojectionP15halide_buffer_tS0_iidS0_S0_S0_S0_S0_.
 *
 * As in Halide's generated headers, halide_buffer_t is only forward-declared here; its definition
 * comes from HalideRuntime.h (or include/halide_buffer_compat.h). The arguments, bounds-query
 * protocol and return codes are those of the C-linkage `backprojection` in include/sarbp_hip.h;
 * both names run the same code. A translation unit may include this header and sarbp_hip.h in
 * either order: whichever comes first declares the name.
 */
#ifndef SARBP_BACKPROJECTION_H
#define SARBP_BACKPROJECTION_H

#include <stdint.h>

struct halide_buffer_t;

#if defined(__cplusplus) && !defined(SARBP_DECLARED_backprojection)
#define SARBP_DECLARED_backprojection 1

int backprojection(struct halide_buffer_t *phs, struct halide_buffer_t *k_r, int32_t taylor_s_l,
                   int32_t N_fft, double delta_r, struct halide_buffer_t *u,
                   struct halide_buffer_t *v, struct halide_buffer_t *pos,
                   struct halide_buffer_t *r, struct halide_buffer_t *output_img);

/* The argv form Halide generates beside each pipeline: args[i] is the halide_buffer_t * of a buffer
 * argument and a pointer to the value of a scalar one (taylor_s_l, N_fft: int32_t; delta_r:
 * double), in the order above. */
int backprojection_argv(void **args);

#endif
#endif
