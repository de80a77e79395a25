/*
 * img_output_u8 as the reference's Halide AOT build declares it: without name mangling
 * (`add_halide_library(img_output_u8 ... FEATURES large_buffers)`, CMakeLists.txt:106-107), so with
 * C linkage; called at sarbp.cpp:316 and img_output_u8_test.cpp:36. Same function as in
 * include/sarbp_hip.h.
 */
#ifndef SARBP_IMG_OUTPUT_U8_H
#define SARBP_IMG_OUTPUT_U8_H

struct halide_buffer_t;

#ifdef __cplusplus
extern "C" {
#endif

int img_output_u8(struct halide_buffer_t *dB, double dB_min, double dB_max,
                  struct halide_buffer_t *out);

#ifdef __cplusplus
}
#endif
#endif
