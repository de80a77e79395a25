/*
 * img_output_to_dB as the reference's Halide AOT build declares it: without name mangling
 * (`add_halide_library(img_output_to_dB ... FEATURES large_buffers)`, CMakeLists.txt:104-105), so
 * with C linkage; called at sarbp.cpp:299. Same function as in include/sarbp_hip.h.
 */
#ifndef SARBP_IMG_OUTPUT_TO_DB_H
#define SARBP_IMG_OUTPUT_TO_DB_H

struct halide_buffer_t;

#ifdef __cplusplus
extern "C" {
#endif

int img_output_to_dB(struct halide_buffer_t *img, struct halide_buffer_t *out);

#ifdef __cplusplus
}
#endif
#endif
