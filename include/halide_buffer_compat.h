/*
 * Binary-compatible declarations of the Halide runtime buffer ABI used by the reference's
 * AOT-generated `backprojection`, `img_output_to_dB` and `img_output_u8` entry points.
 *
 * These mirror the public, documented layout of `halide_buffer_t`, `halide_dimension_t` and
 * `halide_type_t` from Halide's HalideRuntime.h (the reference builds against Halide >= 10,
 * README.md:21), so that a caller compiled against Halide can pass its buffers straight through.
 * When the real HalideRuntime.h is already included (HALIDE_HALIDERUNTIME_H defined), the
 * declarations below are skipped and Halide's own are used.
 */
#ifndef SARBP_HALIDE_BUFFER_COMPAT_H
#define SARBP_HALIDE_BUFFER_COMPAT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef HALIDE_HALIDERUNTIME_H

typedef enum halide_type_code_t {
    halide_type_int = 0,
    halide_type_uint = 1,
    halide_type_float = 2,
    halide_type_handle = 3,
    halide_type_bfloat = 4,
} halide_type_code_t;

struct halide_type_t {
    uint8_t code; /* halide_type_code_t */
    uint8_t bits;
    uint16_t lanes;
};

typedef struct halide_dimension_t {
    int32_t min, extent, stride;
    uint32_t flags;
} halide_dimension_t;

struct halide_device_interface_t;

typedef enum {
    halide_buffer_flag_host_dirty = 1,
    halide_buffer_flag_device_dirty = 2,
} halide_buffer_flags;

typedef struct halide_buffer_t {
    uint64_t device;
    const struct halide_device_interface_t *device_interface;
    uint8_t *host;
    uint64_t flags;
    struct halide_type_t type;
    int32_t dimensions;
    halide_dimension_t *dim;
    void *padding;
} halide_buffer_t;

/* Subset of halide_error_code_t returned by the entry points (values as in HalideRuntime.h). */
enum halide_error_code_t {
    halide_error_code_success = 0,
    halide_error_code_generic_error = -1,
    halide_error_code_bad_type = -3,
    halide_error_code_access_out_of_bounds = -4,
    halide_error_code_buffer_allocation_too_large = -5,
    halide_error_code_buffer_extents_too_large = -6,
    halide_error_code_constraint_violated = -8,
    halide_error_code_param_too_small = -9,
    halide_error_code_param_too_large = -10,
    halide_error_code_out_of_memory = -11,
    halide_error_code_buffer_argument_is_null = -12,
    halide_error_code_copy_to_host_failed = -14,
    halide_error_code_copy_to_device_failed = -15,
    halide_error_code_device_malloc_failed = -16,
    halide_error_code_no_device_interface = -19,
    halide_error_code_device_run_failed = -23,
    halide_error_code_buffer_extents_negative = -28,
    halide_error_code_bad_dimensions = -43,
};

#endif /* HALIDE_HALIDERUNTIME_H */

#ifdef __cplusplus
}
#endif
#endif
