// A caller compiled the way the reference's driver is (sarbp.cpp:17-34): the pipeline headers as
// Halide's AOT build declares them (backprojection and backprojection_cuda with C++ linkage,
// img_output_* with C linkage) and dft_init_fftw / dft_destroy_fftw declared with C++ linkage as in
// the reference's dft.h:5-6. It makes a bounds query through each backprojection name (no GPU
// needed: bounds queries return before any device work) and prints the required regions.
#include <cstddef>
#include <cstdio>
#include <cstring>

#include "halide_buffer_compat.h"  // struct definitions (HalideRuntime.h in the reference)
#include "backprojection.h"
#include "backprojection_cuda.h"
#include "img_output_to_dB.h"
#include "img_output_u8.h"

// dft.h:5-6 of the reference
void dft_init_fftw(size_t N_fft);
void dft_destroy_fftw(void);

namespace {

struct Buf {
    halide_buffer_t b;
    halide_dimension_t d[3];
    Buf(int dims, uint8_t code, uint8_t bits, const int *ext) {
        std::memset(&b, 0, sizeof b);
        std::memset(d, 0, sizeof d);
        b.dimensions = dims;
        b.dim = d;
        b.type.code = code;
        b.type.bits = bits;
        b.type.lanes = 1;
        for (int i = 0; i < dims; ++i) d[i].extent = ext ? ext[i] : 0;
    }
};

int query(int which) {
    const int ns = 24, np = 15, nu = 32, nv = 16;
    const int e_phs[3] = {2, ns, np}, e_u[1] = {nu}, e_v[1] = {nv}, e_out[3] = {2, nu, nv};
    Buf phs(3, halide_type_float, 32, e_phs), k_r(1, halide_type_float, 32, nullptr),
        u(1, halide_type_float, 64, e_u), v(1, halide_type_float, 64, e_v),
        pos(2, halide_type_float, 32, nullptr), r(2, halide_type_float, 64, nullptr),
        out(3, halide_type_float, 64, e_out);
    int32_t taylor = 17, nfft = 64;
    double delta_r = 0.5;
    int rv;
    if (which == 0) {
        rv = backprojection(&phs.b, &k_r.b, taylor, nfft, delta_r, &u.b, &v.b, &pos.b, &r.b,
                            &out.b);
    } else if (which == 1) {
        rv = backprojection_cuda(&phs.b, &k_r.b, taylor, nfft, delta_r, &u.b, &v.b, &pos.b, &r.b,
                                 &out.b);
    } else {
        void *args[10] = {&phs.b, &k_r.b, &taylor, &nfft, &delta_r, &u.b, &v.b, &pos.b, &r.b,
                          &out.b};
        rv = backprojection_argv(args);
    }
    std::printf("query %d rv %d k_r %d pos %d,%d r %d,%d\n", which, rv, k_r.d[0].extent,
                pos.d[0].extent, pos.d[1].extent, r.d[0].extent, r.d[1].extent);
    return rv;
}

}  // namespace

int main() {
    // volatile function pointers make the linker resolve every name the reference's driver binds
    int (*volatile to_dB)(halide_buffer_t *, halide_buffer_t *) = img_output_to_dB;
    int (*volatile to_u8)(halide_buffer_t *, double, double, halide_buffer_t *) = img_output_u8;
    int (*volatile cuda_argv)(void **) = backprojection_cuda_argv;
    void (*volatile init)(size_t) = dft_init_fftw;
    if (!to_dB || !to_u8 || !cuda_argv || !init) return 3;
    int bad = 0;
    for (int w = 0; w < 3; ++w) bad |= query(w);
    dft_destroy_fftw();
    std::printf(bad ? "FAIL\n" : "OK\n");
    return bad ? 1 : 0;
}
