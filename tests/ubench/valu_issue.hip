// Issue cost of single VALU instructions on gfx950, measured on the device.
//
// Each lane runs 8 independent chains of one instruction, written as inline asm on register
// operands only (no memory operand). Dependent instructions of one chain are 8 apart, and several
// waves share each SIMD, so the SIMD issues them back to back. One launch holds 8192 workgroups of
// 256 threads. The figure per instruction is
//     cycles per wave64 instruction per SIMD = duration x clock x SIMDs / (waves x instructions per wave)
// with the duration from HIP events around the launch and the clock from each wave's clock64() /
// wall_clock64() ratio (the shader clock against the constant-rate wall clock). The probe also
// reports the error of the hardware sin/cos (argument in revolutions) against sin(2 pi x) and cos(2 pi x),
// and of v_rsq_f64 against 1/sqrt(x).
//
// Build and run (repository root; prints a markdown table):
//     make -C tests/ubench && tests/ubench/valu_issue
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define CHECK(x)                                                                   \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #x,         \
                         hipGetErrorString(e_));                                   \
            std::exit(1);                                                          \
        }                                                                          \
    } while (0)

constexpr int kThreads = 256;
constexpr int kBlocks = 8192;
constexpr int kChains = 8;

// One op: A is the type each chain holds (written by the instruction), X the type of the extra
// operands. ASM names the chain as %0 and the extra operands as %1 and %2.
#define DEF_OP(NAME, TA, TX, ASM)                                                  \
    struct NAME {                                                                  \
        using A = TA;                                                              \
        using X = TX;                                                              \
        static constexpr const char *text = ASM;                                   \
        static __device__ __forceinline__ void step(A (&a)[kChains], X x, X y) {   \
            _Pragma("unroll") for (int k = 0; k < kChains; ++k)                    \
                asm volatile(ASM : "+v"(a[k]) : "v"(x), "v"(y));                   \
        }                                                                          \
    };

DEF_OP(FmaF64, double, double, "v_fma_f64 %0, %0, %1, %2")
DEF_OP(MulF64, double, double, "v_mul_f64 %0, %0, %1")
DEF_OP(AddF64, double, double, "v_add_f64 %0, %0, %1")
DEF_OP(RsqF64, double, double, "v_rsq_f64 %0, %0")
DEF_OP(RcpF64, double, double, "v_rcp_f64 %0, %0")
DEF_OP(FractF64, double, double, "v_fract_f64 %0, %0")
DEF_OP(FloorF64, double, double, "v_floor_f64 %0, %0")
DEF_OP(FmaF32, float, float, "v_fma_f32 %0, %0, %1, %2")
DEF_OP(FmacF32, float, float, "v_fmac_f32 %0, %1, %2")
DEF_OP(MulF32, float, float, "v_mul_f32 %0, %0, %1")
DEF_OP(RsqF32, float, float, "v_rsq_f32 %0, %0")
DEF_OP(SinF32, float, float, "v_sin_f32 %0, %0")
DEF_OP(CosF32, float, float, "v_cos_f32 %0, %0")
DEF_OP(FractF32, float, float, "v_fract_f32 %0, %0")
DEF_OP(Med3I32, int, int, "v_med3_i32 %0, %0, %1, %2")
DEF_OP(LshlI32, int, int, "v_lshlrev_b32 %0, %1, %0")
DEF_OP(AndB32, int, int, "v_and_b32 %0, %0, %1")
DEF_OP(CvtF32I32, float, int, "v_cvt_f32_i32 %0, %1")
DEF_OP(CvtF32U32, float, int, "v_cvt_f32_u32 %0, %1")
DEF_OP(CvtF32F64, float, double, "v_cvt_f32_f64 %0, %1")
DEF_OP(CvtI32F64, int, double, "v_cvt_i32_f64 %0, %1")
DEF_OP(CvtF64F32, double, float, "v_cvt_f64_f32 %0, %1")
DEF_OP(CvtFlrI32F32, int, float, "v_cvt_flr_i32_f32 %0, %1")

// The correctly rounded sqrt(double) of the device library (not a single instruction): one call
// per chain element.
struct SqrtF64Lib {
    using A = double;
    using X = double;
    static constexpr const char *text = "sqrt(double), device library";
    static __device__ __forceinline__ void step(A (&a)[kChains], X x, X) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) {
            a[k] = sqrt(a[k]) + x;
            asm volatile("" : "+v"(a[k]));
        }
    }
};

template <class Op>
__global__ __launch_bounds__(kThreads) void issue_kernel(int iters, typename Op::X x,
                                                         typename Op::X y,
                                                         unsigned long long *__restrict__ clk,
                                                         double *__restrict__ sink) {
    typename Op::A a[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) a[k] = (typename Op::A)(threadIdx.x + 2 * k + 1);
    const unsigned long long c0 = clock64();
    const unsigned long long w0 = wall_clock64();
    for (int i = 0; i < iters; ++i) Op::step(a, x, y);
    const unsigned long long c1 = clock64();
    const unsigned long long w1 = wall_clock64();
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += (double)a[k];
    const size_t gid = (size_t)blockIdx.x * kThreads + threadIdx.x;
    sink[gid] = s;  // keeps every chain live
    if ((threadIdx.x & 63) == 0) {
        const size_t w = gid >> 6;
        clk[2 * w] = c1 - c0;
        clk[2 * w + 1] = w1 - w0;
    }
}

__global__ void sincos_kernel(int n, float *__restrict__ s, float *__restrict__ c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = -0.5f + (float)i / (float)n;  // exact in f32 for n = 2^20
    s[i] = __builtin_amdgcn_sinf(x);
    c[i] = __builtin_amdgcn_cosf(x);
}

__global__ void rsq_kernel(int n, double *__restrict__ r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    r[i] = __builtin_amdgcn_rsq(1.0 + 3.0 * (double)i / (double)n);
}

struct Dev {
    int simds;
    double wall_hz;
    unsigned long long *clk;
    double *sink;
    hipEvent_t e0, e1;
};

template <class Op>
void measure(Dev &d, const char *label, int iters) {
    const size_t waves = (size_t)kBlocks * kThreads / 64;
    typename Op::X x = (typename Op::X)1, y = (typename Op::X)0;
    hipLaunchKernelGGL(issue_kernel<Op>, dim3(kBlocks), dim3(kThreads), 0, 0, 16, x, y, d.clk,
                       d.sink);  // warm-up
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(d.e0, 0));
    hipLaunchKernelGGL(issue_kernel<Op>, dim3(kBlocks), dim3(kThreads), 0, 0, iters, x, y, d.clk,
                       d.sink);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(d.e1, 0));
    CHECK(hipEventSynchronize(d.e1));
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, d.e0, d.e1));
    std::vector<unsigned long long> clk(2 * waves);
    CHECK(hipMemcpy(clk.data(), d.clk, clk.size() * sizeof(unsigned long long),
                    hipMemcpyDeviceToHost));
    double ratio = 0.0;
    size_t used = 0;
    for (size_t w = 0; w < waves; ++w)
        if (clk[2 * w + 1] > 0) {
            ratio += (double)clk[2 * w] / (double)clk[2 * w + 1];
            ++used;
        }
    const double hz = used ? ratio / (double)used * d.wall_hz : 0.0;
    const double instrs = (double)waves * (double)iters * kChains;
    const double cycles = (double)ms * 1e-3 * hz * d.simds / instrs;
    std::printf("| %s | %.2f | %.2f | %.3f |\n", label, cycles, hz * 1e-9, ms);
    std::fflush(stdout);
}

int main() {
    Dev d{};
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    int wall_khz = 0;
    CHECK(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, 0));
    d.simds = prop.multiProcessorCount * 4;
    d.wall_hz = wall_khz * 1e3;
    const size_t waves = (size_t)kBlocks * kThreads / 64;
    CHECK(hipMalloc(&d.clk, 2 * waves * sizeof(unsigned long long)));
    CHECK(hipMalloc(&d.sink, (size_t)kBlocks * kThreads * sizeof(double)));
    CHECK(hipEventCreate(&d.e0));
    CHECK(hipEventCreate(&d.e1));
    std::printf("device %s (%s), %d CUs, %d SIMDs, wall clock %d kHz\n\n", prop.name,
                prop.gcnArchName, prop.multiProcessorCount, d.simds, wall_khz);
    std::printf("| instruction | cycles per wave64 instruction per SIMD | clock GHz | launch ms |\n");
    std::printf("|---|---|---|---|\n");
    const int it = 2048;
    measure<FmaF64>(d, FmaF64::text, it);
    measure<MulF64>(d, MulF64::text, it);
    measure<AddF64>(d, AddF64::text, it);
    measure<RsqF64>(d, RsqF64::text, it / 4);
    measure<RcpF64>(d, RcpF64::text, it / 4);
    measure<FractF64>(d, FractF64::text, it);
    measure<FloorF64>(d, FloorF64::text, it);
    measure<SqrtF64Lib>(d, SqrtF64Lib::text, it / 8);
    measure<FmaF32>(d, FmaF32::text, it);
    measure<FmacF32>(d, FmacF32::text, it);
    measure<MulF32>(d, MulF32::text, it);
    measure<RsqF32>(d, RsqF32::text, it);
    measure<SinF32>(d, SinF32::text, it);
    measure<CosF32>(d, CosF32::text, it);
    measure<FractF32>(d, FractF32::text, it);
    measure<Med3I32>(d, Med3I32::text, it);
    measure<LshlI32>(d, LshlI32::text, it);
    measure<AndB32>(d, AndB32::text, it);
    measure<CvtF32I32>(d, CvtF32I32::text, it);
    measure<CvtF32U32>(d, CvtF32U32::text, it);
    measure<CvtF32F64>(d, CvtF32F64::text, it);
    measure<CvtI32F64>(d, CvtI32F64::text, it);
    measure<CvtF64F32>(d, CvtF64F32::text, it);
    measure<CvtFlrI32F32>(d, CvtFlrI32F32::text, it);

    // accuracy of the hardware sin / cos (revolutions) and of v_rsq_f64
    const int n = 1 << 20;
    float *s = nullptr, *c = nullptr;
    double *r = nullptr;
    CHECK(hipMalloc(&s, n * sizeof(float)));
    CHECK(hipMalloc(&c, n * sizeof(float)));
    CHECK(hipMalloc(&r, n * sizeof(double)));
    hipLaunchKernelGGL(sincos_kernel, dim3(n / 256), dim3(256), 0, 0, n, s, c);
    hipLaunchKernelGGL(rsq_kernel, dim3(n / 256), dim3(256), 0, 0, n, r);
    CHECK(hipGetLastError());
    std::vector<float> hs(n), hc(n);
    std::vector<double> hr(n);
    CHECK(hipMemcpy(hs.data(), s, n * sizeof(float), hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(hc.data(), c, n * sizeof(float), hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(hr.data(), r, n * sizeof(double), hipMemcpyDeviceToHost));
    const long double two_pi = 6.283185307179586476925286766559L;
    double es = 0.0, ec = 0.0, er = 0.0;
    for (int i = 0; i < n; ++i) {
        const long double x = -0.5L + (long double)i / (long double)n;
        es = std::fmax(es, (double)std::fabs((long double)hs[i] - std::sin(two_pi * x)));
        ec = std::fmax(ec, (double)std::fabs((long double)hc[i] - std::cos(two_pi * x)));
        const long double xr = 1.0L + 3.0L * (long double)i / (long double)n;
        const long double want = 1.0L / std::sqrt(xr);
        er = std::fmax(er, (double)std::fabs(((long double)hr[i] - want) / want));
    }
    std::printf("\nv_sin_f32 max abs error vs sin(2 pi x), x in [-0.5, 0.5): %.3g\n", es);
    std::printf("v_cos_f32 max abs error vs cos(2 pi x), x in [-0.5, 0.5): %.3g\n", ec);
    std::printf("v_rsq_f64 max relative error vs 1/sqrt(x), x in [1, 4): %.3g\n", er);
    CHECK(hipFree(s));
    CHECK(hipFree(c));
    CHECK(hipFree(r));
    CHECK(hipFree(d.clk));
    CHECK(hipFree(d.sink));
    return 0;
}
