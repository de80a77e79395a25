// sarbp_img_output_device_noalloc: the dB / u8 image output of sarbp_img_output_device
// (img_output.cpp:10-52) for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops). It allocates nothing: the max |z| is reduced through `scratch`, a
// file-scope device array, instead of a cell the library allocates per stream.
//
//   partial_absmax_kernel   block b writes max |z|^2 over its pixels into scratch[b]
//   finalize_absmax_kernel  one block reduces the partials and writes the bits of max |z| into
//                           scratch[kNoAllocParts] (sqrt is monotone: the same value as the
//                           largest |z|)
//   db_u8 (kernels.hip)     dB and u8 of every pixel against that maximum
//
// Every call writes each partial it reads, so nothing needs clearing between calls. All calls on
// one GPU share that GPU's scratch: calls that overlap on one GPU (different streams or host
// threads) race, and their results are undefined; calls in stream order on one stream are safe.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "img_math.h"
#include "internal.h"

namespace sarbp {
namespace {

constexpr int kNoAllocThreads = 256;
constexpr int kNoAllocParts = 1024;  // partial maxima per call at most (one per block)

// The scratch of the allocation-free entry point: kNoAllocParts partial maxima (f64 |z|^2) and one
// cell for the bits of max |z|; 8200 bytes. Each GPU the library runs on holds its own copy until
// the process ends. Host code must not pass `scratch` itself to a kernel (on the host the name is a
// stand-in, not a device address): it uses the per-device address from hipGetSymbolAddress.
static __device__ __align__(256) double scratch[kNoAllocParts + 1];

std::mutex g_scratch_mu;
std::vector<double *> g_scratch_ptr;  // per device ordinal: that device's address of `scratch`

// The device address of `scratch` on device `dev` (the current device), looked up once per device.
hipError_t scratch_on(int dev, double **out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

__global__ __launch_bounds__(kNoAllocThreads) void partial_absmax_kernel(
    const double2 *__restrict__ img, int64_t n, double *__restrict__ part) {
    __shared__ double red[kNoAllocThreads / 64];
    double m = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double a2 = abs2(img[i]);
        m = a2 > m ? a2 : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kNoAllocThreads / 64; ++w) m = red[w] > m ? red[w] : m;
        part[blockIdx.x] = m;
    }
}

__global__ __launch_bounds__(kNoAllocThreads) void finalize_absmax_kernel(
    const double *__restrict__ part, int nparts, unsigned long long *__restrict__ maxbits) {
    __shared__ double red[kNoAllocThreads / 64];
    double m = 0.0;
    for (int k = threadIdx.x; k < nparts; k += kNoAllocThreads) m = part[k] > m ? part[k] : m;
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kNoAllocThreads / 64; ++w) m = red[w] > m ? red[w] : m;
        *maxbits = (unsigned long long)__double_as_longlong(sqrt(m));
    }
}

}  // namespace
}  // namespace sarbp

using namespace sarbp;

extern "C" int sarbp_img_output_device_noalloc(const void *img_dev, int64_t npix, double *dB_dev,
                                               uint8_t *u8_dev, double dB_min, double dB_max,
                                               void *stream) {
    if (npix < 0) {
        set_error("sarbp_img_output_device_noalloc: negative pixel count %lld", (long long)npix);
        return halide_error_code_buffer_extents_negative;
    }
    if (npix == 0) return 0;
    if (!img_dev) {
        set_error("sarbp_img_output_device_noalloc: img is NULL");
        return halide_error_code_buffer_argument_is_null;
    }
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    double *ptr = nullptr;
    if (e == hipSuccess) e = scratch_on(dev, &ptr);
    if (e != hipSuccess) {
        set_error("sarbp_img_output_device_noalloc: scratch address on device %d: %s", dev,
                  hipGetErrorString(e));
        return halide_error_code_device_run_failed;
    }
    hipStream_t s = (hipStream_t)stream;
    const int64_t want = (npix + kNoAllocThreads - 1) / kNoAllocThreads;
    const int parts = (int)std::min<int64_t>(want, kNoAllocParts);
    unsigned long long *maxbits = reinterpret_cast<unsigned long long *>(ptr + kNoAllocParts);
    hipLaunchKernelGGL(partial_absmax_kernel, dim3(parts), dim3(kNoAllocThreads), 0, s,
                       (const double2 *)img_dev, npix, ptr);
    hipLaunchKernelGGL(finalize_absmax_kernel, dim3(1), dim3(kNoAllocThreads), 0, s,
                       (const double *)ptr, parts, maxbits);
    e = hipGetLastError();
    if (e == hipSuccess)
        e = launch_db_u8((const double2 *)img_dev, npix, maxbits, dB_dev, u8_dev, dB_min, dB_max, s);
    if (e != hipSuccess) {
        set_error("sarbp_img_output_device_noalloc: launch failed: %s", hipGetErrorString(e));
        return halide_error_code_device_run_failed;
    }
    return 0;
}
