// Measured FMA ceilings of the current device (no counterpart in the reference): the vector
// ALU in FP64 and FP32 (plain, not packed, FMAs) and the matrix cores in FP64 and FP32
// (16x16x4 MFMA), each from a kernel of independent FMA chains that fills the device. The
// element kernels' roofline fractions are quoted against the published peaks; these rates
// show what the instruction mix can reach on the part at hand.
#include <hip/hip_runtime.h>

#include "../../include/adsgpu.h"

namespace {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int kChains = 8, kUnroll = 16;

template <typename T>
__global__ __launch_bounds__(256) void k_valu_rate(T* out, int iters, T s) {
    T a[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) a[c] = threadIdx.x * T(1e-3) + T(c);
    const T m = T(0.999999);
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
#pragma unroll
            for (int c = 0; c < kChains; ++c) a[c] = fma(a[c], m, s);
    T r = 0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) r += a[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = r;
}

__global__ __launch_bounds__(256) void k_mfma64_rate(double* out, int iters) {
    const double a = threadIdx.x * 1e-3, b = 1.0 - threadIdx.x * 1e-4;
    d4v c[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) c[k] = d4v{0, 0, 0, 0};
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int k = 0; k < kChains; ++k) c[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c[k], 0, 0, 0);
    d4v r = c[0];
#pragma unroll
    for (int k = 1; k < kChains; ++k) r += c[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = r[0] + r[1] + r[2] + r[3];
}

__global__ __launch_bounds__(256) void k_mfma32_rate(float* out, int iters) {
    const float a = threadIdx.x * 1e-3f, b = 1.0f - threadIdx.x * 1e-4f;
    f4v c[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) c[k] = f4v{0, 0, 0, 0};
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int k = 0; k < kChains; ++k) c[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c[k], 0, 0, 0);
    f4v r = c[0];
#pragma unroll
    for (int k = 1; k < kChains; ++k) r += c[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = r[0] + r[1] + r[2] + r[3];
}

}  // namespace

extern "C" int adsg_device_rates(double* tflops) {
    if (!tflops) return ADSG_EINVAL;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return ADSG_EHIP;
    const int blocks = 8 * cus, threads = 256;  // eight workgroups of four waves per CU
    void* buf = nullptr;
    if (hipMalloc(&buf, sizeof(double) * blocks * threads) != hipSuccess) return ADSG_ENOMEM;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = ADSG_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = ADSG_EHIP;
    const int iters_v = 2000, iters_m = 1000;
    const double valu_flop = 2.0 * kChains * kUnroll * double(iters_v) * blocks * threads;
    const double mfma_flop = 2.0 * 1024 * kChains * double(iters_m) * blocks * (threads / 64);
    for (int which = 0; which < 4 && rc == ADSG_OK; ++which) {
        float ms = 0;
        for (int rep = 0; rep < 2 && rc == ADSG_OK; ++rep) {  // the second launch is timed
            if (hipEventRecord(e0, nullptr) != hipSuccess) rc = ADSG_EHIP;
            switch (which) {
                case 0:
                    hipLaunchKernelGGL(k_valu_rate<double>, dim3(blocks), dim3(threads), 0, nullptr,
                                       static_cast<double*>(buf), iters_v, 1e-9);
                    break;
                case 1:
                    hipLaunchKernelGGL(k_valu_rate<float>, dim3(blocks), dim3(threads), 0, nullptr,
                                       static_cast<float*>(buf), iters_v, 1e-9f);
                    break;
                case 2:
                    hipLaunchKernelGGL(k_mfma64_rate, dim3(blocks), dim3(threads), 0, nullptr,
                                       static_cast<double*>(buf), iters_m);
                    break;
                default:
                    hipLaunchKernelGGL(k_mfma32_rate, dim3(blocks), dim3(threads), 0, nullptr,
                                       static_cast<float*>(buf), iters_m);
            }
            if (hipGetLastError() != hipSuccess || hipEventRecord(e1, nullptr) != hipSuccess ||
                hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                rc = ADSG_EHIP;
        }
        if (rc == ADSG_OK) tflops[which] = (which < 2 ? valu_flop : mfma_flop) / (ms * 1e-3) / 1e12;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    return rc;
}
