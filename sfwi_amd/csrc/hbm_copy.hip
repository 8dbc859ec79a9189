// Calibration program for HBM traffic counters: a plain device copy of a
// known size (1 GiB read + 1 GiB written per launch, 4 launches).  The bytes
// the FETCH_SIZE / WRITE_SIZE counters report for one launch of this kernel,
// divided by 1 GiB, give the factor that converts those counters into bytes
// for the stencil kernels.  Also prints the copy bandwidth from HIP events.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void __launch_bounds__(256) hbm_copy(const float4 *__restrict__ a, float4 *__restrict__ b,
                                                size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

#define CK(x)                                                                         \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));              \
            return 1;                                                                 \
        }                                                                             \
    } while (0)

int main() {
    const size_t bytes = size_t(1) << 30;
    const size_t n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    CK(hipMalloc(&a, bytes));
    CK(hipMalloc(&b, bytes));
    CK(hipMemset(a, 0, bytes));
    CK(hipMemset(b, 0, bytes));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    for (int k = 0; k < 4; ++k) {
        CK(hipEventRecord(e0));
        hipLaunchKernelGGL(hbm_copy, dim3(8192), dim3(256), 0, 0, a, b, n);
        CK(hipGetLastError());
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms = 0;
        CK(hipEventElapsedTime(&ms, e0, e1));
        std::printf("hbm_copy 1 GiB: %.3f ms, %.1f GB/s read+write\n", ms, 2.0 * bytes / ms / 1e6);
    }
    CK(hipFree(a));
    CK(hipFree(b));
    return 0;
}
