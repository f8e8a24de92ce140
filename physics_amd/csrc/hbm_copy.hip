// hbm_copy.hip — a streaming copy of a known number of bytes, the reference
// point for reading HBM traffic counters (rocprofv3 --pmc FETCH_SIZE /
// WRITE_SIZE): a copy of B bytes compulsorily fetches B and writes B bytes,
// so counter / B is the counter's scale for the other kernels.
//
// Usage: hbm_copy [bytes per copy, default 1 GiB] [launches, default 5]
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ a, float4* __restrict__ b,
                                              size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

__global__ __launch_bounds__(256) void k_fill(float4* a, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        a[i] = make_float4((float)i, 1.f, 2.f, 3.f);
}

static void check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        exit(1);
    }
}

int main(int argc, char** argv) {
    size_t bytes = argc > 1 ? strtoull(argv[1], nullptr, 10) : (size_t)1 << 30;
    int launches = argc > 2 ? atoi(argv[2]) : 5;
    bytes = bytes / sizeof(float4) * sizeof(float4);
    const size_t n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    check(hipMalloc(&a, bytes), "hipMalloc");
    check(hipMalloc(&b, bytes), "hipMalloc");
    const int grid = 256 * 8;
    hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, 0, a, n);
    check(hipDeviceSynchronize(), "fill");
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "event");
    check(hipEventCreate(&e1), "event");
    float total = 0.f;
    for (int k = 0; k < launches; ++k) {
        check(hipEventRecord(e0, 0), "record");
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
        check(hipEventRecord(e1, 0), "record");
        check(hipEventSynchronize(e1), "sync");
        float ms = 0.f;
        check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        if (k > 0) total += ms;
    }
    const double ms = launches > 1 ? total / (launches - 1) : 0.0;
    printf("k_copy: %zu bytes read + %zu bytes written per launch, %d launches, %.4f ms/launch "
           "(after the first), %.1f GB/s\n",
           bytes, bytes, launches, ms, ms > 0 ? 2.0 * bytes / (ms * 1e-3) / 1e9 : 0.0);
    check(hipFree(a), "hipFree");
    check(hipFree(b), "hipFree");
    return 0;
}
