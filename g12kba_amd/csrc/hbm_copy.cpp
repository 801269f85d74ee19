// hbm_copy — plain device copy of a known size, used to calibrate the
// FETCH_SIZE / WRITE_SIZE hardware counters against algorithmic bytes.
//   hbm_copy [MiB=1024] [reps=5]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy(const double2* __restrict__ src, double2* __restrict__ dst, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int main(int argc, char** argv) {
    const long mib = argc > 1 ? std::atol(argv[1]) : 1024;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 5;
    const long n = mib * 1024 * 1024 / (long)sizeof(double2);
    double2 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * sizeof(double2)) != hipSuccess || hipMalloc(&b, n * sizeof(double2)) != hipSuccess) {
        std::fprintf(stderr, "hbm_copy: allocation failed\n");
        return 1;
    }
    if (hipMemset(a, 0, n * sizeof(double2)) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    for (int r = 0; r < reps; ++r) {
        if (hipEventRecord(e0) != hipSuccess) return 1;
        hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, a, b, n);
        if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
        std::printf("copy %ld MiB: %.3f ms, %.1f GB/s (read+write)\n", mib, ms, 2.0 * n * sizeof(double2) / ms / 1e6);
    }
    hipFree(a);
    hipFree(b);
    return 0;
}
