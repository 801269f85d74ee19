// hbm_copy — device copies of a known size: the achievable HBM bandwidth of
// the box (the denominator bench.py's roofline is read against) and the
// calibration of the FETCH_SIZE / WRITE_SIZE counters against bytes moved.
//   hbm_copy [MiB=1024] [reps=5] [variant=all|plain|wide|chunk]
// plain: one 16-byte element per thread per grid-stride iteration (4096
//        blocks; the counter calibration);
// wide:  four independent 16-byte loads in flight per thread, non-temporal;
// chunk: one 16-byte element per thread, one block per 4 KB chunk, no loop
//        (the fastest form measured: the achievable copy rate).
// Built by __graft_entry__.build() into _build/hbm_copy.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

__global__ void k_copy(const double2* __restrict__ src, double2* __restrict__ dst, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

typedef double dv2 __attribute__((ext_vector_type(2)));

__global__ void k_copy_chunk(const double2* __restrict__ src, double2* __restrict__ dst, long n) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void k_copy_wide(const dv2* __restrict__ src, dv2* __restrict__ dst, long n) {
    const long stride = (long)gridDim.x * blockDim.x;
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        dv2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
        for (int u = 0; u < 4; ++u) __builtin_nontemporal_store(v[u], dst + i + u * stride);
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

int main(int argc, char** argv) {
    const long mib = argc > 1 ? std::atol(argv[1]) : 1024;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 5;
    const char* variant = argc > 3 ? argv[3] : "all";
    const long n = mib * 1024 * 1024 / (long)sizeof(double2);
    double2 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * sizeof(double2)) != hipSuccess || hipMalloc(&b, n * sizeof(double2)) != hipSuccess) {
        std::fprintf(stderr, "hbm_copy: allocation failed\n");
        return 1;
    }
    if (hipMemset(a, 0, n * sizeof(double2)) != hipSuccess) return 1;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    for (int v = 0; v < 3; ++v) {
        const char* name = v == 0 ? "plain" : v == 1 ? "wide" : "chunk";
        if (std::strcmp(variant, "all") != 0 && std::strcmp(variant, name) != 0) continue;
        for (int r = 0; r < reps; ++r) {
            if (hipEventRecord(e0) != hipSuccess) return 1;
            if (v == 0)
                hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, a, b, n);
            else if (v == 1)
                hipLaunchKernelGGL(k_copy_wide, dim3(cus * 8), dim3(256), 0, 0, (const dv2*)a, (dv2*)b, n);
            else
                hipLaunchKernelGGL(k_copy_chunk, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, a, b, n);
            if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
            float ms = 0;
            if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
            std::printf("%s copy %ld MiB: %.3f ms, %.1f GB/s (read+write)\n", name, mib, ms,
                        2.0 * n * sizeof(double2) / ms / 1e6);
        }
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
