// copy_calib.hip -- plain device-to-device copy of a known size, used to
// calibrate rocprofv3's FETCH_SIZE / WRITE_SIZE counters (bytes per KiB unit)
// for the traffic figure reported next to the back end's roofline.
// Usage: copy_calib <bytes> <repeats>
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

__global__ void k_copy_calib(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int main(int argc, char** argv)
{
    const size_t bytes = argc > 1 ? strtoull(argv[1], nullptr, 10) : (256ull << 20);
    const int reps = argc > 2 ? atoi(argv[2]) : 5;
    const size_t n = bytes / sizeof(uint4);
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc((void**)&a, n * sizeof(uint4)) != hipSuccess || hipMalloc((void**)&b, n * sizeof(uint4)) != hipSuccess) {
        fprintf(stderr, "allocation failed\n");
        return 1;
    }
    hipMemset(a, 1, n * sizeof(uint4));
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    float best = 1e30f;
    for (int r = 0; r < reps; ++r) {
        hipEventRecord(e0, 0);
        hipLaunchKernelGGL(k_copy_calib, dim3(4096), dim3(256), 0, 0, a, b, n);
        hipEventRecord(e1, 0);
        hipEventSynchronize(e1);
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) best = ms;
    }
    printf("copy_calib bytes=%zu read+write=%zu best_ms=%.4f GBps=%.1f\n", n * sizeof(uint4), 2 * n * sizeof(uint4), best,
           2.0 * n * sizeof(uint4) / (best * 1e-3) / 1e9);
    hipFree(a);
    hipFree(b);
    return 0;
}
