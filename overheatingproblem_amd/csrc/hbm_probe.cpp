// hbm_probe -- a plain device copy of a known size.
//
//   overheatingproblem_amd/lib/hbm_probe [MiB]
//
// Its kernel (copy_kernel) reads and writes exactly `MiB` MiB each, so rocprofv3's
// FETCH_SIZE / WRITE_SIZE for it calibrate what those counters read for the solver's
// kernels, and its event timing is the attainable streaming rate of the card.  Three
// grid shapes are timed after a warm-up; the best is reported.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy_kernel(const double4* __restrict__ a, double4* __restrict__ b, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

static int check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "hbm_probe: %s: %s\n", what, hipGetErrorString(e));
        return 1;
    }
    return 0;
}

int main(int argc, char** argv) {
    const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024;
    if (mib == 0 || mib > 65536) {
        std::fprintf(stderr, "hbm_probe: size must be 1..65536 MiB\n");
        return 2;
    }
    const size_t bytes = mib << 20;
    const size_t n = bytes / sizeof(double4);
    double4 *a = nullptr, *b = nullptr;
    if (check(hipMalloc(&a, bytes), "hipMalloc") || check(hipMalloc(&b, bytes), "hipMalloc"))
        return 1;
    if (check(hipMemset(a, 0, bytes), "hipMemset")) return 1;
    hipEvent_t e0, e1;
    if (check(hipEventCreate(&e0), "event") || check(hipEventCreate(&e1), "event")) return 1;
    const unsigned grids[3] = {4096, 16384, 65536};
    double best = 0.0;
    for (unsigned g : grids) {
        hipLaunchKernelGGL(copy_kernel, dim3(g), dim3(256), 0, 0, a, b, n);
        if (check(hipEventRecord(e0, 0), "record")) return 1;
        hipLaunchKernelGGL(copy_kernel, dim3(g), dim3(256), 0, 0, a, b, n);
        if (check(hipEventRecord(e1, 0), "record") || check(hipEventSynchronize(e1), "sync"))
            return 1;
        float ms = 0.0f;
        if (check(hipEventElapsedTime(&ms, e0, e1), "elapsed")) return 1;
        const double gbs = 2.0 * bytes / (ms * 1e-3) / 1e9;
        std::printf("copy %zu MiB, grid %u x 256: %.4f ms, %.1f GB/s (read+write)\n", mib, g, ms,
                    gbs);
        if (gbs > best) best = gbs;
    }
    std::printf("best %.1f GB/s; dispatches 6, each reads and writes %zu bytes\n", best, bytes);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
