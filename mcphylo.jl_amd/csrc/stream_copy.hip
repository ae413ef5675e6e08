// Achievable-HBM-bandwidth probe and counter calibration: copies a buffer of
// known size with a vectorised grid-stride kernel and reports GB/s
// (read + write bytes per second).  Usage: stream_copy [MiB] [reps]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void __launch_bounds__(256) k_copy(const double2* __restrict__ src,
                                              double2* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = src[i];
}

#define CHECK(x)                                                                   \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));           \
            return 1;                                                              \
        }                                                                          \
    } while (0)

int main(int argc, char** argv) {
    const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 4096;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
    const size_t bytes = mib << 20;
    const size_t n = bytes / sizeof(double2);
    double2 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 0, bytes));
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const int grid = prop.multiProcessorCount * 16;
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    float best = 1e30f, total = 0.f;
    for (int r = 0; r < reps; ++r) {
        CHECK(hipEventRecord(e0, 0));
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
        CHECK(hipEventRecord(e1, 0));
        CHECK(hipEventSynchronize(e1));
        float ms;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        best = ms < best ? ms : best;
        total += ms;
    }
    std::printf("{\"copy_MiB\": %zu, \"bytes_per_launch\": %zu, \"best_GBs\": %.1f, "
                "\"mean_GBs\": %.1f, \"reps\": %d}\n",
                mib, 2 * bytes, 2.0 * bytes / (best * 1e-3) / 1e9,
                2.0 * bytes * reps / (total * 1e-3) / 1e9, reps);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
