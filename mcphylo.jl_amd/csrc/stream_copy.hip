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

// four independent 16-byte loads in flight per thread before the stores
__global__ void __launch_bounds__(256) k_copy4(const double2* __restrict__ src,
                                               double2* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const double2 a = src[i], b = src[i + stride], c = src[i + 2 * stride],
                      d = src[i + 3 * stride];
        dst[i] = a;
        dst[i + stride] = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < n; i += stride) dst[i] = src[i];
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
    auto run = [&](int variant, int g, float& best, float& mean) -> int {
        auto launch = [&]() {
            if (variant == 0)
                hipLaunchKernelGGL(k_copy, dim3(g), dim3(256), 0, 0, a, b, n);
            else
                hipLaunchKernelGGL(k_copy4, dim3(g), dim3(256), 0, 0, a, b, n);
        };
        launch();
        CHECK(hipGetLastError());
        CHECK(hipDeviceSynchronize());
        best = 1e30f;
        float total = 0.f;
        for (int r = 0; r < reps; ++r) {
            CHECK(hipEventRecord(e0, 0));
            launch();
            CHECK(hipEventRecord(e1, 0));
            CHECK(hipEventSynchronize(e1));
            float ms;
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            best = ms < best ? ms : best;
            total += ms;
        }
        mean = total / reps;
        return 0;
    };
    std::printf("{\"copy_MiB\": %zu, \"bytes_per_launch\": %zu, \"reps\": %d, \"runs\": [",
                mib, 2 * bytes, reps);
    const int grids[3] = {grid, prop.multiProcessorCount * 64, prop.multiProcessorCount * 4};
    bool first = true;
    for (int variant = 0; variant < 2; ++variant)
        for (int gi = 0; gi < 3; ++gi) {
            float best, mean;
            if (run(variant, grids[gi], best, mean)) return 1;
            std::printf("%s{\"kernel\": \"%s\", \"grid\": %d, \"best_GBs\": %.1f, "
                        "\"mean_GBs\": %.1f}",
                        first ? "" : ", ", variant ? "k_copy4" : "k_copy", grids[gi],
                        2.0 * bytes / (best * 1e-3) / 1e9, 2.0 * bytes / (mean * 1e-3) / 1e9);
            first = false;
        }
    std::printf("]}\n");
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
