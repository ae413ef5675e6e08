// Device helpers shared by the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>

namespace mcp {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over the block; the result is valid in thread 0.  `lds` holds one
// double per wavefront.
__device__ __forceinline__ double block_sum(double v, double* lds) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < nw; ++w) s += lds[w];
    return s;
}

// Pointers read out of job records are generic; casting them to the global
// address space lets the compiler emit global (not flat) loads and stores,
// which wait on the vector-memory counter only.
#define MCP_GLOBAL __attribute__((address_space(1)))

template <class T>
__device__ __forceinline__ const MCP_GLOBAL T* gptr(const T* p) {
    return (const MCP_GLOBAL T*)p;
}
template <class T>
__device__ __forceinline__ MCP_GLOBAL T* gptr_w(T* p) {
    return (MCP_GLOBAL T*)p;
}

typedef double dvec2 __attribute__((ext_vector_type(2)));

template <int S>
__device__ __forceinline__ void load_row(const MCP_GLOBAL double* __restrict__ p, double (&v)[S]) {
    if constexpr (S % 2 == 0) {
        const MCP_GLOBAL dvec2* q = reinterpret_cast<const MCP_GLOBAL dvec2*>(p);
#pragma unroll
        for (int i = 0; i < S / 2; ++i) {
            const dvec2 t = q[i];
            v[2 * i] = t.x;
            v[2 * i + 1] = t.y;
        }
    } else {
#pragma unroll
        for (int i = 0; i < S; ++i) v[i] = p[i];
    }
}

template <int S>
__device__ __forceinline__ void store_row(MCP_GLOBAL double* __restrict__ p, const double (&v)[S]) {
    if constexpr (S % 2 == 0) {
        MCP_GLOBAL dvec2* q = reinterpret_cast<MCP_GLOBAL dvec2*>(p);
#pragma unroll
        for (int i = 0; i < S / 2; ++i) {
            dvec2 t;
            t.x = v[2 * i];
            t.y = v[2 * i + 1];
            q[i] = t;
        }
    } else {
#pragma unroll
        for (int i = 0; i < S; ++i) p[i] = v[i];
    }
}

}  // namespace mcp
