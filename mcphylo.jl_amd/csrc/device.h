// Device helpers shared by the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>

namespace mcp {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// One DPP move of a double (both 32-bit halves with the same pattern); lanes
// of rows outside ROW_MASK receive 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROW_MASK, 0xf, false);
    const unsigned hi =
        (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROW_MASK, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum over the wavefront's 64 lanes, returned to every lane.  Every lane
// must be active.  Register-only DPP steps (no LDS traffic, unlike lane
// shuffles): quad butterflies, half-row and row mirrors give every lane its
// 16-lane row sum, row broadcasts 15 and 31 carry the row sums up to lane
// 63, whose total is read into a scalar register.  Fixed order, so the
// result is the same on every run.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_move<0xB1, 0xf>(v);   // quad_perm [1,0,3,2]
    v += dpp_move<0x4E, 0xf>(v);   // quad_perm [2,3,0,1]
    v += dpp_move<0x141, 0xf>(v);  // row_half_mirror
    v += dpp_move<0x140, 0xf>(v);  // row_mirror
    v += dpp_move<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_move<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 63);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
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

// A value the caller knows to be uniform across the wavefront, moved to a
// scalar register.  Integer division is computed on the vector unit, so a
// quotient of uniform operands lands in a VGPR; without this the addresses
// derived from it turn wave-uniform loads into vector loads.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

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

// Makes the listed wave-uniform values (fields of a job record) live in
// scalar registers at this point.  The loads producing them cannot be sunk
// to their first uses, so they issue together and are waited on once
// instead of one dependent round trip per field.  Emits no instruction.
#define MCP_PIN_SCALAR(...) asm volatile("" ::__VA_ARGS__)

// Read-only operands of a launch (transition matrices, root frequencies,
// written by an earlier launch): the constant address space tells the
// compiler nothing in this launch stores to them, so wave-uniform reads
// become scalar loads (SGPR operands) instead of vector loads of one address.
#define MCP_CONST __attribute__((address_space(4)))

template <class T>
__device__ __forceinline__ const MCP_CONST T* cptr(const T* p) {
    return (const MCP_CONST T*)p;
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
