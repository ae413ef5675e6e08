// Level kernels for 20-state models on FP64 matrix cores (v_mfma_f64_16x16x4f64).
//
// For a 20-state model the per-node work is a dense 20 x 20 contraction per
// pattern column: Down = P . F (post-order), dP . F and G = P^T . T
// (pre-order).  A task is one wavefront over 64 pattern columns of one
// (node, category), processed as four 16-column MFMA tiles.  With i = lane % 16
// and g = lane / 16 the operand layouts of v_mfma_f64_16x16x4f64 on gfx950 are
//   A (16 x 4):  lane holds A[i][g]
//   B (4 x 16):  lane holds B[g][i]
//   D (16 x 16): lane holds D[g + 4 j][i] in result register j (j < 4)
// (checked on the device with a probe before this kernel was written).  A
// 20-row product is two 16-row blocks; the lane then holds rows g + 4 t,
// t < 5, of its column i (rows 0..15 from block 0, row 16 + g from block 1).
// Those are exactly the rows the same lane supplies as the B operand of the
// next contraction (k-chunk t covers rows 4 t .. 4 t + 3), so T = G_p (.) Down
// feeds P^T . T without any lane exchange.  Partials keep the engine's
// [r][c][s] slab layout; column reductions over states combine the four
// lanes of a column (xor 16, xor 32).
#include <hip/hip_runtime.h>
#include <math.h>

#include "device.h"
#include "engine.h"

namespace mcp {

typedef double dvec4 __attribute__((ext_vector_type(4)));

constexpr int kS20 = 20;

__device__ __forceinline__ const ChildRef& child_of20(const LevelJob& J, int q) {
    return q < kInlineChildren ? J.kid[q] : J.more[q - kInlineChildren];
}

// A-operand fragments of a 20 x 20 row-major matrix M (or of M^T):
// a[5 mb + kc] = M[16 mb + i][4 kc + g] (rows >= 20 are zero).
__device__ __forceinline__ void afrag20(const MCP_GLOBAL double* __restrict__ M, bool transpose,
                                        int lane, double (&a)[10]) {
    const int i = lane & 15, g = lane >> 4;
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int kc = 0; kc < 5; ++kc) {
            const int row = 16 * mb + i, col = 4 * kc + g;
            a[5 * mb + kc] = row < kS20 ? (transpose ? M[col * kS20 + row] : M[row * kS20 + col])
                                        : 0.0;
        }
}

// The same fragments of a matrix staged in LDS (row-major, 20 x 20).
__device__ __forceinline__ void afrag20_lds(const double* M, bool transpose, int lane,
                                            double (&a)[10]) {
    const int i = lane & 15, g = lane >> 4;
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int kc = 0; kc < 5; ++kc) {
            const int row = 16 * mb + i, col = 4 * kc + g;
            a[5 * mb + kc] = row < kS20 ? (transpose ? M[col * kS20 + row] : M[row * kS20 + col])
                                        : 0.0;
        }
}

// y[t] = row g + 4t of (M . X) for the lane's column, from B fragments
// xb[t] = X[4t + g][column].
__device__ __forceinline__ void mm20(const double (&a)[10], const double (&xb)[5], double (&y)[5]) {
    dvec4 c0 = {0.0, 0.0, 0.0, 0.0};
    dvec4 c1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kc = 0; kc < 5; ++kc) {
        c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kc], xb[kc], c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[5 + kc], xb[kc], c1, 0, 0, 0);
    }
    y[0] = c0[0];
    y[1] = c0[1];
    y[2] = c0[2];
    y[3] = c0[3];
    y[4] = c1[0];
}

__device__ __forceinline__ double col_sum(double v) {
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

__device__ __forceinline__ double col_max(double v) {
    double o = __shfl_xor(v, 16, 64);
    v = (v < o) ? o : v;
    o = __shfl_xor(v, 32, 64);
    v = (v < o) ? o : v;
    return v;
}

// Message of one child for the lane's column c (rows g + 4t): the leaf's
// P.tip row, or P . F with F loaded as B fragments into f.
__device__ __forceinline__ void msg20(const ChildRef& ch, const LevelJob& J, int r, int c, int g,
                                      const double (&afr)[10], double (&f)[5], double (&m)[5]) {
    if (ch.codes) {
        const int code = (int)gptr(ch.codes)[c];
        const MCP_GLOBAL double* row = gptr(ch.P) + ((size_t)r * J.K + code) * kS20;
#pragma unroll
        for (int t = 0; t < 5; ++t) m[t] = row[g + 4 * t];
    } else {
        const MCP_GLOBAL double* src = gptr(ch.F) + ((size_t)r * J.Cp + c) * kS20;
#pragma unroll
        for (int t = 0; t < 5; ++t) f[t] = src[4 * t + g];
        mm20(afr, f, m);
    }
}

// post-order task: 64 columns of one (node, category)
__device__ __forceinline__ void fwd_task20(const LevelJob& J, int r, int tile,
                                           double* __restrict__ parts) {
    const int lane = (int)(threadIdx.x & 63), i = lane & 15, g = lane >> 4;
    double acc[4][5];
#pragma unroll
    for (int st = 0; st < 4; ++st)
#pragma unroll
        for (int t = 0; t < 5; ++t) acc[st][t] = 1.0;
    for (int q = 0; q < J.nchild; ++q) {
        const ChildRef& ch = child_of20(J, q);
        double afr[10];
        if (!ch.codes) afrag20(gptr(ch.P) + (size_t)r * kS20 * kS20, false, lane, afr);
#pragma unroll
        for (int st = 0; st < 4; ++st) {
            const int c = tile * 64 + 16 * st + i;
            double f[5], m[5];
            msg20(ch, J, r, c, g, afr, f, m);
#pragma unroll
            for (int t = 0; t < 5; ++t) acc[st][t] *= m[t];
        }
    }
    const MCP_GLOBAL double* wts = gptr(J.weights);
    double contrib = 0.0;
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        const int c = tile * 64 + 16 * st + i;
        const bool own = g == 0 && c < J.C;
        if (J.is_root) {
            // root_sum (VectorizedFunctions.jl:76-87)
            const MCP_GLOBAL double* pi = gptr(J.pi);
            double s = 0.0;
#pragma unroll
            for (int t = 0; t < 5; ++t) s = fma(acc[st][t], pi[g + 4 * t], s);
            s = col_sum(s);
            const double v = wts[c] * log(s);
            contrib += own ? v : 0.0;
        } else {
            // by_max! (VectorizedFunctions.jl:31-49)
            double mx = -INFINITY;
#pragma unroll
            for (int t = 0; t < 5; ++t) mx = (mx < acc[st][t]) ? acc[st][t] : mx;
            mx = col_max(mx);
            const double inv = 1.0 / mx;
            MCP_GLOBAL double* dst = gptr_w(J.F) + ((size_t)r * J.Cp + c) * kS20;
#pragma unroll
            for (int t = 0; t < 5; ++t) dst[g + 4 * t] = acc[st][t] * inv;
            const double v = wts[c] * log(mx);
            contrib += own ? v : 0.0;
        }
    }
    const double bs = wave_sum(contrib);
    if (lane == 0) parts[J.lpart + r * J.ntiles + tile] = bs;
}

// Pre-order contribution of one child for the lane's column, given
// T = G_p (.) siblings' messages, the child's F (internal) and message m:
// num = T . (dP F), col = T . m, G = P^T T rescaled by its column maximum.
__device__ __forceinline__ double child_adj20(const ChildRef& ch, const LevelJob& J, int r, int c,
                                              int g, const double (&T)[5], const double (&f)[5],
                                              const double (&m)[5], const double (&dpfr)[10],
                                              const double (&ptfr)[10]) {
    double num = 0.0, col = 0.0;
#pragma unroll
    for (int t = 0; t < 5; ++t) col = fma(T[t], m[t], col);
    if (ch.codes) {
        const int code = (int)gptr(ch.codes)[c];
        const MCP_GLOBAL double* row = gptr(ch.dP) + ((size_t)r * J.K + code) * kS20;
#pragma unroll
        for (int t = 0; t < 5; ++t) num = fma(T[t], row[g + 4 * t], num);
    } else {
        double dpf[5], G[5];
        mm20(dpfr, f, dpf);
#pragma unroll
        for (int t = 0; t < 5; ++t) num = fma(T[t], dpf[t], num);
        mm20(ptfr, T, G);
        double mx = -INFINITY;
#pragma unroll
        for (int t = 0; t < 5; ++t) mx = (mx < G[t]) ? G[t] : mx;
        mx = col_max(mx);
        const double inv = 1.0 / mx;
        MCP_GLOBAL double* dst = gptr_w(ch.G) + ((size_t)r * J.Cp + c) * kS20;
#pragma unroll
        for (int t = 0; t < 5; ++t) dst[g + 4 * t] = G[t] * inv;
    }
    num = col_sum(num);
    col = col_sum(col);
    return num / col;
}

__device__ __forceinline__ void child_frags20(const ChildRef& ch, int r, int lane, double (&p)[10],
                                              double (&dp)[10], double (&pt)[10]) {
    if (ch.codes) return;
    const MCP_GLOBAL double* P = gptr(ch.P) + (size_t)r * kS20 * kS20;
    const MCP_GLOBAL double* dP = gptr(ch.dP) + (size_t)r * kS20 * kS20;
    afrag20(P, false, lane, p);
    afrag20(dP, false, lane, dp);
    afrag20(P, true, lane, pt);
}

__device__ __forceinline__ void gp_load20(const LevelJob& J, int r, int c, int g, double (&gp)[5]) {
    if (J.G) {
        const MCP_GLOBAL double* src = gptr(J.G) + ((size_t)r * J.Cp + c) * kS20;
#pragma unroll
        for (int t = 0; t < 5; ++t) gp[t] = src[g + 4 * t];
    } else {
        const MCP_GLOBAL double* pi = gptr(J.pi);
#pragma unroll
        for (int t = 0; t < 5; ++t) gp[t] = pi[g + 4 * t];
    }
}

// The two children's P and dP of one (parent, category) staged in LDS by
// k_level20 (smat: Pa, dPa, Pb, dPb, 400 doubles each; a leaf child's
// slots are unused).  The pre-order task needs P, dP and P^T fragments of
// both children for each of its four 16-column subtiles; the registers do
// not hold them across subtiles (two waves per SIMD), and re-reading them
// from LDS replaces about 60 scattered global loads per subtile.
constexpr int kStageDoubles = 4 * kS20 * kS20;

__device__ __forceinline__ void stage_mats20(const LevelJob& J, int r, double* smat) {
    const int lane = (int)(threadIdx.x & 63);
    const ChildRef& a = J.kid[0];
    const ChildRef& b = J.kid[1];
    const size_t off = (size_t)r * kS20 * kS20;
    for (int i = lane; i < kS20 * kS20; i += 64) {
        if (!a.codes) {
            smat[i] = gptr(a.P)[off + i];
            smat[kS20 * kS20 + i] = gptr(a.dP)[off + i];
        }
        if (!b.codes) {
            smat[2 * kS20 * kS20 + i] = gptr(b.P)[off + i];
            smat[3 * kS20 * kS20 + i] = gptr(b.dP)[off + i];
        }
    }
}

// pre-order task: 64 columns of one (parent, category); smat holds the
// children's matrices for (J, r) when J has two children
__device__ __forceinline__ void adj_task20(const LevelJob& J, int r, int tile,
                                           double* __restrict__ parts, const double* smat) {
    const int lane = (int)(threadIdx.x & 63), i = lane & 15, g = lane >> 4;
    const MCP_GLOBAL double* wts = gptr(J.weights);
    if (J.nchild == 2) {
        const ChildRef& a = J.kid[0];
        const ChildRef& b = J.kid[1];
        const double* Pa = smat;
        const double* dPa = smat + kS20 * kS20;
        const double* Pb = smat + 2 * kS20 * kS20;
        const double* dPb = smat + 3 * kS20 * kS20;
        double ca = 0.0, cb = 0.0;
#pragma unroll 1
        for (int st = 0; st < 4; ++st) {
            // matrix fragments are re-read per subtile from LDS, to keep the
            // live register set small
            const int c = tile * 64 + 16 * st + i;
            double gp[5], fa[5], ma[5], fb[5], mb[5], T[5];
            gp_load20(J, r, c, g, gp);
            {
                double fr[10];
                if (!a.codes) afrag20_lds(Pa, false, lane, fr);
                msg20(a, J, r, c, g, fr, fa, ma);
                if (!b.codes) afrag20_lds(Pb, false, lane, fr);
                msg20(b, J, r, c, g, fr, fb, mb);
            }
            const bool own = g == 0 && c < J.C;
            double dfr[10], tfr[10];
#pragma unroll
            for (int t = 0; t < 5; ++t) T[t] = gp[t] * mb[t];
            if (!a.codes) {
                afrag20_lds(dPa, false, lane, dfr);
                afrag20_lds(Pa, true, lane, tfr);
            }
            const double va = wts[c] * child_adj20(a, J, r, c, g, T, fa, ma, dfr, tfr);
#pragma unroll
            for (int t = 0; t < 5; ++t) T[t] = gp[t] * ma[t];
            if (!b.codes) {
                afrag20_lds(dPb, false, lane, dfr);
                afrag20_lds(Pb, true, lane, tfr);
            }
            const double vb = wts[c] * child_adj20(b, J, r, c, g, T, fb, mb, dfr, tfr);
            ca += own ? va : 0.0;
            cb += own ? vb : 0.0;
        }
        const double sa = wave_sum(ca);
        const double sb = wave_sum(cb);
        if (lane == 0) {
            parts[a.gpart + r * J.ntiles + tile] = sa;
            parts[b.gpart + r * J.ntiles + tile] = sb;
        }
        return;
    }
    // any other arity: siblings' messages recomputed per child
    for (int q = 0; q < J.nchild; ++q) {
        const ChildRef& u = child_of20(J, q);
        double pu[10], dpu[10], ptu[10];
        child_frags20(u, r, lane, pu, dpu, ptu);
        double cu = 0.0;
#pragma unroll 1
        for (int st = 0; st < 4; ++st) {
            const int c = tile * 64 + 16 * st + i;
            double T[5], fu[5], mu_[5];
            gp_load20(J, r, c, g, T);
            for (int q2 = 0; q2 < J.nchild; ++q2) {
                if (q2 == q) continue;
                const ChildRef& w = child_of20(J, q2);
                double pw[10], fw[5], mw[5];
                if (!w.codes) afrag20(gptr(w.P) + (size_t)r * kS20 * kS20, false, lane, pw);
                msg20(w, J, r, c, g, pw, fw, mw);
#pragma unroll
                for (int t = 0; t < 5; ++t) T[t] *= mw[t];
            }
            msg20(u, J, r, c, g, pu, fu, mu_);
            const double v = wts[c] * child_adj20(u, J, r, c, g, T, fu, mu_, dpu, ptu);
            cu += (g == 0 && c < J.C) ? v : 0.0;
        }
        const double su = wave_sum(cu);
        if (lane == 0) parts[u.gpart + r * J.ntiles + tile] = su;
    }
}

// Same launch contract as k_level (single-wave blocks, contiguous task chunks).
#ifndef MCP_ADJ20_WAVES_PER_EU
#define MCP_ADJ20_WAVES_PER_EU 2
#endif

template <bool ADJ>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(ADJ ? MCP_ADJ20_WAVES_PER_EU : 1))) k_level20(const LevelJob* __restrict__ jobs,
                                                const int32_t* __restrict__ task_off,
                                                const int32_t* __restrict__ blk_job, int total,
                                                int chunk, double* __restrict__ parts) {
    const int t0 = blockIdx.x * chunk;
    const int t1 = min(total, t0 + chunk);
    int j = blk_job[blockIdx.x];
    int jbeg = task_off[j], jend = task_off[j + 1];
    __shared__ double smat[ADJ ? kStageDoubles : 1];
    int sj = -1, sr = -1;  // (job, category) whose matrices smat holds
    for (int t = t0; t < t1; ++t) {
        while (t >= jend) {
            ++j;
            jbeg = jend;
            jend = task_off[j + 1];
        }
        const LevelJob& J = jobs[j];
        const int local = t - jbeg;
        const int r = uniform(local / J.ntiles);
        const int tile = uniform(local - r * J.ntiles);
        if constexpr (ADJ) {
            if (J.nchild == 2 && (j != sj || r != sr)) {
                __syncthreads();  // the previous task's fragment reads are done
                stage_mats20(J, r, smat);
                __syncthreads();
                sj = j;
                sr = r;
            }
            adj_task20(J, r, tile, parts, smat);
        } else {
            fwd_task20(J, r, tile, parts);
        }
    }
}

hipError_t launch_level20(bool adjoint, const LevelJob* jobs, const int32_t* task_off,
                          const int32_t* blk_job, int total, int chunk, int grid, double* parts,
                          hipStream_t st) {
    if (total <= 0 || grid <= 0) return hipSuccess;
    if (adjoint)
        hipLaunchKernelGGL(k_level20<true>, dim3(grid), dim3(64), 0, st, jobs, task_off, blk_job,
                           total, chunk, parts);
    else
        hipLaunchKernelGGL(k_level20<false>, dim3(grid), dim3(64), 0, st, jobs, task_off, blk_job,
                           total, chunk, parts);
    return hipGetLastError();
}

}  // namespace mcp
