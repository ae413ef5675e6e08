// Pruning likelihood + branch-length adjoint kernels for gfx950 (FP64).
//
// Kernel map (reference loop nest -> kernel):
//   parallel_transition_prob / R_gemmturbo_large! / L_gemmturbo_large! /
//   R_gemmturbo_large_ptg!  (VectorizedFunctions.jl:89-168)   -> k_build_P, k_tip_tables
//   comb_sum_product_loop! + by_max! + root_sum
//   (VectorizedFunctions.jl:31-49,76-87,187-213;
//    LikelihoodCalculator_Node.jl:82-114)                     -> k_fwd_level
//   pre-order loop of FelsensteinFunction/6 with turbo_mul!,
//   mygemmturbo!, bymax! (LikelihoodCalculator_Node.jl:18-74;
//   VectorizedFunctions.jl:2-11,51-74,170-177)                -> k_adj_level
//   sums over (pattern, category) of log scalers / gradients  -> k_reduce
//
// Work decomposition of the level kernels: a job is one node of one item; it
// expands into R * ntiles tasks, one per (category r, tile of `tile`
// consecutive patterns).  One thread owns one pattern column of one
// category and keeps the S-state vectors of that column in registers; the
// transition matrices of a task are block-uniform, so they are read through
// scalar loads and enter the FMAs as SGPR operands.  Blocks are persistent:
// each strides over the task list of the launch.
//
// Determinism: per-task sums are reduced with a fixed wavefront shuffle tree
// plus a fixed LDS order, written to one slot per task, and summed by
// k_reduce in a fixed order, so results are bitwise reproducible run to run.
#include <hip/hip_runtime.h>
#include <math.h>

#include "engine.h"

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

template <int S>
__device__ __forceinline__ void load_row(const double* __restrict__ p, double (&v)[S]) {
    if constexpr (S % 2 == 0) {
        const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
        for (int i = 0; i < S / 2; ++i) {
            double2 t = q[i];
            v[2 * i] = t.x;
            v[2 * i + 1] = t.y;
        }
    } else {
#pragma unroll
        for (int i = 0; i < S; ++i) v[i] = p[i];
    }
}

template <int S>
__device__ __forceinline__ void store_row(double* __restrict__ p, const double (&v)[S]) {
    if constexpr (S % 2 == 0) {
        double2* q = reinterpret_cast<double2*>(p);
#pragma unroll
        for (int i = 0; i < S / 2; ++i) q[i] = make_double2(v[2 * i], v[2 * i + 1]);
    } else {
#pragma unroll
        for (int i = 0; i < S; ++i) p[i] = v[i];
    }
}

// Locate the job owning task t: task_off[j] <= t < task_off[j+1].
__device__ __forceinline__ int find_job(const int32_t* __restrict__ task_off, int njobs, int t) {
    int lo = 0, hi = njobs;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (task_off[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

// Message a child u sends up its branch in column (c, r):
// msg[s] = sum_s1 P_u[s][s1] f_u[s1]; for a leaf it is the precomputed row
// PT[leaf][r][code].  `f` receives the child's partial (internal children).
template <int S>
__device__ __forceinline__ void child_message(const Item& it, const NodeInfo& nu, int u, int r,
                                              int c, double (&msg)[S], double (&f)[S]) {
    if (nu.leaf >= 0) {
        const int code = it.codes[(size_t)nu.leaf * it.Cp + c];
        const double* pt = it.PT + (((size_t)nu.leaf * it.R + r) * it.K + code) * S;
#pragma unroll
        for (int s = 0; s < S; ++s) { msg[s] = pt[s]; f[s] = 0.0; }
    } else {
        load_row<S>(it.F + (((size_t)nu.fslot * it.R + r) * it.Cp + c) * S, f);
        const double* __restrict__ P = it.P + ((size_t)u * it.R + r) * (S * S);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            double a = 0.0;
#pragma unroll
            for (int s1 = 0; s1 < S; ++s1) a = fma(f[s1], P[s * S + s1], a);
            msg[s] = a;
        }
    }
}

// ---------------------------------------------------------------------------
// transition matrices: P = U diag(exp(mu t D rho)) Uinv and dP/dt
// ---------------------------------------------------------------------------

__device__ __forceinline__ int find_item(const int32_t* __restrict__ off, int n, long long e) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (off[mid] <= e) lo = mid; else hi = mid;
    }
    return lo;
}

// One thread per element (item, branch b, r, m, n).  elem_off[i] is the
// first element of item i ((N-1)*R*S*S elements each).
__global__ void k_build_P(const ModelDev* __restrict__ models, const int32_t* __restrict__ elem_off,
                          int nitems) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= elem_off[nitems]) return;
    const int i = find_item(elem_off, nitems, e);
    const ModelDev& md = models[i];
    const int S = md.S;
    long long k = e - elem_off[i];
    const int n = (int)(k % S); k /= S;
    const int m = (int)(k % S); k /= S;
    const int r = (int)(k % md.R);
    const int b = (int)(k / md.R);
    const double t = md.blv[b];
    const double rho = md.rates[r];
    double p = 0.0, dp = 0.0;
    for (int q = 0; q < S; ++q) {
        // argument order of VectorizedFunctions.jl:106,131: mu * blv * D * rate
        const double ex = exp(md.mu * t * md.D[q] * rho);
        const double u = md.U[m + S * q];
        const double ui = md.Uinv[q + S * n];
        p = fma(u * ex, ui, p);
        dp = fma(u * md.D[q] * rho * md.mu * ex, ui, dp);
    }
    const size_t o = (((size_t)b * md.R + r) * S + m) * S + n;
    md.P[o] = p;
    md.dP[o] = dp;
}

// One thread per element (item, branch b, r, code k, state s); only leaf
// branches write: PT[leaf][r][k][s] = sum_s1 P_b[r][s][s1] tiptab[k][s1].
__global__ void k_tip_tables(const ModelDev* __restrict__ models, const int32_t* __restrict__ elem_off,
                             int nitems) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= elem_off[nitems]) return;
    const int i = find_item(elem_off, nitems, e);
    const ModelDev& md = models[i];
    const int S = md.S;
    long long k = e - elem_off[i];
    const int s = (int)(k % S); k /= S;
    const int code = (int)(k % md.K); k /= md.K;
    const int r = (int)(k % md.R);
    const int b = (int)(k / md.R);
    const int leaf = md.leaf_of[b];
    if (leaf < 0) return;
    const double* P = md.P + ((size_t)b * md.R + r) * S * S + (size_t)s * S;
    const double* dP = md.dP + ((size_t)b * md.R + r) * S * S + (size_t)s * S;
    const double* tv = md.tiptab + (size_t)code * S;
    double a = 0.0, da = 0.0;
    for (int s1 = 0; s1 < S; ++s1) {
        a = fma(P[s1], tv[s1], a);
        da = fma(dP[s1], tv[s1], da);
    }
    const size_t o = (((size_t)leaf * md.R + r) * md.K + code) * S + s;
    md.PT[o] = a;
    md.dPT[o] = da;
}

// ---------------------------------------------------------------------------
// post-order level: partial of every internal node of the level
// ---------------------------------------------------------------------------

template <int S>
__device__ __forceinline__ void fwd_task(const Item& it, int v, int r, int tile, double* red,
                                         double* __restrict__ parts) {
    const NodeInfo nv = it.nodes[v];
    const int c = tile * it.tile + (int)threadIdx.x;
    const bool valid = c < it.C;
    double contrib = 0.0;
    if (valid) {
        double acc[S];
#pragma unroll
        for (int s = 0; s < S; ++s) acc[s] = 1.0;
        for (int q = 0; q < nv.nchild; ++q) {
            const int u = it.children[nv.first_child + q];
            const NodeInfo nu = it.nodes[u];
            double msg[S], f[S];
            child_message<S>(it, nu, u, r, c, msg, f);
#pragma unroll
            for (int s = 0; s < S; ++s) acc[s] *= msg[s];
        }
        if (nv.is_root) {
            // root_sum: log sum_s pi_s Ftilde_s (VectorizedFunctions.jl:76-87)
            double t = 0.0;
#pragma unroll
            for (int s = 0; s < S; ++s) t = fma(acc[s], it.pi[s], t);
            contrib = it.weights[c] * log(t);
        } else {
            // by_max!: rescale by the state maximum, keep its log (VectorizedFunctions.jl:31-49)
            double m = -INFINITY;
#pragma unroll
            for (int s = 0; s < S; ++s) m = (m < acc[s]) ? acc[s] : m;
            const double inv = 1.0 / m;
#pragma unroll
            for (int s = 0; s < S; ++s) acc[s] *= inv;
            store_row<S>(it.F + (((size_t)nv.fslot * it.R + r) * it.Cp + c) * S, acc);
            contrib = it.weights[c] * log(m);
        }
    }
    const double bs = block_sum(contrib, red);
    if (threadIdx.x == 0) parts[nv.lpart + r * it.ntiles + tile] = bs;
}

template <int S>
__global__ void __launch_bounds__(256) k_fwd_level(const Item* __restrict__ items,
                                                   const Job* __restrict__ jobs,
                                                   const int32_t* __restrict__ task_off, int njobs,
                                                   double* __restrict__ parts) {
    __shared__ double red[16];
    const int total = task_off[njobs];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int j = find_job(task_off, njobs, t);
        const Job jb = jobs[j];
        const Item& it = items[jb.item];
        const int local = t - task_off[j];
        const int r = local / it.ntiles;
        const int tile = local - r * it.ntiles;
        fwd_task<S>(it, jb.node, r, tile, red, parts);
    }
}

// ---------------------------------------------------------------------------
// pre-order level: for every internal parent p of the level and each child u
//   T_u = G_p (.) prod_{w != u} msg_w                      (:28-37)
//   g_u = sum_m T_u[m] (dP_u f_u)[m]                        (:39-48)
//   G_u[k] = sum_m P_u[m][k] T_u[m]                         (:50-57)
//   col = sum_k f_u[k] G_u[k];  grad_u += w_c g_u / col     (:59-69)
//   G_u /= max_k G_u[k] (internal u)                        (:71-73)
// (line numbers: LikelihoodCalculator_Node.jl).  For a leaf u, g_u and col
// are dot products of T_u with the leaf's dP.tip and P.tip rows.
// ---------------------------------------------------------------------------

template <int S>
__device__ __forceinline__ double child_adjoint(const Item& it, const NodeInfo& nu, int u, int r,
                                                int c, const double (&T)[S], const double (&f)[S],
                                                const double (&msg)[S]) {
    double num = 0.0, col = 0.0;
    if (nu.leaf >= 0) {
        const int code = it.codes[(size_t)nu.leaf * it.Cp + c];
        const double* dpt = it.dPT + (((size_t)nu.leaf * it.R + r) * it.K + code) * S;
#pragma unroll
        for (int m = 0; m < S; ++m) {
            num = fma(T[m], dpt[m], num);
            col = fma(T[m], msg[m], col);
        }
    } else {
        const double* __restrict__ P = it.P + ((size_t)u * it.R + r) * (S * S);
        const double* __restrict__ dP = it.dP + ((size_t)u * it.R + r) * (S * S);
        double g[S];
#pragma unroll
        for (int k = 0; k < S; ++k) g[k] = 0.0;
#pragma unroll
        for (int m = 0; m < S; ++m) {
            double dm = 0.0;
#pragma unroll
            for (int k = 0; k < S; ++k) {
                dm = fma(dP[m * S + k], f[k], dm);
                g[k] = fma(P[m * S + k], T[m], g[k]);
            }
            num = fma(T[m], dm, num);
        }
        double mx = -INFINITY;
#pragma unroll
        for (int k = 0; k < S; ++k) {
            col = fma(f[k], g[k], col);
            mx = (mx < g[k]) ? g[k] : mx;
        }
        const double inv = 1.0 / mx;
#pragma unroll
        for (int k = 0; k < S; ++k) g[k] *= inv;
        double* G = nu.gbuf ? it.G1 : it.G0;
        store_row<S>(G + (((size_t)nu.gslot * it.R + r) * it.Cp + c) * S, g);
    }
    return it.weights[c] * (num / col);
}

template <int S>
__device__ __forceinline__ void adj_task(const Item& it, int p, int r, int tile, double* red,
                                         double* __restrict__ parts) {
    const NodeInfo np = it.nodes[p];
    const int c = tile * it.tile + (int)threadIdx.x;
    const bool valid = c < it.C;
    const int cc = valid ? c : 0;
    double gp[S];
    if (np.is_root) {
#pragma unroll
        for (int s = 0; s < S; ++s) gp[s] = it.pi[s];
    } else {
        const double* G = np.gbuf ? it.G1 : it.G0;
        if (valid) load_row<S>(G + (((size_t)np.gslot * it.R + r) * it.Cp + cc) * S, gp);
        else {
#pragma unroll
            for (int s = 0; s < S; ++s) gp[s] = 0.0;
        }
    }
    const int A = np.nchild;
    if (A == 2) {
        const int u0 = it.children[np.first_child];
        const int u1 = it.children[np.first_child + 1];
        const NodeInfo n0 = it.nodes[u0], n1 = it.nodes[u1];
        double c0 = 0.0, c1 = 0.0;
        if (valid) {
            double m0[S], f0[S], m1[S], f1[S], T[S];
            child_message<S>(it, n0, u0, r, cc, m0, f0);
            child_message<S>(it, n1, u1, r, cc, m1, f1);
#pragma unroll
            for (int s = 0; s < S; ++s) T[s] = gp[s] * m1[s];
            c0 = child_adjoint<S>(it, n0, u0, r, cc, T, f0, m0);
#pragma unroll
            for (int s = 0; s < S; ++s) T[s] = gp[s] * m0[s];
            c1 = child_adjoint<S>(it, n1, u1, r, cc, T, f1, m1);
        }
        const double s0 = block_sum(c0, red);
        if (threadIdx.x == 0) parts[n0.gpart + r * it.ntiles + tile] = s0;
        const double s1 = block_sum(c1, red);
        if (threadIdx.x == 0) parts[n1.gpart + r * it.ntiles + tile] = s1;
    } else {
        // any other arity: recompute the siblings' messages per child
        for (int q = 0; q < A; ++q) {
            const int u = it.children[np.first_child + q];
            const NodeInfo nu = it.nodes[u];
            double cu = 0.0;
            if (valid) {
                double T[S], mu_[S], fu[S];
#pragma unroll
                for (int s = 0; s < S; ++s) T[s] = gp[s];
                for (int q2 = 0; q2 < A; ++q2) {
                    if (q2 == q) continue;
                    const int w = it.children[np.first_child + q2];
                    const NodeInfo nw = it.nodes[w];
                    double mw[S], fw[S];
                    child_message<S>(it, nw, w, r, cc, mw, fw);
#pragma unroll
                    for (int s = 0; s < S; ++s) T[s] *= mw[s];
                }
                child_message<S>(it, nu, u, r, cc, mu_, fu);
                cu = child_adjoint<S>(it, nu, u, r, cc, T, fu, mu_);
            }
            const double su = block_sum(cu, red);
            if (threadIdx.x == 0) parts[nu.gpart + r * it.ntiles + tile] = su;
        }
    }
}

template <int S>
__global__ void __launch_bounds__(256) k_adj_level(const Item* __restrict__ items,
                                                   const Job* __restrict__ jobs,
                                                   const int32_t* __restrict__ task_off, int njobs,
                                                   double* __restrict__ parts) {
    __shared__ double red[16];
    const int total = task_off[njobs];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int j = find_job(task_off, njobs, t);
        const Job jb = jobs[j];
        const Item& it = items[jb.item];
        const int local = t - task_off[j];
        const int r = local / it.ntiles;
        const int tile = local - r * it.ntiles;
        adj_task<S>(it, jb.node, r, tile, red, parts);
    }
}

// ---------------------------------------------------------------------------
// final sums: one block per output, fixed order.  Each thread sums a
// strided slice with 8 independent accumulators (the loads of an unrolled
// step are independent, so long ranges are not latency-bound), then the
// block reduces in a fixed tree.
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_reduce(const RedOp* __restrict__ ops, int nops,
                                                const double* __restrict__ parts,
                                                double* __restrict__ out) {
    __shared__ double red[16];
    const RedOp op = ops[blockIdx.x];
    const int nt = blockDim.x;
    const double* p = parts + op.start;
    double a[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = 0.0;
    int i = threadIdx.x;
    for (; i + 7 * nt < op.len; i += 8 * nt) {
#pragma unroll
        for (int k = 0; k < 8; ++k) a[k] += p[i + k * nt];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (i + k * nt < op.len) a[k] += p[i + k * nt];
    const double s = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    const double tot = block_sum(s, red);
    if (threadIdx.x == 0) out[op.dst] = tot;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

hipError_t launch_build_P(const ModelDev* models, const int32_t* elem_off, int nitems,
                          long long nelem, hipStream_t st) {
    if (nelem <= 0) return hipSuccess;
    const int tpb = 256;
    hipLaunchKernelGGL(k_build_P, dim3((unsigned)((nelem + tpb - 1) / tpb)), dim3(tpb), 0, st,
                       models, elem_off, nitems);
    return hipGetLastError();
}

hipError_t launch_tip_tables(const ModelDev* models, const int32_t* elem_off, int nitems,
                             long long nelem, hipStream_t st) {
    if (nelem <= 0) return hipSuccess;
    const int tpb = 256;
    hipLaunchKernelGGL(k_tip_tables, dim3((unsigned)((nelem + tpb - 1) / tpb)), dim3(tpb), 0, st,
                       models, elem_off, nitems);
    return hipGetLastError();
}

#define MCP_FOR_EACH_S(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(16) X(20)

bool supported_states(int S) {
    switch (S) {
#define X(n) case n: return true;
        MCP_FOR_EACH_S(X)
#undef X
        default: return false;
    }
}

hipError_t launch_level(bool adjoint, int S, const Item* items, const Job* jobs,
                        const int32_t* task_off, int njobs, double* parts, int grid, int tpb,
                        hipStream_t st) {
    if (njobs <= 0 || grid <= 0) return hipSuccess;
    switch (S) {
#define X(n)                                                                                    \
    case n:                                                                                     \
        if (adjoint)                                                                            \
            hipLaunchKernelGGL(k_adj_level<n>, dim3(grid), dim3(tpb), 0, st, items, jobs,       \
                               task_off, njobs, parts);                                         \
        else                                                                                    \
            hipLaunchKernelGGL(k_fwd_level<n>, dim3(grid), dim3(tpb), 0, st, items, jobs,       \
                               task_off, njobs, parts);                                         \
        break;
        MCP_FOR_EACH_S(X)
#undef X
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_reduce(const RedOp* ops, int nops, const double* parts, double* out,
                         hipStream_t st) {
    if (nops <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce, dim3(nops), dim3(256), 0, st, ops, nops, parts, out);
    return hipGetLastError();
}

}  // namespace mcp
