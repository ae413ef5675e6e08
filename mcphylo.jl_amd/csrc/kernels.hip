// Pruning likelihood + branch-length adjoint kernels for gfx950 (FP64).
//
// Kernel map (reference loop nest -> kernel):
//   parallel_transition_prob / R_gemmturbo_large! / L_gemmturbo_large! /
//   R_gemmturbo_large_ptg!  (VectorizedFunctions.jl:89-168)   -> k_build_P_tips
//   comb_sum_product_loop! + by_max! + root_sum
//   (VectorizedFunctions.jl:31-49,76-87,187-213;
//    LikelihoodCalculator_Node.jl:82-114)                     -> k_fwd_level
//   pre-order loop of FelsensteinFunction/6 with turbo_mul!,
//   mygemmturbo!, bymax! (LikelihoodCalculator_Node.jl:18-74;
//   VectorizedFunctions.jl:2-11,51-74,170-177)                -> k_adj_level
//   sums over (pattern, category) of log scalers / gradients  -> k_reduce
//
// Work decomposition: a job is one internal node of one item (a LevelJob
// record with every operand address resolved on the host); it expands into
// R * ntiles tasks, one per (category r, tile of 64 * CPT consecutive
// patterns).  A task is executed by one wavefront: lane l owns columns
// tile_base + l + 64 k (k < CPT), so every load and store instruction covers
// one contiguous span of slab rows, and the S-state vectors of a column stay
// in registers.  The transition matrices of a task are wave-uniform: they
// are read with scalar loads and enter the FMAs as SGPR operands.
// Two schedules run the same tasks:
//   k_level  one launch per tree level (post-order by height, pre-order by
//            depth); each single-wave block takes one contiguous chunk of
//            the launch's task list, locates its first job once and then
//            walks the jobs in order;
//   k_fused  one launch per evaluation; a block owns one (item, category,
//            tile) and walks every step of its item (see below).
//
// Determinism: per-task sums are reduced with a fixed DPP tree per
// wavefront, written to one slot per task (a fused block folds its log
// scalers and writes one block sum, its wavefront sums added in a fixed
// order), and summed by k_reduce in a fixed order, so results are bitwise
// reproducible run to run.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stddef.h>

#include "device.h"
#include "engine.h"

namespace mcp {

// Locate the job owning task t: task_off[j] <= t < task_off[j+1].
__device__ __forceinline__ int find_job(const int32_t* __restrict__ task_off, int njobs, int t) {
    int lo = 0, hi = njobs;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (task_off[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
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

// One launch builds every item's P, dP (one thread per element (item, branch
// b, r, m, n); elemP[i] is the first element of item i, (N-1)*R*S*S each) and
// the leaves' tip tables (one thread per element (item, branch b, r, code k,
// state s), indices past the P elements; elemT[i] as elemP, only leaf
// branches write): PT[leaf][r][k][s] = sum_s1 P_b[r][s][s1] tiptab[k][s1],
// evaluated as sum_q U[s][q] e^{a_q} (Uinv tiptab_k)[q] so that no thread waits
// for another's P.
__global__ void k_build_P_tips(const ModelDev* __restrict__ models,
                               const int32_t* __restrict__ elemP,
                               const int32_t* __restrict__ elemT, int nitems) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nP = elemP[nitems];
    if (e < nP) {
        const int i = find_item(elemP, nitems, e);
        const ModelDev& md = models[i];
        const int S = md.S;
        long long k = e - elemP[i];
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
        return;
    }
    e -= nP;
    if (e >= elemT[nitems]) return;
    const int i = find_item(elemT, nitems, e);
    const ModelDev& md = models[i];
    const int S = md.S;
    long long k = e - elemT[i];
    const int s = (int)(k % S); k /= S;
    const int code = (int)(k % md.K); k /= md.K;
    const int r = (int)(k % md.R);
    const int b = (int)(k / md.R);
    const int leaf = md.leaf_of[b];
    if (leaf < 0) return;
    const double t = md.blv[b];
    const double rho = md.rates[r];
    const double* tv = md.tiptab + (size_t)code * S;
    double a = 0.0, da = 0.0;
    for (int q = 0; q < S; ++q) {
        double w = 0.0;
        for (int s1 = 0; s1 < S; ++s1) w = fma(md.Uinv[q + S * s1], tv[s1], w);
        const double ex = exp(md.mu * t * md.D[q] * rho);
        const double u = md.U[s + S * q];
        a = fma(u * ex, w, a);
        da = fma(u * md.D[q] * rho * md.mu * ex, w, da);
    }
    const size_t o = (((size_t)leaf * md.R + r) * md.K + code) * S + s;
    md.PT[o] = a;
    md.dPT[o] = da;
}

// ---------------------------------------------------------------------------
// level kernels
// ---------------------------------------------------------------------------

__device__ __forceinline__ const ChildRef& child_of(const LevelJob& J, int q) {
    return q < kInlineChildren ? J.kid[q] : J.more[q - kInlineChildren];
}

// Row `code` of a wave-uniform table tab[K][S] for each lane: a gather from
// a table of a few hundred bytes that stays in the L1/L2 caches.
template <int S>
__device__ __forceinline__ void select_row(const MCP_GLOBAL double* __restrict__ tab, int code,
                                           double (&out)[S]) {
    const MCP_GLOBAL double* row = tab + code * S;
#pragma unroll
    for (int s = 0; s < S; ++s) out[s] = row[s];
}

// Tip codes the fused kernel stages in LDS: [leaf][64 * CPT] bytes, the
// block's tile of every code row of its item's dataset.  A task's leaf child
// then reads its code bytes from LDS (lrow = 64 * CPT) instead of global
// memory (lrow = 0), which takes a cold HBM round trip off every level step
// that has a leaf child.
extern __shared__ uint8_t lds_codes[];
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Slab stage of the fused walk in LDS (the item's lds_g >= 0): slot q is
// [S][64 * CPT] doubles at byte lgo + q * 512 * S * CPT, so a lane's
// accesses (lane + 64 k + 64 CPT s) are consecutive across lanes.  The
// post-order pass keeps copies of its frontier's partials there, the
// pre-order pass its pre-order slabs.
template <int S, int CPT>
__device__ __forceinline__ double* lds_slab(int lgo, int slot) {
    return reinterpret_cast<double*>(lds_codes + lgo) + slot * (64 * CPT * S);
}

template <int CPT>
__device__ __forceinline__ void leaf_codes(const ChildRef& ch, int c0, int lrow, int (&code)[CPT]) {
    if (lrow) {
        const int lane = (int)(threadIdx.x & 63);
#pragma unroll
        for (int k = 0; k < CPT; ++k) code[k] = (int)lds_codes[ch.leaf * lrow + lane + 64 * k];
    } else {
        const MCP_GLOBAL uint8_t* codes = gptr(ch.codes);
#pragma unroll
        for (int k = 0; k < CPT; ++k) code[k] = (int)codes[c0 + 64 * k];
    }
}

// Columns of a task: lane l owns c0 + 64 k (k < CPT), c0 = tile base + l.
// Slabs, codes and weights are allocated over the padded pattern range, so
// loads and stores of padding columns stay in bounds and are issued
// unconditionally; only the sums mask them (select, so padding garbage never
// reaches a result).

// Message each of a thread's CPT columns receives from child `ch` in
// category r: msg[s] = sum_s1 P[s][s1] f[s1] (internal child, f loaded into
// `f`: from the slab stage's slot ch.fslot when lfo >= 0, else from its
// partial slab), or the leaf's precomputed row PT[r][code] (`f` untouched).
template <int S, int CPT>
__device__ __forceinline__ void child_fetch(const ChildRef& ch, const LevelJob& J, int r, int c0,
                                            int lrow, double (&f)[CPT][S], double (&m)[CPT][S],
                                            int lfo = -1) {
    if (ch.codes) {
        const MCP_GLOBAL double* pt = gptr(ch.P) + (size_t)r * J.K * S;
        int code[CPT];
        leaf_codes<CPT>(ch, c0, lrow, code);
#pragma unroll
        for (int k = 0; k < CPT; ++k) select_row<S>(pt, code[k], m[k]);
    } else {
        if (lfo >= 0) {
            const double* src = lds_slab<S, CPT>(lfo, ch.fslot) + (int)(threadIdx.x & 63);
#pragma unroll
            for (int k = 0; k < CPT; ++k)
#pragma unroll
                for (int s = 0; s < S; ++s) f[k][s] = src[64 * k + 64 * CPT * s];
        } else {
            const MCP_GLOBAL double* src = gptr(ch.F) + (size_t)r * J.Cp * S;
#pragma unroll
            for (int k = 0; k < CPT; ++k) load_row<S>(src + (size_t)(c0 + 64 * k) * S, f[k]);
        }
        const MCP_CONST double* __restrict__ P = cptr(ch.P) + (size_t)r * (S * S);
#pragma unroll
        for (int k = 0; k < CPT; ++k)
#pragma unroll
            for (int s = 0; s < S; ++s) {
                double a = 0.0;
#pragma unroll
                for (int s1 = 0; s1 < S; ++s1) a = fma(f[k][s1], P[s * S + s1], a);
                m[k][s] = a;
            }
    }
}

// Pre-order form of child_fetch: a leaf child also gets its dP.tip rows
// dPT[r][code] (dm), loaded beside its P.tip rows rather than after the
// task's arithmetic.  Used where the registers allow it (S * CPT <= 8).
template <int S, int CPT>
constexpr bool kLeafRowsEarly = S * CPT <= 8;

template <int S, int CPT>
__device__ __forceinline__ void child_fetch_pre(const ChildRef& ch, const LevelJob& J, int r,
                                                int c0, int lrow, double (&f)[CPT][S],
                                                double (&m)[CPT][S], double (&dm)[CPT][S]) {
    if (ch.codes) {
        const MCP_GLOBAL double* pt = gptr(ch.P) + (size_t)r * J.K * S;
        const MCP_GLOBAL double* dpt = gptr(ch.dP) + (size_t)r * J.K * S;
        int code[CPT];
        leaf_codes<CPT>(ch, c0, lrow, code);
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            select_row<S>(pt, code[k], m[k]);
            select_row<S>(dpt, code[k], dm[k]);
        }
    } else {
        child_fetch<S, CPT>(ch, J, r, c0, lrow, f, m);
    }
}

// Per lane and column, the product of the log-scaler factors of every
// post-order task a fused wavefront runs, kept as mant * 2^ex with mant in
// [0.5, 1) (renormalised after every factor, so it neither underflows nor
// overflows).  sum_v log(f_v) = log(mant) + ex log 2 replaces a log, a
// weight load and a wavefront sum per task by one of each per pass.
template <int CPT>
struct ScaleProd {
    double mant[CPT];
    int ex[CPT];
};

// Partials are rescaled by 2^-e, e the binary exponent of their state
// maximum (the maximum lands in [0.5, 1)): the reference's by_max! divides
// by the maximum itself.  Both keep the partial in range; the power of two
// is exact and needs no division, and its log-scaler term is e log 2
// instead of log(max).  Every log-scaler, ratio and gradient is invariant
// to the choice.  frexp_exp is 0 for 0, inf and NaN, so impossible
// columns still give -inf at the root and NaN still propagates.
__device__ __forceinline__ int scale_exp(double mx) { return __builtin_amdgcn_frexp_exp(mx); }

template <int CPT>
__device__ __forceinline__ void scale_fold(ScaleProd<CPT>& p, int k, double f) {
    int e1, e2;
    const double m = frexp(f, &e1);
    p.mant[k] = frexp(p.mant[k] * m, &e2);
    p.ex[k] += e1 + e2;
}

// post-order: partial of the job's node for the thread's columns.  Its
// log-scaler sum goes to its own slot, or (prod != nullptr) is folded into
// *prod and the slot gets 0.  (lfo >= 0: internal children are read from the
// fused block's slab stage at byte lfo, and the partial is also copied there,
// slot J.fslot.)
template <int S, int CPT>
__device__ __forceinline__ void fwd_task(const LevelJob& J, int r, int tile,
                                         double* __restrict__ parts, int lrow = 0,
                                         ScaleProd<CPT>* prod = nullptr, int lfo = -1) {
    const int lane = (int)(threadIdx.x & 63);
    const int c0 = tile * 64 * CPT + lane;
    // the record fields the task reads, fetched as one batch of scalar loads
    const ChildRef a = J.kid[0];
    const ChildRef b = J.kid[1];
    MCP_PIN_SCALAR("s"(a.F), "s"(a.codes), "s"(a.P), "s"(a.leaf), "s"(b.F), "s"(b.codes),
                   "s"(b.P), "s"(b.leaf), "s"(J.F), "s"(J.weights), "s"(J.nchild),
                   "s"(J.is_root), "s"(J.lpart), "s"(J.Cp), "s"(J.K), "s"(J.C), "s"(J.ntiles));
    if (lfo >= 0) MCP_PIN_SCALAR("s"(J.fslot), "s"(a.fslot), "s"(b.fslot));
    double acc[CPT][S];
    if (J.nchild == 2) {
        double fa[CPT][S], ma[CPT][S], fb[CPT][S], mb[CPT][S];
        child_fetch<S, CPT>(a, J, r, c0, lrow, fa, ma, lfo);
        child_fetch<S, CPT>(b, J, r, c0, lrow, fb, mb, lfo);
#pragma unroll
        for (int k = 0; k < CPT; ++k)
#pragma unroll
            for (int s = 0; s < S; ++s) acc[k][s] = ma[k][s] * mb[k][s];
    } else {
#pragma unroll
        for (int k = 0; k < CPT; ++k)
#pragma unroll
            for (int s = 0; s < S; ++s) acc[k][s] = 1.0;
        for (int q = 0; q < J.nchild; ++q) {
            double f[CPT][S], m[CPT][S];
            child_fetch<S, CPT>(child_of(J, q), J, r, c0, lrow, f, m, lfo);
#pragma unroll
            for (int k = 0; k < CPT; ++k)
#pragma unroll
                for (int s = 0; s < S; ++s) acc[k][s] *= m[k][s];
        }
    }
    const MCP_GLOBAL double* wts = gptr(J.weights);
    double contrib = 0.0;
    if (J.is_root) {
        // root_sum: log sum_s pi_s Ftilde_s (VectorizedFunctions.jl:76-87)
        const MCP_CONST double* pi = cptr(J.pi);
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const int c = c0 + 64 * k;
            double t = 0.0;
#pragma unroll
            for (int s = 0; s < S; ++s) t = fma(acc[k][s], pi[s], t);
            if (prod) {
                scale_fold<CPT>(*prod, k, t);
            } else {
                const double v = wts[c] * log(t);
                contrib += c < J.C ? v : 0.0;
            }
        }
    } else {
        // by_max! (VectorizedFunctions.jl:31-49), rescaling by the power of
        // two at the state maximum; keep its log
        MCP_GLOBAL double* dst = gptr_w(J.F) + (size_t)r * J.Cp * S;
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const int c = c0 + 64 * k;
            double mx = -INFINITY;
#pragma unroll
            for (int s = 0; s < S; ++s) mx = (mx < acc[k][s]) ? acc[k][s] : mx;
            const int e = scale_exp(mx);
#pragma unroll
            for (int s = 0; s < S; ++s) acc[k][s] = ldexp(acc[k][s], -e);
            store_row<S>(dst + (size_t)c * S, acc[k]);
            if (lfo >= 0) {
                double* fl = lds_slab<S, CPT>(lfo, J.fslot) + lane + 64 * k;
#pragma unroll
                for (int s = 0; s < S; ++s) fl[64 * CPT * s] = acc[k][s];
            }
            if (prod) {
                prod->ex[k] += e;
            } else {
                const double v = wts[c] * ((double)e * M_LN2);
                contrib += c < J.C ? v : 0.0;
            }
        }
    }
    if (prod) {
        if (lane == 0) parts[J.lpart + r * J.ntiles + tile] = 0.0;
        return;
    }
    const double bs = wave_sum(contrib);
    if (lane == 0) parts[J.lpart + r * J.ntiles + tile] = bs;
}

// pre-order, one child u of the job's node p, given T_u = G_p (.) prod of
// the siblings' messages:
//   g_u = sum_m T_u[m] (dP_u f_u)[m]                        (:39-48)
//   G_u[k] = sum_m P_u[m][k] T_u[m]                         (:50-57)
//   col = sum_k f_u[k] G_u[k];  grad_u += w_c g_u / col     (:59-69)
//   G_u /= max_k G_u[k] (internal u; here the power of two   (:71-73)
//   at the maximum, see scale_exp)
// (LikelihoodCalculator_Node.jl).  For a leaf u, g_u and col are dot
// products of T_u with the leaf's dP.tip and P.tip rows.
template <int S, int CPT, bool DM = false>
__device__ __forceinline__ double child_adjoint(const ChildRef& ch, const LevelJob& J, int r, int c0,
                                                const double (&T)[CPT][S],
                                                const double (&f)[CPT][S],
                                                const double (&m)[CPT][S],
                                                double (&colv)[CPT], int lrow,
                                                const double (&dm)[CPT][S], int lgo) {
    const MCP_GLOBAL double* wts = gptr(J.weights);
    double contrib = 0.0;
    if (ch.codes) {
        const MCP_GLOBAL double* dpt = gptr(ch.dP) + (size_t)r * J.K * S;
        int code[CPT];
        if constexpr (!DM) leaf_codes<CPT>(ch, c0, lrow, code);
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const int c = c0 + 64 * k;
            double row[S];
            if constexpr (DM) {
#pragma unroll
                for (int s = 0; s < S; ++s) row[s] = dm[k][s];
            } else {
                select_row<S>(dpt, code[k], row);
            }
            double num = 0.0, col = 0.0;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                num = fma(T[k][s], row[s], num);
                col = fma(T[k][s], m[k][s], col);
            }
            colv[k] = col;
            const double v = wts[c] * (num / col);
            contrib += c < J.C ? v : 0.0;
        }
    } else {
        const MCP_CONST double* __restrict__ P = cptr(ch.P) + (size_t)r * (S * S);
        const MCP_CONST double* __restrict__ dP = cptr(ch.dP) + (size_t)r * (S * S);
        MCP_GLOBAL double* Gdst = gptr_w(ch.G) + (size_t)r * J.Cp * S;
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const int c = c0 + 64 * k;
            double g[S];
#pragma unroll
            for (int s = 0; s < S; ++s) g[s] = 0.0;
            double num = 0.0;
#pragma unroll
            for (int mm = 0; mm < S; ++mm) {
                double dm = 0.0;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    dm = fma(dP[mm * S + s], f[k][s], dm);
                    g[s] = fma(P[mm * S + s], T[k][mm], g[s]);
                }
                num = fma(T[k][mm], dm, num);
            }
            double col = 0.0, mx = -INFINITY;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                col = fma(f[k][s], g[s], col);
                mx = (mx < g[s]) ? g[s] : mx;
            }
            // rescale by the power of two at the maximum (see scale_exp)
            const int e = scale_exp(mx);
#pragma unroll
            for (int s = 0; s < S; ++s) g[s] = ldexp(g[s], -e);
            if (lgo >= 0) {
                double* gl = lds_slab<S, CPT>(lgo, ch.gslot) + (int)(threadIdx.x & 63) + 64 * k;
#pragma unroll
                for (int s = 0; s < S; ++s) gl[64 * CPT * s] = g[s];
            } else {
                store_row<S>(Gdst + (size_t)c * S, g);
            }
            colv[k] = col;
            const double v = wts[c] * (num / col);
            contrib += c < J.C ? v : 0.0;
        }
    }
    return contrib;
}

// Model gradient (the substitution model's parameters enter through every
// P): d logL / d P_u[m][s] for child u in category r is
//   A[m][s] = sum_c w_c T_u[c][m] f_u[c][s] / col_c
// (f_u: the child's partial, or its tip vector), since d col_c = sum_{m,s}
// T_u[m] dP[m][s] f_u[s] and d logL = sum_c w_c d col_c / col_c.  The
// wavefront's sums go to mparts[mpart + (r S^2 + e) ntiles + tile].
template <int S, int CPT>
__device__ __forceinline__ void mg_branch(const ChildRef& ch, const LevelJob& J, int r, int tile,
                                          int c0, const double (&T)[CPT][S],
                                          const double (&f)[CPT][S], const double (&colv)[CPT],
                                          double* __restrict__ mparts) {
    const MCP_GLOBAL double* wts = gptr(J.weights);
    double A[S][S];
#pragma unroll
    for (int m = 0; m < S; ++m)
#pragma unroll
        for (int s = 0; s < S; ++s) A[m][s] = 0.0;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int c = c0 + 64 * k;
        const double wc = c < J.C ? wts[c] / colv[k] : 0.0;
        double fv[S];
        if (ch.codes) {
            select_row<S>(gptr(J.tiptab), (int)gptr(ch.codes)[c], fv);
        } else {
#pragma unroll
            for (int s = 0; s < S; ++s) fv[s] = f[k][s];
        }
#pragma unroll
        for (int m = 0; m < S; ++m) {
            const double tm = wc * T[k][m];
#pragma unroll
            for (int s = 0; s < S; ++s) A[m][s] = fma(tm, fv[s], A[m][s]);
        }
    }
    double* dst = mparts + ch.mpart + ((size_t)r * S * S) * J.ntiles + tile;
    const int lane = (int)(threadIdx.x & 63);
#pragma unroll
    for (int m = 0; m < S; ++m)
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const double v = wave_sum(A[m][s]);
            if (lane == 0) dst[(size_t)(m * S + s) * J.ntiles] = v;
        }
}

// Root term of the model gradient: d logL / d pi_s = sum_c w_c Ft[c][s] /
// sum_s' pi_s' Ft[c][s'] with Ft the product of the root's children messages
// (VectorizedFunctions.jl:76-87); sums to mparts[rpart + (s R + r) ntiles + tile].
template <int S, int CPT>
__device__ __forceinline__ void mg_root(const LevelJob& J, int r, int tile, int c0,
                                        const double (&Ft)[CPT][S],
                                        double* __restrict__ mparts) {
    const MCP_GLOBAL double* wts = gptr(J.weights);
    const MCP_CONST double* pi = cptr(J.pi);
    double a[S];
#pragma unroll
    for (int s = 0; s < S; ++s) a[s] = 0.0;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int c = c0 + 64 * k;
        double t = 0.0;
#pragma unroll
        for (int s = 0; s < S; ++s) t = fma(pi[s], Ft[k][s], t);
        const double wc = c < J.C ? wts[c] / t : 0.0;
#pragma unroll
        for (int s = 0; s < S; ++s) a[s] = fma(wc, Ft[k][s], a[s]);
    }
    const int lane = (int)(threadIdx.x & 63);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const double v = wave_sum(a[s]);
        if (lane == 0) mparts[J.rpart + ((size_t)s * J.R + r) * J.ntiles + tile] = v;
    }
}

// (lgo >= 0: the item's pre-order slabs live in the fused block's LDS stage
// at byte lgo, slots J.gslot / ch.gslot)
template <int S, int CPT, bool MG = false>
__device__ __forceinline__ void adj_task(const LevelJob& J, int r, int tile,
                                         double* __restrict__ parts,
                                         double* __restrict__ mparts = nullptr, int lrow = 0,
                                         int lgo = -1) {
    const int lane = (int)(threadIdx.x & 63);
    const int c0 = tile * 64 * CPT + lane;
    // the record fields the task reads, fetched as one batch of scalar loads
    const ChildRef a = J.kid[0];
    const ChildRef b = J.kid[1];
    MCP_PIN_SCALAR("s"(a.F), "s"(a.codes), "s"(a.P), "s"(a.dP), "s"(a.G), "s"(a.gpart),
                   "s"(a.leaf), "s"(b.F), "s"(b.codes), "s"(b.P), "s"(b.dP), "s"(b.G),
                   "s"(b.gpart), "s"(b.leaf), "s"(J.G), "s"(J.pi), "s"(J.weights), "s"(J.nchild),
                   "s"(J.Cp), "s"(J.K), "s"(J.C), "s"(J.ntiles));
    if (lgo >= 0) MCP_PIN_SCALAR("s"(J.gslot), "s"(a.gslot), "s"(b.gslot));
    double gp[CPT][S];
    if (J.G && lgo >= 0) {
        const double* src = lds_slab<S, CPT>(lgo, J.gslot) + lane;
#pragma unroll
        for (int k = 0; k < CPT; ++k)
#pragma unroll
            for (int s = 0; s < S; ++s) gp[k][s] = src[64 * k + 64 * CPT * s];
    } else if (J.G) {
        const MCP_GLOBAL double* src = gptr(J.G) + (size_t)r * J.Cp * S;
#pragma unroll
        for (int k = 0; k < CPT; ++k) load_row<S>(src + (size_t)(c0 + 64 * k) * S, gp[k]);
    } else {
        const MCP_CONST double* pi = cptr(J.pi);
#pragma unroll
        for (int k = 0; k < CPT; ++k)
#pragma unroll
            for (int s = 0; s < S; ++s) gp[k][s] = pi[s];
    }
    if (J.nchild == 2) {
        double fa[CPT][S], ma[CPT][S], fb[CPT][S], mb[CPT][S], T[CPT][S];
        constexpr bool E = kLeafRowsEarly<S, CPT>;
        double da[E ? CPT : 1][S], db[E ? CPT : 1][S];
        if constexpr (E) {
            child_fetch_pre<S, CPT>(a, J, r, c0, lrow, fa, ma, da);
            child_fetch_pre<S, CPT>(b, J, r, c0, lrow, fb, mb, db);
        } else {
            child_fetch<S, CPT>(a, J, r, c0, lrow, fa, ma);
            child_fetch<S, CPT>(b, J, r, c0, lrow, fb, mb);
        }
#pragma unroll
        for (int k = 0; k < CPT; ++k)
#pragma unroll
            for (int s = 0; s < S; ++s) T[k][s] = gp[k][s] * mb[k][s];
        double colv[CPT];
        double ca, cb;
        if constexpr (E) ca = child_adjoint<S, CPT, true>(a, J, r, c0, T, fa, ma, colv, lrow, da, lgo);
        else ca = child_adjoint<S, CPT>(a, J, r, c0, T, fa, ma, colv, lrow, ma, lgo);
        if constexpr (MG) mg_branch<S, CPT>(a, J, r, tile, c0, T, fa, colv, mparts);
#pragma unroll
        for (int k = 0; k < CPT; ++k)
#pragma unroll
            for (int s = 0; s < S; ++s) T[k][s] = gp[k][s] * ma[k][s];
        if constexpr (E) cb = child_adjoint<S, CPT, true>(b, J, r, c0, T, fb, mb, colv, lrow, db, lgo);
        else cb = child_adjoint<S, CPT>(b, J, r, c0, T, fb, mb, colv, lrow, mb, lgo);
        if constexpr (MG) {
            mg_branch<S, CPT>(b, J, r, tile, c0, T, fb, colv, mparts);
            if (!J.G) {
#pragma unroll
                for (int k = 0; k < CPT; ++k)
#pragma unroll
                    for (int s = 0; s < S; ++s) T[k][s] = ma[k][s] * mb[k][s];
                mg_root<S, CPT>(J, r, tile, c0, T, mparts);
            }
        }
        const double sa = wave_sum(ca);
        const double sb = wave_sum(cb);
        if (lane == 0) {
            parts[a.gpart + r * J.ntiles + tile] = sa;
            parts[b.gpart + r * J.ntiles + tile] = sb;
        }
        return;
    }
    // any other arity: T_u from the siblings' messages, recomputed per child
    for (int q = 0; q < J.nchild; ++q) {
        const ChildRef& u = child_of(J, q);
        double T[CPT][S], fu[CPT][S], mu_[CPT][S];
#pragma unroll
        for (int k = 0; k < CPT; ++k)
#pragma unroll
            for (int s = 0; s < S; ++s) T[k][s] = gp[k][s];
        for (int q2 = 0; q2 < J.nchild; ++q2) {
            if (q2 == q) continue;
            double fw[CPT][S], mw[CPT][S];
            child_fetch<S, CPT>(child_of(J, q2), J, r, c0, lrow, fw, mw);
#pragma unroll
            for (int k = 0; k < CPT; ++k)
#pragma unroll
                for (int s = 0; s < S; ++s) T[k][s] *= mw[k][s];
        }
        child_fetch<S, CPT>(u, J, r, c0, lrow, fu, mu_);
        double colv[CPT];
        const double cu = child_adjoint<S, CPT>(u, J, r, c0, T, fu, mu_, colv, lrow, mu_, lgo);
        if constexpr (MG) mg_branch<S, CPT>(u, J, r, tile, c0, T, fu, colv, mparts);
        const double su = wave_sum(cu);
        if (lane == 0) parts[u.gpart + r * J.ntiles + tile] = su;
    }
    if constexpr (MG) {
        if (!J.G) {
            double Ft[CPT][S];
#pragma unroll
            for (int k = 0; k < CPT; ++k)
#pragma unroll
                for (int s = 0; s < S; ++s) Ft[k][s] = 1.0;
            for (int q = 0; q < J.nchild; ++q) {
                double fw[CPT][S], mw[CPT][S];
                child_fetch<S, CPT>(child_of(J, q), J, r, c0, lrow, fw, mw);
#pragma unroll
                for (int k = 0; k < CPT; ++k)
#pragma unroll
                    for (int s = 0; s < S; ++s) Ft[k][s] *= mw[k][s];
            }
            mg_root<S, CPT>(J, r, tile, c0, Ft, mparts);
        }
    }
}

// One launch = one level (post-order: nodes of one height; pre-order: nodes
// of one depth) of every item with S states.  Blocks are single wavefronts;
// block b takes the contiguous task chunk [b*chunk, (b+1)*chunk) and starts
// at job blk_job[b] (computed on the host), so a block reaches its first
// operands after two rounds of scalar loads.
#ifndef MCP_ADJ_WAVES_PER_EU
#define MCP_ADJ_WAVES_PER_EU 1
#endif
#ifndef MCP_FWD_WAVES_PER_EU
#define MCP_FWD_WAVES_PER_EU 1
#endif

template <int S, int CPT, bool ADJ>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(ADJ ? MCP_ADJ_WAVES_PER_EU : MCP_FWD_WAVES_PER_EU))) k_level(const LevelJob* __restrict__ jobs,
                                              const int32_t* __restrict__ task_off,
                                              const int32_t* __restrict__ blk_job, int total,
                                              int chunk, double* __restrict__ parts) {
    const int t0 = blockIdx.x * chunk;
    const int t1 = min(total, t0 + chunk);
    int j = blk_job[blockIdx.x];
    int jbeg = task_off[j], jend = task_off[j + 1];
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
        if constexpr (ADJ) adj_task<S, CPT>(J, r, tile, parts);
        else fwd_task<S, CPT>(J, r, tile, parts);
    }
}

// Pre-order level with the model gradient (d logL / dP per branch and
// category, d logL / d pi at the root) accumulated beside the branch-length
// gradient; same task decomposition as k_level.
template <int S, int CPT>
__global__ void __launch_bounds__(64) k_level_mg(const LevelJob* __restrict__ jobs,
                                                 const int32_t* __restrict__ task_off,
                                                 const int32_t* __restrict__ blk_job, int total,
                                                 int chunk, double* __restrict__ parts,
                                                 double* __restrict__ mparts) {
    const int t0 = blockIdx.x * chunk;
    const int t1 = min(total, t0 + chunk);
    int j = blk_job[blockIdx.x];
    int jbeg = task_off[j], jend = task_off[j + 1];
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
        adj_task<S, CPT, true>(J, r, tile, parts, mparts);
    }
}

// Fused schedule: every (item, category, pattern tile) is independent of all
// other tiles through the whole evaluation, so one block walks all
// post-order steps of its tile and then all pre-order steps, with a block
// barrier between steps.  A step (a "level" of lvl_f / lvl_a) is a set of
// nodes whose inputs were produced by earlier steps: the host builds Hu
// steps of at most one node per wavefront under the fused schedule, height
// and depth levels for the narrow top of the hybrid schedule.  Within a step
// the block's wavefronts take its nodes round robin (the wavefront's lanes
// covering the tile's columns).  A tile's partials are written and read back
// by the same block, hence on the same XCD and its L2.  When the item has a
// slab stage in LDS (lds_g >= 0), the slabs passed from one step to a later
// one stay on chip: a mother reads her internal children's partials from
// their copies there, and the pre-order slabs live only there; HBM then sees
// the partials' writes and their reads in the pre-order pass.  Every barrier
// orders a slot's writer before its readers, and a slot's reuse after them
// (the host's interval colouring).
template <int S, int CPT, bool GRAD>
__global__ void __launch_bounds__(512) k_fused(const ItemSched* __restrict__ items, int nitems,
                                               const LevelJob* __restrict__ jobs_f,
                                               const int32_t* __restrict__ lvl_f,
                                               const LevelJob* __restrict__ jobs_a,
                                               const int32_t* __restrict__ lvl_a,
                                               double* __restrict__ parts) {
    int lo = 0, hi = nitems;
    const int b = (int)blockIdx.x;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (items[mid].block_off <= b) lo = mid; else hi = mid;
    }
    const ItemSched I = items[lo];
    const int local = b - I.block_off;
    const int r = uniform(local / I.ntiles);
    const int tile = uniform(local - r * I.ntiles);
    // the wavefront index is uniform within a wavefront: readfirstlane keeps
    // the job indices (and the job-record loads) scalar
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int W = (int)(blockDim.x >> 6);
    // stage the tile's code bytes of every leaf row (16-byte chunks; rows are
    // Cp bytes apart, Cp and the tile base are multiples of 64 * CPT)
    const int lrow = I.lds_leaves > 0 ? 64 * CPT : 0;
    if (lrow) {
        constexpr int kChunks = 64 * CPT / 16;
        const MCP_GLOBAL uint8_t* src = gptr(I.codes) + (size_t)tile * 64 * CPT;
        const int n = I.lds_leaves * kChunks;
        for (int i = (int)threadIdx.x; i < n; i += (int)blockDim.x) {
            const int l = i / kChunks;
            const int o = (i - l * kChunks) * 16;
            *reinterpret_cast<u32x4*>(lds_codes + l * 64 * CPT + o) =
                *reinterpret_cast<const MCP_GLOBAL u32x4*>(src + (size_t)l * I.Cp + o);
        }
        __syncthreads();
    }
    // the post-order tasks fold their log scalers into prod; the pass ends
    // with one sum per block (see below)
    ScaleProd<CPT> prod;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        prod.mant[k] = 1.0;
        prod.ex[k] = 0;
    }
    for (int k = I.lf0; k < I.lf1; ++k) {
        const int j1 = lvl_f[k + 1];
        for (int j = lvl_f[k] + w; j < j1; j += W)
            fwd_task<S, CPT>(jobs_f[j], r, tile, parts, lrow, &prod, I.lds_g);
        __syncthreads();
    }
    if (I.lf0 < I.lf1) {
        // sum_c w_c log(prod_c) per wavefront, the wavefront sums added in a
        // fixed order into the slot of the pass's first job (zeroed by
        // wavefront 0's own task, earlier in its program order)
        __shared__ double wsum[8];
        const LevelJob& J0 = jobs_f[lvl_f[I.lf0]];
        const int lane = (int)(threadIdx.x & 63);
        const int c0 = tile * 64 * CPT + lane;
        const MCP_GLOBAL double* wts = gptr(J0.weights);
        double contrib = 0.0;
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const int c = c0 + 64 * k;
            const double v = wts[c] * (log(prod.mant[k]) + (double)prod.ex[k] * M_LN2);
            contrib += c < J0.C ? v : 0.0;
        }
        const double ws = wave_sum(contrib);
        if (lane == 0) wsum[w] = ws;
        __syncthreads();
        if (w == 0 && lane == 0) {
            double t = 0.0;
            for (int i = 0; i < W; ++i) t += wsum[i];
            parts[J0.lpart + r * J0.ntiles + tile] = t;
        }
    }
    if constexpr (GRAD) {
        for (int k = I.la0; k < I.la1; ++k) {
            const int j1 = lvl_a[k + 1];
            for (int j = lvl_a[k] + w; j < j1; j += W)
                adj_task<S, CPT>(jobs_a[j], r, tile, parts, nullptr, lrow, I.lds_g);
            __syncthreads();
        }
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
                                                double* __restrict__ out,
                                                double* __restrict__ tmp) {
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
    if (threadIdx.x == 0) (op.to_tmp ? tmp : out)[op.dst] = tot;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

hipError_t launch_build_P_tips(const ModelDev* models, const int32_t* elemP,
                              const int32_t* elemT, int nitems, long long nelemP,
                              long long nelemT, hipStream_t st) {
    const long long nelem = nelemP + nelemT;
    if (nelem <= 0) return hipSuccess;
    const int tpb = 256;
    hipLaunchKernelGGL(k_build_P_tips, dim3((unsigned)((nelem + tpb - 1) / tpb)), dim3(tpb), 0,
                       st, models, elemP, elemT, nitems);
    return hipGetLastError();
}

// State counts with built kernels and their default columns per thread (CPT:
// a task is one wavefront over a tile of 64 * CPT consecutive patterns).
#ifndef MCP_CPT_S2
#define MCP_CPT_S2 4
#endif
#ifndef MCP_CPT_S4
#define MCP_CPT_S4 2
#endif
#define MCP_FOR_EACH_S(X)                                                                  \
    X(1, 2) X(2, MCP_CPT_S2) X(3, 2) X(4, MCP_CPT_S4) X(5, 1) X(6, 1) X(7, 1) X(8, 1) X(16, 1) \
    X(20, 1)

// Every (S, CPT) pair with built kernels: the candidates of the tile autotuner.
#define MCP_FOR_EACH_SC(X)                                                                   \
    X(1, 2) X(2, 1) X(2, 2) X(2, 4) X(3, 1) X(3, 2) X(3, 4) X(4, 1) X(4, 2) X(4, 4) X(5, 1) \
    X(6, 1) X(7, 1) X(8, 1) X(16, 1) X(20, 1)

// The (S, CPT) pairs with model-gradient kernels (S <= 8: the per-lane
// S x S accumulator stays in registers).
#define MCP_FOR_EACH_SC_MG(X)                                                                \
    X(1, 2) X(2, 1) X(2, 2) X(2, 4) X(3, 1) X(3, 2) X(3, 4) X(4, 1) X(4, 2) X(4, 4) X(5, 1) \
    X(6, 1) X(7, 1) X(8, 1)

#ifndef MCP_MFMA20
#define MCP_MFMA20 1
#endif

bool supported_states(int S) {
    switch (S) {
#define X(n, c) case n: return true;
        MCP_FOR_EACH_S(X)
#undef X
        default: return false;
    }
}

// Default CPT of S; for S = 20, 0 selects the FP64-MFMA level kernels (their
// tile is 64 patterns, like CPT 1).
int default_cpt(int S) {
    if (S == 20) return MCP_MFMA20 ? 0 : 1;
    switch (S) {
#define X(n, c) case n: return c;
        MCP_FOR_EACH_S(X)
#undef X
        default: return 1;
    }
}

bool cpt_available(int S, int cpt) {
    if (S == 20 && cpt == 0) return true;
#define X(n, c) if (S == n && cpt == c) return true;
    MCP_FOR_EACH_SC(X)
#undef X
    return false;
}

// Widest tile over the candidates of S: the pattern axis is padded to it so
// that every candidate fits the same layout.
int max_tile_cols(int S) {
    int m = 1;
#define X(n, c) if (S == n && c > m) m = c;
    MCP_FOR_EACH_SC(X)
#undef X
    return 64 * m;
}

int tile_cols(int cpt) { return 64 * (cpt < 1 ? 1 : cpt); }

bool mg_available(int S, int cpt) {
#define X(n, c) if (S == n && cpt == c) return true;
    MCP_FOR_EACH_SC_MG(X)
#undef X
    return false;
}

hipError_t launch_level_mg(int S, int cpt, const LevelJob* jobs, const int32_t* task_off,
                           const int32_t* blk_job, int total, int chunk, int grid, double* parts,
                           double* mparts, hipStream_t st) {
    if (total <= 0 || grid <= 0) return hipSuccess;
#define X(n, c)                                                                                  \
    if (S == n && cpt == c) {                                                                    \
        hipLaunchKernelGGL((k_level_mg<n, c>), dim3(grid), dim3(64), 0, st, jobs, task_off,      \
                           blk_job, total, chunk, parts, mparts);                                \
        return hipGetLastError();                                                                \
    }
    MCP_FOR_EACH_SC_MG(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_level20(bool adjoint, const LevelJob* jobs, const int32_t* task_off,
                          const int32_t* blk_job, int total, int chunk, int grid, double* parts,
                          hipStream_t st);

hipError_t launch_level(bool adjoint, int S, int cpt, const LevelJob* jobs,
                        const int32_t* task_off, const int32_t* blk_job, int total, int chunk,
                        int grid, double* parts, hipStream_t st) {
    if (total <= 0 || grid <= 0) return hipSuccess;
    if (S == 20 && cpt == 0)
        return launch_level20(adjoint, jobs, task_off, blk_job, total, chunk, grid, parts, st);
#define X(n, c)                                                                                  \
    if (S == n && cpt == c) {                                                                    \
        if (adjoint)                                                                             \
            hipLaunchKernelGGL((k_level<n, c, true>), dim3(grid), dim3(64), 0, st, jobs,         \
                               task_off, blk_job, total, chunk, parts);                          \
        else                                                                                     \
            hipLaunchKernelGGL((k_level<n, c, false>), dim3(grid), dim3(64), 0, st, jobs,        \
                               task_off, blk_job, total, chunk, parts);                          \
        return hipGetLastError();                                                                \
    }
    MCP_FOR_EACH_SC(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_fused(bool grad, int S, int cpt, const ItemSched* items, int nitems,
                        int nblocks, int waves, size_t lds_bytes, const LevelJob* jobs_f,
                        const int32_t* lvl_f, const LevelJob* jobs_a, const int32_t* lvl_a,
                        double* parts, hipStream_t st) {
    if (nitems <= 0 || nblocks <= 0) return hipSuccess;
    if (cpt == 0) cpt = 1;  // the fused schedule has no MFMA variant; same 64-pattern tile
#define X(n, c)                                                                                  \
    if (S == n && cpt == c) {                                                                    \
        if (grad)                                                                                \
            hipLaunchKernelGGL((k_fused<n, c, true>), dim3(nblocks), dim3(64 * waves), lds_bytes, \
                               st,                                                               \
                               items, nitems, jobs_f, lvl_f, jobs_a, lvl_a, parts);              \
        else                                                                                     \
            hipLaunchKernelGGL((k_fused<n, c, false>), dim3(nblocks), dim3(64 * waves),          \
                               lds_bytes, st,                                                    \
                               items, nitems, jobs_f, lvl_f, jobs_a, lvl_a, parts);              \
        return hipGetLastError();                                                                \
    }
    MCP_FOR_EACH_SC(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_reduce(const RedOp* ops, int nops, const double* parts, double* out,
                         double* tmp, hipStream_t st) {
    if (nops <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce, dim3(nops), dim3(256), 0, st, ops, nops, parts, out, tmp);
    return hipGetLastError();
}

}  // namespace mcp

namespace mcp {

// ---------------------------------------------------------------------------
// device leapfrog (one block per item)
// ---------------------------------------------------------------------------

__device__ __forceinline__ double molifier(double x, double delta) {
    // auxilliary.jl:117-119
    return x >= delta ? x : (x * x + delta * delta) / (2.0 * delta);
}

__device__ __forceinline__ double scale_fac(double x, double delta) {
    // auxilliary.jl:113-115
    return x < delta ? x / delta : 1.0;
}

// Block-wide argmin of v over [0, n) (ties: lowest index); result valid in
// every thread.
__device__ __forceinline__ void block_argmin(double v, int idx, double* sv, int* si, double& best,
                                             int& besti) {
    // v/idx of threads without an element are +inf / INT_MAX
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off, 64);
        const int oi = __shfl_xor(idx, off, 64);
        if (ov < v || (ov == v && oi < idx)) {
            v = ov;
            idx = oi;
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 0) {
        sv[wid] = v;
        si[wid] = idx;
    }
    __syncthreads();
    best = sv[0];
    besti = si[0];
    for (int w = 1; w < nw; ++w)
        if (sv[w] < best || (sv[w] == best && si[w] < besti)) {
            best = sv[w];
            besti = si[w];
        }
}

// First half of a leapfrog for every item:
//   r += eps/2 g;  y = x + eps r;  reflect at zero in time order
//   (refraction.jl:35-91 without the NNI attempt);  x = molifier(y, delta);
//   blv = x (input of the likelihood).
// The reflection loop is bounded by `max_reflect` rounds; a coordinate still
// non-positive after that is reflected with abs() (every block exits).
__global__ void __launch_bounds__(256) k_lf_pre(const HmcItem* __restrict__ items,
                                                double* __restrict__ x, double* __restrict__ r,
                                                const double* __restrict__ g,
                                                const double* __restrict__ eps_delta,
                                                double* __restrict__ blv, int max_reflect) {
    __shared__ double sv[4];
    __shared__ int si[4];
    __shared__ int s_any;
    const HmcItem it = items[blockIdx.x];
    const double eps = eps_delta[2 * blockIdx.x];
    const double delta = eps_delta[2 * blockIdx.x + 1];
    double* xi = x + it.off;
    double* ri = r + it.off;
    const double* gi = g + it.off;
    for (int k = threadIdx.x; k < it.n; k += blockDim.x) ri[k] += (eps * 0.5) * gi[k];
    __syncthreads();
    // y = x + eps r (kept in x until the end; blv vector of the reflections in blv)
    double* b = blv + it.off;
    for (int k = threadIdx.x; k < it.n; k += blockDim.x) b[k] = xi[k];
    __syncthreads();
    double t = 0.0;
    for (int round = 0; round < max_reflect; ++round) {
        // tmpB = b + (eps - t) r ; stop when all > 0
        double best = INFINITY;
        int besti = 0x7fffffff;
        if (threadIdx.x == 0) s_any = 0;
        __syncthreads();
        for (int k = threadIdx.x; k < it.n; k += blockDim.x) {
            const double tb = b[k] + (eps - t) * ri[k];
            if (tb <= 0.0) s_any = 1;
            const double tl = tb / fabs(ri[k]);
            if (tl < best) {
                best = tl;
                besti = k;
            }
        }
        __syncthreads();
        if (!s_any) break;
        double tmin;
        int kmin;
        block_argmin(best, besti, sv, si, tmin, kmin);
        // move to the first crossing, reflect that coordinate
        const double temp = eps - t + tmin;
        for (int k = threadIdx.x; k < it.n; k += blockDim.x) b[k] = fabs(b[k] + temp * ri[k]);
        __syncthreads();
        if (threadIdx.x == 0) ri[kmin] = -ri[kmin];
        t = eps + tmin;
        __syncthreads();
    }
    for (int k = threadIdx.x; k < it.n; k += blockDim.x) {
        const double y = fabs(b[k] + (eps - t) * ri[k]);
        const double xm = molifier(y, delta);
        xi[k] = xm;
        b[k] = xm;
    }
}

// Second half: g = (dll + dlogprior) * scale_fac(x, delta); r += eps/2 g;
// lp = ll + logprior(x).  The CompoundDirichlet prior (Prior.jl:1-37) needs
// block sums first: total length, sums of log lengths over leaf and over
// internal branches (`internal` flags, blv layout), and the leaf count.
__global__ void __launch_bounds__(256) k_lf_post(const HmcItem* __restrict__ items,
                                                 const double* __restrict__ x,
                                                 double* __restrict__ r, double* __restrict__ g,
                                                 const double* __restrict__ eps_delta,
                                                 const double* __restrict__ out,
                                                 double* __restrict__ lp,
                                                 const uint8_t* __restrict__ internal) {
    __shared__ double red[16];
    __shared__ double bc[4];
    const HmcItem it = items[blockIdx.x];
    const double eps = eps_delta[2 * blockIdx.x];
    const double delta = eps_delta[2 * blockIdx.x + 1];
    const double* xi = x + it.off;
    double* ri = r + it.off;
    double* gi = g + it.off;
    const double* grad = out + it.out + 1;
    const uint8_t* inl = internal ? internal + it.off : nullptr;
    double lpr = 0.0, cd_lp = 0.0, cd_k = 0.0;
    if (it.prior == 2) {
        double sx = 0.0, sll = 0.0, sli = 0.0, nl = 0.0;
        for (int k = threadIdx.x; k < it.n; k += blockDim.x) {
            const double xk = xi[k], lk = log(xk);
            sx += xk;
            if (inl && inl[k]) {
                sli += lk;
            } else {
                sll += lk;
                nl += 1.0;
            }
        }
        sx = block_sum(sx, red);
        if (threadIdx.x == 0) bc[0] = sx;
        sll = block_sum(sll, red);
        if (threadIdx.x == 0) bc[1] = sll;
        sli = block_sum(sli, red);
        if (threadIdx.x == 0) bc[2] = sli;
        nl = block_sum(nl, red);
        if (threadIdx.x == 0) bc[3] = nl;
        __syncthreads();
        const double alpha = it.p[0], a = it.p[1], beta = it.p[2], c = it.p[3];
        const double tl = bc[0], n_int = bc[3] - 3.0;
        cd_k = alpha - a * bc[3] - a * c * n_int;
        cd_lp = alpha * log(beta) - lgamma(alpha) - tl * beta - lgamma(a) - lgamma(c) +
                lgamma(a + c) + bc[1] * (a - 1.0) + bc[2] * (a * c - 1.0) + cd_k * log(tl);
        cd_k /= tl;
    }
    for (int k = threadIdx.x; k < it.n; k += blockDim.x) {
        double gk = grad[k];
        if (it.prior == 1) {
            gk += -1.0 / it.prior_scale;
            lpr += -log(it.prior_scale) - xi[k] / it.prior_scale;
        } else if (it.prior == 2) {
            const double a = it.p[1], c = it.p[3];
            gk += -it.p[2] + ((inl && inl[k]) ? a * c - 1.0 : a - 1.0) / xi[k] + cd_k;
        }
        gk *= scale_fac(xi[k], delta);
        gi[k] = gk;
        ri[k] += (eps * 0.5) * gk;
    }
    const double s = block_sum(lpr, red);
    if (threadIdx.x == 0) lp[blockIdx.x] = out[it.out] + s + cd_lp;
}

hipError_t launch_lf_pre(const HmcItem* items, int nitems, double* x, double* r, const double* g,
                         const double* eps_delta, double* blv, int max_reflect, hipStream_t st) {
    hipLaunchKernelGGL(k_lf_pre, dim3(nitems), dim3(256), 0, st, items, x, r, g, eps_delta, blv,
                       max_reflect);
    return hipGetLastError();
}

hipError_t launch_lf_post(const HmcItem* items, int nitems, const double* x, double* r, double* g,
                          const double* eps_delta, const double* out, double* lp,
                          const uint8_t* internal, hipStream_t st) {
    hipLaunchKernelGGL(k_lf_post, dim3(nitems), dim3(256), 0, st, items, x, r, g, eps_delta, out,
                       lp, internal);
    return hipGetLastError();
}

}  // namespace mcp
