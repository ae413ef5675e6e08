// Learned surrogate of the log-likelihood, first stage of delayed-acceptance
// MCMC.
//
// Network (per state, z = standardised log branch lengths, d inputs padded to D):
//   h1 = tanh(W1 z + b1)                    W1: H x D   (H = hidden width)
//   h2 = tanh(W2 h1 + b2)                   W2: H x H
//   y  = w3 . h2 + b3 + v . z + q . (z*z)   separable quadratic skip v, q: D
// y is the standardised log-likelihood: logL~ = y * y_sd + y_mean.  Input
// column c belongs to one split (bipartition) of the tree; for the fixed
// topology of a sampler run the split embeddings are the columns of W1, v, q.
//
// Matrix products run on BF16 matrix cores with FP32 accumulation
// (v_mfma_f32_16x16x32_bf16; operand layouts checked on gfx950 with a probe:
// A lane l holds A[l%16][8(l/16)+e], B lane l holds B[8(l/16)+e][l%16], the
// result lane l register j holds C[4(l/16)+j][l%16]).  Master weights,
// activations, gradients and Adam moments are FP32; GEMM operands are rounded
// to BF16 (nearest even) when staged into LDS.  Every GEMM dimension is padded
// to the 64 x 64 block tile (K to the 32-wide step) with zeros, and padded
// rows/columns stay zero through training (their gradients are zero).
#include <hip/hip_runtime.h>

#include <math.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "device.h"
#include "engine.h"
#include "mcphylo_gpu.h"

namespace mcp {
hipError_t launch_lf_pre(const HmcItem*, int, double*, double*, const double*, const double*,
                         double*, int, hipStream_t);
hipError_t launch_lf_post(const HmcItem*, int, const double*, double*, double*, const double*,
                          const double*, double*, const uint8_t*, hipStream_t);
int set_last_error(int code, const char* msg);
}  // namespace mcp

using namespace mcp;

namespace {

typedef float fvec4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kTile = 64;    // block tile (M and N)
constexpr int kKStep = 32;   // K per MFMA step
constexpr int kKChunk = 128; // K staged per barrier pair
constexpr int kLdsK = kKChunk + 8;  // LDS row length in bf16 (16-byte aligned rows)

// GEMM epilogues (o = m * ldc + n):
//   EPI_STORE      C = acc
//   EPI_BIAS_TANH  C = tanh(acc + aux[n])
//   EPI_MUL_DTANH  C = acc (1 - aux[o]^2)
//   EPI_SOB_T1     C = acc;  out2 = acc (1 - aux[o]^2)                 (T1, delta1)
//   EPI_SOB_K1     C = acc (1 - h^2);  out2 = -2 h acc aux2[o], h = aux[o]   (mu1, Gam1)
//   EPI_SOB_K2     h = aux[o], s2 = 1 - h^2, g = vm[m]:
//                  out2 = g aux2[o] - 2 h acc vn[n] s2;  C = g h + acc s2   (A2tot, w3 rows)
//   EPI_SOB_A1     C = (acc + aux2[o]) (1 - aux[o]^2)                   (A1tot)
enum {
    EPI_STORE = 0,
    EPI_BIAS_TANH = 1,
    EPI_MUL_DTANH = 2,
    EPI_SOB_T1 = 4,
    EPI_SOB_K1 = 5,
    EPI_SOB_K2 = 6,
    EPI_SOB_A1 = 7
};

struct Epi {
    const float* aux = nullptr;
    const float* aux2 = nullptr;
    float* out2 = nullptr;
    const float* vn = nullptr;  // per output column
    const float* vm = nullptr;  // per output row
};

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

__device__ inline bf16x4 to_bf16x4(float a, float b, float c, float d) {
    return bf16x4{(__bf16)a, (__bf16)b, (__bf16)c, (__bf16)d};
}

// One 64 (rows of the tile) x kc (K) operand chunk into registers, 8 float4
// per thread, always along the contiguous memory direction.
//   KC (K contiguous):   element (r, k) at P[(r0 + r) * ld + k0 + k]
//   !KC (r contiguous):  element (r, k) at P[(k0 + k) * ld + r0 + r]; a thread
//                        reads 4 consecutive k rows of one 4-wide r group
template <bool KC>
__device__ inline void load_chunk(const float* __restrict__ P, int ld, int r0, int k0, int kc,
                                  fvec4 (&reg)[8], int tid) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        reg[e] = fvec4{0.f, 0.f, 0.f, 0.f};
        if (KC) {
            const int q = tid + 256 * e;
            const int r = q >> 5, k = (q & 31) * 4;
            if (k < kc) reg[e] = *(const fvec4*)(P + (size_t)(r0 + r) * ld + k0 + k);
        } else {
            const int q = tid + 256 * (e >> 2);
            const int r = (q & 15) * 4, k = (q >> 4) * 4 + (e & 3);
            if (k < kc) reg[e] = *(const fvec4*)(P + (size_t)(k0 + k) * ld + r0 + r);
        }
    }
}

// registers of load_chunk -> LDS tile S[r][k] in BF16 (8-byte stores)
template <bool KC>
__device__ inline void store_chunk(__bf16 (*S)[kLdsK], int kc, const fvec4 (&reg)[8], int tid) {
    if (KC) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int q = tid + 256 * e;
            const int r = q >> 5, k = (q & 31) * 4;
            if (k < kc) *(bf16x4*)&S[r][k] = to_bf16x4(reg[e][0], reg[e][1], reg[e][2], reg[e][3]);
        }
    } else {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int q = tid + 256 * h;
            const int r = (q & 15) * 4, k = (q >> 4) * 4;
            if (k < kc) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    *(bf16x4*)&S[r + i][k] = to_bf16x4(reg[4 * h][i], reg[4 * h + 1][i],
                                                       reg[4 * h + 2][i], reg[4 * h + 3][i]);
            }
        }
    }
}

// C = epilogue(opA . opB [+ opA2 . opB2]), opA: M x K, opB: K x N, all FP32 in
// memory; M, N, K, lda, ldb, ldc are multiples of 64 and the operands 16-byte
// aligned.  A second segment (A2, B2 with the same shapes and strides, A2 !=
// nullptr) continues the same accumulation.
//   opA[m][k] = TA ? A[k*lda + m] : A[m*lda + k]
//   opB[k][n] = TB ? B[n*ldb + k] : B[k*ldb + n]
// Grid (N/64, M/64), 256 threads: four waves, each a 32 x 32 quarter of the
// tile as 2 x 2 MFMA 16 x 16 blocks.  K advances in chunks of 128 staged
// through LDS as BF16 ([m][k] and [n][k]); the global loads of the next chunk
// are in flight while the waves run the 4 MFMA steps of the current one.
template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) k_gemm_bf16(const float* __restrict__ A,
                                                   const float* __restrict__ A2, int lda,
                                                   const float* __restrict__ B,
                                                   const float* __restrict__ B2, int ldb,
                                                   float* __restrict__ C, int ldc, int K, Epi ep) {
    __shared__ __attribute__((aligned(16))) __bf16 As[kTile][kLdsK];  // [m][k]
    __shared__ __attribute__((aligned(16))) __bf16 Bs[kTile][kLdsK];  // [n][k]
    const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    const int li = lane & 15, lg = lane >> 4;
    const int nck = (K + kKChunk - 1) / kKChunk;
    const int nchunks = (A2 ? 2 : 1) * nck;
    fvec4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = fvec4{0.f, 0.f, 0.f, 0.f};
    fvec4 ra[8], rb[8];
    load_chunk<!TA>(A, lda, m0, 0, min(kKChunk, K), ra, tid);
    load_chunk<TB>(B, ldb, n0, 0, min(kKChunk, K), rb, tid);
    for (int c = 0; c < nchunks; ++c) {
        const int k0 = (c % nck) * kKChunk;
        const int kc = min(kKChunk, K - k0);
        store_chunk<!TA>(As, kc, ra, tid);
        store_chunk<TB>(Bs, kc, rb, tid);
        __syncthreads();
        if (c + 1 < nchunks) {
            const bool second = (c + 1) >= nck;
            const int k1 = ((c + 1) % nck) * kKChunk;
            const int kc1 = min(kKChunk, K - k1);
            load_chunk<!TA>(second ? A2 : A, lda, m0, k1, kc1, ra, tid);
            load_chunk<TB>(second ? B2 : B, ldb, n0, k1, kc1, rb, tid);
        }
        for (int kk = 0; kk < kc; kk += kKStep) {
            bf16x8 af[2], bfr[2];
#pragma unroll
            for (int a = 0; a < 2; ++a)
                af[a] = *(const bf16x8*)&As[wm + 16 * a + li][kk + 8 * lg];
#pragma unroll
            for (int b = 0; b < 2; ++b)
                bfr[b] = *(const bf16x8*)&Bs[wn + 16 * b + li][kk + 8 * lg];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] =
                        __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int m = m0 + wm + 16 * a + 4 * lg + j;
                const int n = n0 + wn + 16 * b + li;
                const size_t o = (size_t)m * ldc + n;
                const float v = acc[a][b][j];
                if (EPI == EPI_STORE) C[o] = v;
                if (EPI == EPI_BIAS_TANH) C[o] = tanhf(v + ep.aux[n]);
                if (EPI == EPI_MUL_DTANH) {
                    const float h = ep.aux[o];
                    C[o] = v * (1.f - h * h);
                }
                if (EPI == EPI_SOB_T1) {
                    const float h = ep.aux[o];
                    C[o] = v;
                    ep.out2[o] = v * (1.f - h * h);
                }
                if (EPI == EPI_SOB_K1) {
                    const float h = ep.aux[o];
                    C[o] = v * (1.f - h * h);
                    ep.out2[o] = -2.f * h * v * ep.aux2[o];
                }
                if (EPI == EPI_SOB_K2) {
                    const float h = ep.aux[o], s2 = 1.f - h * h, g = ep.vm[m];
                    ep.out2[o] = g * ep.aux2[o] - 2.f * h * v * ep.vn[n] * s2;
                    C[o] = g * h + v * s2;
                }
                if (EPI == EPI_SOB_A1) {
                    const float h = ep.aux[o];
                    C[o] = (v + ep.aux2[o]) * (1.f - h * h);
                }
            }
}

__device__ inline float wave_sum_f(float s) {
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

// Output head, one wavefront per row:
//   y[r] = w3 . h2[r] + b3 + v . z[r] + q . z[r]^2
// with dA2 != nullptr also the input-gradient seed of dy = 1:
//   dA2[r][k] = w3[k] (1 - h2[r][k]^2)
__global__ void __launch_bounds__(256) k_head(const float* __restrict__ h2, int H,
                                              const float* __restrict__ z, int D,
                                              const float* __restrict__ w3,
                                              const float* __restrict__ b3,
                                              const float* __restrict__ v,
                                              const float* __restrict__ q, float* __restrict__ y,
                                              float* __restrict__ dA2, int rows) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= rows) return;
    const float* hr = h2 + (size_t)row * H;
    const float* zr = z + (size_t)row * D;
    float s = 0.f;
    for (int k = lane; k < H; k += 64) {
        const float h = hr[k];
        s += w3[k] * h;
        if (dA2) dA2[(size_t)row * H + k] = w3[k] * (1.f - h * h);
    }
    for (int k = lane; k < D; k += 64) {
        const float zz = zr[k];
        s += (v[k] + q[k] * zz) * zz;
    }
    s = wave_sum_f(s);
    if (lane == 0) y[row] = s + b3[0];
}

// Training head, one wavefront per row (n valid rows, padding rows give zeros):
//   y = w3 . h2 + b3 + v . z + q . z^2;  e = y - t;  dy = e / n;
//   lrow = e^2 / (2n);  dA2[r][k] = dy w3[k] (1 - h2^2)
// With D2 != nullptr (gradient matching) D2[r][k] = w3[k] (1 - h2^2) instead
// of dA2 (the second backward pass writes dA2).
__global__ void __launch_bounds__(256) k_head_train(const float* __restrict__ h2, int H,
                                                    const float* __restrict__ z, int D,
                                                    const float* __restrict__ w3,
                                                    const float* __restrict__ b3,
                                                    const float* __restrict__ v,
                                                    const float* __restrict__ q,
                                                    const float* __restrict__ t,
                                                    float* __restrict__ y, float* __restrict__ dy,
                                                    float* __restrict__ lrow,
                                                    float* __restrict__ dA2,
                                                    float* __restrict__ D2, int n, int rows) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= rows) return;
    const float* hr = h2 + (size_t)row * H;
    const float* zr = z + (size_t)row * D;
    float s = 0.f;
    for (int k = lane; k < H; k += 64) s += w3[k] * hr[k];
    for (int k = lane; k < D; k += 64) {
        const float zz = zr[k];
        s += (v[k] + q[k] * zz) * zz;
    }
    s = wave_sum_f(s) + b3[0];
    const float e = row < n ? s - t[row] : 0.f;
    const float g = e / (float)n;
    for (int k = lane; k < H; k += 64) {
        const float h = hr[k];
        const float d2 = w3[k] * (1.f - h * h);
        if (D2)
            D2[(size_t)row * H + k] = d2;
        else
            dA2[(size_t)row * H + k] = g * d2;
    }
    if (lane == 0) {
        y[row] = s;
        dy[row] = g;
        lrow[row] = 0.5f * e * e / (float)n;
    }
}

// Column reductions over the rows of a batch, several jobs per launch, each
// in a fixed order (deterministic):
//   COL_SUM:  out[c] = sum_r s[r] X[r][c]             (s == nullptr: s = 1)
//   COL_VQ:   out[c] = sum_r (s[r] X[r][c] + X2[r][c])          (v gradient)
//             out2[c] = sum_r (s[r] X^2 + 2 X X2)    (X2 == nullptr: 0; q gradient)
//   COL_SCALAR (one block): out[0] = sum_r s[r]; out2[0] += sum_r X2[r] (if
//             out2); with a step counter also cur = rates[2 step .. 2 step + 1]
//             and step + 1 -- the one place a training step advances it.
// Block: 64 columns (one per lane, coalesced rows) x 16 waves; wave w sums rows
// w, w + 16, ... and wave 0 adds the 16 partial sums in order.
constexpr int kColWaves = 16;
enum { COL_SUM = 0, COL_VQ = 1, COL_SCALAR = 2 };

struct ColJob {
    int kind = COL_SUM;
    int cols = 0;
    int block0 = 0;  // first block of the job
    const float* X = nullptr;
    const float* s = nullptr;
    const float* X2 = nullptr;
    float* out = nullptr;
    float* out2 = nullptr;
};

struct ColJobs {
    ColJob job[4];
    int njobs = 0;
    int rows = 0;
    int32_t* step = nullptr;
    const float* rates = nullptr;
    float* cur = nullptr;
};

__global__ void __launch_bounds__(64 * kColWaves) k_colsums(ColJobs jobs) {
    __shared__ float part[2][kColWaves][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int j = 0;
    while (j + 1 < jobs.njobs && (int)blockIdx.x >= jobs.job[j + 1].block0) ++j;
    const ColJob& jb = jobs.job[j];
    const int rows = jobs.rows;
    float a = 0.f, b = 0.f;
    if (jb.kind == COL_SCALAR) {
        for (int r = w * 64 + lane; r < rows; r += 64 * kColWaves) {
            a += jb.s[r];
            if (jb.out2) b += jb.X2[r];
        }
        a = wave_sum_f(a);
        b = wave_sum_f(b);
    } else {
        const int c = ((int)blockIdx.x - jb.block0) * 64 + lane;
        if (c < jb.cols) {
            const int ld = jb.cols;
#pragma unroll 4
            for (int r = w; r < rows; r += kColWaves) {
                const float xv = jb.X[(size_t)r * ld + c];
                const float sv = jb.s ? jb.s[r] : 1.f;
                if (jb.kind == COL_SUM) {
                    a += sv * xv;
                } else {
                    const float p = jb.X2 ? jb.X2[(size_t)r * ld + c] : 0.f;
                    a += sv * xv + p;
                    b += sv * xv * xv + 2.f * xv * p;
                }
            }
        }
    }
    part[0][w][lane] = a;
    part[1][w][lane] = b;
    __syncthreads();
    if (w != 0) return;
    float t = 0.f, t2 = 0.f;
    for (int k = 0; k < kColWaves; ++k) {
        t += part[0][k][lane];
        t2 += part[1][k][lane];
    }
    if (jb.kind == COL_SCALAR) {
        if (lane == 0) {
            jb.out[0] = t;
            if (jb.out2) jb.out2[0] += t2;
            if (jobs.step) {
                const int st = jobs.step[0];
                jobs.cur[0] = jobs.rates[2 * st];
                jobs.cur[1] = jobs.rates[2 * st + 1];
                jobs.step[0] = st + 1;
            }
        }
    } else {
        const int c = ((int)blockIdx.x - jb.block0) * 64 + lane;
        if (c < jb.cols) {
            jb.out[c] = t;
            if (jb.kind == COL_VQ) jb.out2[c] = t2;
        }
    }
}

// Z[r][c] = (log x[r][c] - mean[c]) / sd[c] for r < n, c < d; zero padding
__global__ void k_features(const double* __restrict__ x, int d, const float* __restrict__ mean,
                           const float* __restrict__ sd, float* __restrict__ Z, int rows, int D,
                           int n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)rows * D) return;
    const int r = (int)(i / D), c = (int)(i % D);
    Z[i] = (r < n && c < d) ? (float)((log(x[(size_t)r * d + c]) - (double)mean[c]) / (double)sd[c])
                            : 0.f;
}

// out[r] = (logL~, dlogL~/dx_1..d):
//   logL~ = y y_sd + y_mean
//   dlogL~/dx_c = y_sd (dZ[r][c] + v[c] + 2 q[c] z) / (sd[c] x_c)
__global__ void k_finalize(const float* __restrict__ y, const float* __restrict__ dZ,
                           const float* __restrict__ Z, int D, const float* __restrict__ v,
                           const float* __restrict__ q, const double* __restrict__ x, int d,
                           const float* __restrict__ sd, float y_mean, float y_sd,
                           double* __restrict__ out, int n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n * (d + 1)) return;
    const int r = (int)(i / (d + 1)), c = (int)(i % (d + 1));
    if (c == 0) {
        out[i] = (double)y[r] * (double)y_sd + (double)y_mean;
    } else {
        const int k = c - 1;
        const size_t o = (size_t)r * D + k;
        const double gz = (double)dZ[o] + (double)v[k] + 2.0 * (double)q[k] * (double)Z[o];
        out[i] = (double)y_sd * gz / ((double)sd[k] * x[(size_t)r * d + k]);
    }
}

// ---- gradient-matching (Sobolev) term of the training loss -----------------
// Loss  (1/2n) sum_r (y_r - t_r)^2 + (lambda / 2nd) sum_{r, c<d} (g_rc - G_rc)^2
// with g = dy/dz = W1^T delta1 + v + 2 q z (the input gradient, delta2 =
// w3 (1 - h2^2), delta1 = (W2^T delta2)(1 - h1^2)) and G the standardised
// exact gradient.  Its parameter gradient is a second backward pass through
// the input-gradient graph; its elementwise steps are the EPI_SOB_* GEMM
// epilogues and the row kernel below.

// rho[r][c] = lam_n (dZ[r][c] + v[c] + 2 q[c] z[r][c] - G[r][c]) for r < n,
// c < d (zero elsewhere), lam_n = lambda / (n d); lrow[r] += lam_n/2 sum_c
// (...)^2.  One wavefront per row.
__global__ void __launch_bounds__(256) k_sob_rho(const float* __restrict__ dZ,
                                                 const float* __restrict__ Z,
                                                 const float* __restrict__ G,
                                                 const float* __restrict__ v,
                                                 const float* __restrict__ q,
                                                 float* __restrict__ rho, float* __restrict__ lrow,
                                                 int D, int d, int n, int rows, float lam_n) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= rows) return;
    float acc = 0.f;
    for (int c = lane; c < D; c += 64) {
        const size_t o = (size_t)row * D + c;
        float e = 0.f;
        if (row < n && c < d) e = dZ[o] + v[c] + 2.f * q[c] * Z[o] - G[o];
        rho[o] = lam_n * e;
        acc += e * e;
    }
    acc = wave_sum_f(acc);
    if (lane == 0) lrow[row] += 0.5f * lam_n * acc;
}

// ---- graph-replayed training steps -----------------------------------------
// The minibatch offset and Adam's per-step rate come from device memory
// (a step counter and a per-epoch table), so one captured step serves every
// minibatch of an epoch.

// rows of minibatch *step (size `batch`) from the resident set (padding zero)
__global__ void k_gather_step(const float* __restrict__ Zall, const int32_t* __restrict__ idx,
                              const int32_t* __restrict__ step, int batch,
                              const float* __restrict__ tall, const float* __restrict__ Gall,
                              float* __restrict__ Z, float* __restrict__ t, float* __restrict__ G,
                              int n, int rows, int D) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)rows * D) return;
    const int r = (int)(i / D), c = (int)(i % D);
    const int32_t* id = idx + (size_t)step[0] * batch;
    const bool valid = r < n;
    const size_t src = valid ? (size_t)id[r] * D + c : 0;
    Z[i] = valid ? Zall[src] : 0.f;
    if (G) G[i] = valid ? Gall[src] : 0.f;
    if (c == 0) t[r] = valid ? tall[id[r]] : 0.f;
}

// Adam with the bias corrections folded into lr_t = lr sqrt(1-b2^t)/(1-b1^t)
// and eps_t = eps sqrt(1-b2^t); (lr_t, eps_t) = cur[0], cur[1] (device memory)
__global__ void k_adam_step(float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, long long n,
                            const float* __restrict__ cur, float b1, float b2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float lr_t = cur[0], eps_t = cur[1];
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr_t * mi / (sqrtf(vi) + eps_t);
}

__global__ void k_step_set(int32_t* __restrict__ step, int value) {
    if (threadIdx.x == 0) step[0] = value;
}

int pad_to(int x, int m) { return (x + m - 1) / m * m; }
unsigned nblocks(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace

struct mcp_mlp {
    int device = 0;
    int d = 0;     // inputs (branches)
    int D = 0;     // padded inputs
    int H = 256;   // hidden width
    int rows = 0;  // row capacity of the activation buffers
    hipStream_t stream = nullptr;
    // parameters (flat):
    //   W1 [H][D] | b1 [H] | W2 [H][H] | b2 [H] | w3 [H] | b3 [64] | v [D] | q [D]
    size_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, ow3 = 0, ob3 = 0, ov = 0, oq = 0, nparam = 0;
    float *p = nullptr, *gp = nullptr, *am = nullptr, *av = nullptr;
    long long adam_t = 0;
    // activations (rows x ...)
    float *Z = nullptr, *H1 = nullptr, *H2 = nullptr, *y = nullptr, *t = nullptr, *dy = nullptr,
          *lrow = nullptr, *dA2 = nullptr, *dH1 = nullptr, *dZ = nullptr;
    // gradient-matching training (rows x ...)
    float *Gb = nullptr, *rho = nullptr, *D2 = nullptr, *T1 = nullptr, *Dl1 = nullptr,
          *K1 = nullptr, *K2 = nullptr, *Gam1 = nullptr;
    double *dX = nullptr, *dOut = nullptr;  // rows x d, rows x (d+1)
    float* loss = nullptr;                  // 64 floats
    // feature and target scaling
    float *xmean = nullptr, *xsd = nullptr;
    std::vector<float> h_xmean, h_xsd;
    float ymean = 0.f, ysd = 1.f;
    // resident training set (Gall: standardised exact gradients, optional)
    float *Zall = nullptr, *tall = nullptr, *Gall = nullptr;
    int32_t* idx = nullptr;
    int64_t ntrain = 0;
    // graph-replayed training step
    int32_t* dstep = nullptr;
    float* dcur = nullptr;    // the current step's (lr_t, eps_t)
    float* drates = nullptr;  // [2 * steps per epoch] (lr_t, eps_t)
    size_t drates_cap = 0;
    hipGraphExec_t train_graph = nullptr;
    int train_graph_batch = 0;
    bool train_graph_sob = false;
    bool train_graph_on = true;
    // surrogate leapfrog
    int chains = 0;
    double prior_scale = -1.0;
    double* dstate = nullptr;  // x | r | g | (eps, delta)
    double* dlp = nullptr;
    HmcItem* ditems = nullptr;
    double* hstate = nullptr;  // pinned: x | r | g | (eps, delta) | lp
    hipGraphExec_t lf_graph = nullptr;
    bool use_graph = false;

    void drop_graph() {
        if (lf_graph) (void)hipGraphExecDestroy(lf_graph);
        lf_graph = nullptr;
        drop_train_graph();
    }
    void drop_train_graph() {
        if (train_graph) (void)hipGraphExecDestroy(train_graph);
        train_graph = nullptr;
    }
    ~mcp_mlp() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graph();
        void* ptrs[] = {p,   gp,  am,  av,    Z,   H1,   H2,   y,   t,      dy,  lrow,   dA2,
                        dH1, dZ,  dX,  dOut,  loss, xmean, xsd, Zall, tall, idx, dstate, dlp,
                        ditems, Gb, rho, D2, T1, Dl1, K1, K2, Gam1, Gall, dstep, dcur, drates};
        for (void* ptr : ptrs)
            if (ptr) (void)hipFree(ptr);
        if (hstate) (void)hipHostFree(hstate);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

#define MLP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return set_last_error(e_ == hipErrorOutOfMemory ? MCP_ERR_NOMEM : MCP_ERR_HIP,   \
                                  (std::string(#expr) + ": " + hipGetErrorString(e_)).c_str()); \
    } while (0)

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int d) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != d) (void)hipSetDevice(d);
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// stream-ordered host <-> device copies, complete on return
hipError_t h2d(mcp_mlp* m, void* dst, const void* src, size_t bytes) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, m->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(m->stream);
}
hipError_t d2h(mcp_mlp* m, void* dst, const void* src, size_t bytes) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, m->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(m->stream);
}

int ensure_rows(mcp_mlp* m, int need) {
    const int rows = pad_to(std::max(need, 1), kTile);
    if (rows <= m->rows) return MCP_OK;
    MLP_TRY(hipStreamSynchronize(m->stream));
    m->drop_graph();  // captured launches hold the old addresses
    m->rows = 0;
    const size_t R = rows, D = m->D, H = m->H, d = m->d;
    struct Buf {
        void** p;
        size_t bytes;
    } bufs[] = {{(void**)&m->Z, R * D * 4},    {(void**)&m->H1, R * H * 4},
                {(void**)&m->H2, R * H * 4},   {(void**)&m->y, R * 4},
                {(void**)&m->t, R * 4},        {(void**)&m->dy, R * 4},
                {(void**)&m->lrow, R * 4},     {(void**)&m->dA2, R * H * 4},
                {(void**)&m->dH1, R * H * 4},  {(void**)&m->dZ, R * D * 4},
                {(void**)&m->dX, R * d * 8},   {(void**)&m->dOut, R * (d + 1) * 8},
                {(void**)&m->Gb, R * D * 4},   {(void**)&m->rho, R * D * 4},
                {(void**)&m->D2, R * H * 4},   {(void**)&m->T1, R * H * 4},
                {(void**)&m->Dl1, R * H * 4},  {(void**)&m->K1, R * H * 4},
                {(void**)&m->K2, R * H * 4},   {(void**)&m->Gam1, R * H * 4}};
    for (Buf& b : bufs) {
        if (*b.p) (void)hipFree(*b.p);
        *b.p = nullptr;
        MLP_TRY(hipMalloc(b.p, b.bytes));
        MLP_TRY(hipMemsetAsync(*b.p, 0, b.bytes, m->stream));
    }
    MLP_TRY(hipStreamSynchronize(m->stream));
    m->rows = rows;
    return MCP_OK;
}

template <bool TA, bool TB, int EPI>
void gemm(mcp_mlp* m, int M, int N, int K, const float* A, int lda, const float* B, int ldb,
          float* C, int ldc, Epi ep = Epi{}, const float* A2 = nullptr,
          const float* B2 = nullptr) {
    hipLaunchKernelGGL((k_gemm_bf16<TA, TB, EPI>), dim3(N / kTile, M / kTile), dim3(256), 0,
                       m->stream, A, A2, lda, B, B2, ldb, C, ldc, K, ep);
}

Epi epi(const float* aux, const float* aux2 = nullptr, float* out2 = nullptr,
        const float* vn = nullptr, const float* vm = nullptr) {
    Epi e;
    e.aux = aux;
    e.aux2 = aux2;
    e.out2 = out2;
    e.vn = vn;
    e.vm = vm;
    return e;
}

// h1 = tanh(W1 z + b1), h2 = tanh(W2 h1 + b2) over `rows` rows of Z
void forward_hidden(mcp_mlp* m, int rows) {
    const int H = m->H, D = m->D;
    const float* P = m->p;
    gemm<false, true, EPI_BIAS_TANH>(m, rows, H, D, m->Z, D, P + m->oW1, D, m->H1, H,
                                     epi(P + m->ob1));
    gemm<false, true, EPI_BIAS_TANH>(m, rows, H, H, m->H1, H, P + m->oW2, H, m->H2, H,
                                     epi(P + m->ob2));
}

// h1, h2, y over `rows` rows of Z; with `seed` also the input-gradient seed
// delta2 = w3 (1 - h2^2) into seed
void forward(mcp_mlp* m, int rows, float* seed) {
    const int H = m->H, D = m->D;
    const float* P = m->p;
    forward_hidden(m, rows);
    hipLaunchKernelGGL(k_head, dim3((rows + 3) / 4), dim3(256), 0, m->stream, m->H2, H, m->Z, D,
                       P + m->ow3, P + m->ob3, P + m->ov, P + m->oq, m->y, seed, rows);
}

// input gradient dZ = ((dA2 W2) * (1 - h1^2)) W1 (without the skip terms)
void backward_input(mcp_mlp* m, int rows) {
    const int H = m->H, D = m->D;
    gemm<false, false, EPI_MUL_DTANH>(m, rows, H, H, m->dA2, H, m->p + m->oW2, H, m->dH1, H,
                                      epi(m->H1));
    gemm<false, false, EPI_STORE>(m, rows, D, H, m->dH1, H, m->p + m->oW1, D, m->dZ, D);
}

ColJob coljob(int kind, int cols, const float* X, const float* s, const float* X2, float* out,
              float* out2) {
    ColJob j;
    j.kind = kind;
    j.cols = cols;
    j.X = X;
    j.s = s;
    j.X2 = X2;
    j.out = out;
    j.out2 = out2;
    return j;
}

void launch_colsums(mcp_mlp* m, ColJobs& jobs) {
    int blocks = 0;
    for (int k = 0; k < jobs.njobs; ++k) {
        jobs.job[k].block0 = blocks;
        blocks += jobs.job[k].kind == COL_SCALAR ? 1 : (jobs.job[k].cols + 63) / 64;
    }
    hipLaunchKernelGGL(k_colsums, dim3(blocks), dim3(64 * kColWaves), 0, m->stream, jobs);
}

// Parameter gradients of the batch loss into gp (every entry written) from
// h1, h2 of forward_hidden(); with lam > 0 the gradient-matching term against
// the targets in Gb is included.  loss_acc += batch loss.  With `advance` the
// scalar column job also moves the training step counter and fetches the
// step's Adam rates into dcur.
void backward_params(mcp_mlp* m, int rows, int n, float* loss_acc, float lam, bool advance) {
    const int H = m->H, D = m->D;
    const float* P = m->p;
    const float* W1 = P + m->oW1;
    const float* W2 = P + m->oW2;
    float* G = m->gp;
    hipStream_t st = m->stream;
    const bool sob = lam > 0.f;
    hipLaunchKernelGGL(k_head_train, dim3((rows + 3) / 4), dim3(256), 0, st, m->H2, H, m->Z, D,
                       P + m->ow3, P + m->ob3, P + m->ov, P + m->oq, m->t, m->y, m->dy, m->lrow,
                       m->dA2, sob ? m->D2 : nullptr, n, rows);
    if (sob) {
        // input gradient g = W1^T delta1 + v + 2 q z (T1 = delta2 W2, delta1)
        gemm<false, false, EPI_SOB_T1>(m, rows, H, H, m->D2, H, W2, H, m->T1, H,
                                       epi(m->H1, nullptr, m->Dl1));
        gemm<false, false, EPI_STORE>(m, rows, D, H, m->Dl1, H, W1, D, m->dZ, D);
        hipLaunchKernelGGL(k_sob_rho, dim3((rows + 3) / 4), dim3(256), 0, st, m->dZ, m->Z, m->Gb,
                           P + m->ov, P + m->oq, m->rho, m->lrow, D, m->d, n, rows,
                           lam / ((float)n * (float)m->d));
        // second backward pass through the input-gradient graph
        gemm<false, true, EPI_SOB_K1>(m, rows, H, D, m->rho, D, W1, D, m->K1, H,
                                      epi(m->H1, m->T1, m->Gam1));
        gemm<false, true, EPI_SOB_K2>(m, rows, H, H, m->K1, H, W2, H, m->K2, H,
                                      epi(m->H2, m->D2, m->dA2, P + m->ow3, m->dy));
    }
    ColJobs jobs;
    jobs.rows = rows;
    jobs.job[0] = sob ? coljob(COL_SUM, H, m->K2, nullptr, nullptr, G + m->ow3, nullptr)
                      : coljob(COL_SUM, H, m->H2, m->dy, nullptr, G + m->ow3, nullptr);
    jobs.job[1] = coljob(COL_SUM, H, m->dA2, nullptr, nullptr, G + m->ob2, nullptr);
    jobs.job[2] = coljob(COL_VQ, D, m->Z, m->dy, sob ? m->rho : nullptr, G + m->ov, G + m->oq);
    jobs.job[3] = coljob(COL_SCALAR, 1, nullptr, m->dy, m->lrow, G + m->ob3, loss_acc);
    jobs.njobs = 4;
    if (advance) {
        jobs.step = m->dstep;
        jobs.rates = m->drates;
        jobs.cur = m->dcur;
    }
    launch_colsums(m, jobs);
    // gW2 = A2tot^T h1 (+ delta2^T mu1)
    gemm<true, false, EPI_STORE>(m, H, H, rows, m->dA2, H, m->H1, H, G + m->oW2, H, Epi{},
                                 sob ? m->D2 : nullptr, sob ? m->K1 : nullptr);
    if (sob)
        gemm<false, false, EPI_SOB_A1>(m, rows, H, H, m->dA2, H, W2, H, m->dH1, H,
                                       epi(m->H1, m->Gam1));
    else
        gemm<false, false, EPI_MUL_DTANH>(m, rows, H, H, m->dA2, H, W2, H, m->dH1, H, epi(m->H1));
    ColJobs jb1;
    jb1.rows = rows;
    jb1.job[0] = coljob(COL_SUM, H, m->dH1, nullptr, nullptr, G + m->ob1, nullptr);
    jb1.njobs = 1;
    launch_colsums(m, jb1);
    // gW1 = A1tot^T z (+ delta1^T rho)
    gemm<true, false, EPI_STORE>(m, H, D, rows, m->dH1, H, m->Z, D, G + m->oW1, D, Epi{},
                                 sob ? m->Dl1 : nullptr, sob ? m->rho : nullptr);
}

// one Adam step on minibatch *dstep (nb rows): gather, forward, backward
// (which advances the step counter), update.  Every launch reads the step
// or its rates from device memory.
void enqueue_train_step(mcp_mlp* m, int batch, int nb, float lam) {
    const int rows = pad_to(nb, kTile);
    hipLaunchKernelGGL(k_gather_step, dim3(nblocks((long long)rows * m->D)), dim3(256), 0,
                       m->stream, m->Zall, m->idx, m->dstep, batch, m->tall,
                       lam > 0.f ? m->Gall : nullptr, m->Z, m->t, lam > 0.f ? m->Gb : nullptr,
                       nb, rows, m->D);
    forward_hidden(m, rows);
    backward_params(m, rows, nb, m->loss, lam, true);
    hipLaunchKernelGGL(k_adam_step, dim3(nblocks((long long)m->nparam)), dim3(256), 0, m->stream,
                       m->p, m->gp, m->am, m->av, (long long)m->nparam, m->dcur, 0.9f, 0.999f);
}

// logL~ and its gradient for the n rows of m->dX (already on the device)
void infer(mcp_mlp* m, int n) {
    const int rows = pad_to(n, kTile);
    hipLaunchKernelGGL(k_features, dim3(nblocks((long long)rows * m->D)), dim3(256), 0, m->stream,
                       m->dX, m->d, m->xmean, m->xsd, m->Z, rows, m->D, n);
    forward(m, rows, m->dA2);
    backward_input(m, rows);
    hipLaunchKernelGGL(k_finalize, dim3(nblocks((long long)n * (m->d + 1))), dim3(256), 0, m->stream,
                       m->y, m->dZ, m->Z, m->D, m->p + m->ov, m->p + m->oq, m->dX, m->d, m->xsd,
                       m->ymean, m->ysd, m->dOut, n);
}

int check_rows(mcp_mlp* m, const double* X, int64_t n) {
    for (int64_t i = 0; i < n * m->d; ++i)
        if (!(X[i] > 0.0) || !std::isfinite(X[i]))
            return set_last_error(MCP_ERR_INVALID, "branch lengths must be positive and finite");
    return MCP_OK;
}

// standardised features of host rows (same arithmetic as k_features)
int features_host(mcp_mlp* m, const double* X, int64_t n, std::vector<float>& out) {
    int st = check_rows(m, X, n);
    if (st) return st;
    out.assign((size_t)n * m->D, 0.f);
    for (int64_t r = 0; r < n; ++r)
        for (int c = 0; c < m->d; ++c)
            out[(size_t)r * m->D + c] =
                (float)((std::log(X[r * m->d + c]) - (double)m->h_xmean[c]) / (double)m->h_xsd[c]);
    return MCP_OK;
}

// standardised gradient targets dy/dz = dlogL/dx * x * x_sd / y_sd
int grad_targets_host(mcp_mlp* m, const double* X, const double* grad, int64_t n,
                      std::vector<float>& out) {
    out.assign((size_t)n * m->D, 0.f);
    for (int64_t r = 0; r < n; ++r)
        for (int c = 0; c < m->d; ++c) {
            const double g = grad[r * m->d + c];
            if (!std::isfinite(g)) return set_last_error(MCP_ERR_INVALID, "non-finite gradient");
            out[(size_t)r * m->D + c] =
                (float)(g * X[r * m->d + c] * (double)m->h_xsd[c] / (double)m->ysd);
        }
    return MCP_OK;
}

// H2D [x|r|g|eps,delta] -> k_lf_pre -> surrogate logL~ and gradient ->
// k_lf_post -> D2H [x|r|g], lp
int enqueue_leapfrog(mcp_mlp* m) {
    hipStream_t st = m->stream;
    const int n = m->chains, d = m->d;
    const size_t T = (size_t)n * d;
    double* dx = m->dstate;
    double* dr = dx + T;
    double* dg = dr + T;
    double* ded = dg + T;
    MLP_TRY(hipMemcpyAsync(dx, m->hstate, (3 * T + 2 * (size_t)n) * sizeof(double),
                           hipMemcpyHostToDevice, st));
    MLP_TRY(launch_lf_pre(m->ditems, n, dx, dr, dg, ded, m->dX, 64 * std::max(1, d), st));
    infer(m, n);
    MLP_TRY(hipGetLastError());
    MLP_TRY(launch_lf_post(m->ditems, n, dx, dr, dg, ded, m->dOut, m->dlp, nullptr, st));
    MLP_TRY(hipMemcpyAsync(m->hstate, dx, 3 * T * sizeof(double), hipMemcpyDeviceToHost, st));
    MLP_TRY(hipMemcpyAsync(m->hstate + 3 * T + 2 * (size_t)n, m->dlp, n * sizeof(double),
                           hipMemcpyDeviceToHost, st));
    return MCP_OK;
}

}  // namespace

extern "C" {

int mcp_mlp_create(mcp_mlp** out, int32_t n_inputs, int32_t hidden, int32_t device,
                   uint64_t seed) {
    if (!out) return set_last_error(MCP_ERR_INVALID, "null argument");
    *out = nullptr;
    if (n_inputs < 1 || n_inputs > (1 << 16) || hidden < kTile || hidden % kTile || hidden > 4096)
        return set_last_error(MCP_ERR_INVALID,
                              "mlp: 1 <= n_inputs <= 65536, hidden a multiple of 64 up to 4096");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return set_last_error(MCP_ERR_INVALID, "device ordinal out of range");
    DeviceGuard guard(device);
    mcp_mlp* m = new (std::nothrow) mcp_mlp();
    if (!m) return set_last_error(MCP_ERR_NOMEM, "host allocation failed");
    m->device = device;
    m->d = n_inputs;
    m->D = pad_to(n_inputs, kTile);
    m->H = hidden;
    const int H = m->H, D = m->D;
    m->oW1 = 0;
    m->ob1 = m->oW1 + (size_t)H * D;
    m->oW2 = m->ob1 + H;
    m->ob2 = m->oW2 + (size_t)H * H;
    m->ow3 = m->ob2 + H;
    m->ob3 = m->ow3 + H;
    m->ov = m->ob3 + 64;
    m->oq = m->ov + D;
    m->nparam = m->oq + D;
    m->h_xmean.assign(D, 0.f);
    m->h_xsd.assign(D, 1.f);
    auto fail = [&](int code, const char* msg) {
        delete m;
        return set_last_error(code, msg);
    };
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(MCP_ERR_HIP, "hipStreamCreate failed");
    float** pb[] = {&m->p, &m->gp, &m->am, &m->av};
    for (float** q : pb) {
        if (hipMalloc(q, m->nparam * sizeof(float)) != hipSuccess)
            return fail(MCP_ERR_NOMEM, "mlp parameters");
        if (hipMemsetAsync(*q, 0, m->nparam * sizeof(float), m->stream) != hipSuccess)
            return fail(MCP_ERR_HIP, "mlp parameters");
    }
    if (hipMalloc(&m->loss, 64 * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->xmean, D * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->xsd, D * sizeof(float)) != hipSuccess)
        return fail(MCP_ERR_NOMEM, "mlp buffers");
    // Glorot-uniform weights from a seeded host generator; zero biases and skip
    std::vector<float> host(m->nparam, 0.f);
    std::mt19937_64 rng(seed);
    auto init = [&](size_t off, int nrows, int ld, int fan_in, int valid_cols) {
        const float a = std::sqrt(6.f / (float)(fan_in + nrows));
        std::uniform_real_distribution<float> U(-a, a);
        for (int r = 0; r < nrows; ++r)
            for (int c = 0; c < valid_cols; ++c) host[off + (size_t)r * ld + c] = U(rng);
    };
    init(m->oW1, H, D, n_inputs, n_inputs);
    init(m->oW2, H, H, H, H);
    init(m->ow3, 1, H, H, H);
    if (h2d(m, m->p, host.data(), m->nparam * sizeof(float)) != hipSuccess ||
        h2d(m, m->xmean, m->h_xmean.data(), D * sizeof(float)) != hipSuccess ||
        h2d(m, m->xsd, m->h_xsd.data(), D * sizeof(float)) != hipSuccess)
        return fail(MCP_ERR_HIP, "mlp upload");
    *out = m;
    return MCP_OK;
}

void mcp_mlp_destroy(mcp_mlp* m) {
    if (!m) return;
    DeviceGuard guard(m->device);
    delete m;
}

int mcp_mlp_layout(mcp_mlp* m, int64_t* offsets9) {
    if (!m || !offsets9) return set_last_error(MCP_ERR_INVALID, "null argument");
    const size_t o[9] = {m->oW1, m->ob1, m->oW2, m->ob2, m->ow3, m->ob3, m->ov, m->oq, m->nparam};
    for (int i = 0; i < 9; ++i) offsets9[i] = (int64_t)o[i];
    return MCP_OK;
}

int mcp_mlp_get_params(mcp_mlp* m, float* params) {
    if (!m || !params) return set_last_error(MCP_ERR_INVALID, "null argument");
    DeviceGuard guard(m->device);
    MLP_TRY(d2h(m, params, m->p, m->nparam * sizeof(float)));
    return MCP_OK;
}

int mcp_mlp_set_params(mcp_mlp* m, const float* params) {
    if (!m || !params) return set_last_error(MCP_ERR_INVALID, "null argument");
    DeviceGuard guard(m->device);
    MLP_TRY(h2d(m, m->p, params, m->nparam * sizeof(float)));
    return MCP_OK;
}

int mcp_mlp_set_scaling(mcp_mlp* m, const float* x_mean, const float* x_sd, float y_mean,
                        float y_sd) {
    if (!m || !x_mean || !x_sd) return set_last_error(MCP_ERR_INVALID, "null argument");
    if (!(y_sd > 0.f) || !std::isfinite(y_sd) || !std::isfinite(y_mean))
        return set_last_error(MCP_ERR_INVALID, "target scale must be positive and finite");
    for (int c = 0; c < m->d; ++c)
        if (!(x_sd[c] > 0.f) || !std::isfinite(x_sd[c]) || !std::isfinite(x_mean[c]))
            return set_last_error(MCP_ERR_INVALID, "feature scales must be positive and finite");
    DeviceGuard guard(m->device);
    for (int c = 0; c < m->d; ++c) {
        m->h_xmean[c] = x_mean[c];
        m->h_xsd[c] = x_sd[c];
    }
    MLP_TRY(h2d(m, m->xmean, m->h_xmean.data(), m->D * sizeof(float)));
    MLP_TRY(h2d(m, m->xsd, m->h_xsd.data(), m->D * sizeof(float)));
    m->ymean = y_mean;
    m->ysd = y_sd;
    m->drop_graph();  // y_mean and y_sd are captured launch arguments
    return MCP_OK;
}

int mcp_mlp_set_data(mcp_mlp* m, const double* X, const double* logl, const double* grad,
                     int64_t n) {
    if (!m || !X || !logl) return set_last_error(MCP_ERR_INVALID, "null argument");
    if (n < 1 || n > INT32_MAX) return set_last_error(MCP_ERR_INVALID, "bad number of rows");
    DeviceGuard guard(m->device);
    std::vector<float> Zh, Gh;
    int st = features_host(m, X, n, Zh);
    if (st) return st;
    if (grad && (st = grad_targets_host(m, X, grad, n, Gh))) return st;
    std::vector<float> th(n);
    for (int64_t i = 0; i < n; ++i) {
        if (!std::isfinite(logl[i])) return set_last_error(MCP_ERR_INVALID, "non-finite target");
        th[i] = (float)((logl[i] - (double)m->ymean) / (double)m->ysd);
    }
    MLP_TRY(hipStreamSynchronize(m->stream));
    m->drop_train_graph();  // the captured step holds the old buffer addresses
    float** fb[] = {&m->Zall, &m->tall, &m->Gall};
    for (float** q : fb) {
        if (*q) (void)hipFree(*q);
        *q = nullptr;
    }
    if (m->idx) (void)hipFree(m->idx);
    m->idx = nullptr;
    m->ntrain = 0;
    MLP_TRY(hipMalloc(&m->Zall, Zh.size() * sizeof(float)));
    MLP_TRY(hipMalloc(&m->tall, th.size() * sizeof(float)));
    MLP_TRY(hipMalloc(&m->idx, (size_t)n * sizeof(int32_t)));
    MLP_TRY(h2d(m, m->Zall, Zh.data(), Zh.size() * sizeof(float)));
    MLP_TRY(h2d(m, m->tall, th.data(), th.size() * sizeof(float)));
    if (grad) {
        MLP_TRY(hipMalloc(&m->Gall, Gh.size() * sizeof(float)));
        MLP_TRY(h2d(m, m->Gall, Gh.data(), Gh.size() * sizeof(float)));
    }
    m->ntrain = n;
    return MCP_OK;
}

int mcp_mlp_train(mcp_mlp* m, int32_t batch, int32_t epochs, float lr, float lr_final,
                  float grad_weight, uint64_t seed, float* epoch_loss) {
    if (!m) return set_last_error(MCP_ERR_INVALID, "null argument");
    if (m->ntrain < 1) return set_last_error(MCP_ERR_STATE, "no training data (mcp_mlp_set_data)");
    if (batch < 1 || batch > (1 << 16) || epochs < 0 || !(lr > 0.f) || !(lr_final > 0.f) ||
        !(grad_weight >= 0.f) || !std::isfinite(grad_weight))
        return set_last_error(MCP_ERR_INVALID, "bad training arguments");
    if (grad_weight > 0.f && !m->Gall)
        return set_last_error(MCP_ERR_STATE, "gradient weight > 0 needs gradient targets");
    const float lam = grad_weight;
    DeviceGuard guard(m->device);
    int st = ensure_rows(m, batch);
    if (st) return st;
    const int64_t n = m->ntrain;
    std::mt19937_64 rng(seed);
    std::vector<int32_t> perm(n);
    for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
    const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
    const int64_t nfull = n / batch;
    const int rem = (int)(n % batch);
    const int64_t nbatch = nfull + (rem ? 1 : 0);
    if (!m->dstep) MLP_TRY(hipMalloc(&m->dstep, 64 * sizeof(int32_t)));
    if (!m->dcur) MLP_TRY(hipMalloc(&m->dcur, 64 * sizeof(float)));
    if (m->drates_cap < (size_t)(2 * nbatch)) {
        m->drop_train_graph();
        if (m->drates) (void)hipFree(m->drates);
        m->drates = nullptr;
        m->drates_cap = 0;
        MLP_TRY(hipMalloc(&m->drates, (size_t)(2 * nbatch) * sizeof(float)));
        m->drates_cap = (size_t)(2 * nbatch);
    }
    if (m->train_graph && (m->train_graph_batch != batch || m->train_graph_sob != (lam > 0.f)))
        m->drop_train_graph();
    std::vector<float> rates((size_t)(2 * nbatch));
    for (int ep = 0; ep < epochs; ++ep) {
        // exponential decay from lr to lr_final over the epochs
        const double frac = epochs > 1 ? (double)ep / (double)(epochs - 1) : 0.0;
        const double lr_e = (double)lr * std::pow((double)lr_final / (double)lr, frac);
        for (int64_t k = 0; k < nbatch; ++k) {
            const double tt = (double)(m->adam_t + k + 1);
            const double c1 = 1.0 - std::pow(b1, tt), c2 = 1.0 - std::pow(b2, tt);
            rates[2 * k] = (float)(lr_e * std::sqrt(c2) / c1);
            rates[2 * k + 1] = (float)(eps * std::sqrt(c2));
        }
        std::shuffle(perm.begin(), perm.end(), rng);
        MLP_TRY(h2d(m, m->idx, perm.data(), n * sizeof(int32_t)));
        MLP_TRY(h2d(m, m->drates, rates.data(), rates.size() * sizeof(float)));
        MLP_TRY(hipMemsetAsync(m->loss, 0, sizeof(float), m->stream));
        hipLaunchKernelGGL(k_step_set, dim3(1), dim3(64), 0, m->stream, m->dstep, 0);
        if (nfull > 0 && !m->train_graph_on) {
            for (int64_t k = 0; k < nfull; ++k) enqueue_train_step(m, batch, batch, lam);
        } else if (nfull > 0) {
            if (!m->train_graph) {
                MLP_TRY(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
                enqueue_train_step(m, batch, batch, lam);
                hipGraph_t cap = nullptr;
                const hipError_t ce = hipStreamEndCapture(m->stream, &cap);
                if (ce != hipSuccess) return set_last_error(MCP_ERR_HIP, "graph capture failed");
                const hipError_t ie = hipGraphInstantiate(&m->train_graph, cap, nullptr, nullptr, 0);
                (void)hipGraphDestroy(cap);
                if (ie != hipSuccess) {
                    m->train_graph = nullptr;
                    return set_last_error(MCP_ERR_HIP, "graph instantiation failed");
                }
                m->train_graph_batch = batch;
                m->train_graph_sob = lam > 0.f;
            }
            for (int64_t k = 0; k < nfull; ++k) MLP_TRY(hipGraphLaunch(m->train_graph, m->stream));
        }
        if (rem) enqueue_train_step(m, batch, rem, lam);
        m->adam_t += nbatch;
        MLP_TRY(hipGetLastError());
        if (epoch_loss) {
            float l = 0.f;
            MLP_TRY(d2h(m, &l, m->loss, sizeof(float)));
            epoch_loss[ep] = l / (float)nbatch;
        }
    }
    MLP_TRY(hipStreamSynchronize(m->stream));
    return MCP_OK;
}

int mcp_mlp_param_grad(mcp_mlp* m, const double* X, const double* logl, const double* dlogl,
                       int64_t n, float grad_weight, float* grad, float* loss) {
    if (!m || !X || !logl || !grad) return set_last_error(MCP_ERR_INVALID, "null argument");
    if (n < 1 || n > (1 << 16)) return set_last_error(MCP_ERR_INVALID, "bad number of rows");
    if (!(grad_weight >= 0.f) || !std::isfinite(grad_weight) || (grad_weight > 0.f && !dlogl))
        return set_last_error(MCP_ERR_INVALID, "gradient weight > 0 needs gradient targets");
    const float lam = grad_weight;
    DeviceGuard guard(m->device);
    int st = ensure_rows(m, (int)n);
    if (st) return st;
    const int rows = pad_to((int)n, kTile);
    std::vector<float> Zh;
    if ((st = features_host(m, X, n, Zh))) return st;
    Zh.resize((size_t)rows * m->D, 0.f);
    std::vector<float> th(rows, 0.f);
    for (int64_t i = 0; i < n; ++i) th[i] = (float)((logl[i] - (double)m->ymean) / (double)m->ysd);
    MLP_TRY(h2d(m, m->Z, Zh.data(), Zh.size() * sizeof(float)));
    MLP_TRY(h2d(m, m->t, th.data(), th.size() * sizeof(float)));
    if (lam > 0.f) {
        std::vector<float> Gh;
        if ((st = grad_targets_host(m, X, dlogl, n, Gh))) return st;
        Gh.resize((size_t)rows * m->D, 0.f);
        MLP_TRY(h2d(m, m->Gb, Gh.data(), Gh.size() * sizeof(float)));
    }
    MLP_TRY(hipMemsetAsync(m->loss, 0, sizeof(float), m->stream));
    forward_hidden(m, rows);
    backward_params(m, rows, (int)n, m->loss, lam, false);
    MLP_TRY(hipGetLastError());
    MLP_TRY(d2h(m, grad, m->gp, m->nparam * sizeof(float)));
    if (loss) MLP_TRY(d2h(m, loss, m->loss, sizeof(float)));
    return MCP_OK;
}

int mcp_mlp_predict(mcp_mlp* m, const double* X, int64_t n, double* logl, double* grad) {
    if (!m || !X || !logl) return set_last_error(MCP_ERR_INVALID, "null argument");
    if (n < 1) return set_last_error(MCP_ERR_INVALID, "bad number of rows");
    int st = check_rows(m, X, n);
    if (st) return st;
    DeviceGuard guard(m->device);
    const int chunk = 4096;
    if ((st = ensure_rows(m, (int)std::min<int64_t>(n, chunk)))) return st;
    const int d = m->d;
    std::vector<double> out;
    for (int64_t r0 = 0; r0 < n; r0 += chunk) {
        const int nb = (int)std::min<int64_t>(chunk, n - r0);
        MLP_TRY(h2d(m, m->dX, X + r0 * d, (size_t)nb * d * sizeof(double)));
        infer(m, nb);
        MLP_TRY(hipGetLastError());
        out.resize((size_t)nb * (d + 1));
        MLP_TRY(d2h(m, out.data(), m->dOut, out.size() * sizeof(double)));
        for (int i = 0; i < nb; ++i) {
            logl[r0 + i] = out[(size_t)i * (d + 1)];
            if (grad)
                std::memcpy(grad + (r0 + i) * d, out.data() + (size_t)i * (d + 1) + 1,
                            d * sizeof(double));
        }
    }
    return MCP_OK;
}

int mcp_mlp_leapfrog(mcp_mlp* m, int32_t chains, double prior_scale, const double* eps_delta,
                     double* x, double* r, double* g, double* logpost) {
    if (!m || !eps_delta || !x || !r || !g || !logpost)
        return set_last_error(MCP_ERR_INVALID, "null argument");
    if (chains < 1 || chains > (1 << 16))
        return set_last_error(MCP_ERR_INVALID, "chains must be in 1..65536");
    DeviceGuard guard(m->device);
    const int d = m->d;
    const size_t T = (size_t)chains * d;
    int st;
    if ((st = ensure_rows(m, chains))) return st;
    if (chains != m->chains) {
        MLP_TRY(hipStreamSynchronize(m->stream));
        m->drop_graph();
        if (m->dstate) (void)hipFree(m->dstate);
        if (m->dlp) (void)hipFree(m->dlp);
        if (m->ditems) (void)hipFree(m->ditems);
        if (m->hstate) (void)hipHostFree(m->hstate);
        m->dstate = m->dlp = nullptr;
        m->ditems = nullptr;
        m->hstate = nullptr;
        m->chains = 0;
        MLP_TRY(hipMalloc(&m->dstate, (3 * T + 2 * (size_t)chains) * sizeof(double)));
        MLP_TRY(hipMalloc(&m->dlp, (size_t)chains * sizeof(double)));
        MLP_TRY(hipMalloc(&m->ditems, (size_t)chains * sizeof(HmcItem)));
        MLP_TRY(hipHostMalloc(&m->hstate, (3 * T + 3 * (size_t)chains) * sizeof(double),
                              hipHostMallocDefault));
        m->chains = chains;
        m->prior_scale = std::nan("");  // forces the item upload below
    }
    if (!(prior_scale == m->prior_scale)) {
        std::vector<HmcItem> items(chains);
        for (int i = 0; i < chains; ++i) {
            items[i].n = d;
            items[i].off = i * d;
            items[i].out = i * (d + 1);
            items[i].prior = prior_scale > 0.0 ? 1 : 0;
            items[i].prior_scale = prior_scale > 0.0 ? prior_scale : 1.0;
        }
        MLP_TRY(h2d(m, m->ditems, items.data(), chains * sizeof(HmcItem)));
        m->prior_scale = prior_scale;
    }
    std::memcpy(m->hstate, x, T * sizeof(double));
    std::memcpy(m->hstate + T, r, T * sizeof(double));
    std::memcpy(m->hstate + 2 * T, g, T * sizeof(double));
    std::memcpy(m->hstate + 3 * T, eps_delta, 2 * (size_t)chains * sizeof(double));
    if (m->use_graph) {
        if (!m->lf_graph) {
            MLP_TRY(hipStreamSynchronize(m->stream));
            MLP_TRY(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
            st = enqueue_leapfrog(m);
            hipGraph_t cap = nullptr;
            const hipError_t ce = hipStreamEndCapture(m->stream, &cap);
            if (st) {
                if (cap) (void)hipGraphDestroy(cap);
                return st;
            }
            if (ce != hipSuccess) return set_last_error(MCP_ERR_HIP, "graph capture failed");
            const hipError_t ie = hipGraphInstantiate(&m->lf_graph, cap, nullptr, nullptr, 0);
            (void)hipGraphDestroy(cap);
            if (ie != hipSuccess) {
                m->lf_graph = nullptr;
                return set_last_error(MCP_ERR_HIP, "graph instantiation failed");
            }
        }
        MLP_TRY(hipGraphLaunch(m->lf_graph, m->stream));
    } else if ((st = enqueue_leapfrog(m))) {
        return st;
    }
    MLP_TRY(hipStreamSynchronize(m->stream));
    std::memcpy(x, m->hstate, T * sizeof(double));
    std::memcpy(r, m->hstate + T, T * sizeof(double));
    std::memcpy(g, m->hstate + 2 * T, T * sizeof(double));
    std::memcpy(logpost, m->hstate + 3 * T + 2 * (size_t)chains, chains * sizeof(double));
    return MCP_OK;
}

int mcp_mlp_use_graph(mcp_mlp* m, int32_t on) {
    if (!m) return set_last_error(MCP_ERR_INVALID, "null argument");
    DeviceGuard guard(m->device);
    MLP_TRY(hipStreamSynchronize(m->stream));
    m->use_graph = on != 0;
    if (!m->use_graph) m->drop_graph();
    return MCP_OK;
}

int mcp_mlp_train_graph(mcp_mlp* m, int32_t on) {
    if (!m) return set_last_error(MCP_ERR_INVALID, "null argument");
    DeviceGuard guard(m->device);
    MLP_TRY(hipStreamSynchronize(m->stream));
    m->train_graph_on = on != 0;
    if (!m->train_graph_on) m->drop_train_graph();
    return MCP_OK;
}

}  // extern "C"
