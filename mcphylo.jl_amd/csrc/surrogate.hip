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

constexpr int kTile = 64;   // block tile (M and N)
constexpr int kKStep = 32;  // K per MFMA step
constexpr int kLdsK = 40;   // LDS row length in bf16 (32 + 8 padding)

enum { EPI_STORE = 0, EPI_BIAS_TANH = 1, EPI_MUL_DTANH = 2, EPI_ACCUM = 3 };

// C = epilogue(opA . opB), opA: M x K, opB: K x N, all FP32 in memory.
//   opA[m][k] = TA ? A[k*lda + m] : A[m*lda + k]
//   opB[k][n] = TB ? B[n*ldb + k] : B[k*ldb + n]
// EPI_STORE: C = acc; EPI_BIAS_TANH: C = tanh(acc + aux[n]);
// EPI_MUL_DTANH: C = acc * (1 - aux[m*ldc + n]^2); EPI_ACCUM: C += acc.
// Grid (N/64, M/64), 256 threads: four waves, each a 32 x 32 quarter of the
// tile as 2 x 2 MFMA 16 x 16 blocks.
template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) k_gemm_bf16(const float* __restrict__ A, int lda,
                                                   const float* __restrict__ B, int ldb,
                                                   float* __restrict__ C, int ldc, int K,
                                                   const float* __restrict__ aux) {
    __shared__ __bf16 As[kTile][kLdsK];  // [m][k]
    __shared__ __bf16 Bs[kTile][kLdsK];  // [n][k]
    const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    const int li = lane & 15, lg = lane >> 4;
    fvec4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = fvec4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += kKStep) {
        // stage 64 x 32 of opA and of opB^T, 8 elements per thread each; the
        // fastest index follows the contiguous memory direction
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int idx = tid + 256 * e;  // 0 .. 2047
            const int am = TA ? (idx & 63) : (idx >> 5);
            const int ak = TA ? (idx >> 6) : (idx & 31);
            As[am][ak] = (__bf16)(TA ? A[(size_t)(k0 + ak) * lda + m0 + am]
                                     : A[(size_t)(m0 + am) * lda + k0 + ak]);
            const int bn = TB ? (idx >> 5) : (idx & 63);
            const int bk = TB ? (idx & 31) : (idx >> 6);
            Bs[bn][bk] = (__bf16)(TB ? B[(size_t)(n0 + bn) * ldb + k0 + bk]
                                     : B[(size_t)(k0 + bk) * ldb + n0 + bn]);
        }
        __syncthreads();
        bf16x8 af[2], bfr[2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int e = 0; e < 8; ++e) af[a][e] = As[wm + 16 * a + li][8 * lg + e];
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 8; ++e) bfr[b][e] = Bs[wn + 16 * b + li][8 * lg + e];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
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
                float v = acc[a][b][j];
                if (EPI == EPI_BIAS_TANH) v = tanhf(v + aux[n]);
                if (EPI == EPI_MUL_DTANH) {
                    const float h = aux[o];
                    v *= 1.f - h * h;
                }
                if (EPI == EPI_ACCUM) v += C[o];
                C[o] = v;
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

// Loss head of a training batch (n valid rows), one wavefront per row:
//   e = y - t; dy = e / n; dA2[r][k] = dy w3[k] (1 - h2^2); lrow = e^2 / (2n)
// padding rows get zeros.
__global__ void __launch_bounds__(256) k_head_bwd(const float* __restrict__ y,
                                                  const float* __restrict__ t,
                                                  const float* __restrict__ h2, int H,
                                                  const float* __restrict__ w3,
                                                  float* __restrict__ dy, float* __restrict__ dA2,
                                                  float* __restrict__ lrow, int n, int rows) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= rows) return;
    const float e = row < n ? y[row] - t[row] : 0.f;
    const float g = e / (float)n;
    for (int k = lane; k < H; k += 64) {
        const float h = h2[(size_t)row * H + k];
        dA2[(size_t)row * H + k] = g * w3[k] * (1.f - h * h);
    }
    if (lane == 0) {
        dy[row] = g;
        lrow[row] = 0.5f * e * e / (float)n;
    }
}

// Column reductions over rows (fixed order, one thread per column):
//   out[c] = sum_r s[r] X[r][c]         (s == nullptr: s = 1)
//   out2[c] = sum_r s[r] X[r][c]^2      (out2 != nullptr)
__global__ void k_colsum(const float* __restrict__ X, int ld, const float* __restrict__ s,
                         float* __restrict__ out, float* __restrict__ out2, int rows, int cols) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float a = 0.f, a2 = 0.f;
    for (int r = 0; r < rows; ++r) {
        const float xv = X[(size_t)r * ld + c];
        const float sv = s ? s[r] : 1.f;
        a += sv * xv;
        a2 += sv * xv * xv;
    }
    out[c] = a;
    if (out2) out2[c] = a2;
}

// gb3 = sum dy; loss_acc += sum lrow (single wavefront)
__global__ void k_sum2(const float* __restrict__ dy, const float* __restrict__ lrow,
                       float* __restrict__ gb3, float* __restrict__ loss_acc, int rows) {
    float a = 0.f, b = 0.f;
    for (int i = threadIdx.x; i < rows; i += 64) {
        a += dy[i];
        b += lrow[i];
    }
    a = wave_sum_f(a);
    b = wave_sum_f(b);
    if (threadIdx.x == 0) {
        gb3[0] = a;
        if (loss_acc) loss_acc[0] += b;
    }
}

// Adam with the bias corrections folded into lr_t = lr sqrt(1-b2^t)/(1-b1^t)
// and eps_t = eps sqrt(1-b2^t)
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long long n, float lr_t, float b1, float b2,
                       float eps_t) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr_t * mi / (sqrtf(vi) + eps_t);
}

// minibatch rows from the resident training set (padding rows zero)
__global__ void k_gather_rows(const float* __restrict__ Zall, const int32_t* __restrict__ idx,
                              const float* __restrict__ tall, float* __restrict__ Z,
                              float* __restrict__ t, int n, int rows, int D) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)rows * D) return;
    const int r = (int)(i / D), c = (int)(i % D);
    Z[i] = r < n ? Zall[(size_t)idx[r] * D + c] : 0.f;
    if (c == 0) t[r] = r < n ? tall[idx[r]] : 0.f;
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
// the input-gradient graph; the kernels below are its elementwise steps.

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

// Dl1 = T1 (1 - h1^2)  (delta1 from T1 = delta2 W2)
__global__ void k_sob_delta1(const float* __restrict__ T1, const float* __restrict__ H1,
                             float* __restrict__ Dl1, long long total) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const float h = H1[i];
    Dl1[i] = T1[i] * (1.f - h * h);
}

// from kappa1 = rho W1^T (in K1):  K1 <- mu1 = kappa1 (1 - h1^2);
// Gam1 = -2 h1 kappa1 T1
__global__ void k_sob_h1(float* __restrict__ K1, const float* __restrict__ T1,
                         const float* __restrict__ H1, float* __restrict__ Gam1, long long total) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const float h = H1[i], k1 = K1[i];
    K1[i] = k1 * (1.f - h * h);
    Gam1[i] = -2.f * h * k1 * T1[i];
}

// from kappa2 = mu1 W2^T (in K2), per row r with dy[r]:
//   A2tot = dy delta2 - 2 h2 kappa2 w3 (1 - h2^2)      -> dA2
//   K2   <- dy h2 + kappa2 (1 - h2^2)                  (rows of the w3 gradient)
__global__ void k_sob_h2(float* __restrict__ K2, const float* __restrict__ H2,
                         const float* __restrict__ D2, const float* __restrict__ w3,
                         const float* __restrict__ dy, float* __restrict__ dA2, int H,
                         long long total) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int c = (int)(i % H);
    const float g = dy[i / H];
    const float h = H2[i], k2 = K2[i], s2 = 1.f - h * h;
    dA2[i] = g * D2[i] - 2.f * h * k2 * w3[c] * s2;
    K2[i] = g * h + k2 * s2;
}

// A1 <- (A1 + Gam1) (1 - h1^2)   (A1 = A2tot W2 on entry)
__global__ void k_sob_a1(float* __restrict__ A1, const float* __restrict__ Gam1,
                         const float* __restrict__ H1, long long total) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const float h = H1[i];
    A1[i] = (A1[i] + Gam1[i]) * (1.f - h * h);
}

// gv[c] = sum_r dy_r z + rho;  gq[c] = sum_r dy_r z^2 + 2 z rho  (fixed order)
__global__ void k_colsum_vq(const float* __restrict__ Z, const float* __restrict__ dy,
                            const float* __restrict__ rho, float* __restrict__ gv,
                            float* __restrict__ gq, int rows, int D) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D) return;
    float a = 0.f, b = 0.f;
    for (int r = 0; r < rows; ++r) {
        const float z = Z[(size_t)r * D + c], g = dy[r];
        const float p = rho ? rho[(size_t)r * D + c] : 0.f;
        a += g * z + p;
        b += g * z * z + 2.f * z * p;
    }
    gv[c] = a;
    gq[c] = b;
}

// minibatch rows of the exact-gradient targets (padding rows zero)
__global__ void k_gather_grad(const float* __restrict__ Gall, const int32_t* __restrict__ idx,
                              float* __restrict__ G, int n, int rows, int D) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)rows * D) return;
    const int r = (int)(i / D), c = (int)(i % D);
    G[i] = r < n ? Gall[(size_t)idx[r] * D + c] : 0.f;
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
    }
    ~mcp_mlp() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graph();
        void* ptrs[] = {p,   gp,  am,  av,    Z,   H1,   H2,   y,   t,      dy,  lrow,   dA2,
                        dH1, dZ,  dX,  dOut,  loss, xmean, xsd, Zall, tall, idx, dstate, dlp,
                        ditems, Gb, rho, D2, T1, Dl1, K1, K2, Gam1, Gall};
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
          float* C, int ldc, const float* aux) {
    hipLaunchKernelGGL((k_gemm_bf16<TA, TB, EPI>), dim3(N / kTile, M / kTile), dim3(256), 0,
                       m->stream, A, lda, B, ldb, C, ldc, K, aux);
}

// h1, h2, y over `rows` rows of Z; with `seed` also the input-gradient seed
// delta2 = w3 (1 - h2^2) into seed
void forward(mcp_mlp* m, int rows, float* seed) {
    const int H = m->H, D = m->D;
    const float* P = m->p;
    gemm<false, true, EPI_BIAS_TANH>(m, rows, H, D, m->Z, D, P + m->oW1, D, m->H1, H, P + m->ob1);
    gemm<false, true, EPI_BIAS_TANH>(m, rows, H, H, m->H1, H, P + m->oW2, H, m->H2, H, P + m->ob2);
    hipLaunchKernelGGL(k_head, dim3((rows + 3) / 4), dim3(256), 0, m->stream, m->H2, H, m->Z, D,
                       P + m->ow3, P + m->ob3, P + m->ov, P + m->oq, m->y, seed, rows);
}

// input gradient dZ = ((dA2 W2) * (1 - h1^2)) W1 (without the skip terms)
void backward_input(mcp_mlp* m, int rows) {
    const int H = m->H, D = m->D;
    gemm<false, false, EPI_MUL_DTANH>(m, rows, H, H, m->dA2, H, m->p + m->oW2, H, m->dH1, H, m->H1);
    gemm<false, false, EPI_STORE>(m, rows, D, H, m->dH1, H, m->p + m->oW1, D, m->dZ, D, nullptr);
}

// parameter gradients of the batch loss into gp (every entry written); with
// lam > 0 the gradient-matching term against the targets in Gb is included
// (needs delta2 in D2 from forward()).
void backward_params(mcp_mlp* m, int rows, int n, float* loss_acc, float lam) {
    const int H = m->H, D = m->D;
    const float* W1 = m->p + m->oW1;
    const float* W2 = m->p + m->oW2;
    float* G = m->gp;
    hipStream_t st = m->stream;
    const long long tH = (long long)rows * H;
    const unsigned bH = nblocks(tH);
    hipLaunchKernelGGL(k_head_bwd, dim3((rows + 3) / 4), dim3(256), 0, st, m->y, m->t, m->H2, H,
                       m->p + m->ow3, m->dy, m->dA2, m->lrow, n, rows);
    if (lam > 0.f) {
        // input gradient g = W1^T delta1 + v + 2 q z
        gemm<false, false, EPI_STORE>(m, rows, H, H, m->D2, H, W2, H, m->T1, H, nullptr);
        hipLaunchKernelGGL(k_sob_delta1, dim3(bH), dim3(256), 0, st, m->T1, m->H1, m->Dl1, tH);
        gemm<false, false, EPI_STORE>(m, rows, D, H, m->Dl1, H, W1, D, m->dZ, D, nullptr);
        hipLaunchKernelGGL(k_sob_rho, dim3((rows + 3) / 4), dim3(256), 0, st, m->dZ, m->Z, m->Gb,
                           m->p + m->ov, m->p + m->oq, m->rho, m->lrow, D, m->d, n, rows,
                           lam / ((float)n * (float)m->d));
        // second backward pass through the input-gradient graph
        gemm<false, true, EPI_STORE>(m, rows, H, D, m->rho, D, W1, D, m->K1, H, nullptr);
        hipLaunchKernelGGL(k_sob_h1, dim3(bH), dim3(256), 0, st, m->K1, m->T1, m->H1, m->Gam1, tH);
        gemm<false, true, EPI_STORE>(m, rows, H, H, m->K1, H, W2, H, m->K2, H, nullptr);
        hipLaunchKernelGGL(k_sob_h2, dim3(bH), dim3(256), 0, st, m->K2, m->H2, m->D2, m->p + m->ow3,
                           m->dy, m->dA2, H, tH);
        hipLaunchKernelGGL(k_colsum, dim3(nblocks(H)), dim3(256), 0, st, m->K2, H, nullptr,
                           G + m->ow3, nullptr, rows, H);
    } else {
        hipLaunchKernelGGL(k_colsum, dim3(nblocks(H)), dim3(256), 0, st, m->H2, H, m->dy,
                           G + m->ow3, nullptr, rows, H);
    }
    hipLaunchKernelGGL(k_sum2, dim3(1), dim3(64), 0, st, m->dy, m->lrow, G + m->ob3, loss_acc, rows);
    hipLaunchKernelGGL(k_colsum_vq, dim3(nblocks(D)), dim3(256), 0, st, m->Z, m->dy,
                       lam > 0.f ? m->rho : nullptr, G + m->ov, G + m->oq, rows, D);
    hipLaunchKernelGGL(k_colsum, dim3(nblocks(H)), dim3(256), 0, st, m->dA2, H, nullptr, G + m->ob2,
                       nullptr, rows, H);
    gemm<true, false, EPI_STORE>(m, H, H, rows, m->dA2, H, m->H1, H, G + m->oW2, H, nullptr);
    if (lam > 0.f) {
        gemm<true, false, EPI_ACCUM>(m, H, H, rows, m->D2, H, m->K1, H, G + m->oW2, H, nullptr);
        gemm<false, false, EPI_STORE>(m, rows, H, H, m->dA2, H, W2, H, m->dH1, H, nullptr);
        hipLaunchKernelGGL(k_sob_a1, dim3(bH), dim3(256), 0, st, m->dH1, m->Gam1, m->H1, tH);
    } else {
        gemm<false, false, EPI_MUL_DTANH>(m, rows, H, H, m->dA2, H, W2, H, m->dH1, H, m->H1);
    }
    hipLaunchKernelGGL(k_colsum, dim3(nblocks(H)), dim3(256), 0, st, m->dH1, H, nullptr, G + m->ob1,
                       nullptr, rows, H);
    gemm<true, false, EPI_STORE>(m, H, D, rows, m->dH1, H, m->Z, D, G + m->oW1, D, nullptr);
    if (lam > 0.f)
        gemm<true, false, EPI_ACCUM>(m, H, D, rows, m->Dl1, H, m->rho, D, G + m->oW1, D, nullptr);
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
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const int64_t nbatch = (n + batch - 1) / batch;
    for (int ep = 0; ep < epochs; ++ep) {
        // exponential decay from lr to lr_final over the epochs
        const double frac = epochs > 1 ? (double)ep / (double)(epochs - 1) : 0.0;
        const double lr_e = (double)lr * std::pow((double)lr_final / (double)lr, frac);
        std::shuffle(perm.begin(), perm.end(), rng);
        MLP_TRY(h2d(m, m->idx, perm.data(), n * sizeof(int32_t)));
        MLP_TRY(hipMemsetAsync(m->loss, 0, sizeof(float), m->stream));
        for (int64_t b0 = 0; b0 < n; b0 += batch) {
            const int nb = (int)std::min<int64_t>(batch, n - b0);
            const int rows = pad_to(nb, kTile);
            hipLaunchKernelGGL(k_gather_rows, dim3(nblocks((long long)rows * m->D)), dim3(256), 0,
                               m->stream, m->Zall, m->idx + b0, m->tall, m->Z, m->t, nb, rows, m->D);
            if (lam > 0.f)
                hipLaunchKernelGGL(k_gather_grad, dim3(nblocks((long long)rows * m->D)), dim3(256),
                                   0, m->stream, m->Gall, m->idx + b0, m->Gb, nb, rows, m->D);
            forward(m, rows, lam > 0.f ? m->D2 : nullptr);
            backward_params(m, rows, nb, m->loss, lam);
            ++m->adam_t;
            const double c1 = 1.0 - std::pow((double)b1, (double)m->adam_t);
            const double c2 = 1.0 - std::pow((double)b2, (double)m->adam_t);
            const float lr_t = (float)(lr_e * std::sqrt(c2) / c1);
            const float eps_t = (float)(eps * std::sqrt(c2));
            hipLaunchKernelGGL(k_adam, dim3(nblocks((long long)m->nparam)), dim3(256), 0, m->stream,
                               m->p, m->gp, m->am, m->av, (long long)m->nparam, lr_t, b1, b2, eps_t);
        }
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
    forward(m, rows, lam > 0.f ? m->D2 : nullptr);
    backward_params(m, rows, (int)n, m->loss, lam);
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

}  // extern "C"
