// Learned block preconditioner: a mesh message-passing network that predicts,
// per vertex, a 3x3 correction to the block-Jacobi preconditioner and a log
// scale.  P_i = exp(s_i) R_i D_i^-1 R_i^T with R_i = I + 0.3 tanh(o_i) (every
// row strictly diagonally dominant, so R_i is invertible and P_i is SPD
// whenever D_i^-1 is).  The PCG warm start is a multiple of P b.  The reference has no
// counterpart (its PCG is block-Jacobi from x0 = 0, pcg_solver.cpp:8-75;
// integrator.cpp:187); the plain solve stays the fallback.
//
// Network (weights BF16 on the device, FP32 master copy; activations BF16,
// every product accumulated in FP32 by MFMA):
//   features f_i (kF = 32, 16 used) from the system matrix
//   h0 = relu(f W0 + b0)                               K = 32  -> 64
//   h_l = relu([h_{l-1} | mean_j h_{l-1,j}] W_l + b_l) K = 128 -> 64, l = 1..4
//   o = h4 Wo + bo (FP32)                              64 -> 10
// The warm start is the energy-optimal multiple of y = P b,
//   x0 = alpha y,  alpha = (b^T y) / (y^T A y),
// so it is never worse than x0 = 0 in the A-norm.  Training minimises the
// A-norm error it leaves, normalised by b^T D^-1 b (constant terms dropped):
//   L = -(b^T y)^2 / ((y^T A y) (b^T D^-1 b)),
//   dL/dy = 2 (b^T y) / ((y^T A y) N) (alpha A y - b);
// no solution vector is needed, and L is invariant to the scale of P, as PCG is.
//
// GEMMs use v_mfma_f32_32x32x8_bf16 (one wave computes a 32-row x 64-column
// tile as two 32x32 accumulators).  Operand layout of that instruction, per
// lane l of the wave:
//   A (32 x 8):  row l % 32, columns 4 (l / 32) .. +3 (4 BF16 in two VGPRs)
//   B (8 x 32):  column l % 32, rows 4 (l / 32) .. +3
//   C (32 x 32): register v holds row 8 (v / 4) + 4 (l / 32) + v % 4,
//                column l % 32.
// Weight matrices are packed once per update into that B order, so each lane
// reads 8 contiguous bytes.
#include <stdint.h>

#include "kernels.h"

namespace ando {
namespace nn {

constexpr int kF = 32;      // padded input features
constexpr int kFUsed = 16;  // features actually computed
constexpr int kH = 64;      // hidden width
constexpr int kL = 4;       // message-passing layers
constexpr int kOut = 10;    // 9 entries of the correction + 1 log scale
constexpr float kDeltaScale = 0.3f;
constexpr float kLogScale = 2.0f;

// Offsets into the flat FP32 weight vector (row-major [in x out] matrices).
constexpr int kW0 = 0;
constexpr int kB0 = kW0 + kF * kH;
constexpr int kLayer0 = kB0 + kH;
constexpr int kLayerSize = 2 * kH * kH + kH;  // W [2kH x kH], then b [kH]
constexpr int kWo = kLayer0 + kL * kLayerSize;
constexpr int kBo = kWo + kH * kOut;
constexpr int kNumWeights = kBo + kOut;

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));

__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);                                            // round to nearest even
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// --- weights -----------------------------------------------------------------

// Pack W [K x kH] (FP32, row-major, leading dimension ldw) into MFMA B order:
// packed[((kk * 2 + jt) * 64 + lane) * 4 + q] = W[8 kk + 4 (lane / 32) + q][32 jt + lane % 32].
// With transpose = 1 the source is W^T: element (k, j) is read from W[j * ldw + k].
__global__ void k_nn_pack(int K, const float* __restrict__ W, int ldw, int transpose, uint16_t* __restrict__ packed) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = K * kH;
  if (t >= total) return;
  const int q = t & 3, lane = (t >> 2) & 63, jt = (t >> 8) & 1, kk = t >> 9;
  const int k = 8 * kk + 4 * (lane >> 5) + q, j = 32 * jt + (lane & 31);
  const float w = transpose ? W[(size_t)j * ldw + k] : W[(size_t)k * ldw + j];
  packed[t] = f2bf(w);
}

// Round the FP32 master weights to BF16 and back (the values the device
// network uses), for the output layer, which runs in FP32 on BF16 weights.
__global__ void k_nn_round(int count, const float* __restrict__ w, float* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) out[t] = bf2f(f2bf(w[t]));
}

// --- forward -------------------------------------------------------------------

__device__ inline bool inv3_blockjacobi(const float* B, float* P) {
  auto h = [&](int a, int b, int c) { return B[a] * (B[3 + b] * B[6 + c] - B[3 + c] * B[6 + b]); };
  const float det = (h(0, 1, 2) - h(1, 0, 2)) + h(2, 0, 1);
  if (!((double)fabsf(det) > 1e-16)) {
    for (int e = 0; e < 9; ++e) P[e] = (e % 4 == 0) ? 1.0f : 0.0f;
    return false;
  }
  auto cof = [&](int i, int j) {
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return B[3 * i1 + j1] * B[3 * i2 + j2] - B[3 * i1 + j2] * B[3 * i2 + j1];
  };
  const float d = cof(0, 0) * B[0] + (cof(1, 0) * B[3] + cof(2, 0) * B[6]);
  const float inv = 1.0f / d;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) P[3 * a + b] = cof(b, a) * inv;
  return true;
}

// Per-vertex features (BF16, kF per row) and the block-Jacobi inverse D^-1
// (FP32, the plain preconditioner, same formula as the plain PCG).
//   t = tr(D) / 3 (1 if not positive)
//   f[0..5]   = D (xx, yy, zz, xy, xz, yz) / t
//   f[6..11]  = sym(sum_{j != i} A_ij) / t, same order
//   f[12]     = sum_{j != i} |A_ij|_F / (3 t)
//   f[13]     = log(1 + f[12])
//   f[14]     = number of off-diagonal blocks / 8
//   f[15]     = 1
__global__ void k_nn_features(int n, const float* __restrict__ A, const int32_t* __restrict__ rp,
                              const int32_t* __restrict__ cols, const int32_t* __restrict__ diag,
                              uint16_t* __restrict__ feat, float* __restrict__ Dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* D = A + 9 * (size_t)diag[i];
  float P[9];
  inv3_blockjacobi(D, P);
  for (int e = 0; e < 9; ++e) Dinv[9 * (size_t)i + e] = P[e];
  float t = (D[0] + D[4] + D[8]) / 3.0f;
  if (!(t > 0.0f)) t = 1.0f;
  const float it = 1.0f / t;
  float S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  float cpl = 0.0f;
  int deg = 0;
  for (int b = rp[i]; b < rp[i + 1]; ++b) {
    if (cols[b] == i) continue;
    const float* B = A + 9 * (size_t)b;
    float fro = 0.0f;
    for (int e = 0; e < 9; ++e) {
      S[e] += B[e];
      fro += B[e] * B[e];
    }
    cpl += sqrtf(fro);
    ++deg;
  }
  float f[kF];
  for (int k = 0; k < kF; ++k) f[k] = 0.0f;
  f[0] = D[0] * it; f[1] = D[4] * it; f[2] = D[8] * it;
  f[3] = 0.5f * (D[1] + D[3]) * it; f[4] = 0.5f * (D[2] + D[6]) * it; f[5] = 0.5f * (D[5] + D[7]) * it;
  f[6] = S[0] * it; f[7] = S[4] * it; f[8] = S[8] * it;
  f[9] = 0.5f * (S[1] + S[3]) * it; f[10] = 0.5f * (S[2] + S[6]) * it; f[11] = 0.5f * (S[5] + S[7]) * it;
  f[12] = cpl * it / 3.0f;
  f[13] = log1pf(f[12]);
  f[14] = (float)deg / 8.0f;
  f[15] = 1.0f;
  for (int k = 0; k < kF; ++k) feat[(size_t)i * kF + k] = f2bf(f[k]);
}

// out[M x 64] = act(A[M x K] W + bias); A BF16 row-major (leading dimension
// lda, a multiple of 4), W packed by k_nn_pack, K a multiple of 8.  One wave
// per 32 rows, four waves per workgroup, no barriers.  Output BF16 (out_bf)
// and/or FP32 (out_f), leading dimension ldo, written at column offset col0.
__global__ void __launch_bounds__(256) k_nn_gemm(int M, int K, const uint16_t* __restrict__ A, int lda,
                                                 const uint16_t* __restrict__ Wp, const float* __restrict__ bias,
                                                 int relu, uint16_t* __restrict__ out_bf, float* __restrict__ out_f,
                                                 int ldo, int col0) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int row0 = (blockIdx.x * 4 + wave) * 32;
  if (row0 >= M) return;
  const int ar = row0 + (l & 31);
  const bool aok = ar < M;
  f16v acc0 = {}, acc1 = {};
  for (int kk = 0; kk < K / 8; ++kk) {
    bf4 a = {0, 0, 0, 0};
    if (aok) a = *reinterpret_cast<const bf4*>(A + (size_t)ar * lda + 8 * kk + 4 * (l >> 5));
    const bf4 b0 = *reinterpret_cast<const bf4*>(Wp + ((size_t)(kk * 2 + 0) * 64 + l) * 4);
    const bf4 b1 = *reinterpret_cast<const bf4*>(Wp + ((size_t)(kk * 2 + 1) * 64 + l) * 4);
    acc0 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b0, acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b1, acc1, 0, 0, 0);
  }
  const int c = l & 31;
  const float bb0 = bias ? bias[c] : 0.0f, bb1 = bias ? bias[32 + c] : 0.0f;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int row = row0 + 8 * (v >> 2) + 4 * (l >> 5) + (v & 3);
    if (row >= M) continue;
    float v0 = acc0[v] + bb0, v1 = acc1[v] + bb1;
    if (relu) {
      v0 = fmaxf(v0, 0.0f);
      v1 = fmaxf(v1, 0.0f);
    }
    const size_t o = (size_t)row * ldo + col0;
    if (out_bf) {
      out_bf[o + c] = f2bf(v0);
      out_bf[o + 32 + c] = f2bf(v1);
    }
    if (out_f) {
      out_f[o + c] = v0;
      out_f[o + 32 + c] = v1;
    }
  }
}

// X[i] = [h_i | mean_{j in N(i)} h_j] (BF16, 2 kH per row); one wave per
// vertex, lane c owns feature c; the neighbour sum is FP32 in column order.
__global__ void k_nn_aggregate(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ cols,
                               const uint16_t* __restrict__ h, uint16_t* __restrict__ X) {
  const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int c = threadIdx.x & 63;
  if (i >= n) return;
  float s = 0.0f;
  int deg = 0;
  for (int b = rp[i]; b < rp[i + 1]; ++b) {
    const int j = cols[b];
    if (j == i) continue;
    s += bf2f(h[(size_t)j * kH + c]);
    ++deg;
  }
  X[(size_t)i * 2 * kH + c] = h[(size_t)i * kH + c];
  X[(size_t)i * 2 * kH + kH + c] = f2bf(deg > 0 ? s / (float)deg : 0.0f);
}

// Output layer and preconditioner blocks.  o = h4 Wo + bo in FP32 (Wo, bo
// already rounded to BF16 values); P_i = exp(s) R D^-1 R^T; y_i = P_i b_i.
// o is kept for the backward pass.
__device__ inline void precond_block(const float* o, const float* M, float* P) {
  float R[9];
  for (int e = 0; e < 9; ++e) R[e] = ((e % 4) == 0 ? 1.0f : 0.0f) + kDeltaScale * tanhf(o[e]);
  const float sc = expf(kLogScale * tanhf(o[9]));
  float RM[9];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) RM[3 * a + b] = R[3 * a] * M[b] + R[3 * a + 1] * M[3 + b] + R[3 * a + 2] * M[6 + b];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      P[3 * a + b] = sc * (RM[3 * a] * R[3 * b] + RM[3 * a + 1] * R[3 * b + 1] + RM[3 * a + 2] * R[3 * b + 2]);
}

__global__ void k_nn_output(int n, const uint16_t* __restrict__ h, const float* __restrict__ Wo,
                            const float* __restrict__ bo, const float* __restrict__ Dinv, const float* __restrict__ rhs,
                            int negate, float* __restrict__ o_out, float* __restrict__ P_out, float* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float o[kOut];
  for (int k = 0; k < kOut; ++k) o[k] = bo[k];
  for (int c = 0; c < kH; ++c) {
    const float hv = bf2f(h[(size_t)i * kH + c]);
    for (int k = 0; k < kOut; ++k) o[k] += hv * Wo[c * kOut + k];
  }
  if (o_out)
    for (int k = 0; k < kOut; ++k) o_out[(size_t)i * kOut + k] = o[k];
  float P[9];
  precond_block(o, Dinv + 9 * (size_t)i, P);
  for (int e = 0; e < 9; ++e) P_out[9 * (size_t)i + e] = P[e];
  if (y) {
    float b[3];
    for (int k = 0; k < 3; ++k) b[k] = negate ? -rhs[3 * i + k] : rhs[3 * i + k];
    for (int k = 0; k < 3; ++k) y[3 * i + k] = P[3 * k] * b[0] + (P[3 * k + 1] * b[1] + P[3 * k + 2] * b[2]);
  }
}

// --- loss and backward --------------------------------------------------------------

// Per vertex q_i = (A y)_i, and the per-workgroup partials of y^T A y,
// b^T y and b^T D^-1 b.
__global__ void k_nn_resid(int n, const float* __restrict__ A, const int32_t* __restrict__ rp,
                           const int32_t* __restrict__ cols, const float* __restrict__ y, const float* __restrict__ rhs,
                           int negate, const float* __restrict__ Dinv, float* __restrict__ q,
                           double* __restrict__ part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double vy = 0.0, uy = 0.0, nb = 0.0;
  if (i < n) {
    float Ay[3] = {0.0f, 0.0f, 0.0f};
    for (int b = rp[i]; b < rp[i + 1]; ++b) {
      const int j = cols[b];
      const float* B = A + 9 * (size_t)b;
      for (int r = 0; r < 3; ++r) Ay[r] += B[3 * r] * y[3 * j] + (B[3 * r + 1] * y[3 * j + 1] + B[3 * r + 2] * y[3 * j + 2]);
    }
    float bv[3];
    for (int k = 0; k < 3; ++k) bv[k] = negate ? -rhs[3 * i + k] : rhs[3 * i + k];
    const float* M = Dinv + 9 * (size_t)i;
    for (int k = 0; k < 3; ++k) {
      q[3 * i + k] = Ay[k];
      vy += (double)y[3 * i + k] * Ay[k];
      uy += (double)bv[k] * y[3 * i + k];
      nb += (double)bv[k] * (double)(M[3 * k] * bv[0] + M[3 * k + 1] * bv[1] + M[3 * k + 2] * bv[2]);
    }
  }
  __shared__ double lds[4];
  const double sv = block_reduce(vy, lds, [](double a, double b) { return a + b; });
  const double su = block_reduce(uy, lds, [](double a, double b) { return a + b; });
  const double sn = block_reduce(nb, lds, [](double a, double b) { return a + b; });
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = sv;
    part[3 * blockIdx.x + 1] = su;
    part[3 * blockIdx.x + 2] = sn;
  }
}

// Loss, step and gradient coefficient from the k_nn_resid partials (fixed
// order): out[0] = 2 u / (v N), out[1] = L = -u^2 / (v N), out[2] = N,
// out[3] = alpha = u / v, with u = b^T y, v = y^T A y, N = b^T D^-1 b.  A
// non-positive v (y = 0, or A not positive along y) gives alpha = 0 and a zero
// gradient.
__global__ void k_nn_loss_finalize(int groups, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double lds[4];
  double v = 0.0, u = 0.0, nb = 0.0;
  for (int g = threadIdx.x; g < groups; g += blockDim.x) {
    v += part[3 * g];
    u += part[3 * g + 1];
    nb += part[3 * g + 2];
  }
  const double sv = block_reduce(v, lds, [](double a, double b) { return a + b; });
  const double su = block_reduce(u, lds, [](double a, double b) { return a + b; });
  const double sn = block_reduce(nb, lds, [](double a, double b) { return a + b; });
  if (threadIdx.x == 0) {
    const double N = sn > 1e-300 ? sn : 1.0;
    const bool ok = sv > 1e-300;
    out[0] = ok ? 2.0 * su / (sv * N) : 0.0;
    out[1] = ok ? -su * su / (sv * N) : 0.0;
    out[2] = sn;
    out[3] = ok ? su / sv : 0.0;
  }
}

// Warm start x0 = alpha y in place (alpha = coef[3]).
__global__ void k_nn_scale_warm_start(int count, const double* __restrict__ coef, float* __restrict__ y) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) y[t] = (float)((double)y[t] * coef[3]);
}

// Output backward, per vertex: do = dL/do, dh4 = do Wo^T (FP32), and the
// per-workgroup partial dWo = h4^T do, dbo = sum do (kH * kOut + kOut values).
// With coef from k_nn_loss_finalize and q = A y: gy = coef[0] (coef[3] q - b).  kOutBwdBlock threads per workgroup.
constexpr int kOutBwdBlock = 128;
__global__ void __launch_bounds__(kOutBwdBlock) k_nn_output_bwd(int n, const uint16_t* __restrict__ h,
                                                       const float* __restrict__ Wo, const float* __restrict__ o_in,
                                                       const float* __restrict__ Dinv, const float* __restrict__ rhs,
                                                       int negate, const float* __restrict__ q, const double* __restrict__ coef,
                                                       float* __restrict__ dh, float* __restrict__ part) {
  __shared__ float hs[kOutBwdBlock][kH + 1];
  __shared__ float ds[kOutBwdBlock][kOut];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double c0 = coef[0], alpha = coef[3];
  float dout[kOut];
  for (int k = 0; k < kOut; ++k) dout[k] = 0.0f;
  if (i < n) {
    float o[kOut];
    for (int k = 0; k < kOut; ++k) o[k] = o_in[(size_t)i * kOut + k];
    const float* M = Dinv + 9 * (size_t)i;
    float R[9], th[9];
    for (int e = 0; e < 9; ++e) {
      th[e] = tanhf(o[e]);
      R[e] = ((e % 4) == 0 ? 1.0f : 0.0f) + kDeltaScale * th[e];
    }
    const float ts = tanhf(o[9]);
    const float sc = expf(kLogScale * ts);
    float b[3], gy[3];
    for (int k = 0; k < 3; ++k) {
      b[k] = negate ? -rhs[3 * i + k] : rhs[3 * i + k];
      gy[k] = (float)(c0 * (alpha * (double)q[3 * i + k] - (double)b[k]));
    }
    float G[9];  // dL/dP = gy b^T
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) G[3 * a + c] = gy[a] * b[c];
    float P[9];
    precond_block(o, M, P);
    float dls = 0.0f;
    for (int e = 0; e < 9; ++e) dls += G[e] * P[e];
    // dL/dR = sc (G R M^T + G^T R M)
    float RM[9], RMt[9];
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) {
        RM[3 * a + c] = R[3 * a] * M[c] + R[3 * a + 1] * M[3 + c] + R[3 * a + 2] * M[6 + c];
        RMt[3 * a + c] = R[3 * a] * M[3 * c] + R[3 * a + 1] * M[3 * c + 1] + R[3 * a + 2] * M[3 * c + 2];
      }
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) {
        float v = 0.0f;
        for (int k = 0; k < 3; ++k) v += G[3 * a + k] * RMt[3 * k + c] + G[3 * k + a] * RM[3 * k + c];
        dout[3 * a + c] = sc * v * kDeltaScale * (1.0f - th[3 * a + c] * th[3 * a + c]);
      }
    dout[9] = dls * kLogScale * (1.0f - ts * ts);
    for (int c = 0; c < kH; ++c) {
      float v = 0.0f;
      for (int k = 0; k < kOut; ++k) v += dout[k] * Wo[c * kOut + k];
      dh[(size_t)i * kH + c] = v;
    }
  }
  for (int c = 0; c < kH; ++c) hs[threadIdx.x][c] = i < n ? bf2f(h[(size_t)i * kH + c]) : 0.0f;
  for (int k = 0; k < kOut; ++k) ds[threadIdx.x][k] = dout[k];
  __syncthreads();
  for (int t = threadIdx.x; t < kH * kOut + kOut; t += blockDim.x) {
    float s = 0.0f;
    if (t < kH * kOut) {
      const int c = t / kOut, k = t % kOut;
      for (int r = 0; r < kOutBwdBlock; ++r) s += hs[r][c] * ds[r][k];
    } else {
      const int k = t - kH * kOut;
      for (int r = 0; r < kOutBwdBlock; ++r) s += ds[r][k];
    }
    part[(size_t)blockIdx.x * (kH * kOut + kOut) + t] = s;
  }
}

// dZ = dh * (h > 0), rounded to BF16 for the MFMA products.
__global__ void k_nn_relu_bwd(int count, const float* __restrict__ dh, const uint16_t* __restrict__ h,
                              uint16_t* __restrict__ dz) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  dz[t] = f2bf(bf2f(h[t]) > 0.0f ? dh[t] : 0.0f);
}

// Weight gradient partials: part[g] = X[rows of g]^T dZ[rows of g] (Kin x kH)
// followed by the column sums of dZ (kH), for the workgroup's chunk of
// vertices.  Wave w owns input features 32 w .. 32 w + 31 (Kin / 32 waves
// active); MFMA with A = X^T (features x vertices), B = dZ (vertices x kH).
__global__ void __launch_bounds__(256) k_nn_wgrad(int n, int Kin, const uint16_t* __restrict__ X, int ldx,
                                                  const uint16_t* __restrict__ dz, int chunk, float* __restrict__ part) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int start = blockIdx.x * chunk, end = min(n, start + chunk);
  float* out = part + (size_t)blockIdx.x * (Kin * kH + kH);
  if (threadIdx.x < kH) {
    float s = 0.0f;
    for (int k = start; k < end; ++k) s += bf2f(dz[(size_t)k * kH + threadIdx.x]);
    out[Kin * kH + threadIdx.x] = s;
  }
  if (32 * wave >= Kin) return;
  const int fi = 32 * wave + (l & 31);
  f16v acc0 = {}, acc1 = {};
  for (int k0 = start; k0 < end; k0 += 8) {
    bf4 a, b0, b1;
    for (int q = 0; q < 4; ++q) {
      const int k = k0 + 4 * (l >> 5) + q;
      const bool ok = k < end;
      a[q] = ok ? (short)X[(size_t)k * ldx + fi] : (short)0;
      b0[q] = ok ? (short)dz[(size_t)k * kH + (l & 31)] : (short)0;
      b1[q] = ok ? (short)dz[(size_t)k * kH + 32 + (l & 31)] : (short)0;
    }
    acc0 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b0, acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b1, acc1, 0, 0, 0);
  }
  const int c = l & 31;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int r = 32 * wave + 8 * (v >> 2) + 4 * (l >> 5) + (v & 3);
    out[(size_t)r * kH + c] = acc0[v];
    out[(size_t)r * kH + 32 + c] = acc1[v];
  }
}

// Sum of per-workgroup partials in workgroup order: out[t] = sum_g part[g][t].
__global__ void k_nn_sum_parts(int groups, int size, const float* __restrict__ part, float* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= size) return;
  float s = 0.0f;
  for (int g = 0; g < groups; ++g) s += part[(size_t)g * size + t];
  out[t] = s;
}

// dh_prev[i] = dX[i][0:kH] + sum_{j in N(i)} dX[j][kH:2kH] / deg_j  (the
// neighbour graph is symmetric, so the transpose of the mean is a gather).
__global__ void k_nn_aggregate_bwd(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ cols,
                                   const float* __restrict__ dX, float* __restrict__ dh) {
  const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int c = threadIdx.x & 63;
  if (i >= n) return;
  float s = dX[(size_t)i * 2 * kH + c];
  for (int b = rp[i]; b < rp[i + 1]; ++b) {
    const int j = cols[b];
    if (j == i) continue;
    const int degj = rp[j + 1] - rp[j] - 1;
    if (degj > 0) s += dX[(size_t)j * 2 * kH + kH + c] / (float)degj;
  }
  dh[(size_t)i * kH + c] = s;
}

// Adam on the FP32 master weights (bias-corrected step sizes from the host).
__global__ void k_nn_adam(int count, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                          float* __restrict__ v, float lr_t, float b1, float b2, float eps) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.0f - b1) * gt;
  const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  w[t] -= lr_t * mt / (sqrtf(vt) + eps);
}

}  // namespace nn
}  // namespace ando
