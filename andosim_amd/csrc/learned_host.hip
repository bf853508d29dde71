// Host side of the learned preconditioner (kernels in learned.hip): weight
// storage and packing, forward pass, loss and backward pass, Adam.  Included
// by ctx.hip after DBuf / HIPCHK / nblk.

namespace nnh {

using namespace ando::nn;

inline int layer_off(int l) { return kLayer0 + l * kLayerSize; }  // l = 0 .. kL-1

// Network weights on the device: FP32 master copy, its BF16-rounded values
// (used by the FP32 output layer and as epilogue biases), Adam moments,
// gradient, and the MFMA-packed BF16 matrices (forward, and transposed halves
// for the input-gradient products).
struct Net {
  DBuf<float> w, wr, m, v, grad;
  DBuf<uint16_t> wp0, wpl[kL], wpt[kL][2];
  int adam_t = 0;
  bool ready = false;
};

// Per-graph activations and scratch (grown on demand).
struct Work {
  int cap = 0;
  DBuf<uint16_t> feat, h[kL + 1], X[kL], dz;
  DBuf<float> Dinv, o, P, y, q, dh, dX, part;
  DBuf<double> lpart, lfin;
  void ensure(int n) {
    if (n <= cap) return;
    const size_t N = (size_t)n;
    feat.ensure(N * kF);
    for (auto& b : h) b.ensure(N * kH);
    for (auto& b : X) b.ensure(N * 2 * kH);
    dz.ensure(N * kH);
    Dinv.ensure(9 * N);
    o.ensure(N * kOut);
    P.ensure(9 * N);
    y.ensure(3 * N);
    q.ensure(3 * N);
    dh.ensure(N * kH);
    dX.ensure(N * 2 * kH);
    const size_t gw = (size_t)nblk(n, 256), go = (size_t)nblk(n, kOutBwdBlock);
    part.ensure(std::max(gw * (size_t)kLayerSize, go * (size_t)(kH * kOut + kOut)));
    lpart.ensure(3 * (size_t)nblk(n) + 3);
    lfin.ensure(4);
    cap = n;
  }
};

inline void pack(Net& net, hipStream_t s) {
  auto launch_pack = [&](int K, const float* W, int ldw, int tr, uint16_t* out) {
    hipLaunchKernelGGL(k_nn_pack, dim3(nblk((long long)K * kH)), dim3(256), 0, s, K, W, ldw, tr, out);
  };
  launch_pack(kF, net.w.p + kW0, kH, 0, net.wp0.p);
  for (int l = 0; l < kL; ++l) {
    const float* W = net.w.p + layer_off(l);
    launch_pack(2 * kH, W, kH, 0, net.wpl[l].p);
    launch_pack(kH, W, kH, 1, net.wpt[l][0].p);
    launch_pack(kH, W + kH * kH, kH, 1, net.wpt[l][1].p);
  }
  hipLaunchKernelGGL(k_nn_round, dim3(nblk(kNumWeights)), dim3(256), 0, s, kNumWeights, (const float*)net.w.p,
                     net.wr.p);
  HIPCHK(hipGetLastError());
}

inline void upload(Net& net, const float* w, hipStream_t s) {
  net.w.upload(w, kNumWeights, s);
  net.wr.ensure(kNumWeights);
  net.grad.ensure(kNumWeights);
  net.m.ensure(kNumWeights);
  net.v.ensure(kNumWeights);
  HIPCHK(hipMemsetAsync(net.m.p, 0, kNumWeights * sizeof(float), s));
  HIPCHK(hipMemsetAsync(net.v.p, 0, kNumWeights * sizeof(float), s));
  net.wp0.ensure((size_t)kF * kH);
  for (int l = 0; l < kL; ++l) {
    net.wpl[l].ensure((size_t)2 * kH * kH);
    net.wpt[l][0].ensure((size_t)kH * kH);
    net.wpt[l][1].ensure((size_t)kH * kH);
  }
  net.adam_t = 0;
  pack(net, s);
  net.ready = true;
}

// He-normal hidden layers (SplitMix64 + Box-Muller), zero biases, zero output
// layer: the untrained network gives P = D^-1 exactly (block-Jacobi).
inline std::vector<float> init_weights(uint64_t seed) {
  std::vector<float> w(kNumWeights, 0.0f);
  uint64_t st = seed;
  auto next = [&]() {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  auto unif = [&]() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); };
  auto gauss = [&]() {
    const double u1 = unif(), u2 = unif();  // drawn in this order
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  };
  for (int k = 0; k < kFUsed; ++k)
    for (int j = 0; j < kH; ++j) w[kW0 + k * kH + j] = (float)(gauss() * std::sqrt(2.0 / kFUsed));
  for (int l = 0; l < kL; ++l)
    for (int k = 0; k < 2 * kH * kH; ++k) w[layer_off(l) + k] = (float)(gauss() * std::sqrt(2.0 / (2 * kH)));
  return w;
}

inline void gemm(int M, int K, const uint16_t* A, int lda, const uint16_t* Wp, const float* bias, int relu,
                 uint16_t* out_bf, float* out_f, int ldo, int col0, hipStream_t s) {
  hipLaunchKernelGGL(k_nn_gemm, dim3(nblk(M, 128)), dim3(256), 0, s, M, K, A, lda, Wp, bias, relu, out_bf, out_f,
                     ldo, col0);
}

// Forward pass: features, five GEMM layers, output blocks P and warm start y.
inline void forward(Net& net, Work& wk, int n, const float* A, const int32_t* rp, const int32_t* cols,
                    const int32_t* diag, const float* rhs, int negate, bool keep_o, hipStream_t s) {
  wk.ensure(n);
  hipLaunchKernelGGL(k_nn_features, dim3(nblk(n)), dim3(256), 0, s, n, A, rp, cols, diag, wk.feat.p, wk.Dinv.p);
  gemm(n, kF, wk.feat.p, kF, net.wp0.p, net.wr.p + kB0, 1, wk.h[0].p, nullptr, kH, 0, s);
  for (int l = 0; l < kL; ++l) {
    hipLaunchKernelGGL(k_nn_aggregate, dim3(nblk(n, 4)), dim3(256), 0, s, n, rp, cols, (const uint16_t*)wk.h[l].p,
                       wk.X[l].p);
    gemm(n, 2 * kH, wk.X[l].p, 2 * kH, net.wpl[l].p, net.wr.p + layer_off(l) + 2 * kH * kH, 1, wk.h[l + 1].p,
         nullptr, kH, 0, s);
  }
  hipLaunchKernelGGL(k_nn_output, dim3(nblk(n)), dim3(256), 0, s, n, (const uint16_t*)wk.h[kL].p,
                     (const float*)(net.wr.p + kWo), (const float*)(net.wr.p + kBo), (const float*)wk.Dinv.p, rhs,
                     negate, keep_o ? wk.o.p : nullptr, wk.P.p, wk.y.p);
  HIPCHK(hipGetLastError());
}

// q = A y and the loss coefficients (wk.lfin) for y = P b from forward().
inline void energy(Work& wk, int n, const float* A, const int32_t* rp, const int32_t* cols, const float* rhs,
                   int negate, hipStream_t s) {
  const int G = nblk(n);
  hipLaunchKernelGGL(k_nn_resid, dim3(G), dim3(256), 0, s, n, A, rp, cols, (const float*)wk.y.p, rhs, negate,
                     (const float*)wk.Dinv.p, wk.q.p, wk.lpart.p);
  hipLaunchKernelGGL(k_nn_loss_finalize, dim3(1), dim3(256), 0, s, G, (const double*)wk.lpart.p, wk.lfin.p);
}

// The PCG warm start: wk.y becomes alpha P b (forward() first).
inline void warm_start(Work& wk, int n, const float* A, const int32_t* rp, const int32_t* cols, const float* rhs,
                       int negate, hipStream_t s) {
  energy(wk, n, A, rp, cols, rhs, negate, s);
  hipLaunchKernelGGL(k_nn_scale_warm_start, dim3(nblk(3LL * n)), dim3(256), 0, s, 3 * n, (const double*)wk.lfin.p,
                     wk.y.p);
  HIPCHK(hipGetLastError());
}

// Loss (into wk.lfin[1]) and gradient (into net.grad); forward(keep_o) first.
inline void backward(Net& net, Work& wk, int n, const float* A, const int32_t* rp, const int32_t* cols,
                     const float* rhs, int negate, hipStream_t s) {
  energy(wk, n, A, rp, cols, rhs, negate, s);
  const int go = nblk(n, kOutBwdBlock);
  hipLaunchKernelGGL(k_nn_output_bwd, dim3(go), dim3(kOutBwdBlock), 0, s, n, (const uint16_t*)wk.h[kL].p,
                     (const float*)(net.wr.p + kWo), (const float*)wk.o.p, (const float*)wk.Dinv.p, rhs, negate,
                     (const float*)wk.q.p, (const double*)wk.lfin.p, wk.dh.p, wk.part.p);
  const int osz = kH * kOut + kOut;
  hipLaunchKernelGGL(k_nn_sum_parts, dim3(nblk(osz)), dim3(256), 0, s, go, osz, (const float*)wk.part.p,
                     net.grad.p + kWo);
  const int gw = nblk(n, 256);
  for (int l = kL; l >= 0; --l) {
    hipLaunchKernelGGL(k_nn_relu_bwd, dim3(nblk((long long)n * kH)), dim3(256), 0, s, n * kH,
                       (const float*)wk.dh.p, (const uint16_t*)wk.h[l].p, wk.dz.p);
    const uint16_t* Xin = l > 0 ? wk.X[l - 1].p : wk.feat.p;
    const int Kin = l > 0 ? 2 * kH : kF;
    hipLaunchKernelGGL(k_nn_wgrad, dim3(gw), dim3(256), 0, s, n, Kin, Xin, Kin, (const uint16_t*)wk.dz.p, 256,
                       wk.part.p);
    const int psz = Kin * kH + kH;
    float* gdst = net.grad.p + (l > 0 ? layer_off(l - 1) : kW0);
    hipLaunchKernelGGL(k_nn_sum_parts, dim3(nblk(psz)), dim3(256), 0, s, gw, psz, (const float*)wk.part.p, gdst);
    if (l == 0) break;
    gemm(n, kH, wk.dz.p, kH, net.wpt[l - 1][0].p, nullptr, 0, nullptr, wk.dX.p, 2 * kH, 0, s);
    gemm(n, kH, wk.dz.p, kH, net.wpt[l - 1][1].p, nullptr, 0, nullptr, wk.dX.p, 2 * kH, kH, s);
    hipLaunchKernelGGL(k_nn_aggregate_bwd, dim3(nblk(n, 4)), dim3(256), 0, s, n, rp, cols, (const float*)wk.dX.p,
                       wk.dh.p);
  }
  HIPCHK(hipGetLastError());
}

inline void adam(Net& net, float lr, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  ++net.adam_t;
  const double c1 = 1.0 - std::pow((double)b1, net.adam_t), c2 = 1.0 - std::pow((double)b2, net.adam_t);
  const float lr_t = (float)(lr * std::sqrt(c2) / c1);
  hipLaunchKernelGGL(k_nn_adam, dim3(nblk(kNumWeights)), dim3(256), 0, s, kNumWeights, net.w.p,
                     (const float*)net.grad.p, net.m.p, net.v.p, lr_t, b1, b2, eps);
  pack(net, s);
}

}  // namespace nnh
