// Host side of the learned block preconditioner: weight storage, forward and
// backward passes as sequences of the BF16-MFMA GEMM and elementwise kernels,
// mini-batch Adam training.
#include "precond.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <vector>

namespace lm {

LearnedPreconditioner::LearnedPreconditioner(int kind) : kind_(kind), layout_(precond_layout(kind)) {
  if (kind != 0 && kind != 1) throw Error("[libmobility_amd] block size must be 3 or 6");
  w_.reserve(sizeof(float) * layout_.total);
  g_.reserve(sizeof(float) * layout_.total);
  m_.reserve(sizeof(float) * layout_.total);
  v_.reserve(sizeof(float) * layout_.total);
  initWeights(1);
}

void LearnedPreconditioner::initWeights(uint64_t seed) {
  // He-style normal init for the hidden layers, small last layer; biases 0,
  // except the diagonal outputs, started at softplus^{-1}(1.3) (the typical
  // diagonal of the target blocks).
  std::mt19937_64 rng(seed);
  std::vector<float> w(layout_.total, 0.f);
  auto fill = [&](int64_t off, int fan_in, int fan_out, double scale) {
    std::normal_distribution<double> nd(0.0, scale / std::sqrt((double)fan_in));
    for (int64_t k = 0; k < (int64_t)fan_in * fan_out; ++k) w[off + k] = (float)nd(rng);
  };
  fill(layout_.W1, kPrecondIn, kPrecondH, std::sqrt(2.0));
  fill(layout_.W2, kPrecondH, kPrecondH, std::sqrt(2.0));
  fill(layout_.W3, kPrecondH, outputs(), 0.1);
  const float sp_inv = (float)std::log(std::expm1(1.3));
  for (int d = 0; d < blockDim(); ++d) w[layout_.B3 + d] = sp_inv;
  setWeights(w.data(), layout_.total);
}

void LearnedPreconditioner::setWeights(const float* w, int64_t n) {
  if (n != layout_.total) throw Error("[libmobility_amd] wrong number of preconditioner weights");
  LM_HIP(hipMemcpy(w_.as<float>(), w, sizeof(float) * n, hipMemcpyHostToDevice));
  LM_HIP(hipMemset(m_.as<void>(), 0, sizeof(float) * n));
  LM_HIP(hipMemset(v_.as<void>(), 0, sizeof(float) * n));
  step_ = 0;
  ++version_;
}

void LearnedPreconditioner::getWeights(float* w, int64_t n) {
  if (n != layout_.total) throw Error("[libmobility_amd] wrong number of preconditioner weights");
  LM_HIP(hipMemcpy(w, w_.as<float>(), sizeof(float) * n, hipMemcpyDeviceToHost));
}

void LearnedPreconditioner::ensure_batch(int Bn) {
  if (Bn <= cap_) return;
  const size_t h = sizeof(float) * (size_t)Bn * kPrecondH;
  z1_.reserve(h);
  h1_.reserve(h);
  z2_.reserve(h);
  h2_.reserve(h);
  dh_.reserve(h);
  dz_.reserve(h);
  out_.reserve(sizeof(float) * (size_t)Bn * kPrecondMaxOut);
  dout_.reserve(sizeof(float) * (size_t)Bn * kPrecondMaxOut);
  loss_part_.reserve(sizeof(float) * ((size_t)Bn / 256 + 2));
  cap_ = Bn;
}

void LearnedPreconditioner::forward(const float* X, int Bn, float* out, hipStream_t s) {
  ensure_batch(Bn);
  const float* w = w_.as<float>();
  const int no = outputs();
  const PrecondLayout& L = layout_;
  launch_gemm_bf16(Bn, kPrecondH, kPrecondIn, X, kPrecondIn, 1, w + L.W1, kPrecondH, 1,
                   h1_.as<float>(), kPrecondH, w + L.B1, 1, z1_.as<float>(), 0, s);
  launch_gemm_bf16(Bn, kPrecondH, kPrecondH, h1_.as<float>(), kPrecondH, 1, w + L.W2, kPrecondH, 1,
                   h2_.as<float>(), kPrecondH, w + L.B2, 1, z2_.as<float>(), 0, s);
  launch_gemm_bf16(Bn, no, kPrecondH, h2_.as<float>(), kPrecondH, 1, w + L.W3, no, 1, out, no,
                   w + L.B3, 0, nullptr, 0, s);
}

float LearnedPreconditioner::lossGrad(const float* X, const float* T, int Bn, hipStream_t s) {
  ensure_batch(Bn);
  forward(X, Bn, out_.as<float>(), s);
  const float* w = w_.as<float>();
  float* g = g_.as<float>();
  const int no = outputs();
  const PrecondLayout& L = layout_;
  launch_head_loss(out_.as<float>(), T, Bn, kind_, dout_.as<float>(), loss_part_.as<float>(), s);
  // layer 3
  launch_gemm_bf16(kPrecondH, no, Bn, h2_.as<float>(), 1, kPrecondH, dout_.as<float>(), no, 1,
                   g + L.W3, no, nullptr, 0, nullptr, 0, s);
  launch_bias_grad(dout_.as<float>(), Bn, no, g + L.B3, s);
  launch_gemm_bf16(Bn, kPrecondH, no, dout_.as<float>(), no, 1, w + L.W3, 1, no, dh_.as<float>(),
                   kPrecondH, nullptr, 0, nullptr, 0, s);
  launch_gelu_back(dh_.as<float>(), z2_.as<float>(), (int64_t)Bn * kPrecondH, dz_.as<float>(), s);
  // layer 2
  launch_gemm_bf16(kPrecondH, kPrecondH, Bn, h1_.as<float>(), 1, kPrecondH, dz_.as<float>(),
                   kPrecondH, 1, g + L.W2, kPrecondH, nullptr, 0, nullptr, 0, s);
  launch_bias_grad(dz_.as<float>(), Bn, kPrecondH, g + L.B2, s);
  launch_gemm_bf16(Bn, kPrecondH, kPrecondH, dz_.as<float>(), kPrecondH, 1, w + L.W2, 1, kPrecondH,
                   dh_.as<float>(), kPrecondH, nullptr, 0, nullptr, 0, s);
  launch_gelu_back(dh_.as<float>(), z1_.as<float>(), (int64_t)Bn * kPrecondH, dz_.as<float>(), s);
  // layer 1
  launch_gemm_bf16(kPrecondIn, kPrecondH, Bn, X, 1, kPrecondIn, dz_.as<float>(), kPrecondH, 1,
                   g + L.W1, kPrecondH, nullptr, 0, nullptr, 0, s);
  launch_bias_grad(dz_.as<float>(), Bn, kPrecondH, g + L.B1, s);
  const int nparts = (Bn + 255) / 256;
  std::vector<float> parts(nparts);
  LM_HIP(hipMemcpyAsync(parts.data(), loss_part_.as<float>(), sizeof(float) * nparts,
                        hipMemcpyDeviceToHost, s));
  LM_HIP(hipStreamSynchronize(s));
  double l = 0.0;
  for (float p : parts) l += p;
  return (float)(l / Bn);
}

void LearnedPreconditioner::train(const float* X_host, const float* T_host, int64_t n, int epochs,
                                  int batch, float lr, uint64_t seed, float* loss_history,
                                  hipStream_t s) {
  if (n <= 0 || batch <= 0) throw Error("[libmobility_amd] empty training set");
  batch = (int)std::min<int64_t>(batch, n);
  const int td = blockDim() * blockDim();
  DevBuf X, T, bx, bt, idx;
  X.reserve(sizeof(float) * n * kPrecondIn);
  T.reserve(sizeof(float) * n * td);
  idx.reserve(sizeof(int64_t) * n);
  LM_HIP(hipMemcpy(X.as<float>(), X_host, sizeof(float) * n * kPrecondIn, hipMemcpyHostToDevice));
  LM_HIP(hipMemcpy(T.as<float>(), T_host, sizeof(float) * n * td, hipMemcpyHostToDevice));
  bx.reserve(sizeof(float) * batch * kPrecondIn);
  bt.reserve(sizeof(float) * batch * td);
  std::vector<int64_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::mt19937_64 rng(seed);
  const int64_t nb = n / batch;
  for (int e = 0; e < epochs; ++e) {
    std::shuffle(order.begin(), order.end(), rng);
    LM_HIP(hipMemcpyAsync(idx.as<int64_t>(), order.data(), sizeof(int64_t) * n,
                          hipMemcpyHostToDevice, s));
    double el = 0.0;
    for (int64_t b = 0; b < nb; ++b) {
      const int64_t* ib = idx.as<int64_t>() + b * batch;
      launch_gather_rows(X.as<float>(), ib, batch, kPrecondIn, bx.as<float>(), s);
      launch_gather_rows(T.as<float>(), ib, batch, td, bt.as<float>(), s);
      el += lossGrad(bx.as<float>(), bt.as<float>(), batch, s);
      ++step_;
      launch_adam(w_.as<float>(), g_.as<float>(), m_.as<float>(), v_.as<float>(), layout_.total,
                  lr, step_, s);
    }
    if (loss_history) loss_history[e] = (float)(el / std::max<int64_t>(nb, 1));
  }
  ++version_;
  LM_HIP(hipStreamSynchronize(s));
}

}  // namespace lm
