// Host side of the coarse level: cell assignment (deterministic counting sort),
// buffers, the build sequence and the three applications.
#include "coarse.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>

#include "kernels.hpp"

namespace lm {

void coarse_cells(const double* xyz, int64_t stride, int64_t N, int ppc, CoarseCells& out) {
  if (N <= 0) throw Error("[libmobility_amd] coarse level: no particles");
  double lo[3], hi[3], ext[3];
  for (int k = 0; k < 3; ++k) lo[k] = hi[k] = xyz[k];
  for (int64_t i = 1; i < N; ++i)
    for (int k = 0; k < 3; ++k) {
      const double x = xyz[i * stride + k];
      lo[k] = std::min(lo[k], x);
      hi[k] = std::max(hi[k], x);
    }
  double vol = 1.0;
  for (int k = 0; k < 3; ++k) {
    ext[k] = hi[k] - lo[k];
    vol *= std::max(ext[k], 1.0);
  }
  const double want = std::max(1.0, (double)N / (double)std::max(ppc, 1));
  double side = std::cbrt(vol / want);
  std::vector<int64_t> count;
  for (;;) {
    int64_t total = 1;
    for (int k = 0; k < 3; ++k) {
      out.g[k] = (int)std::max<long long>(1, std::llround(std::max(ext[k], 1.0) / side));
      total *= out.g[k];
    }
    out.cell_of.assign(N, 0);
    for (int64_t i = 0; i < N; ++i) {
      int q[3];
      for (int k = 0; k < 3; ++k) {
        q[k] = 0;
        if (ext[k] > 0.0) {
          const double f = std::floor((xyz[i * stride + k] - lo[k]) / ext[k] * out.g[k]);
          q[k] = (int)std::min<double>(out.g[k] - 1, std::max(0.0, f));
        }
      }
      out.cell_of[i] = (q[0] * out.g[1] + q[1]) * out.g[2] + q[2];
    }
    count.assign(total, 0);
    for (int64_t i = 0; i < N; ++i) ++count[out.cell_of[i]];
    int C = 0;
    for (int64_t c = 0; c < total; ++c) C += count[c] > 0;
    if (3 * (int64_t)C <= CoarseLevel::kMaxDim) break;
    side *= 1.03;  // too many cells: coarser grid
  }
  // compact the non-empty cells in raw order, then a stable counting sort
  std::vector<int> compact(count.size(), -1);
  int C = 0;
  for (size_t c = 0; c < count.size(); ++c)
    if (count[c] > 0) compact[c] = C++;
  out.C = C;
  out.start.assign(C + 1, 0);
  for (int64_t i = 0; i < N; ++i) {
    out.cell_of[i] = compact[out.cell_of[i]];
    ++out.start[out.cell_of[i] + 1];
  }
  for (int c = 0; c < C; ++c) out.start[c + 1] += out.start[c];
  out.perm.assign(N, 0);
  std::vector<int> fill(out.start.begin(), out.start.end() - 1);
  for (int64_t i = 0; i < N; ++i) out.perm[fill[out.cell_of[i]]++] = (int)i;
}

void CoarseLevel::build(const void* rec64, const void* rec32, int64_t N, bool wall, int ppc_req,
                        hipStream_t s) {
  const auto t_start = std::chrono::steady_clock::now();
  if (N > INT32_MAX) throw Error("[libmobility_amd] coarse level: too many particles");
  std::vector<double> h((size_t)4 * N);
  LM_HIP(hipMemcpyAsync(h.data(), rec64, sizeof(double) * 4 * N, hipMemcpyDeviceToHost, s));
  LM_HIP(hipStreamSynchronize(s));
  coarse_cells(h.data(), 4, N, ppc_req, cells_);
  N_ = N;
  const int C = cells_.C, n3c = 3 * C;
  n_ = (n3c + kCoarseTile - 1) / kCoarseTile * kCoarseTile;
  std::vector<double> scale(C);
  for (int c = 0; c < C; ++c) scale[c] = 1.0 / std::sqrt((double)(cells_.start[c + 1] - cells_.start[c]));
  perm_.reserve(sizeof(int) * N);
  cell_of_.reserve(sizeof(int) * N);
  start_.reserve(sizeof(int) * (C + 1));
  scale_.reserve(sizeof(double) * C);
  LM_HIP(hipMemcpyAsync(perm_.as<int>(), cells_.perm.data(), sizeof(int) * N,
                        hipMemcpyHostToDevice, s));
  LM_HIP(hipMemcpyAsync(cell_of_.as<int>(), cells_.cell_of.data(), sizeof(int) * N,
                        hipMemcpyHostToDevice, s));
  LM_HIP(hipMemcpyAsync(start_.as<int>(), cells_.start.data(), sizeof(int) * (C + 1),
                        hipMemcpyHostToDevice, s));
  LM_HIP(hipMemcpyAsync(scale_.as<double>(), scale.data(), sizeof(double) * C,
                        hipMemcpyHostToDevice, s));
  sorted_.reserve(sizeof(float4) * N);
  launch_coarse_gather(static_cast<const float4*>(rec32), perm_.as<int>(), N, sorted_.as<float4>(),
                       s);
  const size_t mat = sizeof(double) * (size_t)n_ * n_;
  A_.reserve(mat);
  X_.reserve(mat);
  Xt_.reserve(mat);
  Dinv_.reserve(sizeof(double) * (size_t)n_ * kCoarseTile);
  info_.reserve(sizeof(unsigned int));
  cvec_.reserve(sizeof(double) * n_);
  yvec_.reserve(sizeof(double) * n_);
  launch_zero_u32(info_.as<unsigned int>(), 1, s);
  launch_coarse_matrix(sorted_.as<float4>(), start_.as<int>(), scale_.as<double>(), C, wall,
                       A_.as<double>(), n_, n3c, n_, s);
  launch_coarse_factor(A_.as<double>(), X_.as<double>(), Xt_.as<double>(), Dinv_.as<double>(),
                       info_.as<int>(), n_, n_, s);
  int info = 0;
  LM_HIP(hipMemcpyAsync(&info, info_.as<int>(), sizeof(int), hipMemcpyDeviceToHost, s));
  LM_HIP(hipStreamSynchronize(s));
  if (info != 0) throw Error("[libmobility_amd] coarse level: the coarse matrix is not positive definite");
  build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start)
                 .count();
}

void CoarseLevel::apply(Op op, const double* in, double* out, int64_t n, hipStream_t s) {
  if (cells_.C == 0) throw Error("[libmobility_amd] coarse level not built");
  if (n < 3 * N_) throw Error("[libmobility_amd] coarse level: vector too short");
  const double* T = op == kG ? X_.as<double>() : (op == kGt ? Xt_.as<double>() : A_.as<double>());
  launch_coarse_apply(T, n_, 3 * cells_.C, op != kGt, in, out, N_, n, perm_.as<int>(),
                      start_.as<int>(), cell_of_.as<int>(), scale_.as<double>(), cells_.C,
                      cvec_.as<double>(), yvec_.as<double>(), s);
}

void CoarseLevel::factors(double* L, double* X, hipStream_t s) const {
  const int d = 3 * cells_.C;
  if (d == 0) throw Error("[libmobility_amd] coarse level not built");
  const size_t pitch = sizeof(double) * n_;
  LM_HIP(hipMemcpy2DAsync(L, sizeof(double) * d, A_.as<double>(), pitch, sizeof(double) * d, d,
                          hipMemcpyDeviceToHost, s));
  LM_HIP(hipMemcpy2DAsync(X, sizeof(double) * d, X_.as<double>(), pitch, sizeof(double) * d, d,
                          hipMemcpyDeviceToHost, s));
  LM_HIP(hipStreamSynchronize(s));
  for (int r = 0; r < d; ++r)
    for (int c = r + 1; c < d; ++c) L[(size_t)r * d + c] = 0.0;
}

}  // namespace lm
