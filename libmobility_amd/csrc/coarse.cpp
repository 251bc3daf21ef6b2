// Host side of the coarse level: cell assignment (deterministic counting sort),
// buffers, the build sequence and the three applications.
#include "coarse.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>

#include "kernels.hpp"

namespace lm {

void coarse_cells(const double* xyz, int64_t stride, int64_t N, int ppc, CoarseCells& out,
                  int64_t max_dim) {
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
    if (3 * (int64_t)C <= max_dim) break;
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

SideStream::~SideStream() {
  // Teardown: nothing can act on an error here.  The owner synchronised
  // with every use of the stream.
  if (block) (void)hipEventDestroy(block);
  if (done) (void)hipEventDestroy(done);
  if (stream) (void)hipStreamDestroy(stream);
}

void SideStream::ensure() {
  if (stream) return;
  LM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  LM_HIP(hipEventCreateWithFlags(&block, hipEventDisableTiming));
  LM_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
}

void CoarseLevel::build(const void* rec64, const void* rec32, const std::vector<int64_t>& off,
                        const std::vector<int>& wall, int ppc_req, int64_t max_dim,
                        bool force_fp64, hipStream_t s) {
  const auto t_start = std::chrono::steady_clock::now();
  const int nsys = (int)off.size() - 1;
  const int64_t N = off[nsys];
  if (nsys < 1 || N <= 0) throw Error("[libmobility_amd] coarse level: no particles");
  if (N > INT32_MAX) throw Error("[libmobility_amd] coarse level: too many particles");
  std::vector<double> h((size_t)4 * N);
  LM_HIP(hipMemcpyAsync(h.data(), rec64, sizeof(double) * 4 * N, hipMemcpyDeviceToHost, s));
  LM_HIP(hipStreamSynchronize(s));
  // cells of every system, concatenated; a system's particles (and so its
  // sorted list) occupy [off[s], off[s+1])
  cells_.cell_of.assign(N, 0);
  cells_.perm.assign(N, 0);
  cells_.start.clear();
  sys_.assign(nsys, CoarseSys{0, 0, 0, 0, 0, 0});
  int C = 0;
  int64_t moff = 0, doff = 0;
  bool have_grid = false;
  for (int q = 0; q < nsys; ++q) {
    const int64_t n_q = off[q + 1] - off[q];
    CoarseSys& S = sys_[q];
    S.moff = moff;
    S.doff = doff;
    S.c0 = C;
    S.wall = wall[q];
    if (n_q == 0) continue;
    CoarseCells cq;
    coarse_cells(h.data() + 4 * off[q], 4, n_q, ppc_req, cq, max_dim);
    for (int64_t i = 0; i < n_q; ++i) {
      cells_.cell_of[off[q] + i] = C + cq.cell_of[i];
      cells_.perm[off[q] + i] = (int)(off[q] + cq.perm[i]);
    }
    for (int c = 0; c < cq.C; ++c) cells_.start.push_back((int)(off[q] + cq.start[c]));
    if (!have_grid) {
      for (int k = 0; k < 3; ++k) cells_.g[k] = cq.g[k];
      have_grid = true;
    }
    S.C = cq.C;
    S.n = (3 * cq.C + kCoarseTile - 1) / kCoarseTile * kCoarseTile;
    moff += (int64_t)S.n * S.n;
    doff += (int64_t)(S.n / kCoarseTile) * kCoarseTile * kCoarseTile;
    C += cq.C;
  }
  cells_.start.push_back((int)N);
  cells_.C = C;
  max_row_ = 0;
  for (const CoarseSys& S : sys_) max_row_ = std::max(max_row_, 3 * S.C);
  N_ = N;
  std::vector<double> scale(C);
  for (int c = 0; c < C; ++c)
    scale[c] = 1.0 / std::sqrt((double)(cells_.start[c + 1] - cells_.start[c]));
  std::vector<int> row_sys(3 * (size_t)C);
  std::vector<int64_t> pair_off(nsys + 1, 0);
  std::vector<int> small, large;
  for (int q = 0; q < nsys; ++q) {
    const CoarseSys& S = sys_[q];
    for (int r = 0; r < 3 * S.C; ++r) row_sys[3 * (size_t)S.c0 + r] = q;
    pair_off[q + 1] = pair_off[q] + (int64_t)S.C * (S.C + 1) / 2;
    if (S.C == 0) continue;
    (S.n / kCoarseTile <= kSmallTiles ? small : large).push_back(q);
  }
  if ((int)small.size() < kSmallBatch) {  // few small blocks: the per-step launches
    large.insert(large.end(), small.begin(), small.end());
    small.clear();
  }
  auto upload = [&](DevBuf& b, const void* src, size_t bytes) {
    b.reserve(bytes);
    if (bytes) LM_HIP(hipMemcpyAsync(b.as<void>(), src, bytes, hipMemcpyHostToDevice, s));
  };
  upload(perm_, cells_.perm.data(), sizeof(int) * N);
  upload(cell_of_, cells_.cell_of.data(), sizeof(int) * N);
  upload(start_, cells_.start.data(), sizeof(int) * (C + 1));
  upload(scale_, scale.data(), sizeof(double) * C);
  upload(sys_dev_, sys_.data(), sizeof(CoarseSys) * nsys);
  upload(row_sys_, row_sys.data(), sizeof(int) * row_sys.size());
  upload(pair_off_, pair_off.data(), sizeof(int64_t) * (nsys + 1));
  upload(small_, small.data(), sizeof(int) * small.size());
  A_.reserve(sizeof(double) * moff);
  X_.reserve(sizeof(double) * moff);
  Xt_.reserve(sizeof(double) * moff);
  Dinv_.reserve(sizeof(double) * doff);
  info_.reserve(sizeof(unsigned int));
  cvec_.reserve(sizeof(double) * 3 * C);
  yvec_.reserve(sizeof(double) * 3 * C);
  // Q^T M Q from FP32 pair terms; when rounding leaves it indefinite (a
  // non-positive pivot), once more from FP64 pair terms.
  int info = 0;
  for (int fp64 = force_fp64 ? 1 : 0; fp64 < 2; ++fp64) {
    sorted_.reserve((fp64 ? sizeof(double4) : sizeof(float4)) * N);
    launch_coarse_gather(fp64 ? rec64 : rec32, perm_.as<int>(), N, sorted_.as<void>(), fp64 != 0,
                         s);
    launch_zero_u32(info_.as<unsigned int>(), 1, s);
    launch_coarse_matrix(sorted_.as<void>(), fp64 != 0, start_.as<int>(), scale_.as<double>(),
                         sys_dev_.as<CoarseSys>(), pair_off_.as<int64_t>(), nsys, pair_off[nsys],
                         A_.as<double>(), moff, s);
    if (!large.empty()) side_.ensure();
    for (int q : large) {
      const CoarseSys& S = sys_[q];
      launch_coarse_factor(A_.as<double>() + S.moff, X_.as<double>() + S.moff,
                           Xt_.as<double>() + S.moff, Dinv_.as<double>() + S.doff,
                           info_.as<int>(), S.n, S.n, s, side_);
    }
    launch_coarse_factor_small(A_.as<double>(), X_.as<double>(), Xt_.as<double>(),
                               Dinv_.as<double>(), info_.as<int>(), sys_dev_.as<CoarseSys>(),
                               small_.as<int>(), (int)small.size(), s);
    LM_HIP(hipMemcpyAsync(&info, info_.as<int>(), sizeof(int), hipMemcpyDeviceToHost, s));
    LM_HIP(hipStreamSynchronize(s));
    fp64_pairs_ = fp64 != 0;
    if (info == 0) break;
  }
  if (info != 0) throw Error("[libmobility_amd] coarse level: the coarse matrix is not positive definite");
  build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start)
                 .count();
}

void CoarseLevel::apply(Op op, const double* in, double* out, int64_t n, hipStream_t s) {
  if (cells_.C == 0) throw Error("[libmobility_amd] coarse level not built");
  if (n < 3 * N_) throw Error("[libmobility_amd] coarse level: vector too short");
  const double* T = op == kG ? X_.as<double>() : (op == kGt ? Xt_.as<double>() : A_.as<double>());
  launch_coarse_apply(T, sys_dev_.as<CoarseSys>(), row_sys_.as<int>(), 3 * cells_.C, max_row_,
                      op != kGt, in, out, N_, n, perm_.as<int>(), start_.as<int>(),
                      cell_of_.as<int>(), scale_.as<double>(), cells_.C, cvec_.as<double>(),
                      yvec_.as<double>(), s);
}

void CoarseLevel::factors(int q, double* L, double* X, hipStream_t st) const {
  if (cells_.C == 0) throw Error("[libmobility_amd] coarse level not built");
  if (q < 0 || q >= (int)sys_.size()) throw Error("[libmobility_amd] coarse level: no such system");
  const CoarseSys& S = sys_[q];
  const int d = 3 * S.C;
  if (d == 0) return;
  const size_t pitch = sizeof(double) * S.n;
  LM_HIP(hipMemcpy2DAsync(L, sizeof(double) * d, A_.as<double>() + S.moff, pitch,
                          sizeof(double) * d, d, hipMemcpyDeviceToHost, st));
  LM_HIP(hipMemcpy2DAsync(X, sizeof(double) * d, X_.as<double>() + S.moff, pitch,
                          sizeof(double) * d, d, hipMemcpyDeviceToHost, st));
  LM_HIP(hipStreamSynchronize(st));
  for (int r = 0; r < d; ++r)
    for (int c = r + 1; c < d; ++c) L[(size_t)r * d + c] = 0.0;
}

}  // namespace lm
