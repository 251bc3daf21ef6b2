// Coarse level of the two-level sqrtMdotW preconditioner (no counterpart in the
// reference).
//
// Particles of one system are grouped into a grid of spatial cells.  With Q
// the 3N x 3C matrix of orthonormal cell columns (Q[3i+a, 3c+a] = 1/sqrt(n_c)
// for particle i in cell c) and Mc = Q^T M Q = L L^T (translational block, in
// units of t0), the coarse operator
//     Gc = I + Q (L^{-1} - I) Q^T
// whitens M on the span of Q (Q^T Gc M Gc^T Q = I) and leaves the rest alone.
// The long-wavelength collective modes, which set the Lanczos iteration count
// of M^{1/2} W, lie close to that span.  Its exact inverse is
//     Gc^{-1} = I + Q (L - I) Q^T.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "lm_common.hpp"

namespace lm {

constexpr int kCoarseTile = 64;

void launch_coarse_gather(const float4* rec, const int* perm, int64_t N, float4* out,
                          hipStream_t s);
// Mc (n x n, leading dimension ld, zero-filled here) from cell-sorted FP32
// records; rows and columns [n3c, n) are padding with unit diagonal.
void launch_coarse_matrix(const float4* rec, const int* start, const double* scale, int C,
                          bool wall, double* Mc, int64_t ld, int n3c, int n, hipStream_t s);
// In place A -> L (lower; tiles above the diagonal keep stale values), X = L^{-1},
// Xt = X^T, Dinv = inverses of the diagonal tiles; info[0] = 1 if A is not
// positive definite.  n is a multiple of kCoarseTile.
void launch_coarse_factor(double* A, double* X, double* Xt, double* Dinv, int* info, int64_t ld,
                          int n, hipStream_t s);
// out = in + Q (T - I) Q^T in on entries [0, 3N), out = in on [3N, n);
// T lower (m <= r) or upper (m >= r) triangular.  in may equal out.
void launch_coarse_apply(const double* T, int64_t ld, int n3c, bool lower, const double* in,
                         double* out, int64_t N, int64_t n, const int* perm, const int* start,
                         const int* cell_of, const double* scale, int C, double* cvec,
                         double* yvec, hipStream_t s);

// Cell assignment of positions in units of a (x, y, z per particle, host):
// g_k cells along axis k, cell edge ~ (volume / (N / ppc))^{1/3}; empty cells
// are dropped.  cell_of[i] is the compacted cell of particle i, perm lists
// the particles cell by cell (ascending index within a cell), start[c] is the
// first position of cell c in perm.
struct CoarseCells {
  int C = 0;
  int g[3] = {1, 1, 1};
  std::vector<int> cell_of, perm, start;
};
void coarse_cells(const double* xyz, int64_t stride, int64_t N, int ppc, CoarseCells& out);

class CoarseLevel {
 public:
  enum Op : int { kG = 0, kGt = 1, kGinv = 2 };
  // Largest coarse dimension 3C; larger requests get bigger cells.
  static constexpr int kMaxDim = 16384;

  // rec64 / rec32: the solver's position records of one system (N particles).
  void build(const void* rec64, const void* rec32, int64_t N, bool wall, int ppc, hipStream_t s);
  // out = Gc in (kG), Gc^T in (kGt) or Gc^{-1} in (kGinv); vectors of length
  // n >= 3N whose entries past 3N pass through.  in may equal out.
  void apply(Op op, const double* in, double* out, int64_t n, hipStream_t s);

  int cells() const { return cells_.C; }
  int dim() const { return 3 * cells_.C; }
  const CoarseCells& host_cells() const { return cells_; }
  // L and X = L^{-1} (3C x 3C row-major, host) of the last build
  void factors(double* L, double* X, hipStream_t s) const;

  // what it was built for: the solver's layout and structure ids, radius and
  // cell target
  uint64_t layout = 0, structure = 0;
  double a = 0.0;
  int ppc = 0;
  double build_ms = 0.0;

 private:
  CoarseCells cells_;
  int n_ = 0;  // padded dimension
  int64_t N_ = 0;
  DevBuf perm_, start_, cell_of_, scale_, sorted_, A_, X_, Xt_, Dinv_, info_, cvec_, yvec_;
};

}  // namespace lm
