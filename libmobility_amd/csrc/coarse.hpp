// Coarse level of the two-level sqrtMdotW preconditioner (no counterpart in the
// reference).
//
// The particles of every system are grouped into a grid of spatial cells
// (cells never span systems).  With Q the 3N x 3C matrix of orthonormal cell
// columns (Q[3i+a, 3c+a] = 1/sqrt(n_c) for particle i in cell c) and
// Mc = Q^T M Q = L L^T (translational block in units of t0, block-diagonal
// over the systems), the coarse operator
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

// One system's block of the coarse level: its cells [c0, c0 + C) of the
// global cell list, and its n x n matrices (n = 3C rounded up to a multiple
// of kCoarseTile, rows and columns past 3C are padding with unit diagonal)
// at element moff of the packed buffers, its diagonal-tile inverses at doff.
struct CoarseSys {
  int64_t moff, doff;
  int c0, C, n, wall;
};

// out[k] = rec[perm[k]]: float4 records, or double4 with fp64
void launch_coarse_gather(const void* rec, const int* perm, int64_t N, void* out, bool fp64,
                          hipStream_t s);
// Mc (all systems' blocks, `total` elements, zero-filled here) from
// cell-sorted records, FP32 pair terms (float4) or FP64 (double4, fp64);
// pair_off[s] = cell pairs (c >= d) of the systems before s, npairs = all of
// them.
void launch_coarse_matrix(const void* rec, bool fp64, const int* start, const double* scale,
                          const CoarseSys* sys, const int64_t* pair_off, int nsys, int64_t npairs,
                          double* Mc, int64_t total, hipStream_t s);
// A second stream (default priority, non-blocking) and two events, created
// on first use with the build's device current and released with their
// owner.
class SideStream {
 public:
  SideStream() = default;
  SideStream(const SideStream&) = delete;
  SideStream& operator=(const SideStream&) = delete;
  ~SideStream();
  void ensure();
  hipStream_t stream = nullptr;
  hipEvent_t block = nullptr, done = nullptr;
};

// One matrix, one launch per step: in place A -> L (lower; tiles above the
// diagonal keep stale values), X = L^{-1}, Xt = X^T, Dinv = inverses of the
// diagonal tiles; info[0] = 1 if A is not positive definite.  n is a
// multiple of kCoarseTile.  The steps of X run on side.stream, each block of
// X behind an event on s that follows the block's columns of L; s waits for
// the whole of X before it goes on.
void launch_coarse_factor(double* A, double* X, double* Xt, double* Dinv, int* info, int64_t ld,
                          int n, hipStream_t s, const SideStream& side);
// The same for the systems list[0..count), one workgroup each (small blocks).
void launch_coarse_factor_small(double* A, double* X, double* Xt, double* Dinv, int* info,
                                const CoarseSys* sys, const int* list, int count, hipStream_t s);
// out = in + Q (T - I) Q^T in on entries [0, 3N), out = in on [3N, n); T the
// systems' lower (m <= r) or upper (m >= r) triangular blocks; row_sys[r] is
// the system of coarse row r < rows; max_row = the largest 3 C_s (selects one
// wave or one workgroup per row).  in may equal out.
void launch_coarse_apply(const double* T, const CoarseSys* sys, const int* row_sys, int rows,
                         int max_row, bool lower, const double* in, double* out, int64_t N,
                         int64_t n, const int* perm, const int* start, const int* cell_of,
                         const double* scale, int C, double* cvec, double* yvec, hipStream_t s);

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
// max_dim: largest coarse dimension 3C; larger requests get bigger cells
void coarse_cells(const double* xyz, int64_t stride, int64_t N, int ppc, CoarseCells& out,
                  int64_t max_dim);

class CoarseLevel {
 public:
  enum Op : int { kG = 0, kGt = 1, kGinv = 2 };
  // Default largest coarse dimension 3C_s of one system; larger requests get
  // bigger cells.  The build holds three dense (3C_s)^2 FP64 matrices.
  static constexpr int kMaxDim = 24576;
  static constexpr int64_t kMaxDimLimit = 98304;

  // Largest number of 64-tiles of a block factored by the one-workgroup path
  // (used when at least kSmallBatch systems qualify).
  static constexpr int kSmallTiles = 16, kSmallBatch = 8;

  // rec64 / rec32: the solver's position records; off[0..nsys] the systems'
  // particle offsets, wall[s] 1 for a single-wall system.  Every system gets
  // its own cells and coarse block.
  // fp64: the coarse matrix from FP64 pair terms (otherwise FP32, and FP64
  // only when the FP32 one turns out indefinite)
  void build(const void* rec64, const void* rec32, const std::vector<int64_t>& off,
             const std::vector<int>& wall, int ppc, int64_t max_dim, bool fp64, hipStream_t s);
  // out = Gc in (kG), Gc^T in (kGt) or Gc^{-1} in (kGinv); vectors of length
  // n >= 3N whose entries past 3N pass through.  in may equal out.
  void apply(Op op, const double* in, double* out, int64_t n, hipStream_t s);

  int cells() const { return cells_.C; }
  int dim() const { return 3 * cells_.C; }
  int systems() const { return (int)sys_.size(); }
  // global cells (cell_of, perm, start over all systems; g of the first
  // system with particles)
  const CoarseCells& host_cells() const { return cells_; }
  // L and X = L^{-1} of system s's block (3C_s x 3C_s row-major, host)
  void factors(int s, double* L, double* X, hipStream_t st) const;
  int system_cells(int s) const { return sys_.at(s).C; }

  // what it was built for: the solver's layout and structure ids, radius and
  // cell target
  uint64_t layout = 0, structure = 0;
  double a = 0.0;
  int ppc = 0;
  int64_t max_dim = 0;
  bool fp64_requested = false;
  double build_ms = 0.0;
  // the last build's coarse matrix came from FP64 pair terms (the FP32 one
  // was indefinite)
  bool fp64_pairs() const { return fp64_pairs_; }

 private:
  bool fp64_pairs_ = false;
  CoarseCells cells_;
  std::vector<CoarseSys> sys_;
  int64_t N_ = 0;
  int max_row_ = 0;  // largest 3 C_s
  DevBuf perm_, start_, cell_of_, scale_, sorted_, A_, X_, Xt_, Dinv_, info_, cvec_, yvec_;
  DevBuf sys_dev_, row_sys_, pair_off_, small_;
  SideStream side_;
};

}  // namespace lm
