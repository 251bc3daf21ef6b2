// Coarse level of the two-level sqrtMdotW preconditioner (no counterpart in the
// reference): particles are grouped into spatial cells, separately for every
// system of an ensemble; Q (3N x 3C) holds one orthonormal column per cell
// and direction, Q[3i+a, 3c+a] = 1/sqrt(n_c) for particle i in cell c.  The
// kernels build Mc = Q^T M Q from the pair terms (block-diagonal over the
// systems, one dense block each), factor each block Mc_s = L_s L_s^T, form
// X_s = L_s^{-1}, and apply
//     out = in + Q (T - I) Q^T in,   T = X, X^T or L,
// on the translational entries.  Everything is FP64 except the pair terms of
// Mc (FP32 geometry and coefficients, FP32 sums per lane over one source cell,
// FP64 beyond), which only shape the preconditioner: exactness of the square
// root rests on X and L being inverses of each other, which holds to FP64
// rounding whatever Mc is.
#include <hip/hip_runtime.h>

#include "coarse.hpp"
#include "kernels.hpp"
#include "lm_common.hpp"
#include "pair_math.hpp"

namespace lm {

namespace {

constexpr int kT = kCoarseTile;  // 64: dense tile edge
constexpr int kP = kT + 1;       // LDS row pitch (doubles)
constexpr int kKB = 4;           // tile columns per block of the factorisation updates
constexpr int kWaveRowMax = 2048;  // longest coarse row of the one-wave-per-row matvec
// trailing updates of at least this many tiles use the two-by-two-tile kernels
constexpr int64_t kQuadTiles = 1024;

// ---- cell-sorted records ---------------------------------------------------

template <class R4>
__global__ void k_gather_records(const R4* __restrict__ rec, const int* __restrict__ perm,
                                 int64_t N, R4* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < N) out[k] = rec[perm[k]];
}

// ---- Mc = Q^T M Q ------------------------------------------------------------

// p -> (c, d) with c >= d, p = c (c + 1) / 2 + d
__device__ inline void tri_index(int64_t p, int& c, int& d) {
  int64_t cc = (int64_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
  while (cc * (cc + 1) / 2 > p) --cc;
  while ((cc + 1) * (cc + 2) / 2 <= p) ++cc;
  c = (int)cc;
  d = (int)(p - cc * (cc + 1) / 2);
}

// s += the 3x3 translational block M_ij / t0 of the pair (ri, rj),
// row-major (open: the six entries of the upper triangle, times 4/3: the
// caller scales the open sums by 3/4)
template <bool WALL, class R, class R4>
__device__ inline void pair_block_sum(const R4 ri, const R4 rj, R s[9]) {
  if (WALL) {
    UFCoef<R> k;
    uf_coef_wall<R>(ri.x, ri.y, ri.z, ri.w, rj.x, rj.y, rj.z, rj.w, k);
    const R dv[3] = {k.dx, k.dy, k.dz};
    const R ev[3] = {k.ex, k.ey, k.ez};
    const R gv[3] = {k.gx, k.gy, k.gz};
    const R hv[3] = {k.hx, k.hy, k.hz};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
        s[3 * a + b] += (a == b ? k.a0 : R(0)) + k.cd * dv[a] * dv[b] + gv[a] * ev[b] +
                        (b == 2 ? hv[a] : R(0));
  } else {
    const R dx = ri.x - rj.x, dy = ri.y - rj.y, dz = ri.z - rj.z;
    const R r2 = dx * dx + dy * dy + dz * dz;
    const R ir = lm_rsqrt(r2);
    RpyCoef<R> cf{};
    rpy_far_uf34<R>(ir, cf);
    if (__builtin_expect(r2 < R(4), 0)) {
      lm_rare_path();
      rpy_near_select_uf34<R>(r2, ir, cf);
    }
    const R fuf = cf.fuf, guf = cf.guf;
    s[0] += fuf + guf * dx * dx;
    s[1] += guf * dx * dy;
    s[2] += guf * dx * dz;
    s[4] += fuf + guf * dy * dy;
    s[5] += guf * dy * dz;
    s[8] += fuf + guf * dz * dz;
  }
}

// s += sum over targets i in [i0, i1) and sources j in [j0, j1) of the
// pair blocks, this lane's share.  The two cells go through the wave's LDS
// (sh: 128 records) in chunks of up to 64 x 64 particles, and the 64 lanes
// walk a chunk's ni x nj pairs in row-major order, lane l taking pairs
// l, l + 64, ...: every lane works whatever the cell sizes.
template <bool WALL, class R, class R4>
__device__ inline void cell_pair_sum(const R4* __restrict__ rec, int i0, int i1, int j0, int j1,
                                     int lane, R4* sh, R s[9]) {
  for (int ib = i0; ib < i1; ib += 64) {
    const int ni = min(64, i1 - ib);
    for (int jb = j0; jb < j1; jb += 64) {
      const int nj = min(64, j1 - jb);
      // The wave's reads of the previous chunk precede these writes in its
      // LDS order.
      if (lane < ni) sh[lane] = rec[ib + lane];
      if (lane < nj) sh[64 + lane] = rec[jb + lane];
      __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
      __builtin_amdgcn_wave_barrier();
      const int di = 64 / nj, dj = 64 - di * nj;  // a step of 64 pairs
      int ii = lane / nj, jj = lane - ii * nj;
      for (int p = lane; p < ni * nj; p += 64) {
        pair_block_sum<WALL, R, R4>(sh[ii], sh[64 + jj], s);
        ii += di;
        jj += dj;
        if (jj >= nj) {
          jj -= nj;
          ++ii;
        }
      }
    }
  }
  if (!WALL) {
    s[3] = s[1];
    s[6] = s[2];
    s[7] = s[5];
  }
}

// One wave per cell pair (c >= d) of one system: the lanes share the pairs
// of the two cells, then a butterfly sums the lanes.  The block and its
// transpose go to the system's matrix (Mc is symmetric because
// M_ji = M_ij^T).  pair_off[s] counts the cell pairs of the systems before s.
template <class R, class R4>
__global__ __launch_bounds__(256) void k_coarse_matrix(const R4* __restrict__ rec,
                                                       const int* __restrict__ start,
                                                       const double* __restrict__ scale,
                                                       const CoarseSys* __restrict__ sys,
                                                       const int64_t* __restrict__ pair_off,
                                                       int nsys, int64_t npairs,
                                                       double* __restrict__ Mc) {
  __shared__ R4 sh[4][128];
  const int lane = threadIdx.x & 63;
  // grid-stride over the cell pairs: one launch may not exceed 2^32 work-items
  for (int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); p < npairs;
       p += (int64_t)gridDim.x * 4) {  // uniform per wave
    int lo = 0, hi = nsys;  // the last system with pair_off <= p
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (pair_off[mid] <= p) lo = mid; else hi = mid;
    }
    const CoarseSys S = sys[lo];
    int c, d;
    tri_index(p - pair_off[lo], c, d);
    const int gc = S.c0 + c, gd = S.c0 + d;
    R s[9];
  #pragma unroll
    for (int k = 0; k < 9; ++k) s[k] = R(0);
    R4* w = sh[threadIdx.x >> 6];
    if (S.wall)
      cell_pair_sum<true, R, R4>(rec, start[gc], start[gc + 1], start[gd], start[gd + 1], lane, w, s);
    else
      cell_pair_sum<false, R, R4>(rec, start[gc], start[gc + 1], start[gd], start[gd + 1], lane, w,
                                  s);
    double v[9];
  #pragma unroll
    for (int k = 0; k < 9; ++k) {
      v[k] = (double)s[k];
  #pragma unroll
      for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o);
    }
    if (lane == 0) {
      const double f = scale[gc] * scale[gd] * (S.wall ? 1.0 : 0.75);  // open sums carry 4/3
      double* M = Mc + S.moff;
  #pragma unroll
      for (int a = 0; a < 3; ++a)
  #pragma unroll
        for (int b = 0; b < 3; ++b) {
          M[(int64_t)(3 * c + a) * S.n + 3 * d + b] = f * v[3 * a + b];
          if (c != d) M[(int64_t)(3 * d + b) * S.n + 3 * c + a] = f * v[3 * a + b];
        }
    }
  }
}

// unit diagonal on each system's padding rows [3C, n)
__global__ void k_pad_identity(double* __restrict__ A, const CoarseSys* __restrict__ sys) {
  const CoarseSys S = sys[blockIdx.x];
  for (int r = 3 * S.C + (int)threadIdx.x; r < S.n; r += blockDim.x)
    A[S.moff + (int64_t)r * S.n + r] = 1.0;
}

// ---- dense 64 x 64 tiles in LDS ------------------------------------------------

__device__ inline void tile_load(double (*S)[kP], const double* __restrict__ g, int64_t ld) {
  for (int e = threadIdx.x; e < kT * kT; e += blockDim.x) S[e >> 6][e & 63] = g[(e >> 6) * ld + (e & 63)];
}

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ inline void tile_zero(d4 acc[2][2]) {
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
}

// acc += A * (TB ? B^T : B) for 64 x 64 tiles in LDS on the FP64 matrix cores
// (256 threads): wave w computes the 32 x 32 quadrant (w >> 1, w & 1) as
// 2 x 2 tiles of v_mfma_f64_16x16x4f64.  Operand lanes: A[i][k] from lane
// 16 k + i, B[k][j] from lane 16 k + j; result register r of lane l holds
// row 4 r + (l >> 4) of column l & 15 (layout measured on gfx950).  Every
// index stays in the 64 x 64 tile.
template <bool TB>
__device__ inline void tile_mm_mfma(const double (*A)[kP], const double (*B)[kP], d4 acc[2][2]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int R0 = (w >> 1) * 32, C0 = (w & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll 4
  for (int kk = 0; kk < kT; kk += 4) {
    const double a0 = A[R0 + li][kk + lk], a1 = A[R0 + 16 + li][kk + lk];
    const double b0 = TB ? B[C0 + li][kk + lk] : B[kk + lk][C0 + li];
    const double b1 = TB ? B[C0 + 16 + li][kk + lk] : B[kk + lk][C0 + 16 + li];
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
  }
}

// out[row][col] (= or -=) the quadrant result of tile_mm_mfma
template <bool SUB>
__device__ inline void tile_store_mfma(double* __restrict__ out, int64_t ld, const d4 acc[2][2]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int R0 = (w >> 1) * 32, C0 = (w & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double* o = out + (int64_t)(R0 + 16 * a + 4 * r + lk) * ld + C0 + 16 * b + li;
        *o = SUB ? *o - acc[a][b][r] : acc[a][b][r];
      }
}

// acc += A * (TB ? B^T : B) for 64 x 64 tiles in LDS, the whole tile by one
// wave: 4 x 4 tiles of v_mfma_f64_16x16x4f64, 16 products per 8 operand
// reads (same lane layout as tile_mm_mfma).
template <bool TB>
__device__ inline void tile_mm_wave(const double (*A)[kP], const double (*B)[kP], d4 acc[4][4]) {
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll 2
  for (int kk = 0; kk < kT; kk += 4) {
    double a[4], b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a[q] = A[16 * q + li][kk + lk];
      b[q] = TB ? B[16 * q + li][kk + lk] : B[kk + lk][16 * q + li];
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        acc[p][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[p], b[q], acc[p][q], 0, 0, 0);
  }
}

// out[row][col] -= the tile result of tile_mm_wave
__device__ inline void tile_sub_wave(double* __restrict__ out, int64_t ld, const d4 acc[4][4]) {
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double* o = out + (int64_t)(16 * p + 4 * r + lk) * ld + 16 * q + li;
        *o -= acc[p][q][r];
      }
}

// ---- tile operations of the factorisation (256 threads each) -------------------
// Every operation starts with a workgroup barrier, so that one workgroup can
// run a sequence of them on the same LDS tiles and matrices.  The Cholesky
// is right-looking, blocked by kKB tile columns: inside a block [kb, kend)
// the step k updates only the block's own columns, and the tiles right of
// the block receive the block's kKB products in one read-modify-write.
// X = L^{-1} is the block forward substitution on L X = I, blocked the same
// way.

// value of v in lane `lane` (uniform) of the wave
__device__ inline double lane_value(double v, int lane) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// Factor the diagonal tile k, write L_kk (upper part zeroed) and its inverse
// Dinv_k.  One wave (lanes 0..63 of the caller): lane r holds row r in
// registers, and column j of the right-looking factorisation reaches the
// other lanes by readlane.  Then lane c forms column c of L_kk^{-1} by
// forward substitution, the rows of L_kk read from Ls (same address in every
// lane).  info[0] = 1 when a pivot is not positive.
__device__ __attribute__((always_inline)) inline void chol_diag_wave(double* __restrict__ A,
                                                                     int64_t ld, int k,
                                                                     double* __restrict__ Dinv,
                                                                     int* __restrict__ info,
                                                                     double (*Ls)[kP]) {
  const int r = threadIdx.x & 63;
  double* Akk = A + (int64_t)k * kT * ld + (int64_t)k * kT;
  double a[kT];
#pragma unroll
  for (int c = 0; c < kT; ++c) a[c] = Akk[(int64_t)r * ld + c];
  bool bad = false;
#pragma unroll
  for (int j = 0; j < kT; ++j) {
    double d = lane_value(a[j], j);
    if (!(d > 0.0)) {
      bad = true;
      d = 1.0;
    }
    const double piv = sqrt(d);
    const double ipiv = 1.0 / piv;
    a[j] = r == j ? piv : (r > j ? a[j] * ipiv : a[j]);
    // rows r >= c: A_rc -= L_rj L_cj (rows r < c hold the unused upper part)
#pragma unroll
    for (int c = j + 1; c < kT; ++c) a[c] = __builtin_fma(-a[j], lane_value(a[j], c), a[c]);
  }
#pragma unroll
  for (int c = 0; c < kT; ++c) {
    const double v = c <= r ? a[c] : 0.0;
    Ls[r][c] = v;
    Akk[(int64_t)r * ld + c] = v;
  }
  if (bad && r == 0) info[0] = 1;
  // The wave's LDS writes complete before its reads.  No workgroup barrier:
  // the caller may be one wave of a larger workgroup.
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  double x[kT];
#pragma unroll
  for (int i = 0; i < kT; ++i) {
    double s[4] = {i == r ? 1.0 : 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int m = 0; m < i; ++m) s[m & 3] = __builtin_fma(-Ls[i][m], x[m], s[m & 3]);
    x[i] = ((s[0] + s[1]) + (s[2] + s[3])) * (1.0 / Ls[i][i]);
  }
  double* D = Dinv + (int64_t)k * kT * kT;
#pragma unroll
  for (int i = 0; i < kT; ++i) D[i * kT + r] = x[i];
}

// L_ik = A_ik L_kk^{-T}
__device__ inline void op_panel(double* __restrict__ A, int64_t ld, int k, int i,
                                const double* __restrict__ Dinv, double (*P0)[kP],
                                double (*P1)[kP]) {
  __syncthreads();
  double* Aik = A + (int64_t)i * kT * ld + (int64_t)k * kT;
  tile_load(P0, Aik, ld);
  tile_load(P1, Dinv + (int64_t)k * kT * kT, kT);
  __syncthreads();
  d4 acc[2][2];
  tile_zero(acc);
  tile_mm_mfma<true>(P0, P1, acc);
  tile_store_mfma<false>(Aik, ld, acc);
}

// A_ij -= sum_{k in [k0, k1)} L_ik L_jk^T
__device__ inline void op_chol_update(double* __restrict__ A, int64_t ld, int i, int j, int k0,
                                      int k1, double (*P0)[kP], double (*P1)[kP]) {
  d4 acc[2][2];
  tile_zero(acc);
  for (int k = k0; k < k1; ++k) {
    __syncthreads();
    tile_load(P0, A + (int64_t)i * kT * ld + (int64_t)k * kT, ld);
    tile_load(P1, A + (int64_t)j * kT * ld + (int64_t)k * kT, ld);
    __syncthreads();
    tile_mm_mfma<true>(P0, P1, acc);
  }
  tile_store_mfma<true>(A + (int64_t)i * kT * ld + (int64_t)j * kT, ld, acc);
}

// X_km = Dinv_k X_km
__device__ inline void op_trinv_scale(double* __restrict__ X, int64_t ld, int k, int m,
                                      const double* __restrict__ Dinv, double (*P0)[kP],
                                      double (*P1)[kP]) {
  __syncthreads();
  double* Xkm = X + (int64_t)k * kT * ld + (int64_t)m * kT;
  tile_load(P0, Xkm, ld);
  tile_load(P1, Dinv + (int64_t)k * kT * kT, kT);
  __syncthreads();
  d4 acc[2][2];
  tile_zero(acc);
  tile_mm_mfma<false>(P1, P0, acc);
  tile_store_mfma<false>(Xkm, ld, acc);
}

// X_im -= sum_{k in [k0, k1)} L_ik X_km
__device__ inline void op_trinv_update(double* __restrict__ X, const double* __restrict__ L,
                                       int64_t ld, int i, int m, int k0, int k1,
                                       double (*P0)[kP], double (*P1)[kP]) {
  d4 acc[2][2];
  tile_zero(acc);
  for (int k = k0; k < k1; ++k) {
    __syncthreads();
    tile_load(P0, L + (int64_t)i * kT * ld + (int64_t)k * kT, ld);
    tile_load(P1, X + (int64_t)k * kT * ld + (int64_t)m * kT, ld);
    __syncthreads();
    tile_mm_mfma<false>(P0, P1, acc);
  }
  tile_store_mfma<true>(X + (int64_t)i * kT * ld + (int64_t)m * kT, ld, acc);
}

// Xt tile (bj, bi) = (X tile (bi, bj))^T
__device__ inline void op_transpose(const double* __restrict__ X, double* __restrict__ Xt,
                                    int64_t ld, int bi, int bj, double (*P0)[kP]) {
  __syncthreads();
  tile_load(P0, X + (int64_t)bi * kT * ld + (int64_t)bj * kT, ld);
  __syncthreads();
  double* o = Xt + (int64_t)bj * kT * ld + (int64_t)bi * kT;
  for (int e = threadIdx.x; e < kT * kT; e += blockDim.x) o[(e >> 6) * ld + (e & 63)] = P0[e & 63][e >> 6];
}

// ---- one matrix, one launch per step ---------------------------------------------

__global__ __launch_bounds__(64) void k_chol_diag(double* __restrict__ A, int64_t ld, int k,
                                                  double* __restrict__ Dinv,
                                                  int* __restrict__ info) {
  __shared__ double Ls[kT][kP];
  chol_diag_wave(A, ld, k, Dinv, info, Ls);
}

__global__ __launch_bounds__(256) void k_chol_panel(double* __restrict__ A, int64_t ld, int k,
                                                    const double* __restrict__ Dinv) {
  __shared__ double P0[kT][kP];
  __shared__ double P1[kT][kP];
  op_panel(A, ld, k, k + 1 + blockIdx.x, Dinv, P0, P1);
}

// A_ij -= L_ik L_jk^T for the block's columns k < j < kend, rows j <= i < nb
__global__ __launch_bounds__(256) void k_chol_local(double* __restrict__ A, int64_t ld, int k,
                                                    int kend) {
  __shared__ double P0[kT][kP];
  __shared__ double P1[kT][kP];
  const int jcount = kend - k - 1;
  const int j = k + 1 + (int)(blockIdx.x % jcount);
  const int i = k + 1 + (int)(blockIdx.x / jcount);
  if (i < j) return;  // uniform per workgroup
  op_chol_update(A, ld, i, j, k, k + 1, P0, P1);
}

// A_ij -= sum_{k in [kb, kend)} L_ik L_jk^T for kend <= j <= i < nb
__global__ __launch_bounds__(256) void k_chol_trailing(double* __restrict__ A, int64_t ld, int kb,
                                                       int kend) {
  __shared__ double P0[kT][kP];
  __shared__ double P1[kT][kP];
  int ii, jj;
  tri_index(blockIdx.x, ii, jj);
  op_chol_update(A, ld, kend + ii, kend + jj, kb, kend, P0, P1);
}

// Two tile rows [r0, r0 + 2) by two tile columns [c0, c0 + 2) per workgroup
// (wave w owns tile (r0 + (w >> 1), c0 + (w & 1))), so that every tile
// staged in LDS feeds two output tiles: half the panel bytes per flop of the
// one-tile kernels.  Tiles past the matrix (tile row nb; for X, tile column
// kend) are neither read nor written.
// The four tiles of one k (a0, a1 the row operands, b0, b1 the column
// operands; null for a tile past the matrix) pass through registers: pf is
// loaded for step k + 1 while the matrix cores work on step k.
struct QuadTiles {
  double v[4][kT * kT / 256];
};

__device__ inline void quad_fetch(QuadTiles& pf, const double* __restrict__ a0,
                                  const double* __restrict__ a1, const double* __restrict__ b0,
                                  const double* __restrict__ b1, int64_t ld) {
  const double* src[4] = {a0, a1, b0, b1};
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    if (src[t] == nullptr) continue;  // uniform
#pragma unroll
    for (int q = 0; q < kT * kT / 256; ++q) {
      const int e = (int)threadIdx.x + 256 * q;
      pf.v[t][q] = src[t][(e >> 6) * ld + (e & 63)];
    }
  }
}

__device__ inline void quad_stage(double (*S)[kT][kP], const QuadTiles& pf, bool a1, bool b1) {
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    if ((t == 1 && !a1) || (t == 3 && !b1)) continue;  // uniform
#pragma unroll
    for (int q = 0; q < kT * kT / 256; ++q) {
      const int e = (int)threadIdx.x + 256 * q;
      S[t][e >> 6][e & 63] = pf.v[t][q];
    }
  }
}

// A_ij -= sum_{k in [kb, kend)} L_ik L_jk^T for kend <= j <= i < nb, a 2 x 2
// block of tiles per workgroup (block pair (I, J), I >= J, from blockIdx.x)
__global__ __launch_bounds__(256) void k_chol_trailing2(double* __restrict__ A, int64_t ld, int kb,
                                                        int kend, int nb) {
  __shared__ double S[4][kT][kP];
  int I, J;
  tri_index(blockIdx.x, I, J);
  const int i0 = kend + 2 * I, j0 = kend + 2 * J;
  const int w = threadIdx.x >> 6;
  const int i = i0 + (w >> 1), j = j0 + (w & 1);
  const bool live = i < nb && j <= i;  // uniform per wave
  d4 acc[4][4];
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[p][q] = d4{0.0, 0.0, 0.0, 0.0};
  const bool a1 = i0 + 1 < nb, b1 = j0 + 1 < nb;
  QuadTiles pf;
  const double* ai = A + (int64_t)i0 * kT * ld;
  const double* aj = A + (int64_t)j0 * kT * ld;
  const int64_t row = (int64_t)kT * ld;
  quad_fetch(pf, ai + (int64_t)kb * kT, a1 ? ai + row + (int64_t)kb * kT : nullptr,
             aj + (int64_t)kb * kT, b1 ? aj + row + (int64_t)kb * kT : nullptr, ld);
  for (int k = kb; k < kend; ++k) {
    __syncthreads();  // the previous step is done with S
    quad_stage(S, pf, a1, b1);
    __syncthreads();
    if (k + 1 < kend) {
      const int64_t c = (int64_t)(k + 1) * kT;
      quad_fetch(pf, ai + c, a1 ? ai + row + c : nullptr, aj + c, b1 ? aj + row + c : nullptr, ld);
    }
    if (live) tile_mm_wave<true>(S[w >> 1], S[2 + (w & 1)], acc);
  }
  if (live) tile_sub_wave(A + (int64_t)i * kT * ld + (int64_t)j * kT, ld, acc);
}

// X_im -= sum_{k in [max(kb, m0), kend)} L_ik X_km for kend <= i < nb,
// m < kend, a 2 x 2 block of tiles per workgroup (rows i0, i0 + 1, columns
// m0, m0 + 1).  X_km = 0 for m > k, so the shared k range adds exact zeros
// to column m0 + 1.
__global__ __launch_bounds__(256) void k_trinv_trailing2(double* __restrict__ X,
                                                         const double* __restrict__ L, int64_t ld,
                                                         int kb, int kend, int nb) {
  __shared__ double S[4][kT][kP];
  const int mcols = (kend + 1) / 2;
  const int i0 = kend + 2 * (int)(blockIdx.x / mcols), m0 = 2 * (int)(blockIdx.x % mcols);
  const int w = threadIdx.x >> 6;
  const int i = i0 + (w >> 1), m = m0 + (w & 1);
  const bool live = i < nb && m < kend;  // uniform per wave
  d4 acc[4][4];
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[p][q] = d4{0.0, 0.0, 0.0, 0.0};
  const bool a1 = i0 + 1 < nb, b1 = m0 + 1 < kend;
  const int k0 = m0 > kb ? m0 : kb;
  QuadTiles pf;
  const double* li = L + (int64_t)i0 * kT * ld;  // row operands L_ik, L_(i+1)k
  const double* xm = X + (int64_t)m0 * kT;       // column operands X_km, X_k(m+1)
  const int64_t row = (int64_t)kT * ld;
  quad_fetch(pf, li + (int64_t)k0 * kT, a1 ? li + row + (int64_t)k0 * kT : nullptr,
             xm + (int64_t)k0 * row, b1 ? xm + (int64_t)k0 * row + kT : nullptr, ld);
  for (int k = k0; k < kend; ++k) {
    __syncthreads();  // the previous step is done with S
    quad_stage(S, pf, a1, b1);
    __syncthreads();
    if (k + 1 < kend) {
      const int64_t c = (int64_t)(k + 1) * kT, r = (int64_t)(k + 1) * row;
      quad_fetch(pf, li + c, a1 ? li + row + c : nullptr, xm + r, b1 ? xm + r + kT : nullptr, ld);
    }
    if (live) tile_mm_wave<false>(S[w >> 1], S[2 + (w & 1)], acc);
  }
  if (live) tile_sub_wave(X + (int64_t)i * kT * ld + (int64_t)m * kT, ld, acc);
}

__global__ __launch_bounds__(256) void k_trinv_scale(double* __restrict__ X, int64_t ld, int k,
                                                     const double* __restrict__ Dinv) {
  __shared__ double P0[kT][kP];
  __shared__ double P1[kT][kP];
  op_trinv_scale(X, ld, k, blockIdx.x, Dinv, P0, P1);
}

// X_im -= L_ik X_km for the block's rows k < i < kend, m <= k
__global__ __launch_bounds__(256) void k_trinv_local(double* __restrict__ X,
                                                     const double* __restrict__ L, int64_t ld,
                                                     int k) {
  __shared__ double P0[kT][kP];
  __shared__ double P1[kT][kP];
  const int m = blockIdx.x % (k + 1);
  const int i = k + 1 + blockIdx.x / (k + 1);
  op_trinv_update(X, L, ld, i, m, k, k + 1, P0, P1);
}

// X_im -= sum_{k in [max(kb, m), kend)} L_ik X_km for kend <= i < nb, m < kend
// (X_km = 0 for m > k)
__global__ __launch_bounds__(256) void k_trinv_trailing(double* __restrict__ X,
                                                        const double* __restrict__ L, int64_t ld,
                                                        int kb, int kend) {
  __shared__ double P0[kT][kP];
  __shared__ double P1[kT][kP];
  const int m = blockIdx.x % kend;
  const int i = kend + blockIdx.x / kend;
  op_trinv_update(X, L, ld, i, m, m > kb ? m : kb, kend, P0, P1);
}

__global__ void k_identity(double* __restrict__ X, int64_t ld, int n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < (int64_t)n * n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / n, c = e % n;
    X[r * ld + c] = r == c ? 1.0 : 0.0;
  }
}

__global__ __launch_bounds__(256) void k_transpose(const double* __restrict__ X,
                                                   double* __restrict__ Xt, int64_t ld) {
  __shared__ double P0[kT][kP];
  op_transpose(X, Xt, ld, blockIdx.y, blockIdx.x, P0);
}

// ---- many small matrices: one workgroup per system, the same schedule -----------

__global__ __launch_bounds__(256) void k_factor_small(double* __restrict__ A,
                                                      double* __restrict__ X,
                                                      double* __restrict__ Xt,
                                                      double* __restrict__ Dinv,
                                                      int* __restrict__ info,
                                                      const CoarseSys* __restrict__ sys,
                                                      const int* __restrict__ list) {
  __shared__ double P0[kT][kP];
  __shared__ double P1[kT][kP];
  const CoarseSys S = sys[list[blockIdx.x]];
  const int n = S.n, nb = n / kT;
  double* a = A + S.moff;
  double* x = X + S.moff;
  double* xt = Xt + S.moff;
  double* dinv = Dinv + S.doff;
  for (int kb = 0; kb < nb; kb += kKB) {
    const int kend = kb + kKB < nb ? kb + kKB : nb;
    for (int k = kb; k < kend; ++k) {
      __syncthreads();
      if (threadIdx.x < 64) chol_diag_wave(a, n, k, dinv, info, P0);
      for (int i = k + 1; i < nb; ++i) op_panel(a, n, k, i, dinv, P0, P1);
      for (int j = k + 1; j < kend; ++j)
        for (int i = j; i < nb; ++i) op_chol_update(a, n, i, j, k, k + 1, P0, P1);
    }
    for (int j = kend; j < nb; ++j)
      for (int i = j; i < nb; ++i) op_chol_update(a, n, i, j, kb, kend, P0, P1);
  }
  __syncthreads();
  for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += blockDim.x)
    x[e] = (e / n) == (e % n) ? 1.0 : 0.0;
  for (int kb = 0; kb < nb; kb += kKB) {
    const int kend = kb + kKB < nb ? kb + kKB : nb;
    for (int k = kb; k < kend; ++k) {
      for (int m = 0; m <= k; ++m) op_trinv_scale(x, n, k, m, dinv, P0, P1);
      for (int i = k + 1; i < kend; ++i)
        for (int m = 0; m <= k; ++m) op_trinv_update(x, a, n, i, m, k, k + 1, P0, P1);
    }
    for (int i = kend; i < nb; ++i)
      for (int m = 0; m < kend; ++m) op_trinv_update(x, a, n, i, m, m > kb ? m : kb, kend, P0, P1);
  }
  for (int bi = 0; bi < nb; ++bi)
    for (int bj = 0; bj < nb; ++bj) op_transpose(x, xt, n, bi, bj, P0);
}

// ---- per-application kernels ------------------------------------------------------

// cvec[3c+a] = scale[c] * sum over the particles of cell c of in[3i+a]; one
// wave per cell, lanes stride the cell's sorted list, fixed-order butterfly.
__global__ __launch_bounds__(256) void k_coarse_restrict(const double* __restrict__ in,
                                                         const int* __restrict__ perm,
                                                         const int* __restrict__ start,
                                                         const double* __restrict__ scale, int C,
                                                         double* __restrict__ cvec) {
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= C) return;  // uniform per wave
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int k = start[c] + lane; k < start[c + 1]; k += 64) {
    const int64_t p = 3 * (int64_t)perm[k];
    s0 += in[p];
    s1 += in[p + 1];
    s2 += in[p + 2];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    s0 += __shfl_xor(s0, o);
    s1 += __shfl_xor(s1, o);
    s2 += __shfl_xor(s2, o);
  }
  if (lane == 0) {
    cvec[3 * c] = scale[c] * s0;
    cvec[3 * c + 1] = scale[c] * s1;
    cvec[3 * c + 2] = scale[c] * s2;
  }
}

// y[r] = sum_m T_s[lr][m] x[r0 + m] - x[r] over the system's rows, m <= lr
// (lower) or m >= lr (upper), m < 3 C_s; r = r0 + lr with r0 = 3 c0 of the
// row's system s; one workgroup per row, fixed-order reduction.
__global__ __launch_bounds__(256) void k_coarse_gemv(const double* __restrict__ T,
                                                     const CoarseSys* __restrict__ sys,
                                                     const int* __restrict__ row_sys, int lower,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ y) {
  __shared__ double red[256];
  const int r = blockIdx.x;
  const CoarseSys S = sys[row_sys[r]];
  const int r0 = 3 * S.c0, lr = r - r0;
  const int m0 = lower ? 0 : lr, m1 = lower ? lr + 1 : 3 * S.C;
  const double* row = T + S.moff + (int64_t)lr * S.n;
  const double* xs = x + r0;
  double s = 0.0;
  for (int m = m0 + (int)threadIdx.x; m < m1; m += 256) s = __builtin_fma(row[m], xs[m], s);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) y[r] = red[0] - x[r];
}

// The same with one wave per row (short rows: many small systems);
// fixed-order butterfly.
__global__ __launch_bounds__(256) void k_coarse_gemv_wave(const double* __restrict__ T,
                                                          const CoarseSys* __restrict__ sys,
                                                          const int* __restrict__ row_sys,
                                                          int rows, int lower,
                                                          const double* __restrict__ x,
                                                          double* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= rows) return;  // uniform per wave
  const CoarseSys S = sys[row_sys[r]];
  const int r0 = 3 * S.c0, lr = r - r0;
  const int m0 = lower ? 0 : lr, m1 = lower ? lr + 1 : 3 * S.C;
  const double* row = T + S.moff + (int64_t)lr * S.n;
  const double* xs = x + r0;
  double s = 0.0;
  for (int m = m0 + lane; m < m1; m += 64) s = __builtin_fma(row[m], xs[m], s);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) y[r] = s - x[r];
}

// out[k] = in[k] + scale[c] * y[3c + a] for k = 3i + a < 3N (c = cell of i),
// out[k] = in[k] beyond (in may equal out)
__global__ void k_coarse_prolong(const double* in, const int* __restrict__ cell_of,
                                 const double* __restrict__ scale, const double* __restrict__ y,
                                 int64_t n3, int64_t n, double* out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double v = in[k];
  if (k < n3) {
    const int64_t i = k / 3;
    const int a = (int)(k - 3 * i);
    const int c = cell_of[i];
    v += scale[c] * y[3 * c + a];
  }
  out[k] = v;
}

inline unsigned blocks_for(int64_t n, int nt) { return (unsigned)((n + nt - 1) / nt); }
// Workgroups of a grid-stride launch: a dispatch holds at most 2^32 work-items.
inline unsigned stride_blocks(int64_t n, int per_block) {
  const int64_t b = (n + per_block - 1) / per_block;
  return (unsigned)(b < (int64_t)1 << 20 ? b : (int64_t)1 << 20);
}

}  // namespace

void launch_coarse_gather(const void* rec, const int* perm, int64_t N, void* out, bool fp64,
                          hipStream_t s) {
  if (N <= 0) return;
  if (fp64)
    hipLaunchKernelGGL(k_gather_records<double4>, dim3(blocks_for(N, 256)), dim3(256), 0, s,
                       static_cast<const double4*>(rec), perm, N, static_cast<double4*>(out));
  else
    hipLaunchKernelGGL(k_gather_records<float4>, dim3(blocks_for(N, 256)), dim3(256), 0, s,
                       static_cast<const float4*>(rec), perm, N, static_cast<float4*>(out));
  LM_HIP(hipGetLastError());
}

void launch_coarse_matrix(const void* rec, bool fp64, const int* start, const double* scale,
                          const CoarseSys* sys, const int64_t* pair_off, int nsys, int64_t npairs,
                          double* Mc, int64_t total, hipStream_t s) {
  launch_zero(Mc, total, s);
  if (npairs > 0) {
    if (fp64)
      hipLaunchKernelGGL((k_coarse_matrix<double, double4>), dim3(stride_blocks(npairs, 4)),
                         dim3(256), 0, s, static_cast<const double4*>(rec), start, scale, sys,
                         pair_off, nsys, npairs, Mc);
    else
      hipLaunchKernelGGL((k_coarse_matrix<float, float4>), dim3(stride_blocks(npairs, 4)), dim3(256),
                         0, s, static_cast<const float4*>(rec), start, scale, sys, pair_off, nsys,
                         npairs, Mc);
    LM_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_pad_identity, dim3(nsys), dim3(64), 0, s, Mc, sys);
  LM_HIP(hipGetLastError());
}

void launch_coarse_factor(double* A, double* X, double* Xt, double* Dinv, int* info, int64_t ld,
                          int n, hipStream_t s, const SideStream& side) {
  const int nb = n / kT;
  hipStream_t x = side.stream;
  // X may still be read by work queued on s before this call.
  LM_HIP(hipEventRecord(side.block, s));
  LM_HIP(hipStreamWaitEvent(x, side.block, 0));
  hipLaunchKernelGGL(k_identity, dim3(stride_blocks((int64_t)n * n, 256)), dim3(256), 0, x, X, ld,
                     n);
  for (int kb = 0; kb < nb; kb += kKB) {
    const int kend = kb + kKB < nb ? kb + kKB : nb;
    for (int k = kb; k < kend; ++k) {
      hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(64), 0, s, A, ld, k, Dinv, info);
      const int below = nb - k - 1;
      if (below > 0) hipLaunchKernelGGL(k_chol_panel, dim3(below), dim3(256), 0, s, A, ld, k, Dinv);
      if (kend - k - 1 > 0)
        hipLaunchKernelGGL(k_chol_local, dim3((unsigned)((kend - k - 1) * below)), dim3(256), 0, s,
                           A, ld, k, kend);
    }
    // Columns [kb, kend) of L and their diagonal inverses are final here.
    // Block kb of X reads nothing else of L, and the trailing update of L
    // below writes only columns >= kend.
    LM_HIP(hipEventRecord(side.block, s));
    LM_HIP(hipStreamWaitEvent(x, side.block, 0));
    for (int k = kb; k < kend; ++k) {
      hipLaunchKernelGGL(k_trinv_scale, dim3(k + 1), dim3(256), 0, x, X, ld, k, Dinv);
      if (kend - k - 1 > 0)
        hipLaunchKernelGGL(k_trinv_local, dim3((unsigned)((kend - k - 1) * (k + 1))), dim3(256), 0,
                           x, X, A, ld, k);
    }
    const int T = nb - kend;
    const int64_t xtiles = (int64_t)T * kend;
    if (xtiles >= kQuadTiles)
      hipLaunchKernelGGL(k_trinv_trailing2, dim3((unsigned)((int64_t)((T + 1) / 2) * ((kend + 1) / 2))),
                         dim3(256), 0, x, X, A, ld, kb, kend, nb);
    else if (T > 0)
      hipLaunchKernelGGL(k_trinv_trailing, dim3((unsigned)xtiles), dim3(256), 0, x, X, A, ld, kb, kend);
    const int64_t tiles = (int64_t)T * (T + 1) / 2, T2 = (T + 1) / 2;
    if (tiles >= kQuadTiles)
      hipLaunchKernelGGL(k_chol_trailing2, dim3((unsigned)(T2 * (T2 + 1) / 2)), dim3(256), 0, s, A,
                         ld, kb, kend, nb);
    else if (T > 0)
      hipLaunchKernelGGL(k_chol_trailing, dim3((unsigned)tiles), dim3(256), 0, s, A, ld, kb, kend);
  }
  LM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_transpose, dim3(nb, nb), dim3(256), 0, x, X, Xt, ld);
  LM_HIP(hipGetLastError());
  LM_HIP(hipEventRecord(side.done, x));
  LM_HIP(hipStreamWaitEvent(s, side.done, 0));
}

void launch_coarse_factor_small(double* A, double* X, double* Xt, double* Dinv, int* info,
                                const CoarseSys* sys, const int* list, int count, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_factor_small, dim3(count), dim3(256), 0, s, A, X, Xt, Dinv, info, sys, list);
  LM_HIP(hipGetLastError());
}

void launch_coarse_apply(const double* T, const CoarseSys* sys, const int* row_sys, int rows,
                         int max_row, bool lower, const double* in, double* out, int64_t N,
                         int64_t n, const int* perm, const int* start, const int* cell_of,
                         const double* scale, int C, double* cvec, double* yvec, hipStream_t s) {
  hipLaunchKernelGGL(k_coarse_restrict, dim3(blocks_for(C, 4)), dim3(256), 0, s, in, perm, start,
                     scale, C, cvec);
  if (max_row <= kWaveRowMax)
    hipLaunchKernelGGL(k_coarse_gemv_wave, dim3(blocks_for(rows, 4)), dim3(256), 0, s, T, sys,
                       row_sys, rows, lower ? 1 : 0, cvec, yvec);
  else
    hipLaunchKernelGGL(k_coarse_gemv, dim3(rows), dim3(256), 0, s, T, sys, row_sys, lower ? 1 : 0,
                       cvec, yvec);
  hipLaunchKernelGGL(k_coarse_prolong, dim3(blocks_for(n, 256)), dim3(256), 0, s, in, cell_of,
                     scale, yvec, 3 * N, n, out);
  LM_HIP(hipGetLastError());
}

}  // namespace lm
