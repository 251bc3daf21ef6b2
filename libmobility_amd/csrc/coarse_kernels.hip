// Coarse level of the two-level sqrtMdotW preconditioner (no counterpart in the
// reference): particles are grouped into spatial cells; Q (3N x 3C) holds one
// orthonormal column per cell and direction, Q[3i+a, 3c+a] = 1/sqrt(n_c) for
// particle i in cell c.  The kernels build Mc = Q^T M Q from the pair terms,
// factor Mc = L L^T, form X = L^{-1}, and apply
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

// ---- cell-sorted records ---------------------------------------------------

__global__ void k_gather_records(const float4* __restrict__ rec, const int* __restrict__ perm,
                                 int64_t N, float4* __restrict__ out) {
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

// One wave per cell pair (c >= d): lanes take the targets of cell c, every
// lane walks the sources of cell d, then a butterfly sums the lanes.  The
// block and its transpose are written (Mc is symmetric because M_ji = M_ij^T).
template <bool WALL>
__global__ __launch_bounds__(256) void k_coarse_matrix(const float4* __restrict__ rec,
                                                       const int* __restrict__ start,
                                                       const double* __restrict__ scale,
                                                       int64_t npairs, double* __restrict__ Mc,
                                                       int64_t ld) {
  const int lane = threadIdx.x & 63;
  const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= npairs) return;  // uniform per wave
  int c, d;
  tri_index(p, c, d);
  const int i0 = start[c], i1 = start[c + 1], j0 = start[d], j1 = start[d + 1];
  float s[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) s[k] = 0.0f;
  for (int ib = i0; ib < i1; ib += 64) {
    const int i = ib + lane;
    if (i >= i1) break;
    const float4 ri = rec[i];
    for (int j = j0; j < j1; ++j) {
      const float4 rj = rec[j];
      if (WALL) {
        UFCoef<float> k;
        uf_coef_wall<float>(ri.x, ri.y, ri.z, ri.w, rj.x, rj.y, rj.z, rj.w, k);
        const float dv[3] = {k.dx, k.dy, k.dz};
        const float ev[3] = {k.ex, k.ey, k.ez};
        const float gv[3] = {k.gx, k.gy, k.gz};
        const float hv[3] = {k.hx, k.hy, k.hz};
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int b = 0; b < 3; ++b)
            s[3 * a + b] += (a == b ? k.a0 : 0.0f) + k.cd * dv[a] * dv[b] + gv[a] * ev[b] +
                            (b == 2 ? hv[a] : 0.0f);
      } else {
        const float dx = ri.x - rj.x, dy = ri.y - rj.y, dz = ri.z - rj.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        float fuf, guf, fwt, gwt, cc;
        rpy_coeffs<float, kUF>(r2, fuf, guf, fwt, gwt, cc);
        s[0] += fuf + guf * dx * dx;
        s[1] += guf * dx * dy;
        s[2] += guf * dx * dz;
        s[4] += fuf + guf * dy * dy;
        s[5] += guf * dy * dz;
        s[8] += fuf + guf * dz * dz;
      }
    }
  }
  if (!WALL) {
    s[3] = s[1];
    s[6] = s[2];
    s[7] = s[5];
  }
  double v[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    v[k] = (double)s[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o);
  }
  if (lane == 0) {
    const double f = scale[c] * scale[d];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        Mc[(int64_t)(3 * c + a) * ld + 3 * d + b] = f * v[3 * a + b];
        if (c != d) Mc[(int64_t)(3 * d + b) * ld + 3 * c + a] = f * v[3 * a + b];
      }
  }
}

// Mc[r][r] = 1 for the padding rows r in [n3c, n)
__global__ void k_pad_identity(double* __restrict__ A, int64_t ld, int n3c, int n) {
  const int r = n3c + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r < n) A[(int64_t)r * ld + r] = 1.0;
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
// (256 threads): wave w computes the 32 x 32
// quadrant (w >> 1, w & 1) as 2 x 2 tiles of v_mfma_f64_16x16x4f64.  Operand
// lanes: A[i][k] from lane 16 k + i, B[k][j] from lane 16 k + j; result
// register r of lane l holds row 4 r + (l >> 4) of column l & 15 (layout
// measured on gfx950).  Every index stays in the 64 x 64 tile.
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

// ---- Cholesky Mc = L L^T (lower, right-looking, 64 x 64 tiles) -----------------

// value of v in lane `lane` (uniform) of the wave
__device__ inline double lane_value(double v, int lane) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// Factor the diagonal tile k, write L_kk (upper part zeroed) and its inverse
// Dinv_k.  One wave: lane r holds row r in registers; column j of the
// right-looking factorisation reaches the other lanes by readlane, so no
// barrier is needed.  The inverse: lane c forms column c of L_kk^{-1} by
// forward substitution, L_kk rows read from LDS (same address in every lane).
// info[0] = 1 when a pivot is not positive.
__global__ __launch_bounds__(64) void k_chol_diag(double* __restrict__ A, int64_t ld, int k,
                                                  double* __restrict__ Dinv,
                                                  int* __restrict__ info) {
  __shared__ double Ls[kT][kP];
  const int r = threadIdx.x;
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
  __syncthreads();
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

// L_ik = A_ik L_kk^{-T} for the tiles below the diagonal tile k
__global__ __launch_bounds__(256) void k_chol_panel(double* __restrict__ A, int64_t ld, int k,
                                                    const double* __restrict__ Dinv) {
  __shared__ double P[kT][kP];
  __shared__ double Dk[kT][kP];
  const int i = k + 1 + blockIdx.x;
  double* Aik = A + (int64_t)i * kT * ld + (int64_t)k * kT;
  tile_load(P, Aik, ld);
  tile_load(Dk, Dinv + (int64_t)k * kT * kT, kT);
  __syncthreads();
  d4 acc[2][2];
  tile_zero(acc);
  tile_mm_mfma<true>(P, Dk, acc);
  tile_store_mfma<false>(Aik, ld, acc);
}

// The updates are blocked by kKB tile columns: inside a block of columns
// [kb, kend) the step k updates only the block's own columns, and the tiles
// right of the block receive the block's kKB products in one read-modify-write.

// A_ij -= L_ik L_jk^T for the block's columns k < j < kend, rows j <= i < nb
__global__ __launch_bounds__(256) void k_chol_local(double* __restrict__ A, int64_t ld, int k,
                                                    int kend) {
  __shared__ double Pi[kT][kP];
  __shared__ double Pj[kT][kP];
  const int jcount = kend - k - 1;
  const int j = k + 1 + (int)(blockIdx.x % jcount);
  const int i = k + 1 + (int)(blockIdx.x / jcount);
  if (i < j) return;  // uniform per workgroup
  tile_load(Pi, A + (int64_t)i * kT * ld + (int64_t)k * kT, ld);
  tile_load(Pj, A + (int64_t)j * kT * ld + (int64_t)k * kT, ld);
  __syncthreads();
  d4 acc[2][2];
  tile_zero(acc);
  tile_mm_mfma<true>(Pi, Pj, acc);
  tile_store_mfma<true>(A + (int64_t)i * kT * ld + (int64_t)j * kT, ld, acc);
}

// A_ij -= sum_{k in [kb, kend)} L_ik L_jk^T for kend <= j <= i < nb
__global__ __launch_bounds__(256) void k_chol_trailing(double* __restrict__ A, int64_t ld, int kb,
                                                       int kend) {
  __shared__ double Pi[kT][kP];
  __shared__ double Pj[kT][kP];
  int ii, jj;
  tri_index(blockIdx.x, ii, jj);
  const int i = kend + ii, j = kend + jj;
  d4 acc[2][2];
  tile_zero(acc);
  for (int k = kb; k < kend; ++k) {
    if (k > kb) __syncthreads();  // the previous products have read the tiles
    tile_load(Pi, A + (int64_t)i * kT * ld + (int64_t)k * kT, ld);
    tile_load(Pj, A + (int64_t)j * kT * ld + (int64_t)k * kT, ld);
    __syncthreads();
    tile_mm_mfma<true>(Pi, Pj, acc);
  }
  tile_store_mfma<true>(A + (int64_t)i * kT * ld + (int64_t)j * kT, ld, acc);
}

// ---- X = L^{-1}: block forward substitution on L X = I ----------------------------
// Step k: X_km <- Dinv_k X_km (m <= k), then X_im -= L_ik X_km (i > k, m <= k).
// X starts as the identity.

__global__ __launch_bounds__(256) void k_trinv_scale(double* __restrict__ X, int64_t ld, int k,
                                                     const double* __restrict__ Dinv) {
  __shared__ double R[kT][kP];
  __shared__ double Dk[kT][kP];
  const int m = blockIdx.x;
  double* Xkm = X + (int64_t)k * kT * ld + (int64_t)m * kT;
  tile_load(R, Xkm, ld);
  tile_load(Dk, Dinv + (int64_t)k * kT * kT, kT);
  __syncthreads();
  d4 acc[2][2];
  tile_zero(acc);
  tile_mm_mfma<false>(Dk, R, acc);
  tile_store_mfma<false>(Xkm, ld, acc);
}

// X_im -= L_ik X_km for the block's rows k < i < kend, m <= k
__global__ __launch_bounds__(256) void k_trinv_local(double* __restrict__ X,
                                                     const double* __restrict__ L, int64_t ld,
                                                     int k) {
  __shared__ double Lik[kT][kP];
  __shared__ double R[kT][kP];
  const int m = blockIdx.x % (k + 1);
  const int i = k + 1 + blockIdx.x / (k + 1);
  tile_load(Lik, L + (int64_t)i * kT * ld + (int64_t)k * kT, ld);
  tile_load(R, X + (int64_t)k * kT * ld + (int64_t)m * kT, ld);
  __syncthreads();
  d4 acc[2][2];
  tile_zero(acc);
  tile_mm_mfma<false>(Lik, R, acc);
  tile_store_mfma<true>(X + (int64_t)i * kT * ld + (int64_t)m * kT, ld, acc);
}

// X_im -= sum_{k in [max(kb, m), kend)} L_ik X_km for kend <= i < nb, m < kend
// (X_km = 0 for m > k)
__global__ __launch_bounds__(256) void k_trinv_trailing(double* __restrict__ X,
                                                        const double* __restrict__ L, int64_t ld,
                                                        int kb, int kend) {
  __shared__ double Lik[kT][kP];
  __shared__ double R[kT][kP];
  const int m = blockIdx.x % kend;
  const int i = kend + blockIdx.x / kend;
  d4 acc[2][2];
  tile_zero(acc);
  const int k0 = m > kb ? m : kb;
  for (int k = k0; k < kend; ++k) {
    if (k > k0) __syncthreads();
    tile_load(Lik, L + (int64_t)i * kT * ld + (int64_t)k * kT, ld);
    tile_load(R, X + (int64_t)k * kT * ld + (int64_t)m * kT, ld);
    __syncthreads();
    tile_mm_mfma<false>(Lik, R, acc);
  }
  tile_store_mfma<true>(X + (int64_t)i * kT * ld + (int64_t)m * kT, ld, acc);
}

__global__ void k_identity(double* __restrict__ X, int64_t ld, int n) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  const int64_t r = e / n, c = e % n;
  X[r * ld + c] = r == c ? 1.0 : 0.0;
}

__global__ __launch_bounds__(256) void k_transpose(const double* __restrict__ X,
                                                   double* __restrict__ Xt, int64_t ld) {
  __shared__ double S[kT][kP];
  const int bi = blockIdx.y, bj = blockIdx.x;
  tile_load(S, X + (int64_t)bi * kT * ld + (int64_t)bj * kT, ld);
  __syncthreads();
  double* o = Xt + (int64_t)bj * kT * ld + (int64_t)bi * kT;
  for (int e = threadIdx.x; e < kT * kT; e += 256) o[(e >> 6) * ld + (e & 63)] = S[e & 63][e >> 6];
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

// y[r] = sum_m T[r][m] x[m] - x[r] over m <= r (lower) or m >= r (upper), m < n3c;
// one workgroup per row, fixed-order reduction.
__global__ __launch_bounds__(256) void k_coarse_gemv(const double* __restrict__ T, int64_t ld,
                                                     int n3c, int lower,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ y) {
  __shared__ double red[256];
  const int r = blockIdx.x;
  const int m0 = lower ? 0 : r, m1 = lower ? r + 1 : n3c;
  const double* row = T + (int64_t)r * ld;
  double s = 0.0;
  for (int m = m0 + (int)threadIdx.x; m < m1; m += 256) s = __builtin_fma(row[m], x[m], s);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) y[r] = red[0] - x[r];
}

// out[k] = in[k] + scale[c] * y[3c + a] for k = 3i + a < 3N (c = cell of i),
// out[k] = in[k] beyond
// (in may equal out)
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

}  // namespace

void launch_coarse_gather(const float4* rec, const int* perm, int64_t N, float4* out,
                          hipStream_t s) {
  if (N <= 0) return;
  hipLaunchKernelGGL(k_gather_records, dim3(blocks_for(N, 256)), dim3(256), 0, s, rec, perm, N,
                     out);
  LM_HIP(hipGetLastError());
}

void launch_coarse_matrix(const float4* rec, const int* start, const double* scale, int C,
                          bool wall, double* Mc, int64_t ld, int n3c, int n, hipStream_t s) {
  launch_zero(Mc, (int64_t)n * ld, s);
  const int64_t npairs = (int64_t)C * (C + 1) / 2;
  const dim3 grid(blocks_for(npairs, 4));
  if (wall)
    hipLaunchKernelGGL(k_coarse_matrix<true>, grid, dim3(256), 0, s, rec, start, scale, npairs, Mc,
                       ld);
  else
    hipLaunchKernelGGL(k_coarse_matrix<false>, grid, dim3(256), 0, s, rec, start, scale, npairs,
                       Mc, ld);
  LM_HIP(hipGetLastError());
  if (n > n3c) {
    hipLaunchKernelGGL(k_pad_identity, dim3(blocks_for(n - n3c, 64)), dim3(64), 0, s, Mc, ld, n3c,
                       n);
    LM_HIP(hipGetLastError());
  }
}

void launch_coarse_factor(double* A, double* X, double* Xt, double* Dinv, int* info, int64_t ld,
                          int n, hipStream_t s) {
  const int nb = n / kT;
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
    const int T = nb - kend;
    if (T > 0)
      hipLaunchKernelGGL(k_chol_trailing, dim3((unsigned)((int64_t)T * (T + 1) / 2)), dim3(256), 0,
                         s, A, ld, kb, kend);
  }
  LM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_identity, dim3(blocks_for((int64_t)n * n, 256)), dim3(256), 0, s, X, ld, n);
  for (int kb = 0; kb < nb; kb += kKB) {
    const int kend = kb + kKB < nb ? kb + kKB : nb;
    for (int k = kb; k < kend; ++k) {
      hipLaunchKernelGGL(k_trinv_scale, dim3(k + 1), dim3(256), 0, s, X, ld, k, Dinv);
      if (kend - k - 1 > 0)
        hipLaunchKernelGGL(k_trinv_local, dim3((unsigned)((kend - k - 1) * (k + 1))), dim3(256), 0,
                           s, X, A, ld, k);
    }
    const int T = nb - kend;
    if (T > 0)
      hipLaunchKernelGGL(k_trinv_trailing, dim3((unsigned)((int64_t)T * kend)), dim3(256), 0, s, X,
                         A, ld, kb, kend);
  }
  LM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_transpose, dim3(nb, nb), dim3(256), 0, s, X, Xt, ld);
  LM_HIP(hipGetLastError());
}

void launch_coarse_apply(const double* T, int64_t ld, int n3c, bool lower, const double* in,
                         double* out, int64_t N, int64_t n, const int* perm, const int* start,
                         const int* cell_of, const double* scale, int C, double* cvec,
                         double* yvec, hipStream_t s) {
  hipLaunchKernelGGL(k_coarse_restrict, dim3(blocks_for(C, 4)), dim3(256), 0, s, in, perm, start,
                     scale, C, cvec);
  hipLaunchKernelGGL(k_coarse_gemv, dim3(n3c), dim3(256), 0, s, T, ld, n3c, lower ? 1 : 0, cvec,
                     yvec);
  hipLaunchKernelGGL(k_coarse_prolong, dim3(blocks_for(n, 256)), dim3(256), 0, s, in, cell_of,
                     scale, yvec, 3 * N, n, out);
  LM_HIP(hipGetLastError());
}

}  // namespace lm
