// Learned block preconditioner for the Lanczos square root (no counterpart in
// the reference).
//
//   features : for each particle, the positions of its K = 16 nearest
//              neighbours (same system) relative to it, in units of a, sorted
//              by distance (48 floats), then its wall features u = a/max(z, a),
//              b = clamp(z/a, 0, 1), u^3 (open systems: 0, 1, 0) and zeros:
//              64 floats.
//   network  : 64 -> 128 -> 128 -> d(d+1)/2, GELU, BF16 MFMA products with FP32
//              accumulation; the outputs are a Cholesky factor L of a d x d
//              SPD block (diagonal through softplus): d = 3 on the linear
//              entries of a particle, d = 6 on its linear and angular entries.
//   targets  : center d x d block of (M_local / t0)^{-1/2}, M_local the open or
//              bottom-wall mobility of the particle and its 16 (d = 3) or 8
//              (d = 6) nearest neighbours, linear and angular entries scaled
//              by sqrt(t0) and sqrt(r0), computed here in FP64 (coupled
//              Newton-Schulz iteration in LDS).
//   use      : Lanczos runs on A = G M G^T (G block diagonal) and returns
//              G^{-1} A^{1/2} W, whose covariance is M.
#include <algorithm>
#include <cmath>
#include <vector>

#include "kernels.hpp"
#include "lm_common.hpp"
#include "pair_math.hpp"
#include "precond.hpp"

namespace lm {

// ---- k nearest neighbours over a cell grid ---------------------------------------
// One thread per target.  Shell R holds the cells at Chebyshev distance R
// from the target's cell; after each shell the K-th best squared distance is
// compared with the distance to the nearest face of the searched block that
// is not a face of the grid (a 1e-4 margin covers the rounding of the cell
// assignment).  Candidates are ordered by (FP32 squared distance, index), so
// the result is the same whatever the scan order.
template <int K>
__global__ __launch_bounds__(256) void k_knn_cells(
    const double4* __restrict__ rec, const int64_t* __restrict__ sys_off, int nsys,
    const KnnSys* __restrict__ sys, const int* __restrict__ cell_of, const int* __restrict__ perm,
    const int64_t* __restrict__ start, int64_t N, const int* __restrict__ kinds, int wall_all,
    float* __restrict__ feat, int* __restrict__ nbr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int lo = 0, hi = nsys;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (sys_off[mid] <= i) lo = mid; else hi = mid;
  }
  const KnnSys S = sys[lo];
  const int c = cell_of[i];
  const int q[3] = {c / (S.g[1] * S.g[2]), (c / S.g[2]) % S.g[1], c % S.g[2]};
  const double4 pi = rec[i];
  const double x[3] = {pi.x, pi.y, pi.z};
  float bd[K];
  int bi[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    bd[k] = 3.0e38f;
    bi[k] = -1;
  }
  for (int R = 0;; ++R) {
    for (int cx = q[0] - R; cx <= q[0] + R; ++cx) {
      if (cx < 0 || cx >= S.g[0]) continue;
      for (int cy = q[1] - R; cy <= q[1] + R; ++cy) {
        if (cy < 0 || cy >= S.g[1]) continue;
        const bool edge_xy = cx == q[0] - R || cx == q[0] + R || cy == q[1] - R || cy == q[1] + R;
        for (int cz = q[2] - R; cz <= q[2] + R; cz += (edge_xy || R == 0) ? 1 : 2 * R) {
          if (cz < 0 || cz >= S.g[2]) continue;
          const int64_t cell = S.cell0 + ((int64_t)cx * S.g[1] + cy) * S.g[2] + cz;
          for (int64_t p = start[cell]; p < start[cell + 1]; ++p) {
            const int jj = perm[p];
            if (jj == i) continue;
            const double4 pj = rec[jj];
            const float dx = (float)(pj.x - pi.x), dy = (float)(pj.y - pi.y),
                        dz = (float)(pj.z - pi.z);
            const float d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < bd[K - 1] || (d2 == bd[K - 1] && jj < bi[K - 1])) {
              float cd = d2;
              int ci = jj;
#pragma unroll
              for (int k = 0; k < K; ++k) {  // insert keeping (bd, bi) ascending
                const bool sw = cd < bd[k] || (cd == bd[k] && (unsigned)ci < (unsigned)bi[k]);
                const float td = bd[k];
                const int ti = bi[k];
                bd[k] = sw ? cd : td;
                bi[k] = sw ? ci : ti;
                cd = sw ? td : cd;
                ci = sw ? ti : ci;
              }
            }
          }
        }
      }
    }
    // distance to the nearest unsearched cell
    double dmin = 1.0e300;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (q[k] - R > 0) dmin = fmin(dmin, x[k] - (S.lo[k] + (q[k] - R) * S.edge[k]));
      if (q[k] + R + 1 < S.g[k]) dmin = fmin(dmin, S.lo[k] + (q[k] + R + 1) * S.edge[k] - x[k]);
    }
    if (dmin >= 1.0e299) break;  // the block covers the grid
    if (dmin > 0.0 && (double)bd[K - 1] < dmin * dmin * (1.0 - 1e-4)) break;
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    nbr[i * K + k] = bi[k];
    float fx = 0.f, fy = 0.f, fz = 0.f;
    if (bi[k] >= 0) {
      const double4 pj = rec[bi[k]];
      fx = (float)(pj.x - pi.x);
      fy = (float)(pj.y - pi.y);
      fz = (float)(pj.z - pi.z);
    }
    feat[i * kPrecondIn + 3 * k + 0] = fx;
    feat[i * kPrecondIn + 3 * k + 1] = fy;
    feat[i * kPrecondIn + 3 * k + 2] = fz;
  }
  // wall features of the particle (records: z = max(z, a)/a >= 1, w = b)
  const bool wall = kinds ? kinds[lo] != 0 : wall_all != 0;
  const float u = wall ? (float)(1.0 / pi.z) : 0.f;
  float* f = feat + i * kPrecondIn + 3 * K;
  f[0] = u;
  f[1] = wall ? (float)pi.w : 1.f;
  f[2] = u * u * u;
  for (int k = 3 * K + 3; k < kPrecondIn; ++k) feat[i * kPrecondIn + k] = 0.f;
}

void launch_knn_cells(const double4* rec, const int64_t* sys_off, int nsys, const KnnSys* sys,
                      const int* cell_of, const int* perm, const int64_t* start, int64_t N,
                      const int* kinds, int wall, float* feat, int* nbr, hipStream_t s) {
  if (N <= 0) return;
  static_assert(3 * kPrecondK + 3 <= kPrecondIn, "feature layout");
  hipLaunchKernelGGL(k_knn_cells<kPrecondK>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s,
                     rec, sys_off, nsys, sys, cell_of, perm, start, N, kinds, wall, feat, nbr);
  LM_HIP(hipGetLastError());
}

void knn_grid(const double* xyz4, const int64_t* sys_off, int nsys, std::vector<KnnSys>& sys,
              std::vector<int>& cell_of, std::vector<int>& perm, std::vector<int64_t>& start) {
  const int64_t N = sys_off[nsys];
  sys.assign(nsys, KnnSys{});
  cell_of.assign(N, 0);
  perm.assign(N, 0);
  start.clear();
  int64_t cells = 0;
  for (int q = 0; q < nsys; ++q) {
    const int64_t a = sys_off[q], b = sys_off[q + 1], n = b - a;
    KnnSys& S = sys[q];
    S.cell0 = cells;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, ext[3];
    if (n > 0)
      for (int k = 0; k < 3; ++k) lo[k] = hi[k] = xyz4[4 * a + k];
    for (int64_t i = a; i < b; ++i)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], xyz4[4 * i + k]);
        hi[k] = std::max(hi[k], xyz4[4 * i + k]);
      }
    double vol = 1.0;
    for (int k = 0; k < 3; ++k) {
      ext[k] = hi[k] - lo[k];
      vol *= std::max(ext[k], 1.0);
    }
    const double side = std::cbrt(vol / std::max(1.0, (double)n / 8.0));
    int64_t ncell = 1;
    for (int k = 0; k < 3; ++k) {
      S.g[k] = (int)std::max<long long>(1, std::min<long long>(1024, std::llround(std::max(ext[k], 1.0) / side)));
      S.lo[k] = lo[k];
      S.edge[k] = ext[k] > 0.0 ? ext[k] / S.g[k] : 1.0;
      ncell *= S.g[k];
    }
    std::vector<int64_t> count(ncell + 1, 0);
    for (int64_t i = a; i < b; ++i) {
      int qq[3];
      for (int k = 0; k < 3; ++k) {
        qq[k] = 0;
        if (ext[k] > 0.0) {
          const double f = std::floor((xyz4[4 * i + k] - lo[k]) / S.edge[k]);
          qq[k] = (int)std::min<double>(S.g[k] - 1, std::max(0.0, f));
        }
      }
      cell_of[i] = (qq[0] * S.g[1] + qq[1]) * S.g[2] + qq[2];
      ++count[cell_of[i] + 1];
    }
    for (int64_t c = 0; c < ncell; ++c) count[c + 1] += count[c];
    for (int64_t c = 0; c < ncell; ++c) start.push_back(a + count[c]);
    std::vector<int64_t> fill(count.begin(), count.end() - 1);
    for (int64_t i = a; i < b; ++i) perm[a + fill[cell_of[i]]++] = (int)i;
    cells += ncell;
  }
  start.push_back(N);
}


// ---- FP64 targets: center block of (M_local/t0)^{-1/2} -------------------------
template <int KD>
__device__ inline void lds_matmul(const double* A, const double* B, double* C, int tid, int nt) {
  constexpr int LD = KD + 1;
  for (int e = tid; e < KD * KD; e += nt) {
    const int r = e / KD, c = e % KD;
    double s = 0.0;
    for (int k = 0; k < KD; ++k) s += A[r * LD + k] * B[k * LD + c];
    C[r * LD + c] = s;
  }
}

__device__ inline double block_max(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double m = red[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, red[w]);
  return m;
}

// One workgroup per center: the dimensionless local mobility of the center
// and its KC - 1 nearest neighbours (entries [U; W] per particle with D = 6:
// UF, sqrt(3)/2 UT, sqrt(3)/2 WF, WT -- the physical blocks scaled by
// 1/sqrt(t0 t0), 1/sqrt(t0 r0), 1/sqrt(r0 r0); D = 3: UF), open or wall.
// Coupled Newton-Schulz: Y0 = A/c, Z0 = I;  T = (3I - Z Y)/2;  Y <- Y T,
// Z <- T Z.  Z -> (A/c)^{-1/2} when the eigenvalues of A/c lie in (0, 1]
// (c = Gershgorin bound).  Stops when max|Z Y - I| < 1e-13 (status 0) or
// after 100 steps (status 1).
template <int KC, int D>
__global__ __launch_bounds__(256) void k_local_targets(const double4* __restrict__ rec,
                                                       const int* __restrict__ nbr,
                                                       const int64_t* __restrict__ centers,
                                                       const int* __restrict__ wall_of,
                                                       float* __restrict__ target,
                                                       int* __restrict__ status) {
  constexpr int KD = D * KC, LD = KD + 1;
  constexpr int BLK = D == 6 ? (kUF | kUT | kWT | kWF) : kUF;
  __shared__ double buf[4][KD * LD];
  __shared__ double4 P[KC];
  __shared__ double red[4];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int64_t ctr = centers[blockIdx.x];
  const bool wall = wall_of[blockIdx.x] != 0;
  if (tid < KC) {
    int64_t j = (tid == 0) ? ctr : (int64_t)nbr[ctr * kPrecondK + tid - 1];
    if (j < 0) j = ctr;
    P[tid] = rec[j];
  }
  __syncthreads();
  double* Y = buf[0];
  double* Z = buf[1];
  double* T = buf[2];
  double* S = buf[3];
  const double s3 = 0.86602540378443865;  // sqrt(3)/2
  for (int e = tid; e < KC * KC; e += nt) {
    const int a = e / KC, b = e % KC;
    const double4 pa = P[a], pb = P[b];
    for (int col = 0; col < D; ++col) {
      double v[6] = {0, 0, 0, 0, 0, 0};
      v[col] = 1.0;
      PairAcc<double> acc;
      acc_zero(acc);
      if (wall)
        wall_pair<double, BLK>(pa.x, pa.y, pa.z, pa.w, pb.x, pb.y, pb.z, pb.w, v[0], v[1], v[2],
                               v[3], v[4], v[5], acc);
      else
        open_pair<double, BLK>(pa.x - pb.x, pa.y - pb.y, pa.z - pb.z, v[0], v[1], v[2], v[3],
                               v[4], v[5], acc);
      double* c = Y + (D * a) * LD + D * b + col;
      if (D == 3) {
        c[0 * LD] = acc.uf.x;
        c[1 * LD] = acc.uf.y;
        c[2 * LD] = acc.uf.z;
      } else if (col < 3) {  // a force: U from UF, W from WF
        c[0 * LD] = acc.uf.x;
        c[1 * LD] = acc.uf.y;
        c[2 * LD] = acc.uf.z;
        c[3 * LD] = s3 * acc.wf.x;
        c[4 * LD] = s3 * acc.wf.y;
        c[5 * LD] = s3 * acc.wf.z;
      } else {  // a torque: U from UT, W from WT
        c[0 * LD] = s3 * acc.ut.x;
        c[1 * LD] = s3 * acc.ut.y;
        c[2 * LD] = s3 * acc.ut.z;
        c[3 * LD] = acc.wt.x;
        c[4 * LD] = acc.wt.y;
        c[5 * LD] = acc.wt.z;
      }
    }
  }
  __syncthreads();
  double g = 0.0;
  for (int r = tid; r < KD; r += nt) {
    double sum = 0.0;
    for (int c = 0; c < KD; ++c) sum += fabs(Y[r * LD + c]);
    g = fmax(g, sum);
  }
  const double cb = block_max(g, red);
  const double ic = 1.0 / cb;
  for (int e = tid; e < KD * KD; e += nt) {
    const int r = e / KD, c = e % KD;
    Y[r * LD + c] *= ic;
    Z[r * LD + c] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  int converged = 0;
  for (int it = 0; it < 100; ++it) {
    lds_matmul<KD>(Z, Y, T, tid, nt);  // T = Z Y
    __syncthreads();
    double err = 0.0;
    for (int e = tid; e < KD * KD; e += nt) {
      const int r = e / KD, c = e % KD;
      const double v = T[r * LD + c];
      const double id = (r == c) ? 1.0 : 0.0;
      err = fmax(err, fabs(v - id));
      T[r * LD + c] = 0.5 * (3.0 * id - v);
    }
    const double m = block_max(err, red);  // includes the barrier for T
    if (m < 1e-13) {
      converged = 1;
      break;
    }
    lds_matmul<KD>(Y, T, S, tid, nt);  // S = Y T
    __syncthreads();
    {
      double* t = Y;
      Y = S;
      S = t;
    }
    lds_matmul<KD>(T, Z, S, tid, nt);  // S = T Z
    __syncthreads();
    {
      double* t = Z;
      Z = S;
      S = t;
    }
  }
  // (A/t0)^{-1/2} = c^{-1/2} (A/(t0 c))^{-1/2}: scale the center block
  const double sc = sqrt(ic);
  if (tid < D * D) target[blockIdx.x * D * D + tid] = (float)(sc * Z[(tid / D) * LD + (tid % D)]);
  if (tid == 0) status[blockIdx.x] = converged ? 0 : 1;
}

void launch_local_targets(const double4* rec, const int* nbr, const int64_t* centers,
                          const int* wall, int64_t ncenters, int kind, float* target, int* status,
                          hipStream_t s) {
  if (ncenters <= 0) return;
  if (kind == 0)
    hipLaunchKernelGGL((k_local_targets<kPrecondK + 1, 3>), dim3((unsigned)ncenters), dim3(256), 0,
                       s, rec, nbr, centers, wall, target, status);
  else
    hipLaunchKernelGGL((k_local_targets<kPrecondK6 + 1, 6>), dim3((unsigned)ncenters), dim3(256),
                       0, s, rec, nbr, centers, wall, target, status);
  LM_HIP(hipGetLastError());
}

// ---- BF16 MFMA GEMM with fused epilogues ---------------------------------------
// C[m][n] = sum_k A(m,k) B(k,n)   (FP32 in memory, BF16 in the MFMA, FP32 sums)
//   A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn], C row-major ldc.
// Epilogue: + bias[n]; if gelu: Zout = pre-activation, C = GELU(pre); if
// accumulate: C += result.  Workgroup tile 64x64 (4 waves of 32x32, each 2x2
// MFMA 16x16x16), K staged through LDS in slices of 16.
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline short to_bf16(float x) {
  // round to nearest even
  unsigned u = __float_as_uint(x);
  const unsigned lsb = (u >> 16) & 1u;
  u += 0x7FFFu + lsb;
  return (short)(u >> 16);
}

__device__ inline float gelu_f(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }

__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const float* __restrict__ A,
                                                   int64_t sam, int64_t sak,
                                                   const float* __restrict__ B, int64_t sbk,
                                                   int64_t sbn, float* __restrict__ C, int64_t ldc,
                                                   const float* __restrict__ bias, int gelu,
                                                   float* __restrict__ Zout, int accumulate) {
  __shared__ short sA[64][16 + 4];  // [m][k]
  __shared__ short sB[16][64 + 4];  // [k][n]
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    __syncthreads();
    for (int e = tid; e < 64 * 16; e += 256) {
      const int r = e / 16, c = e % 16;  // A tile: row r (m), col c (k)
      const int gm = m0 + r, gk = k0 + c;
      sA[r][c] = to_bf16((gm < M && gk < K) ? A[(int64_t)gm * sam + (int64_t)gk * sak] : 0.f);
      const int kk = e / 64, nn = e % 64;  // B tile: row kk (k), col nn (n)
      const int gk2 = k0 + kk, gn = n0 + nn;
      sB[kk][nn] = to_bf16((gk2 < K && gn < N) ? B[(int64_t)gk2 * sbk + (int64_t)gn * sbn] : 0.f);
    }
    __syncthreads();
    // MFMA 16x16x16 bf16 operand layout: lane l supplies A[l%16][4(l/16)+0..3]
    // and B[4(l/16)+0..3][l%16]; it receives D[4(l/16)+0..3][l%16].
    const int r16 = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      bf16x4 av;
#pragma unroll
      for (int q = 0; q < 4; ++q) av[q] = sA[wm + a * 16 + r16][kq + q];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        bf16x4 bv;
#pragma unroll
        for (int q = 0; q < 4; ++q) bv[q] = sB[kq + q][wn + b * 16 + r16];
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[a][b], 0, 0, 0);
      }
    }
  }
  const int c16 = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int gm = m0 + wm + a * 16 + rq + q;
        const int gn = n0 + wn + b * 16 + c16;
        if (gm < M && gn < N) {
          float v = acc[a][b][q];
          if (bias) v += bias[gn];
          const int64_t o = (int64_t)gm * ldc + gn;
          if (gelu) {
            Zout[o] = v;
            v = gelu_f(v);
          }
          C[o] = accumulate ? C[o] + v : v;
        }
      }
}

void launch_gemm_bf16(int M, int N, int K, const float* A, int64_t sam, int64_t sak, const float* B,
                      int64_t sbk, int64_t sbn, float* C, int64_t ldc, const float* bias, int gelu,
                      float* Zout, int accumulate, hipStream_t s) {
  if (M <= 0 || N <= 0) return;
  dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64));
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc,
                     bias, gelu, Zout, accumulate);
  LM_HIP(hipGetLastError());
}

// ---- network head: outputs -> L, G, loss and its gradient ----------------------
__device__ inline float softplus_f(float x) { return x > 20.f ? x : log1pf(expf(x)); }
__device__ inline float sigmoid_f(float x) { return 1.f / (1.f + expf(-x)); }

// L lower triangular d x d: L_ii = softplus(o_i) + eps, then the strictly
// lower entries in row order: o_d = L_10, o_{d+1} = L_20, o_{d+2} = L_21, ...
template <int D>
__device__ inline void head_L(const float* o, float L[D][D]) {
  int q = D;
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) {
      if (j == i)
        L[i][j] = softplus_f(o[i]) + kPrecondEps;
      else if (j < i)
        L[i][j] = o[q++];
      else
        L[i][j] = 0.f;
    }
}

// loss = mean_b sum_ij (G - T)_ij^2 ; dO = d loss / d outputs (divided by B)
template <int D>
__global__ void k_head_loss(const float* __restrict__ out, const float* __restrict__ tgt, int Bn,
                            float* __restrict__ dout, float* __restrict__ loss_part) {
  constexpr int NO = D * (D + 1) / 2;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (b < Bn) {
    const float* o = out + NO * b;
    float L[D][D], R[D][D];
    head_L<D>(o, L);
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) {
        float s = 0.f;
        for (int k = 0; k < D; ++k) s += L[i][k] * L[j][k];
        R[i][j] = s - tgt[D * D * b + D * i + j];
        l += R[i][j] * R[i][j];
      }
    // dloss/dG = 2 R ; dloss/dL = 2 (R + R^T) L
    float* d = dout + NO * b;
    int q = D;
    for (int i = 0; i < D; ++i)
      for (int k = 0; k <= i; ++k) {
        float s = 0.f;
        for (int j = 0; j < D; ++j) s += (R[i][j] + R[j][i]) * L[j][k];
        const float dl = 2.f * s / (float)Bn;
        if (k == i)
          d[i] = dl * sigmoid_f(o[i]);
        else
          d[q++] = dl;
      }
  }
  // block sum of the loss
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
  __shared__ float sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sh[w];
    loss_part[blockIdx.x] = t;
  }
}

void launch_head_loss(const float* out, const float* tgt, int Bn, int kind, float* dout,
                      float* loss_part, hipStream_t s) {
  const dim3 grid((unsigned)((Bn + 255) / 256)), block(256);
  if (kind == 0)
    hipLaunchKernelGGL(k_head_loss<3>, grid, block, 0, s, out, tgt, Bn, dout, loss_part);
  else
    hipLaunchKernelGGL(k_head_loss<6>, grid, block, 0, s, out, tgt, Bn, dout, loss_part);
  LM_HIP(hipGetLastError());
}

// dZ = dH * GELU'(Z), GELU'(z) = Phi(z) + z phi(z)
__global__ void k_gelu_back(const float* __restrict__ dH, const float* __restrict__ Z, int64_t n,
                            float* __restrict__ dZ) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float z = Z[i];
  const float Phi = 0.5f * (1.f + erff(z * 0.70710678118654752f));
  const float phi = 0.39894228040143268f * expf(-0.5f * z * z);
  dZ[i] = dH[i] * (Phi + z * phi);
}

void launch_gelu_back(const float* dH, const float* Z, int64_t n, float* dZ, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gelu_back, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dH, Z, n, dZ);
  LM_HIP(hipGetLastError());
}

// db[n] = sum_b dZ[b][n]  (one thread per column, fixed order)
__global__ void k_bias_grad(const float* __restrict__ dZ, int Bn, int Nc, float* __restrict__ db) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= Nc) return;
  float s = 0.f;
  for (int b = 0; b < Bn; ++b) s += dZ[(int64_t)b * Nc + n];
  db[n] = s;
}

void launch_bias_grad(const float* dZ, int Bn, int Nc, float* db, hipStream_t s) {
  hipLaunchKernelGGL(k_bias_grad, dim3((unsigned)((Nc + 127) / 128)), dim3(128), 0, s, dZ, Bn, Nc, db);
  LM_HIP(hipGetLastError());
}

// Adam (Kingma & Ba), FP32 master weights
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

void launch_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, int step,
                 hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w, g, m, v, n, lr,
                     b1, b2, 1e-8f, c1, c2);
  LM_HIP(hipGetLastError());
}

// dst[q][c] = src[idx[q]][c], q < rows (mini-batch gather)
__global__ void k_gather_rows(const float* __restrict__ src, const int64_t* __restrict__ idx,
                              int rows, int cols, float* __restrict__ dst) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)rows * cols) return;
  const int64_t q = e / cols, c = e % cols;
  dst[e] = src[idx[q] * cols + c];
}

void launch_gather_rows(const float* src, const int64_t* idx, int rows, int cols, float* dst,
                        hipStream_t s) {
  const int64_t n = (int64_t)rows * cols;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, idx,
                     rows, cols, dst);
  LM_HIP(hipGetLastError());
}

// ---- per-particle blocks (FP64) --------------------------------------------------
// 3x3: G = L L^T, G^{-1} = L^{-T} L^{-1}.  6x6: G = L L^T S, G^{-1} = S^{-1}
// L^{-T} L^{-1}, S = diag(1, 1, 1, c, c, c).  Systems smaller than K + 1
// particles (no full neighbourhood) get G = I.
template <int D>
__global__ void k_make_blocks(const float* __restrict__ out, const int* __restrict__ nbr, int64_t N,
                              double c, double* __restrict__ G, double* __restrict__ Gi) {
  constexpr int NO = D * (D + 1) / 2;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double L[D][D], X[D][D];  // X = L^{-1}
  const bool full = nbr[i * kPrecondK + kPrecondK - 1] >= 0;
  if (full) {
    float Lf[D][D];
    head_L<D>(out + NO * i, Lf);
    for (int a = 0; a < D; ++a)
      for (int b = 0; b < D; ++b) L[a][b] = (double)Lf[a][b];
  } else {
    for (int a = 0; a < D; ++a)
      for (int b = 0; b < D; ++b) L[a][b] = (a == b) ? 1.0 : 0.0;
  }
  // forward substitution, column by column: L X = I
  for (int j = 0; j < D; ++j)
    for (int r = 0; r < D; ++r) {
      double s = (r == j) ? 1.0 : 0.0;
      for (int k = 0; k < r; ++k) s -= L[r][k] * X[k][j];
      X[r][j] = (r >= j) ? s / L[r][r] : 0.0;
    }
  for (int a = 0; a < D; ++a)
    for (int b = 0; b < D; ++b) {
      double g = 0.0, gi = 0.0;
      for (int k = 0; k < D; ++k) {
        g += L[a][k] * L[b][k];
        gi += X[k][a] * X[k][b];
      }
      if (D == 6) {
        if (b >= 3) g *= c;
        if (a >= 3) gi /= c;
      }
      G[(int64_t)D * D * i + D * a + b] = g;
      Gi[(int64_t)D * D * i + D * a + b] = gi;
    }
}

void launch_make_blocks(const float* out, const int* nbr, int64_t N, int kind, double a,
                        double* G, double* Gi, hipStream_t s) {
  if (N <= 0) return;
  const dim3 grid((unsigned)((N + 255) / 256)), block(256);
  if (kind == 0)
    hipLaunchKernelGGL(k_make_blocks<3>, grid, block, 0, s, out, nbr, N, 1.0, G, Gi);
  else
    hipLaunchKernelGGL(k_make_blocks<6>, grid, block, 0, s, out, nbr, N,
                       2.0 * a / 1.7320508075688772, G, Gi);
  LM_HIP(hipGetLastError());
}

// 3x3: y[3i..3i+2] = B_i x[3i..3i+2] (or B_i^T) for i < N; entries beyond 3N
// (rotational part of a 6N vector) are copied unchanged.  6x6: the block acts
// on (x[3i..3i+2], x[3N+3i..3N+3i+2]).
template <int D>
__global__ void k_apply_blocks(const double* __restrict__ B, const double* __restrict__ x,
                               int64_t N, int64_t n, int transpose, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) {
    const double* b = B + (int64_t)D * D * i;
    int64_t idx[D];
    double v[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      idx[k] = (k < 3) ? 3 * i + k : 3 * N + 3 * i + (k - 3);
      v[k] = x[idx[k]];
    }
#pragma unroll
    for (int r = 0; r < D; ++r) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) s += (transpose ? b[D * k + r] : b[D * r + k]) * v[k];
      y[idx[r]] = s;
    }
  }
  if (D == 3) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = 3 * N + i; k < n; k += stride) y[k] = x[k];
  }
}

void launch_apply_blocks(const double* B, const double* x, int64_t N, int64_t n, int kind,
                         int transpose, double* y, hipStream_t s) {
  if (N <= 0) return;
  if (kind == 1 && n != 6 * N)
    throw Error("[libmobility_amd] 6x6 preconditioner blocks need linear and angular entries");
  const dim3 grid((unsigned)((N + 255) / 256)), block(256);
  if (kind == 0)
    hipLaunchKernelGGL(k_apply_blocks<3>, grid, block, 0, s, B, x, N, n, transpose, y);
  else
    hipLaunchKernelGGL(k_apply_blocks<6>, grid, block, 0, s, B, x, N, n, transpose, y);
  LM_HIP(hipGetLastError());
}

}  // namespace lm
