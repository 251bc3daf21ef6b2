// Learned block preconditioner for the Lanczos square root (no counterpart in
// the reference).
//
//   features : for each particle, the positions of its K = 16 nearest
//              neighbours (same system) relative to it, in units of a, sorted
//              by distance -> 48 floats.
//   network  : 48 -> 128 -> 128 -> 6, GELU, BF16 MFMA products with FP32
//              accumulation; the 6 outputs are a Cholesky factor L of a 3x3
//              SPD block G = L L^T (diagonal through softplus).
//   targets  : center 3x3 block of (M_local / t0)^{-1/2}, M_local the open-RPY
//              mobility of the particle and its K neighbours, computed here in
//              FP64 (coupled Newton-Schulz iteration in LDS).
//   use      : Lanczos runs on A = G M G (G block diagonal, symmetric) and
//              returns G^{-1} A^{1/2} W, whose covariance is M.
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
    const int64_t* __restrict__ start, int64_t N, float* __restrict__ feat,
    int* __restrict__ nbr) {
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
    feat[i * 3 * K + 3 * k + 0] = fx;
    feat[i * 3 * K + 3 * k + 1] = fy;
    feat[i * 3 * K + 3 * k + 2] = fz;
  }
}

void launch_knn_cells(const double4* rec, const int64_t* sys_off, int nsys, const KnnSys* sys,
                      const int* cell_of, const int* perm, const int64_t* start, int64_t N,
                      float* feat, int* nbr, hipStream_t s) {
  if (N <= 0) return;
  hipLaunchKernelGGL(k_knn_cells<kPrecondK>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s,
                     rec, sys_off, nsys, sys, cell_of, perm, start, N, feat, nbr);
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
constexpr int kC = kPrecondK + 1;  // cluster size
constexpr int kD = 3 * kC;         // 51
constexpr int kLD = kD + 1;        // padded row length

__device__ inline void lds_matmul(const double* A, const double* B, double* C, int tid, int nt) {
  for (int e = tid; e < kD * kD; e += nt) {
    const int r = e / kD, c = e % kD;
    double s = 0.0;
    for (int k = 0; k < kD; ++k) s += A[r * kLD + k] * B[k * kLD + c];
    C[r * kLD + c] = s;
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

// Coupled Newton-Schulz: Y0 = A/c, Z0 = I;  T = (3I - Z Y)/2;  Y <- Y T,
// Z <- T Z.  Z -> (A/c)^{-1/2} when the eigenvalues of A/c lie in (0, 1]
// (c = Gershgorin bound).  Stops when max|Z Y - I| < 1e-13 (status 0) or
// after 100 steps (status = 1).
__global__ __launch_bounds__(256) void k_local_targets(const double4* __restrict__ rec,
                                                       const int* __restrict__ nbr,
                                                       const int64_t* __restrict__ centers,
                                                       float* __restrict__ target9,
                                                       int* __restrict__ status) {
  __shared__ double buf[4][kD * kLD];
  __shared__ double3 P[kC];
  __shared__ double red[4];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int64_t ctr = centers[blockIdx.x];
  if (tid < kC) {
    int64_t j = (tid == 0) ? ctr : (int64_t)nbr[ctr * kPrecondK + tid - 1];
    if (j < 0) j = ctr;
    const double4 r = rec[j];
    P[tid] = double3{r.x, r.y, r.z};
  }
  __syncthreads();
  double* Y = buf[0];
  double* Z = buf[1];
  double* T = buf[2];
  double* S = buf[3];
  for (int e = tid; e < kC * kC; e += nt) {
    const int a = e / kC, b = e % kC;
    const double dx = P[a].x - P[b].x, dy = P[a].y - P[b].y, dz = P[a].z - P[b].z;
    for (int col = 0; col < 3; ++col) {
      PairAcc<double> acc;
      acc_zero(acc);
      open_pair<double, kUF>(dx, dy, dz, col == 0 ? 1.0 : 0.0, col == 1 ? 1.0 : 0.0,
                             col == 2 ? 1.0 : 0.0, 0.0, 0.0, 0.0, acc);
      Y[(3 * a + 0) * kLD + 3 * b + col] = acc.uf.x;
      Y[(3 * a + 1) * kLD + 3 * b + col] = acc.uf.y;
      Y[(3 * a + 2) * kLD + 3 * b + col] = acc.uf.z;
    }
  }
  __syncthreads();
  double g = 0.0;
  for (int r = tid; r < kD; r += nt) {
    double s = 0.0;
    for (int c = 0; c < kD; ++c) s += fabs(Y[r * kLD + c]);
    g = fmax(g, s);
  }
  const double cb = block_max(g, red);
  const double ic = 1.0 / cb;
  for (int e = tid; e < kD * kD; e += nt) {
    const int r = e / kD, c = e % kD;
    Y[r * kLD + c] *= ic;
    Z[r * kLD + c] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  int converged = 0;
  for (int it = 0; it < 100; ++it) {
    lds_matmul(Z, Y, T, tid, nt);  // T = Z Y
    __syncthreads();
    double err = 0.0;
    for (int e = tid; e < kD * kD; e += nt) {
      const int r = e / kD, c = e % kD;
      const double v = T[r * kLD + c];
      const double id = (r == c) ? 1.0 : 0.0;
      err = fmax(err, fabs(v - id));
      T[r * kLD + c] = 0.5 * (3.0 * id - v);
    }
    const double m = block_max(err, red);  // includes the barrier for T
    if (m < 1e-13) {
      converged = 1;
      break;
    }
    lds_matmul(Y, T, S, tid, nt);  // S = Y T
    __syncthreads();
    {
      double* t = Y;
      Y = S;
      S = t;
    }
    lds_matmul(T, Z, S, tid, nt);  // S = T Z
    __syncthreads();
    {
      double* t = Z;
      Z = S;
      S = t;
    }
  }
  // (A/t0)^{-1/2} = c^{-1/2} (A/(t0 c))^{-1/2}: scale the center block
  const double sc = sqrt(ic);
  if (tid < 9) target9[blockIdx.x * 9 + tid] = (float)(sc * Z[(tid / 3) * kLD + (tid % 3)]);
  if (tid == 0) status[blockIdx.x] = converged ? 0 : 1;
}

void launch_local_targets(const double4* rec, const int* nbr, const int64_t* centers,
                          int64_t ncenters, float* target9, int* status, hipStream_t s) {
  if (ncenters <= 0) return;
  hipLaunchKernelGGL(k_local_targets, dim3((unsigned)ncenters), dim3(256), 0, s, rec, nbr,
                     centers, target9, status);
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

// L = [[s0,0,0],[o3,s1,0],[o4,o5,s2]], s_i = softplus(o_i) + eps
__device__ inline void head_L(const float* o, float L[3][3]) {
  L[0][0] = softplus_f(o[0]) + kPrecondEps; L[0][1] = 0.f; L[0][2] = 0.f;
  L[1][0] = o[3]; L[1][1] = softplus_f(o[1]) + kPrecondEps; L[1][2] = 0.f;
  L[2][0] = o[4]; L[2][1] = o[5]; L[2][2] = softplus_f(o[2]) + kPrecondEps;
}

// loss = mean_b sum_ij (G - T)_ij^2 ; dO = d loss / d outputs (divided by B)
__global__ void k_head_loss(const float* __restrict__ out, const float* __restrict__ tgt9, int Bn,
                            float* __restrict__ dout, float* __restrict__ loss_part) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (b < Bn) {
    const float* o = out + 6 * b;
    float L[3][3];
    head_L(o, L);
    float G[3][3], R[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        float s = 0.f;
        for (int k = 0; k < 3; ++k) s += L[i][k] * L[j][k];
        G[i][j] = s;
        R[i][j] = s - tgt9[9 * b + 3 * i + j];
        l += R[i][j] * R[i][j];
      }
    // dloss/dG = 2 R (symmetric) ; dloss/dL = 2 (R + R^T) L = 4 R L
    float dL[3][3];
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) {
        float s = 0.f;
        for (int j = 0; j < 3; ++j) s += (R[i][j] + R[j][i]) * L[j][k];
        dL[i][k] = 2.f * s / (float)Bn;
      }
    float* d = dout + 6 * b;
    d[0] = dL[0][0] * sigmoid_f(o[0]);
    d[1] = dL[1][1] * sigmoid_f(o[1]);
    d[2] = dL[2][2] * sigmoid_f(o[2]);
    d[3] = dL[1][0];
    d[4] = dL[2][0];
    d[5] = dL[2][1];
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

void launch_head_loss(const float* out, const float* tgt9, int Bn, float* dout, float* loss_part,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_head_loss, dim3((unsigned)((Bn + 255) / 256)), dim3(256), 0, s, out, tgt9,
                     Bn, dout, loss_part);
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

// ---- per-particle blocks: G = L L^T and G^{-1} (FP64) ----------------------------
// Systems smaller than K+1 particles (no full neighbourhood) get G = I.
__global__ void k_make_blocks(const float* __restrict__ out, const int* __restrict__ nbr, int64_t N,
                              double* __restrict__ G9, double* __restrict__ Gi9) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double G[3][3];
  const bool full = nbr[i * kPrecondK + kPrecondK - 1] >= 0;
  if (full) {
    float Lf[3][3];
    head_L(out + 6 * i, Lf);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += (double)Lf[a][k] * (double)Lf[b][k];
        G[a][b] = s;
      }
  } else {
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) G[a][b] = (a == b) ? 1.0 : 0.0;
  }
  // inverse of the symmetric 3x3 by cofactors
  const double c00 = G[1][1] * G[2][2] - G[1][2] * G[2][1];
  const double c01 = G[1][2] * G[2][0] - G[1][0] * G[2][2];
  const double c02 = G[1][0] * G[2][1] - G[1][1] * G[2][0];
  const double det = G[0][0] * c00 + G[0][1] * c01 + G[0][2] * c02;
  const double id = 1.0 / det;
  double I[3][3];
  I[0][0] = c00 * id;
  I[0][1] = (G[0][2] * G[2][1] - G[0][1] * G[2][2]) * id;
  I[0][2] = (G[0][1] * G[1][2] - G[0][2] * G[1][1]) * id;
  I[1][0] = c01 * id;
  I[1][1] = (G[0][0] * G[2][2] - G[0][2] * G[2][0]) * id;
  I[1][2] = (G[0][2] * G[1][0] - G[0][0] * G[1][2]) * id;
  I[2][0] = c02 * id;
  I[2][1] = (G[0][1] * G[2][0] - G[0][0] * G[2][1]) * id;
  I[2][2] = (G[0][0] * G[1][1] - G[0][1] * G[1][0]) * id;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      G9[9 * i + 3 * a + b] = G[a][b];
      Gi9[9 * i + 3 * a + b] = I[a][b];
    }
}

void launch_make_blocks(const float* out, const int* nbr, int64_t N, double* G9, double* Gi9,
                        hipStream_t s) {
  if (N <= 0) return;
  hipLaunchKernelGGL(k_make_blocks, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, out, nbr, N,
                     G9, Gi9);
  LM_HIP(hipGetLastError());
}

// y[3i..3i+2] = B_i x[3i..3i+2] for i < N; entries beyond 3N (rotational
// part of a 6N vector) are copied unchanged.
__global__ void k_apply_blocks(const double* __restrict__ B9, const double* __restrict__ x,
                               int64_t N, int64_t n, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) {
    const double* b = B9 + 9 * i;
    const double x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
    y[3 * i] = b[0] * x0 + b[1] * x1 + b[2] * x2;
    y[3 * i + 1] = b[3] * x0 + b[4] * x1 + b[5] * x2;
    y[3 * i + 2] = b[6] * x0 + b[7] * x1 + b[8] * x2;
  }
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = 3 * N + i; k < n; k += stride) y[k] = x[k];
}

void launch_apply_blocks(const double* B9, const double* x, int64_t N, int64_t n, double* y,
                         hipStream_t s) {
  if (N <= 0) return;
  hipLaunchKernelGGL(k_apply_blocks, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, B9, x, N, n, y);
  LM_HIP(hipGetLastError());
}

}  // namespace lm
