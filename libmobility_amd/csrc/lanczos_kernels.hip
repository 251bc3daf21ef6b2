// GPU-resident Lanczos square root (after include/lanczos/LanczosAlgorithm.cu,
// Ando, Chow, Saad, Skolnick, J. Chem. Phys. 137, 064106 (2012)).
//
// Everything stays on the device: the Krylov basis V (column-major, n x m),
// the three-term recurrence scalars alpha/beta, the tridiagonal square root
// and the convergence estimate.  Reductions are single pass: each block
// writes a partial, the last block to arrive (agent-scope counter that it
// resets itself) sums the partials in a fixed order, so results are bitwise
// reproducible and no host round trip is needed until the convergence test.
//
// Tridiagonal square root (replaces host LAPACKE ?steqr + cblas ?gemv,
// LanczosAlgorithm.cu:48-85): for the SPD tridiagonal T,
//     T^{1/2} e1 = T T^{-1/2} e1,   T^{-1/2} = (2/pi) Int_R e^u (e^{2u} I + T)^{-1} du,
// evaluated with the trapezoidal rule after the substitution u = c + s sinh(t),
// which turns the exponential (sech) tails of the integrand into doubly
// exponential ones, so ~70-130 nodes reach ~1e-14.  Each node is one shifted
// tridiagonal solve by UL elimination (a division chain from the bottom, then
// a multiplication chain from the top), one node per thread; the spectrum
// bounds come from Gershgorin (upper) and Sturm-count multisection (lower).  A
// negative eigenvalue (non-SPD operator) yields NaN, which the convergence
// test reports as the reference does (LanczosAlgorithm.cu:217-220).
#include <cmath>

#include "kernels.hpp"
#include "lm_common.hpp"
#include "philox.hpp"

namespace lm {

// ---- Philox normal fill ------------------------------------------------------
__global__ void k_normal_fill(double* __restrict__ out, int64_t n, uint32_t k0, uint32_t k1,
                              uint32_t s0, uint32_t s1) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (2 * g >= n) return;
  U32x4 c;
  c.v[0] = (uint32_t)g;
  c.v[1] = (uint32_t)((uint64_t)g >> 32);
  c.v[2] = s0;
  c.v[3] = s1;
  const U32x4 r = philox4x32_10(c, k0, k1);
  const double u1 = philox_u01(r.v[0], r.v[1]);
  const double u2 = philox_u01(r.v[2], r.v[3]);
  const double rad = sqrt(-2.0 * log(u1));
  double sn, cs;
  sincospi(2.0 * u2, &sn, &cs);
  out[2 * g] = rad * cs;
  if (2 * g + 1 < n) out[2 * g + 1] = rad * sn;
}

void launch_normal_fill(double* out, int64_t n, uint64_t seed, uint64_t stream, hipStream_t s) {
  if (n <= 0) return;
  const int64_t pairs = (n + 1) / 2;
  const int nt = 256;
  hipLaunchKernelGGL(k_normal_fill, dim3((unsigned)((pairs + nt - 1) / nt)), dim3(nt), 0, s, out,
                     n, (uint32_t)seed, (uint32_t)(seed >> 32), (uint32_t)stream,
                     (uint32_t)(stream >> 32));
  LM_HIP(hipGetLastError());
}

// ---- single-pass reduction helpers -----------------------------------------
constexpr int kRT = 256;  // threads per reduction block

// Sum over the block; the result is valid in every thread of wave 0.
__device__ inline double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = (threadIdx.x < (blockDim.x >> 6)) ? sh[threadIdx.x] : 0.0;
  if (wid == 0)
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
  return t;
}

// Publish this block's partials; returns true in every thread of the last
// block to arrive, which then sees all partials.
__device__ inline bool publish_and_arrive(double* partial, int nvals, const double* vals,
                                          unsigned int* counter) {
  __shared__ bool s_last;
  if (threadIdx.x == 0) {
    for (int q = 0; q < nvals; ++q)
      __hip_atomic_store(&partial[q * kReduceBlocks + blockIdx.x], vals[q], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    const unsigned prev =
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  const bool last = s_last;
  if (last) __threadfence();
  return last;
}

// In the last block: fixed-order sum of the nb partials of slot q.
__device__ inline double gather_partials(const double* partial, int q, int nb, double* sh) {
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    v += __hip_atomic_load(&partial[q * kReduceBlocks + b], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
  return block_sum(v, sh);
}

static inline int reduce_blocks(int64_t n) {
  int64_t b = (n + kRT - 1) / kRT;
  if (b > kReduceBlocks) b = kReduceBlocks;
  if (b < 1) b = 1;
  return (int)b;
}

// ---- Lanczos steps ------------------------------------------------------------
// ||z|| -> sc->normz
__global__ __launch_bounds__(kRT) void k_norm_start(const double* __restrict__ z, int64_t n,
                                                   double* partial, unsigned* counter,
                                                   LanczosScalars* sc, int old) {
  __shared__ double sh[kRT / 64];
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRT)
    v += z[i] * z[i];
  v = block_sum(v, sh);
  if (publish_and_arrive(partial, 1, &v, counter)) {
    const double t = gather_partials(partial, 0, gridDim.x, sh);
    if (threadIdx.x == 0) {
      sc->normz = sqrt(t);
      LanczosStatus st{};
      st.conv = 0;
      st.conv_iter = -1;
      st.old = old;
      st.result = -1;
      st.first = -1;
      st.nchk = 0;
      sc->status = st;
      *counter = 0;
    }
  }
}

__global__ void k_scale_start(const double* __restrict__ z, double* __restrict__ v0, int64_t n,
                              const LanczosScalars* sc) {
  const double inv = 1.0 / sc->normz;  // LanczosAlgorithm.cu:115-117
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v0[i] = z[i] * inv;
}

void launch_lanczos_start(const double* z, double* v0, int64_t n, ReduceScratch rs,
                          LanczosScalars* sc, int old, hipStream_t s) {
  const int nb = reduce_blocks(n);
  hipLaunchKernelGGL(k_norm_start, dim3(nb), dim3(kRT), 0, s, z, n, rs.partial, rs.counter, sc,
                     old);
  LM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_scale_start, dim3(nb), dim3(kRT), 0, s, z, v0, n, sc);
  LM_HIP(hipGetLastError());
}

// w -= beta[i-1] v_prev ; alpha[i] = w . v_i   (LanczosAlgorithm.cu:128-134)
__global__ __launch_bounds__(kRT) void k_alpha(double* __restrict__ w,
                                              const double* __restrict__ vi,
                                              const double* __restrict__ vp, int64_t n, int i,
                                              double* partial, unsigned* counter,
                                              LanczosScalars* sc) {
  __shared__ double sh[kRT / 64];
  const double b = (i > 0) ? sc->beta[i - 1] : 0.0;
  double v = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * kRT + threadIdx.x; k < n; k += (int64_t)gridDim.x * kRT) {
    double wk = w[k];
    if (i > 0) {
      wk -= b * vp[k];
      w[k] = wk;
    }
    v += wk * vi[k];
  }
  v = block_sum(v, sh);
  if (publish_and_arrive(partial, 1, &v, counter)) {
    const double t = gather_partials(partial, 0, gridDim.x, sh);
    if (threadIdx.x == 0) {
      sc->alpha[i] = t;
      *counter = 0;
    }
  }
}

void launch_lanczos_alpha(double* w, const double* v_i, const double* v_prev, int64_t n, int i,
                          ReduceScratch rs, LanczosScalars* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_alpha, dim3(reduce_blocks(n)), dim3(kRT), 0, s, w, v_i, v_prev, n, i,
                     rs.partial, rs.counter, sc);
  LM_HIP(hipGetLastError());
}

// w -= alpha[i] v_i ; beta[i] = ||w||, zeroed below 1e-3 alpha/||z||
// (LanczosAlgorithm.cu:136-143)
__global__ __launch_bounds__(kRT) void k_beta(double* __restrict__ w,
                                             const double* __restrict__ vi, int64_t n, int i,
                                             double* partial, unsigned* counter,
                                             LanczosScalars* sc) {
  __shared__ double sh[kRT / 64];
  const double al = sc->alpha[i];
  double v = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * kRT + threadIdx.x; k < n; k += (int64_t)gridDim.x * kRT) {
    const double wk = w[k] - al * vi[k];
    w[k] = wk;
    v += wk * wk;
  }
  v = block_sum(v, sh);
  if (publish_and_arrive(partial, 1, &v, counter)) {
    const double t = gather_partials(partial, 0, gridDim.x, sh);
    if (threadIdx.x == 0) {
      double be = sqrt(t);
      const double tol = 1e-3 * sc->alpha[i] / sc->normz;
      if (be < tol) be = 0.0;
      sc->beta[i] = be;
      *counter = 0;
    }
  }
}

void launch_lanczos_beta(double* w, const double* v_i, int64_t n, int i, ReduceScratch rs,
                         LanczosScalars* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_beta, dim3(reduce_blocks(n)), dim3(kRT), 0, s, w, v_i, n, i, rs.partial,
                     rs.counter, sc);
  LM_HIP(hipGetLastError());
}

// v_{i+1} = w / beta[i], or e1 after a breakdown (LanczosAlgorithm.cu:144-152)
__global__ void k_next(const double* __restrict__ w, double* __restrict__ vn, int64_t n, int i,
                       const LanczosScalars* sc) {
  const double be = sc->beta[i];
  const double inv = be > 0.0 ? 1.0 / be : 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    vn[k] = be > 0.0 ? w[k] * inv : (k == 0 ? 1.0 : 0.0);
}

void launch_lanczos_next(const double* w, double* v_next, int64_t n, int i,
                         const LanczosScalars* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_next, dim3(reduce_blocks(n)), dim3(kRT), 0, s, w, v_next, n, i, sc);
  LM_HIP(hipGetLastError());
}

// Sum over a workgroup of kLanczosFusedThreads in a fixed order; every thread
// gets the total.  `sh` holds one double per wave plus one.
__device__ inline double group_sum_all(double v, double* sh) {
  constexpr int kW = kLanczosFusedThreads / 64;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < kW; ++k) t += sh[k];
    sh[kW] = t;
  }
  __syncthreads();
  const double t = sh[kW];
  __syncthreads();  // sh is reused by the next sum
  return t;
}

// The step of LanczosAlgorithm.cu:128-152 in one launch.  Every workgroup
// forms its entries of w = M v_i (summing the chunk slabs) - beta[i-1]
// v_{i-1}, stores them and publishes its part of alpha = w . v_i; the last
// workgroup to arrive finishes alone (the vectors are short): w -= alpha v_i,
// beta = ||w|| with the breakdown rule, v_next = w / beta or e1.
__global__ __launch_bounds__(kLanczosFusedThreads) void k_lanczos_step(LanczosStepArgs a) {
  __shared__ double sh[kLanczosFusedThreads / 64 + 1];
  constexpr int R = kLanczosFusedPerThread;
  const int t = threadIdx.x;
  const int64_t n = a.n;
  const int i = a.i;
  LanczosScalars* sc = a.sc;
  __shared__ int s_stop;
  if (t == 0) s_stop = a.stop != nullptr ? *(volatile const int*)a.stop : 0;
  __syncthreads();
  const bool skip = s_stop != 0;  // the same in every thread of the workgroup
  const double bprev = i > 0 ? sc->beta[i - 1] : 0.0;
  double p = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * kLanczosFusedThreads + t; !skip && k < n;
       k += (int64_t)gridDim.x * kLanczosFusedThreads) {
    double v = 0.0;
    if (a.w) {
      v = a.w[k];
    } else {
      const double* sl = k < a.n3 ? a.slab_u + k : a.slab_w + (k - a.n3);
      const int nch = a.nch;
      // the first 16 slabs' loads all in flight at once, then added in order
      constexpr int kQ = 16;
      double q[kQ];
#pragma unroll
      for (int c = 0; c < kQ; ++c) q[c] = c < nch ? sl[(int64_t)c * a.n3] : 0.0;
#pragma unroll
      for (int c = 0; c < kQ; ++c)
        if (c < nch) v += q[c];
      for (int c = kQ; c < nch; ++c) v += sl[(int64_t)c * a.n3];
    }
    if (i > 0) v -= bprev * a.v_prev[k];
    a.w_out[k] = v;
    p += v * a.v_i[k];
  }
  const double part = group_sum_all(p, sh);
  __shared__ bool s_last;
  if (t == 0) {
    __hip_atomic_store(&a.partial[blockIdx.x], part, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    const unsigned prev =
        __hip_atomic_fetch_add(a.counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return;
  if (skip) {
    if (t == 0) *a.counter = 0;
    return;
  }
  __threadfence();
  double q = 0.0;
  for (int b = t; b < (int)gridDim.x; b += kLanczosFusedThreads)
    q += __hip_atomic_load(&a.partial[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const double alpha = group_sum_all(q, sh);
  double w[R];
  p = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t k = t + (int64_t)r * kLanczosFusedThreads;
    w[r] = 0.0;
    if (k < n) {
      w[r] = a.w_out[k] - alpha * a.v_i[k];
      p += w[r] * w[r];
    }
  }
  double beta = sqrt(group_sum_all(p, sh));
  if (beta < 1e-3 * alpha / sc->normz) beta = 0.0;  // LanczosAlgorithm.cu:140-142
  if (t == 0) {
    sc->alpha[i] = alpha;
    sc->beta[i] = beta;
    *a.counter = 0;
  }
  const double inv = beta > 0.0 ? 1.0 / beta : 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t k = t + (int64_t)r * kLanczosFusedThreads;
    if (k < n) a.v_next[k] = beta > 0.0 ? w[r] * inv : (k == 0 ? 1.0 : 0.0);
  }
}

// The same step in one workgroup for the shortest vectors (n <= kStep1Max):
// no arrival counter and no device-scope fences (those alone cost about 5 µs
// per launch on gfx950).  Every slab load of a thread's entries is issued
// before the first add, so the sum costs one memory round trip; w stays in
// registers throughout.
constexpr int kStep1R = 4;
constexpr int64_t kStep1Max = (int64_t)kLanczosFusedThreads * kStep1R;

__global__ __launch_bounds__(kLanczosFusedThreads) void k_lanczos_step_1wg(LanczosStepArgs a) {
  __shared__ double sh[kLanczosFusedThreads / 64 + 1];
  __shared__ int s_stop;
  const int t = threadIdx.x;
  if (t == 0) s_stop = a.stop != nullptr ? *(volatile const int*)a.stop : 0;
  __syncthreads();
  if (s_stop != 0) return;  // past convergence (one workgroup: nothing to keep consistent)
  const int64_t n = a.n;
  const int i = a.i;
  LanczosScalars* sc = a.sc;
  constexpr int kQ = 6;  // slabs per round: kStep1R x kQ loads in flight per thread
  double w[kStep1R];
  if (a.w) {
#pragma unroll
    for (int r = 0; r < kStep1R; ++r) {
      const int64_t k = t + (int64_t)r * kLanczosFusedThreads;
      w[r] = k < n ? a.w[k] : 0.0;
    }
  } else {
    const int nch = a.nch;
#pragma unroll
    for (int r = 0; r < kStep1R; ++r) w[r] = 0.0;
    for (int c0 = 0; c0 < nch; c0 += kQ) {
      double q[kStep1R][kQ];
#pragma unroll
      for (int r = 0; r < kStep1R; ++r) {
        const int64_t k = t + (int64_t)r * kLanczosFusedThreads;
        const double* sl = k < a.n3 ? a.slab_u + k : a.slab_w + (k - a.n3);
#pragma unroll
        for (int c = 0; c < kQ; ++c)
          q[r][c] = (k < n && c0 + c < nch) ? sl[(int64_t)(c0 + c) * a.n3] : 0.0;
      }
#pragma unroll
      for (int r = 0; r < kStep1R; ++r)
#pragma unroll
        for (int c = 0; c < kQ; ++c)
          if (c0 + c < nch) w[r] += q[r][c];
    }
  }
  // w -= beta[i-1] v_{i-1}; alpha = w . v_i
  const double bprev = i > 0 ? sc->beta[i - 1] : 0.0;
  double vi[kStep1R];
  double p = 0.0;
#pragma unroll
  for (int r = 0; r < kStep1R; ++r) {
    const int64_t k = t + (int64_t)r * kLanczosFusedThreads;
    vi[r] = 0.0;
    if (k < n) {
      vi[r] = a.v_i[k];
      if (i > 0) w[r] -= bprev * a.v_prev[k];
      p += w[r] * vi[r];
    }
  }
  const double alpha = group_sum_all(p, sh);
  p = 0.0;
#pragma unroll
  for (int r = 0; r < kStep1R; ++r) {
    const int64_t k = t + (int64_t)r * kLanczosFusedThreads;
    if (k < n) {
      w[r] -= alpha * vi[r];
      p += w[r] * w[r];
    }
  }
  double beta = sqrt(group_sum_all(p, sh));
  if (beta < 1e-3 * alpha / sc->normz) beta = 0.0;  // LanczosAlgorithm.cu:140-142
  if (t == 0) {
    sc->alpha[i] = alpha;
    sc->beta[i] = beta;
  }
  const double inv = beta > 0.0 ? 1.0 / beta : 0.0;
#pragma unroll
  for (int r = 0; r < kStep1R; ++r) {
    const int64_t k = t + (int64_t)r * kLanczosFusedThreads;
    if (k < n) a.v_next[k] = beta > 0.0 ? w[r] * inv : (k == 0 ? 1.0 : 0.0);
  }
}

void launch_lanczos_step(const LanczosStepArgs& a, hipStream_t s) {
  if (a.n <= 0 || a.n > kLanczosFusedMaxN)
    throw Error("[libmobility_amd] fused Lanczos step: vector length out of range");
  if (!a.w && (a.nch < 1 || !a.slab_u || (a.n > a.n3 && !a.slab_w)))
    throw Error("[libmobility_amd] fused Lanczos step: no product input");
  if (a.n <= kStep1Max) {
    hipLaunchKernelGGL(k_lanczos_step_1wg, dim3(1), dim3(kLanczosFusedThreads), 0, s, a);
    LM_HIP(hipGetLastError());
    return;
  }
  if (!a.w_out || !a.partial || !a.counter)
    throw Error("[libmobility_amd] fused Lanczos step: no workspace");
  const int grid = (int)((a.n + kLanczosFusedThreads - 1) / kLanczosFusedThreads);
  hipLaunchKernelGGL(k_lanczos_step, dim3(grid), dim3(kLanczosFusedThreads), 0, s, a);
  LM_HIP(hipGetLastError());
}

// ---- full reorthogonalisation (extension) -----------------------------------------
// Four launches: projections c_j = v_j . v (partial dots per row slice, then
// their fixed-order sums), v -= V c with ||v||^2, and the rescale.  All read
// beta[i] first and do nothing after a breakdown.
// Projections: blockIdx.x takes a slice of rows, blockIdx.y a group of
// kRoCols columns; each thread loads its rows' entries of v once and
// multiplies them into kRoCols running dots, so V is read once.
constexpr int kRoCols = 8;

__global__ __launch_bounds__(kRT) void k_reorth_dots(const double* __restrict__ V, int64_t ldv,
                                                     int m, const double* __restrict__ v,
                                                     int64_t n, int i, double* __restrict__ partial,
                                                     const LanczosScalars* sc) {
  __shared__ double red[kRT / 64][kRoCols];
  if (sc->beta[i] == 0.0) return;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int rb = blockIdx.x, nrb = gridDim.x;
  const int j0 = blockIdx.y * kRoCols;
  const int64_t per = (n + nrb - 1) / nrb;
  const int64_t r0 = (int64_t)rb * per;
  const int64_t r1 = r0 + per < n ? r0 + per : n;
  double acc[kRoCols];
#pragma unroll
  for (int q = 0; q < kRoCols; ++q) acc[q] = 0.0;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += kRT) {
    const double x = v[r];
#pragma unroll
    for (int q = 0; q < kRoCols; ++q)
      if (j0 + q < m) acc[q] += V[(int64_t)(j0 + q) * ldv + r] * x;
  }
#pragma unroll
  for (int q = 0; q < kRoCols; ++q) {
    double t = acc[q];
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
    if (lane == 0) red[wid][q] = t;
  }
  __syncthreads();
  if (threadIdx.x < kRoCols && j0 + (int)threadIdx.x < m) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kRT / 64; ++w) t += red[w][threadIdx.x];
    partial[(int64_t)(j0 + threadIdx.x) * kReduceBlocks + rb] = t;
  }
}

// c_j = fixed-order sum of the row slices' partial dots (the launch boundary
// orders it after every slice; no arrival counter, which thousands of
// workgroups would contend on).
__global__ __launch_bounds__(kRT) void k_reorth_sum(const double* __restrict__ partial, int m,
                                                    int nrb, int i, double* __restrict__ c,
                                                    const LanczosScalars* sc) {
  if (sc->beta[i] == 0.0) return;
  for (int j = blockIdx.x * kRT + threadIdx.x; j < m; j += gridDim.x * kRT) {
    double t = 0.0;
    for (int b = 0; b < nrb; ++b) t += partial[(int64_t)j * kReduceBlocks + b];
    c[j] = t;
  }
}

__global__ __launch_bounds__(kRT) void k_reorth_apply(const double* __restrict__ V, int64_t ldv,
                                                      int m, double* __restrict__ v, int64_t n,
                                                      int i, const double* __restrict__ c,
                                                      double* partial, unsigned* counter,
                                                      LanczosScalars* sc) {
  __shared__ double cs[kLanczosMaxIterations];
  __shared__ double red[kRT / 64][64];
  __shared__ double sh[kRT / 64];
  if (sc->beta[i] == 0.0) return;
  for (int j = threadIdx.x; j < m; j += kRT) cs[j] = c[j];
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double p = 0.0;
  for (int64_t tile = blockIdx.x; tile * 64 < n; tile += gridDim.x) {
    const int64_t row = tile * 64 + lane;
    const bool ok = row < n;
    double acc = 0.0;
#pragma unroll 4
    for (int j = wid; j < m; j += kRT / 64) acc += (ok ? V[(int64_t)j * ldv + row] : 0.0) * cs[j];
    red[wid][lane] = acc;
    __syncthreads();
    if (wid == 0 && ok) {
      double y = 0.0;
#pragma unroll
      for (int w = 0; w < kRT / 64; ++w) y += red[w][lane];
      const double x = v[row] - y;
      v[row] = x;
      p += x * x;
    }
    __syncthreads();
  }
  p = block_sum(p, sh);
  if (publish_and_arrive(partial, 1, &p, counter)) {
    const double t = gather_partials(partial, 0, gridDim.x, sh);
    if (threadIdx.x == 0) {
      const double s = sqrt(t);
      double be = sc->beta[i] * s;
      if (be < 1e-3 * sc->alpha[i] / sc->normz) be = 0.0;  // LanczosAlgorithm.cu:140-142
      sc->beta[i] = be;
      sc->rscale = be > 0.0 ? 1.0 / s : 0.0;
      *counter = 0;
    }
  }
}

__global__ void k_reorth_scale(double* __restrict__ v, int64_t n, int i,
                               const LanczosScalars* sc) {
  const double be = sc->beta[i];
  const double f = sc->rscale;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    v[k] = be > 0.0 ? v[k] * f : (k == 0 ? 1.0 : 0.0);
}

void launch_lanczos_reorth(const double* V, int64_t ldv, int i, int64_t n, double* partial,
                           unsigned int* counter, double* c, LanczosScalars* sc, hipStream_t s) {
  const int m = i + 1;
  if (i < 0 || m > kLanczosMaxIterations || n <= 0)
    throw Error("[libmobility_amd] reorthogonalisation: size out of range");
  double* v = const_cast<double*>(V) + (int64_t)m * ldv;
  const int g = reduce_blocks(n);
  const int cg = (m + kRoCols - 1) / kRoCols;
  hipLaunchKernelGGL(k_reorth_dots, dim3(g, cg), dim3(kRT), 0, s, V, ldv, m, v, n, i, partial,
                     sc);
  LM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_reorth_sum, dim3((m + kRT - 1) / kRT), dim3(kRT), 0, s, partial, m, g, i, c,
                     sc);
  LM_HIP(hipGetLastError());
  const int64_t tiles = (n + 63) / 64;
  const int ga = (int)(tiles < kReduceBlocks ? tiles : kReduceBlocks);
  hipLaunchKernelGGL(k_reorth_apply, dim3(ga), dim3(kRT), 0, s, V, ldv, m, v, n, i, c, partial,
                     counter, sc);
  LM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_reorth_scale, dim3(g), dim3(kRT), 0, s, v, n, i, sc);
  LM_HIP(hipGetLastError());
}

// ---- tridiagonal square root (one workgroup) -----------------------------------
constexpr int kTS = kTsqrtNodes;  // threads: bracket points, the first K are quadrature nodes
constexpr int kTSWaves = kTS / 64;
constexpr int kTSChunk = 16;      // multipliers prefetched per step of the top-down pass

// 1/x to within an ulp: the hardware reciprocal refined by two Newton steps
// (each doubles the correct bits).
__device__ inline double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// Number of eigenvalues of T below x (Sturm sequence of the LDL^T pivots).
__device__ inline int sturm_count(const double* a, const double* b2, int m, double x,
                                  double pivmin) {
  int cnt = 0;
  double q = a[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  cnt += q < 0.0;
  for (int i = 1; i < m; ++i) {
    q = (a[i] - x) - b2[i - 1] * rcp64(q);
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
  }
  return cnt;
}

// s_out[0..m) = T^{1/2} e1 for T = tridiag(beta, alpha, beta); one workgroup of
// kTS threads, `work` = tsqrt_scratch(m) doubles.
__device__ void tridiag_sqrt_body(const double* __restrict__ alpha,
                                  const double* __restrict__ beta, int m,
                                  double* __restrict__ s_out, double* __restrict__ work) {
  __shared__ double a[kLanczosMaxIterations], b[kLanczosMaxIterations],
      b2[kLanczosMaxIterations];
  __shared__ double rhi[kTSWaves], rb[kTSWaves];
  __shared__ int ired[kTSWaves];
  __shared__ double s_hi, s_pivmin;
  __shared__ int s_bad;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  for (int i = t; i < m; i += kTS) {
    a[i] = alpha[i];
    b[i] = (i < m - 1) ? beta[i] : 0.0;
    b2[i] = b[i] * b[i];
  }
  __syncthreads();
  // Gershgorin upper bound of the spectrum
  double ghi = -1e300, bmax = 0.0;
  for (int i = t; i < m; i += kTS) {
    const double r = (i > 0 ? fabs(b[i - 1]) : 0.0) + (i < m - 1 ? fabs(b[i]) : 0.0);
    ghi = fmax(ghi, a[i] + r);
    bmax = fmax(bmax, b2[i]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    ghi = fmax(ghi, __shfl_xor(ghi, o));
    bmax = fmax(bmax, __shfl_xor(bmax, o));
  }
  if (lane == 0) {
    rhi[wid] = ghi;
    rb[wid] = bmax;
  }
  __syncthreads();
  if (t == 0) {
    double hi = rhi[0], bm = rb[0];
    for (int w = 1; w < kTSWaves; ++w) {
      hi = fmax(hi, rhi[w]);
      bm = fmax(bm, rb[w]);
    }
    s_hi = hi + 1e-12 * fabs(hi) + 1e-300;
    s_pivmin = 1e-290 * fmax(1.0, bm);
  }
  __syncthreads();
  const double lmax = s_hi, pivmin = s_pivmin;
  // One round brackets the smallest eigenvalue: thread 0 counts the
  // eigenvalues below 0 (a non-SPD T gives NaN), threads 1..kTS-1 the
  // eigenvalues below x_t = lmax 1e-16^{1 - t/(kTS-1)}, log-spaced 7.5% apart
  // from the floor 1e-16 lmax (eigenvalues below it, a numerically singular
  // operator, are treated as lying on it) up to lmax.
  const double floor_l = 1e-16 * lmax;
  const double xt = (t == 0) ? 0.0 : lmax * exp(-36.841361487904734 * (1.0 - (double)t / (kTS - 1)));
  const int cnt = sturm_count(a, b2, m, xt, pivmin);
  int first = (t > 0 && cnt >= 1) ? t : kTS;
  for (int o = 32; o > 0; o >>= 1) first = min(first, __shfl_xor(first, o));
  if (lane == 0) ired[wid] = first;
  if (t == 0) s_bad = cnt > 0;
  __syncthreads();
  if (s_bad) {
    for (int i = t; i < m; i += kTS) s_out[i] = NAN;
    return;
  }
  int f = ired[0];
  for (int w = 1; w < kTSWaves; ++w) f = min(f, ired[w]);
  const double lmin =
      (f <= 1 || f >= kTS) ? floor_l
                            : fmax(floor_l, lmax * exp(-36.841361487904734 * (1.0 - (double)(f - 1) / (kTS - 1))));
  // Nodes u_k = c + s sinh(t_k), t_k = (k - half) h: the map makes the sech
  // tails of the integrand decay double-exponentially; s and h grow and
  // shrink with log10 of the condition number, from 71 nodes at kappa 1e2 to
  // 257 at 1e16, with a scalar quadrature error below 2e-13 over the whole
  // spectrum.
  const double lk = log10(fmax(lmax / lmin, 1.0));
  const double sc_s = 1.2 + 0.2 * lk;
  double h = 0.18 / (1.0 + 0.3 * lk);
  int half = (int)floor(4.0 / h);
  if (2 * half + 1 > kTS) {
    half = (kTS - 1) / 2;
    h = 4.0 / (double)half;
  }
  const int K = 2 * half + 1;
  const int nw = (K + 63) / 64;  // waves holding nodes
  const double c0 = 0.25 * (log(lmin) + log(lmax));
  // T^{1/2} e1 = sum_k w_k T x_k with (sig_k I + T) x_k = e1.  Each node forms
  // its own T x_k: as e1 - sig_k x_k below the geometric middle of the
  // spectrum and as the tridiagonal product above it.  (Summing the x_k first
  // and multiplying by T once amplifies the solves' rounding by about
  // sqrt(kappa): 2e-10 instead of 2e-13 at kappa 1e16.)
  const double tau = sqrt(lmin * lmax);
  double* q = work;                          // [m][kTS]: UL multipliers
  double* part = work + (int64_t)m * kTS;    // [m][kTSWaves]: per-wave node sums
  const bool node = t < K;
  double sig = 0.0, wgt = 0.0, rg = 0.0;
  if (node) {
    const double tk = (double)(t - half) * h;
    const double u = c0 + sc_s * sinh(tk);
    sig = exp(2.0 * u);
    wgt = (2.0 * h / M_PI) * exp(u) * sc_s * cosh(tk);
    // (sig I + T) x = e1 by UL elimination: g_{m-1} = sig + a_{m-1},
    // g_i = sig + a_i - b_i^2 / g_{i+1}; x_0 = 1/g_0, x_i = q_i x_{i-1} with
    // q_i = -b_{i-1} / g_i.
    rg = rcp64(sig + a[m - 1]);
    if (m > 1) q[(int64_t)(m - 1) * kTS + t] = -b[m - 2] * rg;
    for (int i = m - 2; i >= 0; --i) {
      rg = rcp64(sig + a[i] - b2[i] * rg);
      if (i >= 1) q[(int64_t)i * kTS + t] = -b[i - 1] * rg;
    }
  }
  const bool small = sig < tau;
  // Top-down pass: step i forms x_i = q_i x_{i-1} and emits (T x)_{i-1}; the
  // node sums of each step are reduced per wave in a fixed order.
  double xp2 = 0.0, xp = rg;  // x_{i-2}, x_{i-1}
  for (int i0 = 1; i0 <= m; i0 += kTSChunk) {
    double qv[kTSChunk], yv[kTSChunk];
#pragma unroll
    for (int k = 0; k < kTSChunk; ++k) {
      const int i = i0 + k;
      qv[k] = (node && i < m) ? q[(int64_t)i * kTS + t] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < kTSChunk; ++k) {
      const int i = i0 + k;
      yv[k] = 0.0;
      if (i <= m) {
        const int j = i - 1;
        const double xi = (i < m) ? xp * qv[k] : 0.0;
        double y;
        if (small) {
          y = (j == 0 ? 1.0 : 0.0) - sig * xp;
        } else {
          y = a[j] * xp + (i < m ? b[j] * xi : 0.0);
          if (j >= 1) y += b[j - 1] * xp2;
        }
        yv[k] = wgt * y;
        xp2 = xp;
        xp = xi;
      }
    }
#pragma unroll
    for (int k = 0; k < kTSChunk; ++k)
      for (int o = 32; o > 0; o >>= 1) yv[k] += __shfl_xor(yv[k], o);
    if (lane == 0 && wid < nw) {
#pragma unroll
      for (int k = 0; k < kTSChunk; ++k)
        if (i0 + k <= m) part[(int64_t)(i0 + k - 1) * kTSWaves + wid] = yv[k];
    }
  }
  __syncthreads();
  for (int j = t; j < m; j += kTS) {
    double v = 0.0;
    for (int w = 0; w < nw; ++w) v += part[(int64_t)j * kTSWaves + w];
    s_out[j] = v;
  }
}

// ---- convergence checks, up to kCheckBatch per launch ---------------------------
// One workgroup per check j: coef row j = T_{m0+j}^{1/2} e1, zero-padded to
// m0 + c - 1 entries so that every row spans the batch's widest estimate.
__global__ __launch_bounds__(kTS) void k_tridiag_sqrt_batch(const LanczosScalars* __restrict__ sc,
                                                            int m0, int c,
                                                            double* __restrict__ coef,
                                                            double* __restrict__ work,
                                                            int64_t wstride) {
  if (sc->status.conv) return;  // an earlier check converged
  const int j = blockIdx.x;
  const int m = m0 + j;
  double* row = coef + (int64_t)j * kLanczosMaxIterations;
  tridiag_sqrt_body(sc->alpha, sc->beta, m, row, work + (int64_t)j * wstride);
  for (int k = m + threadIdx.x; k < m0 + c - 1; k += kTS) row[k] = 0.0;
}

__global__ __launch_bounds__(kTS) void k_tridiag_sqrt_arrays(const double* __restrict__ alpha,
                                                             const double* __restrict__ beta,
                                                             int m, double* __restrict__ s_out,
                                                             double* __restrict__ work) {
  tridiag_sqrt_body(alpha, beta, m, s_out, work);
}

void tridiag_sqrt_selftest(const double* alpha, const double* beta, int m, double* out) {
  if (m < 1 || m > kLanczosMaxIterations)
    throw Error("[libmobility_amd] tridiagonal size out of range");
  DevBuf da, db, dout, dwork;
  da.reserve(sizeof(double) * m);
  db.reserve(sizeof(double) * m);
  dout.reserve(sizeof(double) * m);
  dwork.reserve(sizeof(double) * tsqrt_scratch(m));
  LM_HIP(hipMemcpy(da.as<double>(), alpha, sizeof(double) * m, hipMemcpyHostToDevice));
  LM_HIP(hipMemcpy(db.as<double>(), beta, sizeof(double) * (m > 1 ? m - 1 : 1),
                   hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_tridiag_sqrt_arrays, dim3(1), dim3(kTS), 0, nullptr, da.as<double>(),
                     db.as<double>(), m, dout.as<double>(), dwork.as<double>());
  LM_HIP(hipGetLastError());
  LM_HIP(hipMemcpy(out, dout.as<double>(), sizeof(double) * m, hipMemcpyDeviceToHost));
}

// One workgroup per listed system (at most kSegTsqrtGroups at a time, each
// with its own scratch slice).
__global__ __launch_bounds__(kTS) void k_seg_tridiag_sqrt(const int* __restrict__ list, int nlist,
                                                          int m, SegScalars sc,
                                                          double* __restrict__ coef,
                                                          double* __restrict__ work) {
  double* my_work = work + (int64_t)blockIdx.x * tsqrt_scratch(m);
  for (int j = blockIdx.x; j < nlist; j += gridDim.x) {
    const int sys = list[j];
    const int64_t row = (int64_t)sys * sc.L;
    tridiag_sqrt_body(sc.alpha + row, sc.beta + row, m, coef + row, my_work);
    __syncthreads();  // shared arrays and the scratch slice are reused
  }
}

void launch_seg_tridiag_sqrt(const int* list, int nlist, int m, const SegScalars& sc,
                             double* coef, double* work, hipStream_t s) {
  if (nlist <= 0) return;
  if (m < 1 || m > sc.L || m > kLanczosMaxIterations)
    throw Error("[libmobility_amd] tridiagonal size out of range");
  const int grid = nlist < kSegTsqrtGroups ? nlist : kSegTsqrtGroups;
  hipLaunchKernelGGL(k_seg_tridiag_sqrt, dim3(grid), dim3(kTS), 0, s, list, nlist, m, sc, coef,
                     work);
  LM_HIP(hipGetLastError());
}

// Estimates of a batch of checks: y_j = normz V[:, 0:m_j] coef_j for the c
// checks at once (each entry of V read once), each compared with the one
// before it (LanczosAlgorithm.cu:158-167 and :202-222), then the decision of
// LanczosAlgorithm.cu:181-197 in check order.  A tile is 64 rows, one per
// lane; the workgroup's four waves split the columns and meet in LDS.
constexpr int kEstRows = 64;
constexpr int kEstWaves = kRT / 64;

__global__ __launch_bounds__(kRT) void k_estimate_batch(CheckArgs a) {
  __shared__ double cs[kCheckBatch][kLanczosMaxIterations];
  __shared__ double red[kEstWaves][kCheckBatch][kEstRows];
  __shared__ double sh[kRT / 64];
  LanczosScalars* sc = a.sc;
  if (sc->status.conv) return;  // an earlier check converged (same value in every block)
  const int c = a.c;
  const int mm = a.i0 + c;  // columns of the widest estimate
  const int o = sc->status.old;
  const int ns = a.nslots;
  const double nz = sc->normz;
  for (int idx = threadIdx.x; idx < kCheckBatch * mm; idx += kRT) {
    const int j = idx / mm, k = idx - j * mm;
    cs[j][k] = j < c ? nz * a.coef[(int64_t)j * kLanczosMaxIterations + k] : 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double num[kCheckBatch], den[kCheckBatch];
#pragma unroll
  for (int j = 0; j < kCheckBatch; ++j) num[j] = den[j] = 0.0;
  const double* __restrict__ yold = a.Y + (int64_t)o * a.ystride;
  for (int64_t tile = blockIdx.x; tile * kEstRows < a.n; tile += gridDim.x) {
    const int64_t row = tile * kEstRows + lane;
    const bool ok = row < a.n;
    double acc[kCheckBatch];
#pragma unroll
    for (int j = 0; j < kCheckBatch; ++j) acc[j] = 0.0;
#pragma unroll 4
    for (int k = wid; k < mm; k += kEstWaves) {
      const double v = ok ? a.V[(int64_t)k * a.ldv + row] : 0.0;
#pragma unroll
      for (int j = 0; j < kCheckBatch; ++j) acc[j] += v * cs[j][k];
    }
#pragma unroll
    for (int j = 0; j < kCheckBatch; ++j) red[wid][j][lane] = acc[j];
    __syncthreads();
    if (wid == 0 && ok) {
      double prev = yold[row];
#pragma unroll
      for (int j = 0; j < kCheckBatch; ++j) {
        if (j < c) {
          double y = 0.0;
#pragma unroll
          for (int w = 0; w < kEstWaves; ++w) y += red[w][j][lane];
          const double d = prev - y;
          num[j] += d * d;
          den[j] += prev * prev;
          a.Y[(int64_t)((o + 1 + j) % ns) * a.ystride + row] = y;
          prev = y;
        }
      }
    }
    __syncthreads();
  }
  double vals[2 * kCheckBatch];
#pragma unroll
  for (int j = 0; j < kCheckBatch; ++j) {
    vals[2 * j] = block_sum(num[j], sh);
    vals[2 * j + 1] = block_sum(den[j], sh);
  }
  // every slot is published (fixed count: no dynamically indexed arrays)
  if (publish_and_arrive(a.rs.partial, 2 * kCheckBatch, vals, a.rs.counter)) {
    double tn[kCheckBatch], td[kCheckBatch];
#pragma unroll
    for (int j = 0; j < kCheckBatch; ++j) {
      tn[j] = gather_partials(a.rs.partial, 2 * j, gridDim.x, sh);
      td[j] = gather_partials(a.rs.partial, 2 * j + 1, gridDim.x, sh);
    }
    if (threadIdx.x == 0) {
      LanczosStatus st = sc->status;
      st.first = a.i0;
      st.nchk = c;
      bool done = false;
#pragma unroll
      for (int j = 0; j < kCheckBatch; ++j) {
        if (j < c && !done) {
          const int it = a.i0 + j;
          const double err = fabs(sqrt(tn[j]) / sqrt(td[j]));
          st.err[j] = err;
          if (it > 0 && isnan(err)) {  // LanczosAlgorithm.cu:217-220
            st.conv = 2;
            st.conv_iter = it;
            st.nchk = j + 1;
            done = true;
          } else if (it > 0 && err <= a.tol) {  // cu:185-192
            st.conv = 1;
            st.conv_iter = it;
            st.result = (o + 1 + j) % ns;
            st.old = j == 0 ? o : (o + j) % ns;  // becomes y_{j-1} - y_j (cu:209-212)
            st.nchk = j + 1;
            done = true;
          }
        }
      }
      if (st.conv == 0) st.old = (o + c) % ns;  // oldBz <- Bz (cu:196)
      sc->status = st;
      if (a.mirror) *a.mirror = st;
      *a.rs.counter = 0;
    }
  }
}

// The batch that converged leaves in slot `old` the reference's oldBz after
// computeError: previous estimate minus result.
__global__ void k_check_post(CheckArgs a) {
  const LanczosStatus& st = a.sc->status;
  if (st.conv != 1 || st.conv_iter < a.i0 || st.conv_iter >= a.i0 + a.c) return;
  double* __restrict__ dst = a.Y + (int64_t)st.old * a.ystride;
  const double* __restrict__ y = a.Y + (int64_t)st.result * a.ystride;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.n;
       k += (int64_t)gridDim.x * blockDim.x)
    dst[k] -= y[k];
}

void launch_lanczos_checks(const CheckArgs& a, hipStream_t s) {
  if (a.c < 1 || a.c > kCheckBatch || a.i0 < 0 || a.i0 + a.c > kLanczosMaxIterations)
    throw Error("[libmobility_amd] convergence check batch out of range");
  if (a.nslots < a.c + 1 || a.ystride < a.n || a.n <= 0)
    throw Error("[libmobility_amd] convergence check: estimate ring too small");
  const int mmax = a.i0 + a.c;
  hipLaunchKernelGGL(k_tridiag_sqrt_batch, dim3(a.c), dim3(kTS), 0, s, a.sc, a.i0 + 1, a.c,
                     a.coef, a.work, tsqrt_scratch(mmax));
  LM_HIP(hipGetLastError());
  const int64_t tiles = (a.n + kEstRows - 1) / kEstRows;
  const int grid = (int)(tiles < kReduceBlocks ? tiles : kReduceBlocks);
  hipLaunchKernelGGL(k_estimate_batch, dim3(grid), dim3(kRT), 0, s, a);
  LM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_check_post, dim3(reduce_blocks(a.n)), dim3(kRT), 0, s, a);
  LM_HIP(hipGetLastError());
}

// ---- independent recurrences per system (segmented kernels) --------------------
// Each workgroup owns one SegBlock: the vector entries of particles [p0, p1)
// of one system.  Partial sums go to partial[2 * block + q]; a finalising
// kernel (one wave per system) adds a system's partials in block order, so the
// sums are reproducible.  Every kernel skips the blocks of inactive systems.

// entry j of a block (0 <= j < 2c for 6N vectors, c = 3 (p1 - p0))
__device__ inline int64_t seg_entry(const SegBlock& b, int64_t j, int64_t N3) {
  const int64_t c = 3 * (b.p1 - b.p0);
  return j < c ? 3 * b.p0 + j : N3 + 3 * b.p0 + (j - c);
}

__device__ inline int64_t seg_count(const SegBlock& b, int angular) {
  return 3 * (b.p1 - b.p0) * (angular ? 2 : 1);
}

// partial[2*blk + 0] = sum z^2
__global__ __launch_bounds__(kRT) void k_seg_norm(const double* __restrict__ z, SegLayout g,
                                                 SegScalars sc) {
  __shared__ double sh[kRT / 64];
  const SegBlock b = g.blocks[blockIdx.x];
  if (!sc.active[b.sys]) return;
  const int64_t cnt = seg_count(b, g.angular);
  double v = 0.0;
  for (int64_t j = threadIdx.x; j < cnt; j += kRT) {
    const double x = z[seg_entry(b, j, g.N3)];
    v += x * x;
  }
  v = block_sum(v, sh);
  if (threadIdx.x == 0) g.partial[2 * blockIdx.x] = v;
}

// Finalising: one wave per system; MODE 0 normz, 1 alpha[i], 2 beta[i] with
// the breakdown rule, 3 err from (num, den).
template <int MODE>
__global__ __launch_bounds__(kRT) void k_seg_finalize(SegLayout g, SegScalars sc, int i) {
  const int sys = blockIdx.x * (kRT / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (sys >= g.nsys) return;
  if (MODE == 3 ? sc.check[sys] == 0 : sc.active[sys] == 0) return;
  const int b0 = g.first[sys], b1 = g.first[sys + 1];
  double v0 = 0.0, v1 = 0.0;
  for (int b = b0 + lane; b < b1; b += 64) {
    v0 += g.partial[2 * b];
    if (MODE == 3) v1 += g.partial[2 * b + 1];
  }
  for (int o = 32; o > 0; o >>= 1) {
    v0 += __shfl_xor(v0, o);
    if (MODE == 3) v1 += __shfl_xor(v1, o);
  }
  if (lane != 0) return;
  const int64_t row = (int64_t)sys * sc.L;
  if (MODE == 0) {
    sc.normz[sys] = sqrt(v0);
  } else if (MODE == 1) {
    sc.alpha[row + i] = v0;
  } else if (MODE == 2) {
    double be = sqrt(v0);
    if (be < 1e-3 * sc.alpha[row + i] / sc.normz[sys]) be = 0.0;  // LanczosAlgorithm.cu:139-143
    sc.beta[row + i] = be;
  } else {
    sc.err[sys] = fabs(sqrt(v0) / sqrt(v1));  // LanczosAlgorithm.cu:215
  }
}

template <int MODE>
static void seg_finalize(const SegLayout& g, const SegScalars& sc, int i, hipStream_t s) {
  const int per = kRT / 64;
  hipLaunchKernelGGL(k_seg_finalize<MODE>, dim3((g.nsys + per - 1) / per), dim3(kRT), 0, s, g, sc,
                     i);
  LM_HIP(hipGetLastError());
}

__global__ __launch_bounds__(kRT) void k_seg_scale(const double* __restrict__ z,
                                                  double* __restrict__ v0, SegLayout g,
                                                  SegScalars sc) {
  const SegBlock b = g.blocks[blockIdx.x];
  if (!sc.active[b.sys]) return;
  const double inv = 1.0 / sc.normz[b.sys];
  const int64_t cnt = seg_count(b, g.angular);
  for (int64_t j = threadIdx.x; j < cnt; j += kRT) {
    const int64_t e = seg_entry(b, j, g.N3);
    v0[e] = z[e] * inv;
  }
}

void launch_seg_start(const double* z, double* v0, const SegLayout& g, const SegScalars& sc,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_seg_norm, dim3(g.nblocks), dim3(kRT), 0, s, z, g, sc);
  LM_HIP(hipGetLastError());
  seg_finalize<0>(g, sc, 0, s);
  hipLaunchKernelGGL(k_seg_scale, dim3(g.nblocks), dim3(kRT), 0, s, z, v0, g, sc);
  LM_HIP(hipGetLastError());
}

// w -= beta_s[i-1] v_prev ; partial w . v_i
__global__ __launch_bounds__(kRT) void k_seg_alpha(double* __restrict__ w,
                                                  const double* __restrict__ vi,
                                                  const double* __restrict__ vp, int i,
                                                  SegLayout g, SegScalars sc) {
  __shared__ double sh[kRT / 64];
  const SegBlock b = g.blocks[blockIdx.x];
  if (!sc.active[b.sys]) return;
  const double be = (i > 0) ? sc.beta[(int64_t)b.sys * sc.L + i - 1] : 0.0;
  const int64_t cnt = seg_count(b, g.angular);
  double v = 0.0;
  for (int64_t j = threadIdx.x; j < cnt; j += kRT) {
    const int64_t e = seg_entry(b, j, g.N3);
    double wk = w[e];
    if (i > 0) {
      wk -= be * vp[e];
      w[e] = wk;
    }
    v += wk * vi[e];
  }
  v = block_sum(v, sh);
  if (threadIdx.x == 0) g.partial[2 * blockIdx.x] = v;
}

void launch_seg_alpha(double* w, const double* vi, const double* vp, int i, const SegLayout& g,
                      const SegScalars& sc, hipStream_t s) {
  hipLaunchKernelGGL(k_seg_alpha, dim3(g.nblocks), dim3(kRT), 0, s, w, vi, vp, i, g, sc);
  LM_HIP(hipGetLastError());
  seg_finalize<1>(g, sc, i, s);
}

// w -= alpha_s[i] v_i ; partial ||w||^2
__global__ __launch_bounds__(kRT) void k_seg_beta(double* __restrict__ w,
                                                 const double* __restrict__ vi, int i, SegLayout g,
                                                 SegScalars sc) {
  __shared__ double sh[kRT / 64];
  const SegBlock b = g.blocks[blockIdx.x];
  if (!sc.active[b.sys]) return;
  const double al = sc.alpha[(int64_t)b.sys * sc.L + i];
  const int64_t cnt = seg_count(b, g.angular);
  double v = 0.0;
  for (int64_t j = threadIdx.x; j < cnt; j += kRT) {
    const int64_t e = seg_entry(b, j, g.N3);
    const double wk = w[e] - al * vi[e];
    w[e] = wk;
    v += wk * wk;
  }
  v = block_sum(v, sh);
  if (threadIdx.x == 0) g.partial[2 * blockIdx.x] = v;
}

void launch_seg_beta(double* w, const double* vi, int i, const SegLayout& g, const SegScalars& sc,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_seg_beta, dim3(g.nblocks), dim3(kRT), 0, s, w, vi, i, g, sc);
  LM_HIP(hipGetLastError());
  seg_finalize<2>(g, sc, i, s);
}

// v_{i+1} = w / beta_s[i], or the system's e1 after a breakdown
__global__ __launch_bounds__(kRT) void k_seg_next(const double* __restrict__ w,
                                                 double* __restrict__ vn, int i, SegLayout g,
                                                 SegScalars sc, const int* __restrict__ first) {
  const SegBlock b = g.blocks[blockIdx.x];
  if (!sc.active[b.sys]) return;
  const double be = sc.beta[(int64_t)b.sys * sc.L + i];
  const double inv = be > 0.0 ? 1.0 / be : 0.0;
  // e1 of a system is its first entry: entry 0 of the system's first block
  const bool head = (int)blockIdx.x == first[b.sys];
  const int64_t cnt = seg_count(b, g.angular);
  for (int64_t j = threadIdx.x; j < cnt; j += kRT) {
    const int64_t e = seg_entry(b, j, g.N3);
    vn[e] = be > 0.0 ? w[e] * inv : ((head && j == 0) ? 1.0 : 0.0);
  }
}

void launch_seg_next(const double* w, double* vn, int i, const SegLayout& g, const SegScalars& sc,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_seg_next, dim3(g.nblocks), dim3(kRT), 0, s, w, vn, i, g, sc, g.first);
  LM_HIP(hipGetLastError());
}

// y = normz_s V[:, :m] coef_s ; old <- old - y ; partials (||old - y||^2, ||old||^2)
__global__ __launch_bounds__(kRT) void k_seg_estimate(const double* __restrict__ V, int64_t ldv,
                                                     int m, const double* __restrict__ coef,
                                                     double* __restrict__ y,
                                                     double* __restrict__ old, SegLayout g,
                                                     SegScalars sc) {
  __shared__ double sc_s[kLanczosMaxIterations];
  __shared__ double sh[kRT / 64];
  const SegBlock b = g.blocks[blockIdx.x];
  if (!sc.check[b.sys]) return;
  const double nz = sc.normz[b.sys];
  const double* cs = coef + (int64_t)b.sys * sc.L;
  for (int k = threadIdx.x; k < m; k += kRT) sc_s[k] = nz * cs[k];
  __syncthreads();
  const int64_t cnt = seg_count(b, g.angular);
  double num = 0.0, den = 0.0;
  for (int64_t j = threadIdx.x; j < cnt; j += kRT) {
    const int64_t e = seg_entry(b, j, g.N3);
    double v = 0.0;
    for (int k = 0; k < m; ++k) v += V[(int64_t)k * ldv + e] * sc_s[k];
    y[e] = v;
    const double o = old[e];
    const double d = o - v;
    old[e] = d;
    num += d * d;
    den += o * o;
  }
  num = block_sum(num, sh);
  den = block_sum(den, sh);
  if (threadIdx.x == 0) {
    g.partial[2 * blockIdx.x] = num;
    g.partial[2 * blockIdx.x + 1] = den;
  }
}

void launch_seg_estimate(const double* V, int64_t ldv, int m, const double* coef, double* y,
                         double* old, const SegLayout& g, const SegScalars& sc, hipStream_t s) {
  if (m < 1 || m > sc.L || m > kLanczosMaxIterations)
    throw Error("[libmobility_amd] Krylov dimension out of range");
  hipLaunchKernelGGL(k_seg_estimate, dim3(g.nblocks), dim3(kRT), 0, s, V, ldv, m, coef, y, old, g,
                     sc);
  LM_HIP(hipGetLastError());
  seg_finalize<3>(g, sc, 0, s);
}

__global__ __launch_bounds__(kRT) void k_seg_store(const double* __restrict__ y,
                                                  double* __restrict__ old, SegLayout g,
                                                  SegScalars sc) {
  const SegBlock b = g.blocks[blockIdx.x];
  if (sc.check[b.sys] != 2) return;
  const int64_t cnt = seg_count(b, g.angular);
  for (int64_t j = threadIdx.x; j < cnt; j += kRT) {
    const int64_t e = seg_entry(b, j, g.N3);
    old[e] = y[e];
  }
}

void launch_seg_store(const double* y, double* old, const SegLayout& g, const SegScalars& sc,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_seg_store, dim3(g.nblocks), dim3(kRT), 0, s, y, old, g, sc);
  LM_HIP(hipGetLastError());
}

__global__ __launch_bounds__(kRT) void k_seg_clear(double* __restrict__ y, SegLayout g,
                                                  SegScalars sc) {
  const SegBlock b = g.blocks[blockIdx.x];
  if (sc.check[b.sys] != 3) return;
  const int64_t cnt = seg_count(b, g.angular);
  for (int64_t j = threadIdx.x; j < cnt; j += kRT) y[seg_entry(b, j, g.N3)] = 0.0;
}

void launch_seg_clear(double* y, const SegLayout& g, const SegScalars& sc, hipStream_t s) {
  hipLaunchKernelGGL(k_seg_clear, dim3(g.nblocks), dim3(kRT), 0, s, y, g, sc);
  LM_HIP(hipGetLastError());
}

// ---- small vector kernels ----------------------------------------------------------
__global__ void k_axpy_acc(double* __restrict__ out, const double* __restrict__ y, double p,
                           int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] += p * y[i];
}
void launch_axpy_accumulate(double* out, const double* y, double pref, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpy_acc, dim3(reduce_blocks(n) * 4), dim3(kRT), 0, s, out, y, pref, n);
  LM_HIP(hipGetLastError());
}

__global__ void k_axpy_to(double* __restrict__ dst, const double* __restrict__ src,
                          const double* __restrict__ w, double c, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i] + c * w[i];
}
void launch_axpy_to(double* dst, const double* src, const double* w, double c, int64_t n,
                    hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpy_to, dim3(reduce_blocks(n) * 4), dim3(kRT), 0, s, dst, src, w, c, n);
  LM_HIP(hipGetLastError());
}

__global__ void k_diff_acc(double* __restrict__ out, const double* __restrict__ a,
                           const double* __restrict__ b, double p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] += p * (a[i] - b[i]);
}
void launch_diff_accumulate(double* out, const double* a, const double* b, double pref, int64_t n,
                            hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_diff_acc, dim3(reduce_blocks(n) * 4), dim3(kRT), 0, s, out, a, b, pref, n);
  LM_HIP(hipGetLastError());
}

__global__ void k_copy(double* __restrict__ dst, const double* __restrict__ src, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
void launch_copy(double* dst, const double* src, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy, dim3(reduce_blocks(n) * 4), dim3(kRT), 0, s, dst, src, n);
  LM_HIP(hipGetLastError());
}

__global__ void k_zero(double* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}
void launch_zero(double* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero, dim3(reduce_blocks(n) * 4), dim3(kRT), 0, s, p, n);
  LM_HIP(hipGetLastError());
}

__global__ void k_zero_u32(unsigned int* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}
void launch_zero_u32(unsigned int* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(kRT), 0, s, p, n);
  LM_HIP(hipGetLastError());
}

}  // namespace lm
