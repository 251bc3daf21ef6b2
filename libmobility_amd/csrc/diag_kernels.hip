// Diagnostic: the VALU issue ceiling of the device, for the roofline of the
// pair kernel.  Each thread runs 8 independent FMA chains (no memory traffic
// beyond one store per thread), optionally with one rsqrt per 3 FMAs of a
// chain, or the same chains in FP64.
#include "kernels.hpp"
#include "lm_common.hpp"

namespace lm {

template <int WITH_RSQ>
__global__ __launch_bounds__(256) void k_diag_f32(float* __restrict__ out, int iters, float seed) {
  float a[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) a[k] = seed + threadIdx.x * 1e-3f + k;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      a[k] = __builtin_fmaf(a[k], 0.999f, 0.001f);
      a[k] = __builtin_fmaf(a[k], 0.999f, 0.001f);
      a[k] = __builtin_fmaf(a[k], 0.999f, 0.001f);
      if (WITH_RSQ) a[k] = __builtin_amdgcn_rsqf(a[k] + 1.0f);
    }
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += a[k];
  out[(int64_t)blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_diag_f64(double* __restrict__ out, int iters, double seed) {
  double a[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) a[k] = seed + threadIdx.x * 1e-3 + k;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      a[k] = __builtin_fma(a[k], 0.999, 0.001);
      a[k] = __builtin_fma(a[k], 0.999, 0.001);
      a[k] = __builtin_fma(a[k], 0.999, 0.001);
    }
  }
  double s = 0.;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += a[k];
  out[(int64_t)blockIdx.x * 256 + threadIdx.x] = s;
}

// ms for the given kernel (0: FP32 FMA, 1: FP32 FMA + rsqrt, 2: FP64 FMA) over
// `blocks` workgroups of 256 threads, `iters` iterations of 24 FMAs each.
float diag_valu_ms(int kind, int blocks, int iters, hipStream_t s) {
  if (kind < 0 || kind > 2 || blocks < 1 || iters < 1)
    throw Error("[libmobility_amd] invalid throughput probe arguments");
  DevBuf out;
  out.reserve(sizeof(double) * (size_t)blocks * 256);
  hipEvent_t e0, e1;
  LM_HIP(hipEventCreate(&e0));
  LM_HIP(hipEventCreate(&e1));
  LM_HIP(hipEventRecord(e0, s));
  if (kind == 0)
    hipLaunchKernelGGL((k_diag_f32<0>), dim3(blocks), dim3(256), 0, s, out.as<float>(), iters, 1.0f);
  else if (kind == 1)
    hipLaunchKernelGGL((k_diag_f32<1>), dim3(blocks), dim3(256), 0, s, out.as<float>(), iters, 1.0f);
  else
    hipLaunchKernelGGL(k_diag_f64, dim3(blocks), dim3(256), 0, s, out.as<double>(), iters, 1.0);
  LM_HIP(hipGetLastError());
  LM_HIP(hipEventRecord(e1, s));
  LM_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  LM_HIP(hipEventElapsedTime(&ms, e0, e1));
  LM_HIP(hipEventDestroy(e0));
  LM_HIP(hipEventDestroy(e1));
  return ms;
}

}  // namespace lm
