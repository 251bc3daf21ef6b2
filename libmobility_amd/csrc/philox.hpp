// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11), written from the published
// description: 10 rounds of two 32x32->64 multiplies with a Weyl-sequence key
// schedule.  Used for the Brownian noise W and the random-finite-difference
// displacement; replaces the reference's Saru generator
// (include/MobilityInterface/lanczos.h:16-22, random_finite_differences.h:15-23).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lm {

struct U32x4 {
  uint32_t v[4];
};

__host__ __device__ inline void philox_mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  const uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

__host__ __device__ inline U32x4 philox4x32_10(U32x4 ctr, uint32_t k0, uint32_t k1) {
  constexpr uint32_t kM0 = 0xD2511F53u, kM1 = 0xCD9E8D57u;
  constexpr uint32_t kW0 = 0x9E3779B9u, kW1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += kW0;
      k1 += kW1;
    }
    uint32_t hi0, lo0, hi1, lo1;
    philox_mulhilo(kM0, ctr.v[0], hi0, lo0);
    philox_mulhilo(kM1, ctr.v[2], hi1, lo1);
    U32x4 n;
    n.v[0] = hi1 ^ ctr.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ ctr.v[3] ^ k1;
    n.v[3] = lo0;
    ctr = n;
  }
  return ctr;
}

// Open-interval uniform double from 64 random bits: (u53 + 1/2) 2^-53.
__host__ __device__ inline double philox_u01(uint32_t a, uint32_t b) {
  const uint64_t u53 = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return ((double)u53 + 0.5) * (1.0 / 9007199254740992.0);
}

}  // namespace lm
