// gfx950 kernels of the O(N^2) mobility product and of setPositions.
//
// Pair kernel layout
//   * One workgroup = one PairItem: a tile of targets (256 x TPL, each lane
//     owns TPL targets held in registers) against one chunk of sources of the
//     same system.
//   * Sources stream through LDS in tiles of 256 records; every lane reads the
//     same source at the same time, so LDS reads are broadcasts.
//   * FP32 mode: pair terms and the per-tile sums are FP32, each tile's sum is
//     added into FP64 accumulators (error grows with the tile, not with N).
//   * Partial results of the chunks of one target go to separate slabs and
//     are summed by k_chunk_reduce in a fixed order (bitwise reproducible).
//   * Items are ordered target-tile-major, chunk-minor.  Workgroups are
//     dispatched round-robin over the 8 XCDs, so with 8 chunks every XCD
//     reads one chunk of the sources and its L2 holds only that chunk.
#include <type_traits>

#include "kernels.hpp"
#include "lm_common.hpp"
#include "pair_math.hpp"

namespace lm {

template <class R>
struct Rec;
template <>
struct Rec<float> {
  using T = float4;
};
template <>
struct Rec<double> {
  using T = double4;
};

// A record far from everything: padding sources (zero force) and idle target
// lanes evaluate finite arithmetic against it and contribute nothing.
template <class R4>
__device__ inline R4 far_record() {
  R4 r;
  r.x = 1.0e6f;
  r.y = 1.0e6f;
  r.z = 1.0e6f;
  r.w = 1.0f;
  return r;
}

// Overlap test of a group of pairs: is any r^2 below 4 (radius units)?
// NearTestCmp compares every r^2 (the mask lives in scalar registers).
// NearTest in FP32 keeps the smallest r^2 by its bit pattern, which orders
// like the value for r^2 >= 0 (NaN sorts above 4 and stays on the far
// formula), so a group costs one unsigned min per pair (the compiler forms
// v_min3_u32) and a single compare; it takes a vector register, which the
// ordered kernel's 4-target tiling cannot spare without losing a wave per
// SIMD, so that kernel keeps the compares.
template <class R>
struct NearTestCmp {
  bool any = false;
  __device__ void add(R r2) { any |= r2 < R(4); }
  __device__ bool hit() const { return any; }
};
template <class R>
struct NearTest : NearTestCmp<R> {};
// FP64: the high words of non-negative doubles order like the values, and
// every double whose high word is that of 4.0 (0x40100000) is >= 4.0, so
// r^2 < 4 exactly when its high word is below 0x40100000.
template <>
struct NearTest<double> {
  uint32_t m = 0xffffffffu;
  __device__ void add(double r2) {
    const uint32_t h = (uint32_t)(__builtin_bit_cast(uint64_t, r2) >> 32);
    m = h < m ? h : m;
  }
  __device__ bool hit() const { return m < 0x40100000u; }
};
template <>
struct NearTest<float> {
  uint32_t m = 0xffffffffu;
  __device__ void add(float r2) {
    const uint32_t b = __builtin_bit_cast(uint32_t, r2);
    m = b < m ? b : m;
  }
  __device__ bool hit() const { return m < 0x40800000u; }  // the bits of 4.0f
};

// One item: targets [tgt0, tgt0 + ntgt) x sources [src0, src1) of one system.
template <class R, bool WALL, int BLK, int TPL>
__device__ inline void pair_item(const PairItem& it, const typename Rec<R>::T* __restrict__ pos,
                                 const double* __restrict__ F, const double* __restrict__ T,
                                 double* __restrict__ U, double* __restrict__ W, int64_t slab,
                                 double t0, double rt0, double r0, typename Rec<R>::T* s_pos,
                                 typename Rec<R>::T* s_f, typename Rec<R>::T* s_t) {
  using R4 = typename Rec<R>::T;
  constexpr bool kNeedF = (BLK & (kUF | kWF)) != 0;
  constexpr bool kNeedT = (BLK & (kUT | kWT)) != 0;
  constexpr bool kAngular = (BLK & (kWT | kWF)) != 0;
  // open translational block: coefficients times 4/3 (one operation less per
  // pair), undone once on the output
  constexpr bool kUF34 = !WALL && BLK == kUF;
  const int tid = threadIdx.x;

  R4 pi[TPL];
#pragma unroll
  for (int k = 0; k < TPL; ++k) {
    const int loc = tid + k * kPairThreads;
    pi[k] = (loc < it.ntgt) ? pos[it.tgt0 + loc] : far_record<R4>();
  }
  PairAcc<R> acc[TPL];
  PairAcc<double> acc64[TPL];
#pragma unroll
  for (int k = 0; k < TPL; ++k) {
    acc_zero(acc[k]);
    acc_zero(acc64[k]);
  }

  for (int64_t base = it.src0; base < it.src1; base += kPairThreads) {
    const int64_t j = base + tid;
    // sources of this tile (a chunk's last tile may be partial; its padding
    // slots hold far records and zero forces and are skipped)
    const int cnt = (int)min((int64_t)kPairThreads, it.src1 - base);
    __syncthreads();
    if (j < it.src1) {
      s_pos[tid] = pos[j];
      if (kNeedF) s_f[tid] = R4{R(F[3 * j]), R(F[3 * j + 1]), R(F[3 * j + 2]), R(0)};
      if (kNeedT) s_t[tid] = R4{R(T[3 * j]), R(T[3 * j + 1]), R(T[3 * j + 2]), R(0)};
    } else {
      s_pos[tid] = far_record<R4>();
      if (kNeedF) s_f[tid] = R4{R(0), R(0), R(0), R(0)};
      if (kNeedT) s_t[tid] = R4{R(0), R(0), R(0), R(0)};
    }
    __syncthreads();
    if constexpr (WALL) {
      for (int s = 0; s < cnt; ++s) {
        const R4 pj = s_pos[s];
        R4 fj, tj;
        if (kNeedF) fj = s_f[s]; else fj = R4{R(0), R(0), R(0), R(0)};
        if (kNeedT) tj = s_t[s]; else tj = R4{R(0), R(0), R(0), R(0)};
#pragma unroll
        for (int k = 0; k < TPL; ++k)
          wall_pair<R, BLK>(pi[k].x, pi[k].y, pi[k].z, pi[k].w, pj.x, pj.y, pj.z, pj.w, fj.x,
                            fj.y, fj.z, tj.x, tj.y, tj.z, acc[k]);
      }
    } else {
      // Open boundaries: a group of SG sources x TPL targets (4 pairs) gets
      // its far-field coefficients first and one overlap test for the whole
      // group; the overlap values (rare: overlapping spheres and the self
      // pair) are patched in by selects inside that single branch.
      constexpr int SG = TPL >= 4 ? 1 : 4 / TPL;
      const int lim = (cnt + SG - 1) / SG * SG;  // padding slots up to a whole group
      for (int s = 0; s < lim; s += SG) {
        R dx[SG][TPL], dy[SG][TPL], dz[SG][TPL], r2[SG][TPL], ir[SG][TPL];
        RpyCoef<R> cf[SG][TPL];
        R4 fj[SG], tj[SG];
        NearTestCmp<R> near;
#pragma unroll
        for (int g = 0; g < SG; ++g) {
          const R4 pj = s_pos[s + g];
          if (kNeedF) fj[g] = s_f[s + g]; else fj[g] = R4{R(0), R(0), R(0), R(0)};
          if (kNeedT) tj[g] = s_t[s + g]; else tj[g] = R4{R(0), R(0), R(0), R(0)};
#pragma unroll
          for (int k = 0; k < TPL; ++k) {
            dx[g][k] = pi[k].x - pj.x;
            dy[g][k] = pi[k].y - pj.y;
            dz[g][k] = pi[k].z - pj.z;
            r2[g][k] = dx[g][k] * dx[g][k] + dy[g][k] * dy[g][k] + dz[g][k] * dz[g][k];
            ir[g][k] = lm_rsqrt(r2[g][k]);
            cf[g][k] = RpyCoef<R>{};
            if constexpr (kUF34)
              rpy_far_uf34<R>(ir[g][k], cf[g][k]);
            else
              rpy_far<R, BLK>(ir[g][k], cf[g][k]);
            near.add(r2[g][k]);
          }
        }
        if (__builtin_expect(near.hit(), 0)) {
          // opaque statement: keep this a branch (not speculated onto every group)
          __asm__ __volatile__("");
#pragma unroll
          for (int g = 0; g < SG; ++g)
#pragma unroll
            for (int k = 0; k < TPL; ++k) {
              if constexpr (kUF34)
                rpy_near_select_uf34<R>(r2[g][k], ir[g][k], cf[g][k]);
              else
                rpy_near_select<R, BLK>(r2[g][k], ir[g][k], cf[g][k]);
            }
        }
#pragma unroll
        for (int g = 0; g < SG; ++g)
#pragma unroll
          for (int k = 0; k < TPL; ++k)
            open_accumulate<R, BLK>(dx[g][k], dy[g][k], dz[g][k], fj[g].x, fj[g].y, fj[g].z,
                                    tj[g].x, tj[g].y, tj[g].z, cf[g][k], acc[k]);
      }
    }
    if constexpr (std::is_same<R, float>::value) {
#pragma unroll
      for (int k = 0; k < TPL; ++k) {
        acc64[k].uf.x += acc[k].uf.x; acc64[k].uf.y += acc[k].uf.y; acc64[k].uf.z += acc[k].uf.z;
        acc64[k].ut.x += acc[k].ut.x; acc64[k].ut.y += acc[k].ut.y; acc64[k].ut.z += acc[k].ut.z;
        acc64[k].wt.x += acc[k].wt.x; acc64[k].wt.y += acc[k].wt.y; acc64[k].wt.z += acc[k].wt.z;
        acc64[k].wf.x += acc[k].wf.x; acc64[k].wf.y += acc[k].wf.y; acc64[k].wf.z += acc[k].wf.z;
        acc_zero(acc[k]);
      }
    }
  }
  if constexpr (std::is_same<R, double>::value) {
#pragma unroll
    for (int k = 0; k < TPL; ++k) acc64[k] = acc[k];
  }
  double* Uo = U + (int64_t)it.chunk * slab;
  double* Wo = kAngular ? W + (int64_t)it.chunk * slab : nullptr;
  const double tuf = kUF34 ? 0.75 * t0 : t0;
#pragma unroll
  for (int k = 0; k < TPL; ++k) {
    const int loc = tid + k * kPairThreads;
    if (loc < it.ntgt) {
      const int64_t i = it.tgt0 + loc;
      const PairAcc<double>& a = acc64[k];
      Uo[3 * i + 0] = tuf * a.uf.x + rt0 * a.ut.x;
      Uo[3 * i + 1] = tuf * a.uf.y + rt0 * a.ut.y;
      Uo[3 * i + 2] = tuf * a.uf.z + rt0 * a.ut.z;
      if (kAngular) {
        Wo[3 * i + 0] = r0 * a.wt.x + rt0 * a.wf.x;
        Wo[3 * i + 1] = r0 * a.wt.y + rt0 * a.wf.y;
        Wo[3 * i + 2] = r0 * a.wt.z + rt0 * a.wf.z;
      }
    }
  }
}

// KIND: 0 open, 1 wall, 2 per item.  PERSIST: the grid loops over items,
// fetching the next one from `counter` (work queue; every workgroup exits
// once the counter passes nitems).
template <class R, int KIND, int BLK, int TPL, bool PERSIST>
__global__ __launch_bounds__(kPairThreads) void k_pair(
    const PairItem* __restrict__ items, int nitems, unsigned int* __restrict__ counter,
    const typename Rec<R>::T* __restrict__ pos, const double* __restrict__ F,
    const double* __restrict__ T, double* __restrict__ U, double* __restrict__ W, int64_t slab,
    const int* __restrict__ active, const int* stop, double t0, double rt0, double r0) {
  using R4 = typename Rec<R>::T;
  constexpr bool kNeedF = (BLK & (kUF | kWF)) != 0;
  constexpr bool kNeedT = (BLK & (kUT | kWT)) != 0;
  __shared__ R4 s_pos[kPairThreads];
  __shared__ R4 s_f[kNeedF ? kPairThreads : 1];
  __shared__ R4 s_t[kNeedT ? kPairThreads : 1];
  __shared__ int s_item;
  int item = blockIdx.x;
  if (!PERSIST && stop != nullptr) {
    // one read per workgroup, so that all its threads take the same branch
    // (the flag may change while the workgroup starts)
    __shared__ int s_stop;
    if (threadIdx.x == 0) s_stop = *(volatile const int*)stop;
    __syncthreads();
    if (s_stop != 0) return;
  }
  while (true) {
    if (PERSIST) {
      __syncthreads();  // previous item finished with s_item and the tiles
      if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
      __syncthreads();
      item = s_item;
    }
    if (item >= nitems) break;
    const PairItem it = items[item];
    if (active != nullptr && active[it.sys] == 0) {  // system already converged
      if (!PERSIST) break;
      continue;
    }
    if (KIND == 2) {
      if (it.wall)
        pair_item<R, true, BLK, TPL>(it, pos, F, T, U, W, slab, t0, rt0, r0, s_pos, s_f, s_t);
      else
        pair_item<R, false, BLK, TPL>(it, pos, F, T, U, W, slab, t0, rt0, r0, s_pos, s_f, s_t);
    } else {
      pair_item<R, KIND == 1, BLK, TPL>(it, pos, F, T, U, W, slab, t0, rt0, r0, s_pos, s_f, s_t);
    }
    if (!PERSIST) break;
  }
}

// ---- symmetric evaluation (open boundaries, translational block) --------------
// Open RPY depends on r only and is even in d, so M_ji F_i uses the pair's
// coefficients unchanged: fuf F_i + guf (d.F_i) d.  An item holds the 256*TPL
// targets of tile I in the lanes and streams each partner tile J through LDS
// in sub-tiles of 256 sources (four blocks of 64, each stored twice).  Lane l
// meets source (l + k) mod 64 of a block at step k (address l + k) and carries
// that source's running sum, which moves one lane down after every step
// (DPP wave_rol:1), so after 64 steps every source has met all targets of the
// wave and its sum sits back in lane (source index).  The four waves' sums are
// added in a fixed order and written once per source to the offset's slab.

// Lane l receives lane (l + 1) mod 64.
template <class R>
__device__ inline R lane_rol1(R v);
template <>
__device__ inline float lane_rol1<float>(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x134, 0xf, 0xf, true));
}
template <>
__device__ inline double lane_rol1<double>(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo =
      (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, 0x134, 0xf, 0xf, true);
  const uint32_t hi =
      (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), 0x134, 0xf, 0xf, true);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

template <class R, int TPL>
__device__ inline void sym_item(const SymItem& it, const SymArgs& a, typename Rec<R>::T* s_pos,
                                typename Rec<R>::T* s_f, R* s_red) {
  using R4 = typename Rec<R>::T;
  constexpr int kTile = kPairThreads * TPL;
  constexpr int SG = TPL >= 4 ? 1 : 4 / TPL;  // steps per overlap test (4 pairs)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const R4* __restrict__ pos = static_cast<const R4*>(a.pos);
  const double* __restrict__ F = a.F;
  const SymSys S = a.sys[it.sys];
  const int64_t N = S.off + S.n;  // end of the system's particles
  const int64_t tgt0 = S.off + (int64_t)it.tile * kTile;
  const double t0s = 0.75 * a.t0;  // the coefficients carry 4/3 (rpy_far_uf34)

  R4 pi[TPL];
  R fix[TPL], fiy[TPL], fiz[TPL];
#pragma unroll
  for (int t = 0; t < TPL; ++t) {
    const int64_t i = tgt0 + tid + t * kPairThreads;
    if (i < N) {
      pi[t] = pos[i];
      fix[t] = R(F[3 * i]);
      fiy[t] = R(F[3 * i + 1]);
      fiz[t] = R(F[3 * i + 2]);
    } else {
      pi[t] = far_record<R4>();
      fix[t] = fiy[t] = fiz[t] = R(0);
    }
  }
  Vec3<R> acc[TPL];
  Vec3<double> acc64[TPL];
#pragma unroll
  for (int t = 0; t < TPL; ++t) {
    acc[t] = {R(0), R(0), R(0)};
    acc64[t] = {0.0, 0.0, 0.0};
  }

  for (int o = it.o0; o < it.o1; ++o) {
    if (2 * o == S.ntiles && 2 * it.tile >= S.ntiles) continue;  // met from the other side
    int J = it.tile + o;
    if (J >= S.ntiles) J -= S.ntiles;
    double* __restrict__ So = a.slab_s + (int64_t)(o - a.obase) * a.n3;
#pragma unroll 1
    for (int sub = 0; sub < TPL; ++sub) {
      const int64_t base = S.off + (int64_t)J * kTile + sub * kPairThreads;
      if (base >= N) break;
      const int64_t j = base + tid;
      __syncthreads();  // the previous sub-tile is done with s_pos, s_f and s_red
      R4 pj = far_record<R4>();
      R4 fj = R4{R(0), R(0), R(0), R(0)};
      if (j < N) {
        pj = pos[j];
        fj = R4{R(F[3 * j]), R(F[3 * j + 1]), R(F[3 * j + 2]), R(0)};
      }
      s_pos[wave * 128 + lane] = pj;
      s_pos[wave * 128 + 64 + lane] = pj;
      s_f[wave * 128 + lane] = fj;
      s_f[wave * 128 + 64 + lane] = fj;
      __syncthreads();
#pragma unroll 1  // one copy of the step loop (the blocks differ only in address)
      for (int blk = 0; blk < 4; ++blk) {
        const R4* sp = s_pos + blk * 128 + lane;
        const R4* sf = s_f + blk * 128 + lane;
        Vec3<R> sa = {R(0), R(0), R(0)};
#pragma unroll 2
        for (int k = 0; k < 64; k += SG) {
          R dx[SG][TPL], dy[SG][TPL], dz[SG][TPL], r2[SG][TPL], ir[SG][TPL];
          RpyCoef<R> cf[SG][TPL];
          R4 pg[SG], fg[SG];
          NearTest<R> near;
#pragma unroll
          for (int g = 0; g < SG; ++g) {
            pg[g] = sp[k + g];
            fg[g] = sf[k + g];
#pragma unroll
            for (int t = 0; t < TPL; ++t) {
              dx[g][t] = pi[t].x - pg[g].x;
              dy[g][t] = pi[t].y - pg[g].y;
              dz[g][t] = pi[t].z - pg[g].z;
              r2[g][t] = dx[g][t] * dx[g][t] + dy[g][t] * dy[g][t] + dz[g][t] * dz[g][t];
              ir[g][t] = lm_rsqrt(r2[g][t]);
              cf[g][t] = RpyCoef<R>{};
              rpy_far_uf34<R>(ir[g][t], cf[g][t]);
              near.add(r2[g][t]);
            }
          }
          if (__builtin_expect(near.hit(), 0)) {
            __asm__ __volatile__("");  // keep this a branch
#pragma unroll
            for (int g = 0; g < SG; ++g)
#pragma unroll
              for (int t = 0; t < TPL; ++t) rpy_near_select_uf34<R>(r2[g][t], ir[g][t], cf[g][t]);
          }
#pragma unroll
          for (int g = 0; g < SG; ++g) {
            // the running sum moves first and the step's terms accumulate
            // into it: one move per component and step, no separate add
            sa.x = lane_rol1(sa.x);
            sa.y = lane_rol1(sa.y);
            sa.z = lane_rol1(sa.z);
#pragma unroll
            for (int t = 0; t < TPL; ++t) {
              const R fuf = cf[g][t].fuf, guf = cf[g][t].guf;
              const R st = guf * (dx[g][t] * fg[g].x + dy[g][t] * fg[g].y + dz[g][t] * fg[g].z);
              acc[t].x = lm_fma(st, dx[g][t], lm_fma(fuf, fg[g].x, acc[t].x));
              acc[t].y = lm_fma(st, dy[g][t], lm_fma(fuf, fg[g].y, acc[t].y));
              acc[t].z = lm_fma(st, dz[g][t], lm_fma(fuf, fg[g].z, acc[t].z));
              const R ss = guf * (dx[g][t] * fix[t] + dy[g][t] * fiy[t] + dz[g][t] * fiz[t]);
              sa.x = lm_fma(ss, dx[g][t], lm_fma(fuf, fix[t], sa.x));
              sa.y = lm_fma(ss, dy[g][t], lm_fma(fuf, fiy[t], sa.y));
              sa.z = lm_fma(ss, dz[g][t], lm_fma(fuf, fiz[t], sa.z));
            }
          }
        }
        // after 64 steps lane l holds source l - 1: one more move puts source l in lane l
        s_red[(wave * 3 + 0) * kPairThreads + blk * 64 + lane] = lane_rol1(sa.x);
        s_red[(wave * 3 + 1) * kPairThreads + blk * 64 + lane] = lane_rol1(sa.y);
        s_red[(wave * 3 + 2) * kPairThreads + blk * 64 + lane] = lane_rol1(sa.z);
      }
      __syncthreads();
      if (j < N) {
        double vx = 0.0, vy = 0.0, vz = 0.0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
          vx += (double)s_red[(w * 3 + 0) * kPairThreads + tid];
          vy += (double)s_red[(w * 3 + 1) * kPairThreads + tid];
          vz += (double)s_red[(w * 3 + 2) * kPairThreads + tid];
        }
        So[3 * j + 0] = t0s * vx;
        So[3 * j + 1] = t0s * vy;
        So[3 * j + 2] = t0s * vz;
      }
      if constexpr (std::is_same<R, float>::value) {
#pragma unroll
        for (int t = 0; t < TPL; ++t) {
          acc64[t].x += acc[t].x;
          acc64[t].y += acc[t].y;
          acc64[t].z += acc[t].z;
          acc[t] = {R(0), R(0), R(0)};
        }
      }
    }
  }
  if constexpr (std::is_same<R, double>::value) {
#pragma unroll
    for (int t = 0; t < TPL; ++t) acc64[t] = acc[t];
  }
  double* __restrict__ Uo = a.slab_t + (int64_t)it.chunk * a.n3;
#pragma unroll
  for (int t = 0; t < TPL; ++t) {
    const int64_t i = tgt0 + tid + t * kPairThreads;
    if (i < N) {
      Uo[3 * i + 0] = t0s * acc64[t].x;
      Uo[3 * i + 1] = t0s * acc64[t].y;
      Uo[3 * i + 2] = t0s * acc64[t].z;
    }
  }
}

// General symmetric item (bottom wall, or open boundaries with angular
// blocks): the same ring, lanes and rotation as sym_item, with pair_both
// giving both directions of every block from one set of coefficients.  The
// source sums carry U (units of t0) and, with angular blocks, W (units of r0).
template <class R, bool WALL, int BLK, int TPL>
__device__ inline void sym_item_gen(const SymItem& it, const SymArgs& a,
                                    typename Rec<R>::T* s_pos, typename Rec<R>::T* s_f,
                                    typename Rec<R>::T* s_t, R* s_red) {
  using R4 = typename Rec<R>::T;
  constexpr int kTile = kPairThreads * TPL;
  constexpr bool kNeedF = (BLK & (kUF | kWF)) != 0;
  constexpr bool kNeedT = (BLK & (kUT | kWT)) != 0;
  constexpr bool kAngular = (BLK & (kWT | kWF)) != 0;
  constexpr int kComp = kAngular ? 6 : 3;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const R4* __restrict__ pos = static_cast<const R4*>(a.pos);
  const SymSys S = a.sys[it.sys];
  const int64_t N = S.off + S.n;  // end of the system's particles
  const int64_t tgt0 = S.off + (int64_t)it.tile * kTile;
  const R kut = R(a.rt0 / a.t0), kwf = R(a.rt0 / a.r0);
  const Vec3<R> zero = {R(0), R(0), R(0)};

  R4 pi[TPL];
  Vec3<R> fi[TPL], ti[TPL];
#pragma unroll
  for (int t = 0; t < TPL; ++t) {
    const int64_t i = tgt0 + tid + t * kPairThreads;
    const bool ok = i < N;
    pi[t] = ok ? pos[i] : far_record<R4>();
    fi[t] = (kNeedF && ok) ? Vec3<R>{R(a.F[3 * i]), R(a.F[3 * i + 1]), R(a.F[3 * i + 2])} : zero;
    ti[t] = (kNeedT && ok) ? Vec3<R>{R(a.T[3 * i]), R(a.T[3 * i + 1]), R(a.T[3 * i + 2])} : zero;
  }
  PairAcc<R> acc[TPL];
  PairAcc<double> acc64[TPL];
#pragma unroll
  for (int t = 0; t < TPL; ++t) {
    acc_zero(acc[t]);
    acc_zero(acc64[t]);
  }

  for (int o = it.o0; o < it.o1; ++o) {
    if (2 * o == S.ntiles && 2 * it.tile >= S.ntiles) continue;  // met from the other side
    int J = it.tile + o;
    if (J >= S.ntiles) J -= S.ntiles;
    for (int sub = 0; sub < TPL; ++sub) {
      const int64_t base = S.off + (int64_t)J * kTile + sub * kPairThreads;
      if (base >= N) break;
      const int64_t j = base + tid;
      __syncthreads();  // the previous sub-tile is done with the staged sources and s_red
      const bool ok = j < N;
      const R4 pj = ok ? pos[j] : far_record<R4>();
      s_pos[wave * 128 + lane] = pj;
      s_pos[wave * 128 + 64 + lane] = pj;
      if (kNeedF) {
        const R4 fj = ok ? R4{R(a.F[3 * j]), R(a.F[3 * j + 1]), R(a.F[3 * j + 2]), R(0)}
                         : R4{R(0), R(0), R(0), R(0)};
        s_f[wave * 128 + lane] = fj;
        s_f[wave * 128 + 64 + lane] = fj;
      }
      if (kNeedT) {
        const R4 tj = ok ? R4{R(a.T[3 * j]), R(a.T[3 * j + 1]), R(a.T[3 * j + 2]), R(0)}
                         : R4{R(0), R(0), R(0), R(0)};
        s_t[wave * 128 + lane] = tj;
        s_t[wave * 128 + 64 + lane] = tj;
      }
      __syncthreads();
      for (int blk = 0; blk < 4; ++blk) {
        const int b0 = blk * 128 + lane;
        Vec3<R> su = zero, sw = zero;
        for (int k = 0; k < 64; ++k) {
          const R4 pj4 = s_pos[b0 + k];
          Vec3<R> fj = zero, tj = zero;
          if (kNeedF) {
            const R4 v = s_f[b0 + k];
            fj = {v.x, v.y, v.z};
          }
          if (kNeedT) {
            const R4 v = s_t[b0 + k];
            tj = {v.x, v.y, v.z};
          }
          Vec3<R> cu = zero, cw = zero;
#pragma unroll
          for (int t = 0; t < TPL; ++t)
            pair_both<R, WALL, BLK>(pi[t].x, pi[t].y, pi[t].z, pi[t].w, pj4.x, pj4.y, pj4.z, pj4.w,
                                    fj, tj, fi[t], ti[t], kut, kwf, acc[t], cu, cw);
          su.x = lane_rol1(su.x) + cu.x;
          su.y = lane_rol1(su.y) + cu.y;
          su.z = lane_rol1(su.z) + cu.z;
          if (kAngular) {
            sw.x = lane_rol1(sw.x) + cw.x;
            sw.y = lane_rol1(sw.y) + cw.y;
            sw.z = lane_rol1(sw.z) + cw.z;
          }
        }
        // lane l holds source l - 1 after 64 steps: one more move
        R* red = s_red + wave * kComp * kPairThreads + blk * 64 + lane;
        red[0 * kPairThreads] = lane_rol1(su.x);
        red[1 * kPairThreads] = lane_rol1(su.y);
        red[2 * kPairThreads] = lane_rol1(su.z);
        if (kAngular) {
          red[3 * kPairThreads] = lane_rol1(sw.x);
          red[4 * kPairThreads] = lane_rol1(sw.y);
          red[5 * kPairThreads] = lane_rol1(sw.z);
        }
      }
      __syncthreads();
      if (ok) {
        double v[kComp];
#pragma unroll
        for (int q = 0; q < kComp; ++q) v[q] = 0.0;
#pragma unroll
        for (int w = 0; w < 4; ++w)
#pragma unroll
          for (int q = 0; q < kComp; ++q) v[q] += (double)s_red[(w * kComp + q) * kPairThreads + tid];
        double* So = a.slab_s + (int64_t)(o - a.obase) * a.n3;
        So[3 * j + 0] = a.t0 * v[0];
        So[3 * j + 1] = a.t0 * v[1];
        So[3 * j + 2] = a.t0 * v[2];
        if (kAngular) {
          double* Sw = a.slab_sw + (int64_t)(o - a.obase) * a.n3;
          Sw[3 * j + 0] = a.r0 * v[kComp - 3];
          Sw[3 * j + 1] = a.r0 * v[kComp - 2];
          Sw[3 * j + 2] = a.r0 * v[kComp - 1];
        }
      }
      if constexpr (std::is_same<R, float>::value) {
#pragma unroll
        for (int t = 0; t < TPL; ++t) {
          acc64[t].uf.x += acc[t].uf.x; acc64[t].uf.y += acc[t].uf.y; acc64[t].uf.z += acc[t].uf.z;
          acc64[t].ut.x += acc[t].ut.x; acc64[t].ut.y += acc[t].ut.y; acc64[t].ut.z += acc[t].ut.z;
          acc64[t].wt.x += acc[t].wt.x; acc64[t].wt.y += acc[t].wt.y; acc64[t].wt.z += acc[t].wt.z;
          acc64[t].wf.x += acc[t].wf.x; acc64[t].wf.y += acc[t].wf.y; acc64[t].wf.z += acc[t].wf.z;
          acc_zero(acc[t]);
        }
      }
    }
  }
  if constexpr (std::is_same<R, double>::value) {
#pragma unroll
    for (int t = 0; t < TPL; ++t) acc64[t] = acc[t];
  }
  double* Uo = a.slab_t + (int64_t)it.chunk * a.n3;
  double* Wo = kAngular ? a.slab_tw + (int64_t)it.chunk * a.n3 : nullptr;
#pragma unroll
  for (int t = 0; t < TPL; ++t) {
    const int64_t i = tgt0 + tid + t * kPairThreads;
    if (i < N) {
      const PairAcc<double>& c = acc64[t];
      Uo[3 * i + 0] = a.t0 * c.uf.x + a.rt0 * c.ut.x;
      Uo[3 * i + 1] = a.t0 * c.uf.y + a.rt0 * c.ut.y;
      Uo[3 * i + 2] = a.t0 * c.uf.z + a.rt0 * c.ut.z;
      if (kAngular) {
        Wo[3 * i + 0] = a.r0 * c.wt.x + a.rt0 * c.wf.x;
        Wo[3 * i + 1] = a.r0 * c.wt.y + a.rt0 * c.wf.y;
        Wo[3 * i + 2] = a.r0 * c.wt.z + a.rt0 * c.wf.z;
      }
    }
  }
}

// One item per workgroup: a diagonal tile (all ordered pairs of tile I, the
// self pairs included) or a range of partner offsets.  KIND: 0 open, 1 wall,
// 2 per system (mixed ensembles).
template <class R, int KIND, int BLK, int TPL>
__global__ __launch_bounds__(kPairThreads) void k_pair_sym(SymArgs a) {
  using R4 = typename Rec<R>::T;
  constexpr bool kNeedF = (BLK & (kUF | kWF)) != 0;
  constexpr bool kNeedT = (BLK & (kUT | kWT)) != 0;
  constexpr int kComp = (BLK & (kWT | kWF)) ? 6 : 3;
  __shared__ R4 s_pos[4 * 128];
  __shared__ R4 s_f[kNeedF ? 4 * 128 : 1];
  __shared__ R4 s_t[kNeedT ? 4 * 128 : 1];
  __shared__ R s_red[4 * kComp * kPairThreads];
  if (a.stop != nullptr) {
    // one read per workgroup, so that all its threads take the same branch
    __shared__ int s_stop;
    if (threadIdx.x == 0) s_stop = *(volatile const int*)a.stop;
    __syncthreads();
    if (s_stop != 0) return;
  }
  const SymItem it = a.items[blockIdx.x];
  if (a.active != nullptr && a.active[it.sys] == 0) return;  // system already converged
  const SymSys S = a.sys[it.sys];
  const bool wall = KIND == 2 ? S.wall != 0 : KIND == 1;
  if (it.o0 == 0) {
    constexpr int kTile = kPairThreads * TPL;
    PairItem p;
    p.tgt0 = S.off + (int64_t)it.tile * kTile;
    const int64_t left = S.off + S.n - p.tgt0;
    p.ntgt = (int32_t)(left < kTile ? left : kTile);
    p.src0 = p.tgt0;
    p.src1 = p.tgt0 + p.ntgt;
    p.chunk = it.chunk;
    p.wall = wall ? 1 : 0;
    p.sys = it.sys;
    const R4* pos = static_cast<const R4*>(a.pos);
    R4* sf = kNeedF ? s_f : s_pos;
    R4* st = kNeedT ? s_t : s_pos;
    if constexpr (KIND != 0) {
      if (wall) {
        pair_item<R, true, BLK, TPL>(p, pos, a.F, a.T, a.slab_t, a.slab_tw, a.n3, a.t0, a.rt0, a.r0,
                                     s_pos, sf, st);
        return;
      }
    }
    if constexpr (KIND != 1)
      pair_item<R, false, BLK, TPL>(p, pos, a.F, a.T, a.slab_t, a.slab_tw, a.n3, a.t0, a.rt0, a.r0,
                                    s_pos, sf, st);
    return;
  }
  if constexpr (KIND != 0) {
    if (wall) {
      sym_item_gen<R, true, BLK, TPL>(it, a, s_pos, s_f, s_t, s_red);
      return;
    }
  }
  if constexpr (KIND != 1) {
    if constexpr (BLK == kUF)
      sym_item<R, TPL>(it, a, s_pos, s_f, s_red);  // the open translational fast path
    else
      sym_item_gen<R, false, BLK, TPL>(it, a, s_pos, s_f, s_t, s_red);
  }
}

// ---- several force vectors per pass (translational block) ----------------------
// Same tiling as pair_item; each pair's geometry and coefficients are computed
// once and applied to the NR force vectors staged side by side in LDS.
template <class R, bool WALL, int NR, int TPL>
__device__ inline void pair_item_multi(const PairItem& it, const typename Rec<R>::T* __restrict__ pos,
                                       const double* __restrict__ F, double* __restrict__ U,
                                       int64_t n3, int nch, double t0, typename Rec<R>::T* s_pos,
                                       typename Rec<R>::T* s_f) {
  using R4 = typename Rec<R>::T;
  const int tid = threadIdx.x;
  R4 pi[TPL];
#pragma unroll
  for (int k = 0; k < TPL; ++k) {
    const int loc = tid + k * kPairThreads;
    pi[k] = (loc < it.ntgt) ? pos[it.tgt0 + loc] : far_record<R4>();
  }
  Vec3<R> acc[TPL][NR];
  Vec3<double> acc64[TPL][NR];
#pragma unroll
  for (int k = 0; k < TPL; ++k)
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      acc[k][r] = {R(0), R(0), R(0)};
      acc64[k][r] = {0.0, 0.0, 0.0};
    }
  for (int64_t base = it.src0; base < it.src1; base += kPairThreads) {
    const int64_t j = base + tid;
    __syncthreads();
    if (j < it.src1) {
      s_pos[tid] = pos[j];
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const double* f = F + (int64_t)r * n3 + 3 * j;
        s_f[r * kPairThreads + tid] = R4{R(f[0]), R(f[1]), R(f[2]), R(0)};
      }
    } else {
      s_pos[tid] = far_record<R4>();
#pragma unroll
      for (int r = 0; r < NR; ++r) s_f[r * kPairThreads + tid] = R4{R(0), R(0), R(0), R(0)};
    }
    __syncthreads();
    for (int s = 0; s < kPairThreads; ++s) {
      const R4 pj = s_pos[s];
      R4 fj[NR];
#pragma unroll
      for (int r = 0; r < NR; ++r) fj[r] = s_f[r * kPairThreads + s];
#pragma unroll
      for (int k = 0; k < TPL; ++k) {
        UFCoef<R> c;
        if constexpr (WALL) {
          uf_coef_wall<R>(pi[k].x, pi[k].y, pi[k].z, pi[k].w, pj.x, pj.y, pj.z, pj.w, c);
#pragma unroll
          for (int r = 0; r < NR; ++r) uf_apply_wall<R>(c, fj[r].x, fj[r].y, fj[r].z, acc[k][r]);
        } else {
          c.dx = pi[k].x - pj.x;
          c.dy = pi[k].y - pj.y;
          c.dz = pi[k].z - pj.z;
          const R r2 = c.dx * c.dx + c.dy * c.dy + c.dz * c.dz;
          const R ir = lm_rsqrt(r2);
          RpyCoef<R> cf{};
          rpy_far_uf34<R>(ir, cf);  // times 4/3, undone on the output
          if (__builtin_expect(r2 < R(4), 0)) {
            lm_rare_path();
            rpy_near_select_uf34<R>(r2, ir, cf);
          }
          c.a0 = cf.fuf;
          c.cd = cf.guf;
#pragma unroll
          for (int r = 0; r < NR; ++r) uf_apply_open<R>(c, fj[r].x, fj[r].y, fj[r].z, acc[k][r]);
        }
      }
    }
    if constexpr (std::is_same<R, float>::value) {
#pragma unroll
      for (int k = 0; k < TPL; ++k)
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          acc64[k][r].x += acc[k][r].x;
          acc64[k][r].y += acc[k][r].y;
          acc64[k][r].z += acc[k][r].z;
          acc[k][r] = {R(0), R(0), R(0)};
        }
    }
  }
  if constexpr (std::is_same<R, double>::value) {
#pragma unroll
    for (int k = 0; k < TPL; ++k)
#pragma unroll
      for (int r = 0; r < NR; ++r) acc64[k][r] = acc[k][r];
  }
#pragma unroll
  for (int k = 0; k < TPL; ++k) {
    const int loc = tid + k * kPairThreads;
    if (loc < it.ntgt) {
      const int64_t i = it.tgt0 + loc;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        double* o = U + ((int64_t)r * nch + it.chunk) * n3;
        const double tuf = WALL ? t0 : 0.75 * t0;
        o[3 * i + 0] = tuf * acc64[k][r].x;
        o[3 * i + 1] = tuf * acc64[k][r].y;
        o[3 * i + 2] = tuf * acc64[k][r].z;
      }
    }
  }
}

template <class R, int KIND, int NR>
__global__ __launch_bounds__(kPairThreads) void k_pair_multi(
    const PairItem* __restrict__ items, const typename Rec<R>::T* __restrict__ pos,
    const double* __restrict__ F, double* __restrict__ U, int64_t n3, int nch, double t0) {
  using R4 = typename Rec<R>::T;
  constexpr int TPL = multi_tpl(NR);
  __shared__ R4 s_pos[kPairThreads];
  __shared__ R4 s_f[NR * kPairThreads];
  const PairItem it = items[blockIdx.x];
  if constexpr (KIND == 0) {
    pair_item_multi<R, false, NR, TPL>(it, pos, F, U, n3, nch, t0, s_pos, s_f);
  } else if constexpr (KIND == 1) {
    pair_item_multi<R, true, NR, TPL>(it, pos, F, U, n3, nch, t0, s_pos, s_f);
  } else {
    if (it.wall)
      pair_item_multi<R, true, NR, TPL>(it, pos, F, U, n3, nch, t0, s_pos, s_f);
    else
      pair_item_multi<R, false, NR, TPL>(it, pos, F, U, n3, nch, t0, s_pos, s_f);
  }
}

template <class R, int KIND>
static void launch_pair_multi_k(const PairMultiArgs& a, hipStream_t s) {
  using R4 = typename Rec<R>::T;
  const R4* pos = static_cast<const R4*>(a.pos);
  if (a.nrhs == 4)
    hipLaunchKernelGGL((k_pair_multi<R, KIND, 4>), dim3(a.nitems), dim3(kPairThreads), 0, s,
                       a.items, pos, a.F, a.U, a.n3, a.nch, a.t0);
  else if (a.nrhs == 2)
    hipLaunchKernelGGL((k_pair_multi<R, KIND, 2>), dim3(a.nitems), dim3(kPairThreads), 0, s,
                       a.items, pos, a.F, a.U, a.n3, a.nch, a.t0);
  else
    throw Error("[libmobility_amd] multi-vector launch handles 2 or 4 vectors");
  LM_HIP(hipGetLastError());
}

void launch_pair_multi(const PairMultiArgs& a, hipStream_t s) {
  if (a.nitems <= 0) return;
  if (a.fp32) {
    if (a.kind == 0) launch_pair_multi_k<float, 0>(a, s);
    else if (a.kind == 1) launch_pair_multi_k<float, 1>(a, s);
    else launch_pair_multi_k<float, 2>(a, s);
  } else {
    if (a.kind == 0) launch_pair_multi_k<double, 0>(a, s);
    else if (a.kind == 1) launch_pair_multi_k<double, 1>(a, s);
    else launch_pair_multi_k<double, 2>(a, s);
  }
}

template <class R, int KIND, int BLK, int TPL>
static void launch_pair_t(const PairArgs& a, hipStream_t s) {
  using R4 = typename Rec<R>::T;
  if (a.persistent) {
    if (!a.counter) throw Error("[libmobility_amd] persistent launch without a work counter");
    launch_zero_u32(a.counter, 1, s);
    hipLaunchKernelGGL((k_pair<R, KIND, BLK, TPL, true>), dim3(a.grid), dim3(kPairThreads), 0, s,
                       a.items, a.nitems, a.counter, static_cast<const R4*>(a.pos), a.F, a.T, a.U,
                       a.W, a.slab, a.active, a.stop, a.t0, a.rt0, a.r0);
  } else {
    hipLaunchKernelGGL((k_pair<R, KIND, BLK, TPL, false>), dim3(a.nitems), dim3(kPairThreads), 0,
                       s, a.items, a.nitems, a.counter, static_cast<const R4*>(a.pos), a.F, a.T,
                       a.U, a.W, a.slab, a.active, a.stop, a.t0, a.rt0, a.r0);
  }
  LM_HIP(hipGetLastError());
}

template <class R, int KIND, int BLK>
static void launch_pair_tpl(const PairArgs& a, hipStream_t s) {
  switch (a.tpl) {
    case 1: launch_pair_t<R, KIND, BLK, 1>(a, s); break;
    case 2: launch_pair_t<R, KIND, BLK, 2>(a, s); break;
    case 4:
      if constexpr (KIND == 0 && BLK == kUF) {
        launch_pair_t<R, KIND, BLK, 4>(a, s);
        break;
      }
      [[fallthrough]];
    default: throw Error("[libmobility_amd] unsupported targets-per-lane for this kernel");
  }
}

template <class R, int KIND>
static void launch_pair_blk(const PairArgs& a, hipStream_t s) {
  switch (a.blocks) {
    case kUF: launch_pair_tpl<R, KIND, kUF>(a, s); break;
    case kUF | kWF: launch_pair_tpl<R, KIND, kUF | kWF>(a, s); break;
    case kUT | kWT: launch_pair_tpl<R, KIND, kUT | kWT>(a, s); break;
    case kUF | kUT | kWT | kWF: launch_pair_tpl<R, KIND, kUF | kUT | kWT | kWF>(a, s); break;
    default: throw Error("[libmobility_amd] unsupported block combination");
  }
}

template <class R>
static void launch_pair_kind(const PairArgs& a, hipStream_t s) {
  switch (a.wall) {
    case 0: launch_pair_blk<R, 0>(a, s); break;
    case 1: launch_pair_blk<R, 1>(a, s); break;
    case 2: launch_pair_blk<R, 2>(a, s); break;
    default: throw Error("[libmobility_amd] invalid kernel kind");
  }
}

void launch_pair(const PairArgs& a, hipStream_t s) {
  if (a.nitems <= 0) return;
  if (a.fp32)
    launch_pair_kind<float>(a, s);
  else
    launch_pair_kind<double>(a, s);
}

template <class R, int KIND, int BLK>
static void launch_pair_sym_tpl(const SymArgs& a, hipStream_t s) {
  const dim3 grid((unsigned)a.nitems), block(kPairThreads);
  switch (a.tpl) {
    case 1: hipLaunchKernelGGL((k_pair_sym<R, KIND, BLK, 1>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_pair_sym<R, KIND, BLK, 2>), grid, block, 0, s, a); break;
    case 4:
      if constexpr (KIND == 0 && BLK == kUF) {
        hipLaunchKernelGGL((k_pair_sym<R, KIND, BLK, 4>), grid, block, 0, s, a);
        break;
      }
      [[fallthrough]];
    default: throw Error("[libmobility_amd] unsupported targets-per-lane for the symmetric kernel");
  }
  LM_HIP(hipGetLastError());
}

template <class R, int KIND>
static void launch_pair_sym_blk(const SymArgs& a, hipStream_t s) {
  switch (a.blocks) {
    case kUF: launch_pair_sym_tpl<R, KIND, kUF>(a, s); break;
    case kUF | kWF: launch_pair_sym_tpl<R, KIND, kUF | kWF>(a, s); break;
    case kUT | kWT: launch_pair_sym_tpl<R, KIND, kUT | kWT>(a, s); break;
    case kUF | kUT | kWT | kWF: launch_pair_sym_tpl<R, KIND, kUF | kUT | kWT | kWF>(a, s); break;
    default: throw Error("[libmobility_amd] unsupported block combination");
  }
}

template <class R>
static void launch_pair_sym_r(const SymArgs& a, hipStream_t s) {
  switch (a.wall) {
    case 0: launch_pair_sym_blk<R, 0>(a, s); break;
    case 1: launch_pair_sym_blk<R, 1>(a, s); break;
    case 2: launch_pair_sym_blk<R, 2>(a, s); break;
    default: throw Error("[libmobility_amd] invalid kernel kind");
  }
}

void launch_pair_sym(const SymArgs& a, hipStream_t s) {
  if (a.nitems <= 0) return;
  if (a.fp32)
    launch_pair_sym_r<float>(a, s);
  else
    launch_pair_sym_r<double>(a, s);
}

// One workgroup per row block (at most 256 particles of one tile of one
// system), one thread per entry (3 x 256).
__global__ __launch_bounds__(3 * kPairThreads) void k_sym_reduce(SymReduceArgs r) {
  const SymRows rb = r.rows[blockIdx.x];
  if (r.active != nullptr && r.active[rb.sys] == 0) return;
  const SymSys S = r.sys[rb.sys];
  const int64_t kTile = (int64_t)kPairThreads * r.tpl;
  const int tile = (int)((rb.p0 - S.off) / kTile);
  const int hi = min(r.o1, S.hmax + 1);  // this system's offsets in the pass: [o0, hi)
  const int width = hi - r.o0;
  const int nchp = width > 0 ? min(r.nch, width) : 0;
  // an even tile count meets offset ntiles/2 only from tiles I < ntiles/2, so
  // that slab holds no rows of tiles J < ntiles/2
  const int skip = (S.ntiles % 2 == 0 && 2 * tile < S.ntiles) ? S.ntiles / 2 : 0;
  const int64_t n = r.n3;
  for (int64_t e = 3 * rb.p0 + threadIdx.x; e < 3 * (rb.p0 + rb.np); e += blockDim.x) {
    double s = r.accumulate ? r.out[e] : 0.0;
    for (int c = 0; c < nchp; ++c) s += r.slab_t[(int64_t)c * n + e];
    if (r.diag) s += r.slab_t[(int64_t)r.nch * n + e];
    for (int o = r.o0; o < hi; ++o)
      if (o != skip) s += r.slab_s[(int64_t)(o - r.o0) * n + e];
    r.out[e] = s;
  }
}

void launch_sym_reduce(const SymReduceArgs& r, hipStream_t s) {
  if (r.nrows <= 0) return;
  hipLaunchKernelGGL(k_sym_reduce, dim3((unsigned)r.nrows), dim3(3 * kPairThreads), 0, s, r);
  LM_HIP(hipGetLastError());
}

__global__ void k_chunk_reduce(const double* __restrict__ part, int nch, int64_t n,
                               double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int c = 0; c < nch; ++c) s += part[(int64_t)c * n + i];
  out[i] = s;
}

void launch_chunk_reduce(const double* part, int nch, int64_t n, double* out, hipStream_t s) {
  if (n <= 0) return;
  const int nt = 256;
  hipLaunchKernelGGL(k_chunk_reduce, dim3((unsigned)((n + nt - 1) / nt)), dim3(nt), 0, s, part,
                     nch, n, out);
  LM_HIP(hipGetLastError());
}

// One workgroup per particle block; blocks of converged systems return at once.
__global__ __launch_bounds__(256) void k_seg_chunk_reduce(const double* __restrict__ part,
                                                          int nch, int64_t n,
                                                          double* __restrict__ out,
                                                          const SegBlock* __restrict__ blocks,
                                                          const int* __restrict__ active) {
  const SegBlock b = blocks[blockIdx.x];
  if (active[b.sys] == 0) return;
  for (int64_t i = 3 * b.p0 + threadIdx.x; i < 3 * b.p1; i += blockDim.x) {
    double s = 0.0;
    for (int c = 0; c < nch; ++c) s += part[(int64_t)c * n + i];
    out[i] = s;
  }
}

void launch_seg_chunk_reduce(const double* part, int nch, int64_t n, double* out,
                             const SegBlock* blocks, int nblocks, const int* active,
                             hipStream_t s) {
  if (n <= 0 || nblocks <= 0) return;
  hipLaunchKernelGGL(k_seg_chunk_reduce, dim3((unsigned)nblocks), dim3(256), 0, s, part, nch, n,
                     out, blocks, active);
  LM_HIP(hipGetLastError());
}

// ---- setPositions -----------------------------------------------------------

__global__ __launch_bounds__(256) void k_system_bbox(const double* __restrict__ pos,
                                                     const int64_t* __restrict__ off,
                                                     double wall_shift_in,
                                                     const int* __restrict__ kinds, int wall,
                                                     double* __restrict__ bbox) {
  const int s = blockIdx.x;
  const int64_t b = off[s], e = off[s + 1];
  const double wall_shift = (kinds ? kinds[s] : wall) ? wall_shift_in : 0.0;
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
    for (int c = 0; c < 3; ++c) {
      double v = pos[3 * i + c];
      if (c == 2) v -= wall_shift;
      mn[c] = fmin(mn[c], v);
      mx[c] = fmax(mx[c], v);
    }
  }
  __shared__ double sh[6][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int c = 0; c < 3; ++c) {
    for (int o = 32; o > 0; o >>= 1) {
      mn[c] = fmin(mn[c], __shfl_xor(mn[c], o));
      mx[c] = fmax(mx[c], __shfl_xor(mx[c], o));
    }
    if (lane == 0) {
      sh[c][wid] = mn[c];
      sh[3 + c][wid] = mx[c];
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int c = threadIdx.x;
    double v = sh[c][0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) v = (c < 3) ? fmin(v, sh[c][w]) : fmax(v, sh[c][w]);
    bbox[6 * s + c] = v;
  }
}

void launch_system_bbox(const double* pos, const int64_t* sys_off, int nsys, double wall_shift,
                        const int* kinds, int wall, double* bbox, hipStream_t s) {
  if (nsys <= 0) return;
  hipLaunchKernelGGL(k_system_bbox, dim3(nsys), dim3(256), 0, s, pos, sys_off, wall_shift, kinds,
                     wall, bbox);
  LM_HIP(hipGetLastError());
}

__global__ void k_make_records(const double* __restrict__ pos, int64_t N,
                               const int64_t* __restrict__ off, int nsys,
                               const double* __restrict__ bbox, const int* __restrict__ kinds,
                               int wall_all, double wall_shift, double a,
                               double* __restrict__ shifted, double4* __restrict__ rec64,
                               float4* __restrict__ rec32) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int lo = 0, hi = nsys;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid;
  }
  const double* bb = bbox + 6 * lo;
  const int wall = kinds ? kinds[lo] : wall_all;
  const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2] - (wall ? wall_shift : 0.0);
  if (shifted) {
    shifted[3 * i] = x;
    shifted[3 * i + 1] = y;
    shifted[3 * i + 2] = z;
  }
  const double ia = 1.0 / a;
  const double cx = 0.5 * (bb[0] + bb[3]), cy = 0.5 * (bb[1] + bb[4]), cz = 0.5 * (bb[2] + bb[5]);
  if (wall) {
    const double h = fmax(z, a) * ia;
    const double b = fmin(fmax(z * ia, 0.0), 1.0);
    rec64[i] = double4{x * ia, y * ia, h, b};
    rec32[i] = float4{(float)((x - cx) * ia), (float)((y - cy) * ia), (float)h, (float)b};
  } else {
    rec64[i] = double4{x * ia, y * ia, z * ia, 1.0};
    rec32[i] = float4{(float)((x - cx) * ia), (float)((y - cy) * ia), (float)((z - cz) * ia), 1.0f};
  }
}

void launch_make_records(const double* pos, int64_t N, const int64_t* sys_off, int nsys,
                         const double* bbox, const int* kinds, int wall, double wall_shift,
                         double a, double* shifted, void* rec64, void* rec32, hipStream_t s) {
  if (N <= 0) return;
  const int nt = 256;
  hipLaunchKernelGGL(k_make_records, dim3((unsigned)((N + nt - 1) / nt)), dim3(nt), 0, s, pos, N,
                     sys_off, nsys, bbox, kinds, wall, wall_shift, a, shifted,
                     static_cast<double4*>(rec64), static_cast<float4*>(rec32));
  LM_HIP(hipGetLastError());
}

__global__ void k_zero_open(double* __restrict__ v, int64_t N, const int64_t* __restrict__ off,
                            int nsys, const int* __restrict__ kinds) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int lo = 0, hi = nsys;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid;
  }
  if (!kinds[lo]) {
    v[3 * i] = 0.0;
    v[3 * i + 1] = 0.0;
    v[3 * i + 2] = 0.0;
  }
}

void launch_zero_open_systems(double* v, int64_t N, const int64_t* sys_off, int nsys,
                              const int* kinds, hipStream_t s) {
  if (N <= 0 || !kinds) return;
  hipLaunchKernelGGL(k_zero_open, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, v, N, sys_off,
                     nsys, kinds);
  LM_HIP(hipGetLastError());
}

}  // namespace lm
