// Pair arithmetic of the NBody mobility operator, shared by the gfx950 kernels
// and by the host-side arithmetic self-check (lm_selftest_pair_block_host).
//
// Geometry is in units of the hydrodynamic radius a.  Position records are
// (x, y, z, b): for open boundaries z is the coordinate and b = 1; for the
// bottom wall z is the clamped height max(z, a)/a and b = clamp(z/a, 0, 1) is
// the damping factor of Usabiaga et al. 2017 (App. 1).
//
// The four 3x3 blocks follow solvers/NBody/extra/hydrodynamicKernels.cuh:
//   open RPY       UF :41-57,143-150   WT :63-86,152-159   UT/WF :89-118,161-173
//   wall terms     UF :187-249  WT :253-282  UT :287-319,416-431  WF :324-359
//   damping/clamp  :362-390
// The wall corrections are regrouped into vector form (fewer operations per
// pair than the reference's entry-by-entry 3x3 application):
//   UF: c = f1 v + f2 e (e.v) + f3 v_z e + z^ (f4 (e.v) + f5 v_z)
//   WT: c = g1 v + g2 e (e.v) + g4 p (e_y, -e_x, 0) + z^ g3 (e.v),  p = e_y v_x - e_x v_y
//   WF: c = (k3 e_y (e.v) - (k1 e_z + k4) v_y + (k1 + k2) e_y v_z,
//            (k1 e_z + k4) v_x - k3 e_x (e.v) - (k1 + k2) e_x v_z,  -k1 p)
//   UT: same factors with R' = (-d_x, -d_y, R_z) and h = h_i
// with R = (d_x, d_y, h_i + h_j), e = R/|R|.  The regrouping is exact algebra;
// the CPU oracle (explicit reference matrices) checks it.
#pragma once
#include <hip/hip_runtime.h>

namespace lm {

enum BlockBits : int { kUF = 1, kUT = 2, kWT = 4, kWF = 8 };

template <class R>
struct Vec3 {
  R x, y, z;
};

template <class R>
struct PairAcc {
  Vec3<R> uf, ut, wt, wf;
};

template <class R>
__host__ __device__ inline void acc_zero(PairAcc<R>& a) {
  a.uf = {R(0), R(0), R(0)};
  a.ut = {R(0), R(0), R(0)};
  a.wt = {R(0), R(0), R(0)};
  a.wf = {R(0), R(0), R(0)};
}

// Marks the branch it is placed in as one that stays a branch: an opaque
// statement, so the compiler does not evaluate that branch's arithmetic for
// every lane and select afterwards.  For the self-pair terms (r == 0), which
// are rare: off the diagonal tiles only coincident distinct particles take them.
__host__ __device__ inline void lm_rare_path() { __asm__ __volatile__(""); }

__host__ __device__ inline float lm_rsqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / __builtin_sqrtf(x);
#endif
}
// FP64: hardware estimate refined by one Newton step, y (1 + h/2) with
// h = 1 - x y^2 (relative error 1.5 e0^2 for an estimate error e0).  No
// special-value handling: x = 0 only occurs for self pairs, whose
// coefficients are replaced by the overlap branch.
__host__ __device__ inline double lm_rsqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double y = __builtin_amdgcn_rsq(x);
  const double h = __builtin_fma(-x * y, y, 1.0);
  return __builtin_fma(0.5 * y, h, y);
#else
  return 1.0 / __builtin_sqrt(x);
#endif
}
__host__ __device__ inline float lm_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}
__host__ __device__ inline double lm_rcp(double x) { return 1.0 / x; }
__host__ __device__ inline float lm_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__host__ __device__ inline double lm_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__host__ __device__ inline float lm_sqrt(float x) { return __builtin_sqrtf(x); }
__host__ __device__ inline double lm_sqrt(double x) { return __builtin_sqrt(x); }

// Open-boundary RPY coefficients for r2 = |d|^2 (scaled units).
//   fuf, guf : f and g/r^2 of M_UF
//   fwt, gwt : f and g/r^2 of M_WT
//   c        : coupling coefficient of U = c (T x d), W = c (F x d)
// Far field (r >= 2) uses rsqrt; the overlap branch (r < 2, including the
// self pair r = 0) is the regularised RPY of hydrodynamicKernels.cuh:51-55,
// :73-84, :97-100.  It is a divergent branch so that waves with no near pair
// skip it.  For r = 0 the far-field values are infinite; they are overwritten
// (never combined) by the overlap branch.
template <class R>
struct RpyCoef {
  R fuf, guf, fwt, gwt, c;
};

// Far-field (r >= 2) coefficients from ir = 1/r.
template <class R, int BLK>
__host__ __device__ inline void rpy_far(R ir, RpyCoef<R>& k) {
  const R ir2 = ir * ir;
  const R ir3 = ir * ir2;
  if (BLK & kUF) {
    k.fuf = ir * (R(0.75) + R(0.5) * ir2);
    k.guf = ir3 * (R(0.75) - R(1.5) * ir2);
  }
  if (BLK & kWT) {
    k.fwt = R(-0.5) * ir3;
    k.gwt = R(1.5) * ir3 * ir2;
  }
  if (BLK & (kUT | kWF)) k.c = ir3;
}

// Overlap (r < 2, including the self pair r = 0) coefficients; r and 1/r come
// from the rsqrt already computed (no sqrt, no division).
template <class R, int BLK>
__host__ __device__ inline void rpy_near(R r2, R ir, RpyCoef<R>& k) {
  const R r = r2 > R(0) ? r2 * ir : R(0);
  const R ir_n = r2 > R(0) ? ir : R(0);
  if (BLK & kUF) {
    k.fuf = R(1) - R(0.28125) * r;
    k.guf = R(0.09375) * ir_n;
  }
  if (BLK & kWT) {
    k.fwt = R(1) - R(0.84375) * r + R(0.078125) * r * r * r;
    k.gwt = r > R(0) ? (R(0.28125) * ir_n - R(0.046875) * r) : R(0);
  }
  if (BLK & (kUT | kWF)) k.c = R(0.5) * (R(1) - R(0.375) * r);
}

// Replace k by the overlap coefficients when r2 < 4, by selects (no branch).
template <class R, int BLK>
__host__ __device__ inline void rpy_near_select(R r2, R ir, RpyCoef<R>& k) {
  RpyCoef<R> n = k;
  rpy_near<R, BLK>(r2, ir, n);
  const bool is_near = r2 < R(4);
  if (BLK & kUF) {
    k.fuf = is_near ? n.fuf : k.fuf;
    k.guf = is_near ? n.guf : k.guf;
  }
  if (BLK & kWT) {
    k.fwt = is_near ? n.fwt : k.fwt;
    k.gwt = is_near ? n.gwt : k.gwt;
  }
  if (BLK & (kUT | kWF)) k.c = is_near ? n.c : k.c;
}

template <class R, int BLK>
__host__ __device__ inline void rpy_coeffs(R r2, R& fuf, R& guf, R& fwt, R& gwt, R& c) {
  const R ir = lm_rsqrt(r2);
  RpyCoef<R> k{};
  rpy_far<R, BLK>(ir, k);
  if (__builtin_expect(r2 < R(4), 0)) {
#if defined(__HIP_DEVICE_COMPILE__)
    // An opaque statement keeps this a branch: speculating it (if-conversion)
    // would put the overlap arithmetic on every pair.
    __asm__ __volatile__("");
#endif
    rpy_near<R, BLK>(r2, ir, k);
  }
  fuf = k.fuf;
  guf = k.guf;
  fwt = k.fwt;
  gwt = k.gwt;
  c = k.c;
}

// Open-boundary contribution of one source (f = force, t = torque) to one
// target with given coefficients, added into a (dimensionless; the caller
// applies t0, rt0, r0).
template <class R, int BLK>
__host__ __device__ inline void open_accumulate(R dx, R dy, R dz, R fx, R fy, R fz, R tx, R ty,
                                                R tz, const RpyCoef<R>& k, PairAcc<R>& a) {
  const R fuf = k.fuf, guf = k.guf, fwt = k.fwt, gwt = k.gwt, c = k.c;
  // Accumulate as two FMAs per component: acc = fma(f, v, acc); acc = fma(s, d, acc).
  if (BLK & kUF) {
    const R s = guf * (dx * fx + dy * fy + dz * fz);
    a.uf.x = lm_fma(s, dx, lm_fma(fuf, fx, a.uf.x));
    a.uf.y = lm_fma(s, dy, lm_fma(fuf, fy, a.uf.y));
    a.uf.z = lm_fma(s, dz, lm_fma(fuf, fz, a.uf.z));
  }
  if (BLK & kWT) {
    const R s = gwt * (dx * tx + dy * ty + dz * tz);
    a.wt.x = lm_fma(s, dx, lm_fma(fwt, tx, a.wt.x));
    a.wt.y = lm_fma(s, dy, lm_fma(fwt, ty, a.wt.y));
    a.wt.z = lm_fma(s, dz, lm_fma(fwt, tz, a.wt.z));
  }
  if (BLK & kUT) {  // c (T x d)
    a.ut.x = lm_fma(c, ty * dz - tz * dy, a.ut.x);
    a.ut.y = lm_fma(c, tz * dx - tx * dz, a.ut.y);
    a.ut.z = lm_fma(c, tx * dy - ty * dx, a.ut.z);
  }
  if (BLK & kWF) {  // c (F x d)
    a.wf.x = lm_fma(c, fy * dz - fz * dy, a.wf.x);
    a.wf.y = lm_fma(c, fz * dx - fx * dz, a.wf.y);
    a.wf.z = lm_fma(c, fx * dy - fy * dx, a.wf.z);
  }
}

template <class R, int BLK>
__host__ __device__ inline void open_pair(R dx, R dy, R dz, R fx, R fy, R fz, R tx, R ty, R tz,
                                          PairAcc<R>& a) {
  const R r2 = dx * dx + dy * dy + dz * dz;
  RpyCoef<R> k{};
  rpy_coeffs<R, BLK>(r2, k.fuf, k.guf, k.fwt, k.gwt, k.c);
  open_accumulate<R, BLK>(dx, dy, dz, fx, fy, fz, tx, ty, tz, k, a);
}

// ---- translational block (UF) split into a per-pair part and a per-vector part
// for products with several force vectors at once.  For one pair
//   U += a0 F + cd (d.F) d + (e.F) g + F_z h
// holds the open RPY part (a0 = b fuf, cd = b guf) and, for the wall, the
// image terms regrouped as above: a0 += b f1, g = b (f2 e + f4 z^),
// h = b (f3 e + f5 z^); self pairs (r = 0) have e = g = 0 and
// a0 = b (fuf + cp), h = (0, 0, b (cz - cp)).  b is the damping b_i b_j.
template <class R>
struct UFCoef {
  R a0, cd, dx, dy, dz, ex, ey, ez, gx, gy, gz, hx, hy, hz;
};

// Open boundaries: coefficients from d and the RPY coefficients (e, g, h unused).
template <class R>
__host__ __device__ inline void uf_apply_open(const UFCoef<R>& k, R fx, R fy, R fz, Vec3<R>& u) {
  const R s = k.cd * (k.dx * fx + k.dy * fy + k.dz * fz);
  u.x = lm_fma(s, k.dx, lm_fma(k.a0, fx, u.x));
  u.y = lm_fma(s, k.dy, lm_fma(k.a0, fy, u.y));
  u.z = lm_fma(s, k.dz, lm_fma(k.a0, fz, u.z));
}

template <class R>
__host__ __device__ inline void uf_apply_wall(const UFCoef<R>& k, R fx, R fy, R fz, Vec3<R>& u) {
  const R s = k.cd * (k.dx * fx + k.dy * fy + k.dz * fz);
  const R ev = k.ex * fx + k.ey * fy + k.ez * fz;
  u.x = lm_fma(fz, k.hx, lm_fma(ev, k.gx, lm_fma(s, k.dx, lm_fma(k.a0, fx, u.x))));
  u.y = lm_fma(fz, k.hy, lm_fma(ev, k.gy, lm_fma(s, k.dy, lm_fma(k.a0, fy, u.y))));
  u.z = lm_fma(fz, k.hz, lm_fma(ev, k.gz, lm_fma(s, k.dz, lm_fma(k.a0, fz, u.z))));
}

// Wall UF coefficients of target record (xi, yi, hi, bi) and source (xj, yj, hj, bj).
template <class R>
__host__ __device__ inline void uf_coef_wall(R xi, R yi, R hi, R bi, R xj, R yj, R hj, R bj,
                                             UFCoef<R>& k) {
  const R dx = xi - xj, dy = yi - yj, dz = hi - hj;
  const R r2 = dx * dx + dy * dy + dz * dz;
  const R b = bi * bj;
  R fuf, guf, fwt, gwt, c;
  rpy_coeffs<R, kUF>(r2, fuf, guf, fwt, gwt, c);
  k.dx = dx;
  k.dy = dy;
  k.dz = dz;
  k.cd = b * guf;
  if (r2 != R(0)) {
    const R Rz = hi + hj;
    const R iR = lm_rsqrt(dx * dx + dy * dy + Rz * Rz);
    const R iR2 = iR * iR;
    const R iR3 = iR2 * iR;
    const R iR5 = iR3 * iR2;
    const R ex = dx * iR, ey = dy * iR, ez = Rz * iR;
    const R ez2 = ez * ez;
    const R hh = hj * lm_rcp(Rz);
    const R t1 = R(5) * ez2 - R(1);
    const R hq = hh * (R(1) - hh) * ez2;
    const R f1 = R(-0.25) * (R(3) * (R(1) + R(2) * hq) * iR + R(2) * (R(1) - R(3) * ez2) * iR3 +
                             R(2) * t1 * iR5);
    const R f2 = R(-0.25) * (R(3) * (R(1) - R(6) * hq) * iR + R(6) * t1 * iR3 +
                             R(10) * (R(1) - R(7) * ez2) * iR5);
    const R f3 = R(0.5) * ez *
                 (R(3) * hh * (R(1) - R(6) * (R(1) - hh) * ez2) * iR + R(6) * t1 * iR3 +
                  R(10) * (R(2) - R(7) * ez2) * iR5);
    const R f4 = R(0.5) * ez * (R(3) * hh * iR - R(10) * iR5);
    const R f5 = -(R(3) * hh * hh * ez2 * iR + R(3) * ez2 * iR3 + (R(2) - R(15) * ez2) * iR5);
    k.a0 = b * (fuf + f1);
    k.ex = ex;
    k.ey = ey;
    k.ez = ez;
    k.gx = b * f2 * ex;
    k.gy = b * f2 * ey;
    k.gz = b * (f2 * ez + f4);
    k.hx = b * f3 * ex;
    k.hy = b * f3 * ey;
    k.hz = b * (f3 * ez + f5);
  } else {
    // self (Swan & Brady B1), as in wall_pair
    lm_rare_path();
    const R ih = R(1) / hj;
    const R ih2 = ih * ih;
    const R ih3 = ih2 * ih;
    const R ih5 = ih3 * ih2;
    const R cp = R(-0.0625) * (R(9) * ih - R(2) * ih3 + ih5);
    const R cz = R(-0.125) * (R(9) * ih - R(4) * ih3 + ih5);
    k.a0 = b * (fuf + cp);
    k.ex = k.ey = k.ez = R(0);
    k.gx = k.gy = k.gz = R(0);
    k.hx = k.hy = R(0);
    k.hz = b * (cz - cp);
  }
}

// Bottom-wall contribution (open RPY + Rotne-Prager-Blake/Swan-Brady image
// terms, damped by b_i b_j).  (xi, yi, hi, bi) and (xj, yj, hj, bj) are
// position records as described at the top of this file.
template <class R, int BLK>
__host__ __device__ inline void wall_pair(R xi, R yi, R hi, R bi, R xj, R yj, R hj, R bj, R fx,
                                          R fy, R fz, R tx, R ty, R tz, PairAcc<R>& a) {
  const R dx = xi - xj, dy = yi - yj, dz = hi - hj;
  const R r2 = dx * dx + dy * dy + dz * dz;
  const R bij = bi * bj;
  R fuf, guf, fwt, gwt, c;
  rpy_coeffs<R, BLK>(r2, fuf, guf, fwt, gwt, c);
  Vec3<R> uf = {R(0), R(0), R(0)}, ut = uf, wt = uf, wf = uf;
  if (r2 != R(0)) {
    // Image system: R = (dx, dy, hi + hj), e = R/|R|, hh = h_j/R_z, hh' = h_i/R_z.
    const R Rz = hi + hj;
    const R iR = lm_rsqrt(dx * dx + dy * dy + Rz * Rz);
    const R iR2 = iR * iR;
    const R iR3 = iR2 * iR;
    const R ex = dx * iR, ey = dy * iR, ez = Rz * iR;
    const R ez2 = ez * ez;
    const R iRz = lm_rcp(Rz);
    if (BLK & kUF) {
      const R hh = hj * iRz;
      const R iR5 = iR3 * iR2;
      const R t1 = R(5) * ez2 - R(1);  // -(1 - 5 ez2)
      const R hq = hh * (R(1) - hh) * ez2;
      const R f1 = R(-0.25) * (R(3) * (R(1) + R(2) * hq) * iR + R(2) * (R(1) - R(3) * ez2) * iR3 +
                               R(2) * t1 * iR5);
      const R f2 = R(-0.25) * (R(3) * (R(1) - R(6) * hq) * iR + R(6) * t1 * iR3 +
                               R(10) * (R(1) - R(7) * ez2) * iR5);
      const R f3 = R(0.5) * ez *
                   (R(3) * hh * (R(1) - R(6) * (R(1) - hh) * ez2) * iR + R(6) * t1 * iR3 +
                    R(10) * (R(2) - R(7) * ez2) * iR5);
      const R f4 = R(0.5) * ez * (R(3) * hh * iR - R(10) * iR5);
      const R f5 = -(R(3) * hh * hh * ez2 * iR + R(3) * ez2 * iR3 + (R(2) - R(15) * ez2) * iR5);
      const R ev = ex * fx + ey * fy + ez * fz;
      const R q = f2 * ev + f3 * fz;
      uf.x = f1 * fx + q * ex;
      uf.y = f1 * fy + q * ey;
      uf.z = f1 * fz + q * ez + f4 * ev + f5 * fz;
    }
    if (BLK & kWT) {
      const R g1 = R(0.5) * (R(1) - R(6) * ez2) * iR3;
      const R g2 = R(-1.5) * iR3;
      const R g3 = R(3) * iR3 * ez;
      const R g4 = R(3) * iR3;
      const R ev = ex * tx + ey * ty + ez * tz;
      const R p = g4 * (ey * tx - ex * ty);
      wt.x = g1 * tx + g2 * ev * ex + p * ey;
      wt.y = g1 * ty + g2 * ev * ey - p * ex;
      wt.z = g1 * tz + (g2 * ez + g3) * ev;
    }
    if (BLK & (kWF | kUT)) {
      const R iR4 = iR2 * iR2;
      const R k1 = iR2;
      if (BLK & kWF) {
        const R hh = hj * iRz;
        const R k2 = R(2) * (R(6) * hh * ez2 * iR2 + (R(1) - R(10) * ez2) * iR4);
        const R k3 = R(-2) * ez * (R(3) * hh * iR2 - R(5) * iR4);
        const R k4 = R(-2) * ez * (hh * iR2 - iR4);
        const R ev = ex * fx + ey * fy + ez * fz;
        const R p = ey * fx - ex * fy;
        const R kz = k1 * ez + k4;
        const R k12 = k1 + k2;
        wf.x = k3 * ey * ev - kz * fy + k12 * ey * fz;
        wf.y = kz * fx - k3 * ex * ev - k12 * ex * fz;
        wf.z = -k1 * p;
      }
      if (BLK & kUT) {
        const R hh = hi * iRz;
        const R k2 = R(2) * (R(6) * hh * ez2 * iR2 + (R(1) - R(10) * ez2) * iR4);
        const R k3 = R(-2) * ez * (R(3) * hh * iR2 - R(5) * iR4);
        const R k4 = R(-2) * ez * (hh * iR2 - iR4);
        const R p = ey * tx - ex * ty;
        const R kz = k1 * ez + k4;
        ut.x = k3 * ex * p + kz * ty + k1 * ey * tz;
        ut.y = k3 * ey * p - kz * tx - k1 * ex * tz;
        ut.z = -(k1 + k2 + k3 * ez) * p;
      }
    }
  } else {
    // Self terms (r == 0), Swan & Brady B1, B2, B3 as in :190-202, :256-261,
    // :290-294, :327-331.  h_i == h_j here.
    lm_rare_path();
    const R ih = R(1) / hj;
    const R ih2 = ih * ih;
    const R ih3 = ih2 * ih;
    if (BLK & kUF) {
      const R ih5 = ih3 * ih2;
      const R cp = R(-0.0625) * (R(9) * ih - R(2) * ih3 + ih5);
      const R cz = R(-0.125) * (R(9) * ih - R(4) * ih3 + ih5);
      uf.x = cp * fx;
      uf.y = cp * fy;
      uf.z = cz * fz;
    }
    if (BLK & kWT) {
      wt.x = R(-0.3125) * ih3 * tx;
      wt.y = R(-0.3125) * ih3 * ty;
      wt.z = R(-0.125) * ih3 * tz;
    }
    const R c4 = R(0.125) * ih3 * ih;
    if (BLK & kWF) {
      wf.x = -c4 * fy;
      wf.y = c4 * fx;
    }
    if (BLK & kUT) {
      ut.x = c4 * ty;
      ut.y = -c4 * tx;
    }
  }
  // open RPY part
  if (BLK & kUF) {
    const R s = guf * (dx * fx + dy * fy + dz * fz);
    a.uf.x += bij * (fuf * fx + s * dx + uf.x);
    a.uf.y += bij * (fuf * fy + s * dy + uf.y);
    a.uf.z += bij * (fuf * fz + s * dz + uf.z);
  }
  if (BLK & kWT) {
    const R s = gwt * (dx * tx + dy * ty + dz * tz);
    a.wt.x += bij * (fwt * tx + s * dx + wt.x);
    a.wt.y += bij * (fwt * ty + s * dy + wt.y);
    a.wt.z += bij * (fwt * tz + s * dz + wt.z);
  }
  if (BLK & kUT) {
    a.ut.x += bij * (c * (ty * dz - tz * dy) + ut.x);
    a.ut.y += bij * (c * (tz * dx - tx * dz) + ut.y);
    a.ut.z += bij * (c * (tx * dy - ty * dx) + ut.z);
  }
  if (BLK & kWF) {
    a.wf.x += bij * (c * (fy * dz - fz * dy) + wf.x);
    a.wf.y += bij * (c * (fz * dx - fx * dz) + wf.y);
    a.wf.z += bij * (c * (fx * dy - fy * dx) + wf.z);
  }
}

// ---- one pair, both directions (symmetric evaluation) ----------------------
// The grand mobility matrix is symmetric: blockwise M_ji = M_ij^T, with UF and
// WT transposing into themselves and the cross blocks swapping,
// M_UT(ji) = M_WF(ij)^T and M_WF(ji) = M_UT(ij)^T.  So the coefficients of the
// pair (i <- j) also give what j receives from i.  In the vector forms above,
// with q = (e_y, -e_x, 0):
//   UF^T v = f1 v + (f2 (e.v) + f4 v_z) e + z^ (f3 (e.v) + f5 v_z)   (f3, f4 swap)
//   WT^T v = g1 v + (g2 (e.v) + g3 v_z) e + g4 (q.v) q
//   WF:  B v = q (k3 (e.v) + (k1 + k2) v_z) + kz (z^ x v) - k1 (q.v) z^
//        B^T v = (k3 e + (k1 + k2) z^) (q.v) - kz (z^ x v) - k1 v_z q
//   UT:  C v = k3 (q.v) e - kz (z^ x v) + k1 v_z q + gam (q.v) z^,
//        gam = -(k1 + k2 + 2 k3 e_z)  (coefficients with h = h_i)
//        C^T v = (k3 (e.v) + gam v_z) q + kz (z^ x v) + k1 (q.v) z^
// and the open parts: UF, WT symmetric; c (v x d) transposes to -c (v x d).
// The target accumulates M_ij [F_j; T_j] into a (as wall_pair / open_pair);
// the source sums are kept in units of t0 (su: UF part + kut * UT part,
// kut = rt0/t0) and r0 (sw: WT part + kwf * WF part, kwf = rt0/r0).
template <class R, bool WALL, int BLK>
__host__ __device__ inline void pair_both(R xi, R yi, R hi, R bi, R xj, R yj, R hj, R bj,
                                          const Vec3<R>& Fj, const Vec3<R>& Tj,
                                          const Vec3<R>& Fi, const Vec3<R>& Ti, R kut, R kwf,
                                          PairAcc<R>& a, Vec3<R>& su, Vec3<R>& sw) {
  constexpr bool kCross = (BLK & (kUT | kWF)) != 0;
  const R dx = xi - xj, dy = yi - yj, dz = hi - hj;
  const R r2 = dx * dx + dy * dy + dz * dz;
  const R bij = WALL ? bi * bj : R(1);
  R fuf, guf, fwt, gwt, c;
  rpy_coeffs<R, BLK | (kCross ? (kUT | kWF) : 0)>(r2, fuf, guf, fwt, gwt, c);
  const Vec3<R> zero = {R(0), R(0), R(0)};
  Vec3<R> uf = zero, ut = zero, wt = zero, wf = zero;      // image part, to i
  Vec3<R> suf = zero, sut = zero, swt = zero, swf = zero;  // image part, to j
  if (WALL) {
    if (r2 != R(0)) {
      const R Rz = hi + hj;
      const R iR = lm_rsqrt(dx * dx + dy * dy + Rz * Rz);
      const R iR2 = iR * iR;
      const R iR3 = iR2 * iR;
      const R ex = dx * iR, ey = dy * iR, ez = Rz * iR;
      const R ez2 = ez * ez;
      const R iRz = lm_rcp(Rz);
      if (BLK & kUF) {
        const R hh = hj * iRz;
        const R iR5 = iR3 * iR2;
        const R t1 = R(5) * ez2 - R(1);
        const R hq = hh * (R(1) - hh) * ez2;
        const R f1 = R(-0.25) * (R(3) * (R(1) + R(2) * hq) * iR + R(2) * (R(1) - R(3) * ez2) * iR3 +
                                 R(2) * t1 * iR5);
        const R f2 = R(-0.25) * (R(3) * (R(1) - R(6) * hq) * iR + R(6) * t1 * iR3 +
                                 R(10) * (R(1) - R(7) * ez2) * iR5);
        const R f3 = R(0.5) * ez *
                     (R(3) * hh * (R(1) - R(6) * (R(1) - hh) * ez2) * iR + R(6) * t1 * iR3 +
                      R(10) * (R(2) - R(7) * ez2) * iR5);
        const R f4 = R(0.5) * ez * (R(3) * hh * iR - R(10) * iR5);
        const R f5 = -(R(3) * hh * hh * ez2 * iR + R(3) * ez2 * iR3 + (R(2) - R(15) * ez2) * iR5);
        const R ev = ex * Fj.x + ey * Fj.y + ez * Fj.z;
        const R q = f2 * ev + f3 * Fj.z;
        uf = {f1 * Fj.x + q * ex, f1 * Fj.y + q * ey, f1 * Fj.z + q * ez + f4 * ev + f5 * Fj.z};
        const R evs = ex * Fi.x + ey * Fi.y + ez * Fi.z;
        const R qs = f2 * evs + f4 * Fi.z;
        suf = {f1 * Fi.x + qs * ex, f1 * Fi.y + qs * ey, f1 * Fi.z + qs * ez + f3 * evs + f5 * Fi.z};
      }
      if (BLK & kWT) {
        const R g1 = R(0.5) * (R(1) - R(6) * ez2) * iR3;
        const R g2 = R(-1.5) * iR3;
        const R g3 = R(3) * iR3 * ez;
        const R g4 = R(3) * iR3;
        const R ev = ex * Tj.x + ey * Tj.y + ez * Tj.z;
        const R p = g4 * (ey * Tj.x - ex * Tj.y);
        wt = {g1 * Tj.x + g2 * ev * ex + p * ey, g1 * Tj.y + g2 * ev * ey - p * ex,
              g1 * Tj.z + (g2 * ez + g3) * ev};
        const R evs = ex * Ti.x + ey * Ti.y + ez * Ti.z;
        const R ps = g4 * (ey * Ti.x - ex * Ti.y);
        const R ss = g2 * evs + g3 * Ti.z;
        swt = {g1 * Ti.x + ss * ex + ps * ey, g1 * Ti.y + ss * ey - ps * ex, g1 * Ti.z + ss * ez};
      }
      if (kCross) {
        const R iR4 = iR2 * iR2;
        const R k1 = iR2;
        // WF factors (h = h_j) and UT factors (h = h_i)
        const R hw = hj * iRz, hu = hi * iRz;
        const R k2w = R(2) * (R(6) * hw * ez2 * iR2 + (R(1) - R(10) * ez2) * iR4);
        const R k3w = R(-2) * ez * (R(3) * hw * iR2 - R(5) * iR4);
        const R kzw = k1 * ez + R(-2) * ez * (hw * iR2 - iR4);
        const R k2u = R(2) * (R(6) * hu * ez2 * iR2 + (R(1) - R(10) * ez2) * iR4);
        const R k3u = R(-2) * ez * (R(3) * hu * iR2 - R(5) * iR4);
        const R kzu = k1 * ez + R(-2) * ez * (hu * iR2 - iR4);
        if (BLK & kWF) {
          // to i: B F_j
          const R ev = ex * Fj.x + ey * Fj.y + ez * Fj.z;
          const R p = ey * Fj.x - ex * Fj.y;
          const R k12 = k1 + k2w;
          wf = {k3w * ey * ev - kzw * Fj.y + k12 * ey * Fj.z,
                kzw * Fj.x - k3w * ex * ev - k12 * ex * Fj.z, -k1 * p};
          // to j: C^T F_i (W_j from F_i is M_UT(ij)^T)
          const R evs = ex * Fi.x + ey * Fi.y + ez * Fi.z;
          const R gam = -(k1 + k2u + R(2) * k3u * ez);
          const R s = k3u * evs + gam * Fi.z;
          swf = {s * ey - kzu * Fi.y, -s * ex + kzu * Fi.x, k1 * (ey * Fi.x - ex * Fi.y)};
        }
        if (BLK & kUT) {
          // to i: C T_j
          const R p = ey * Tj.x - ex * Tj.y;
          ut = {k3u * ex * p + kzu * Tj.y + k1 * ey * Tj.z,
                k3u * ey * p - kzu * Tj.x - k1 * ex * Tj.z, -(k1 + k2u + k3u * ez) * p};
          // to j: B^T T_i (U_j from T_i is M_WF(ij)^T)
          const R ps = ey * Ti.x - ex * Ti.y;
          const R k12 = k1 + k2w;
          sut = {k3w * ex * ps + kzw * Ti.y - k1 * ey * Ti.z,
                 k3w * ey * ps - kzw * Ti.x + k1 * ex * Ti.z, (k3w * ez + k12) * ps};
        }
      }
    } else {
      // Self terms (r == 0, h_i == h_j), Swan & Brady B1-B3 as in wall_pair;
      // each block equals its transpose's counterpart, so j gets the same forms.
      lm_rare_path();
      const R ih = R(1) / hj;
      const R ih2 = ih * ih;
      const R ih3 = ih2 * ih;
      if (BLK & kUF) {
        const R ih5 = ih3 * ih2;
        const R cp = R(-0.0625) * (R(9) * ih - R(2) * ih3 + ih5);
        const R cz = R(-0.125) * (R(9) * ih - R(4) * ih3 + ih5);
        uf = {cp * Fj.x, cp * Fj.y, cz * Fj.z};
        suf = {cp * Fi.x, cp * Fi.y, cz * Fi.z};
      }
      if (BLK & kWT) {
        wt = {R(-0.3125) * ih3 * Tj.x, R(-0.3125) * ih3 * Tj.y, R(-0.125) * ih3 * Tj.z};
        swt = {R(-0.3125) * ih3 * Ti.x, R(-0.3125) * ih3 * Ti.y, R(-0.125) * ih3 * Ti.z};
      }
      const R c4 = R(0.125) * ih3 * ih;
      if (BLK & kWF) {
        wf = {-c4 * Fj.y, c4 * Fj.x, R(0)};
        swf = {-c4 * Fi.y, c4 * Fi.x, R(0)};
      }
      if (BLK & kUT) {
        ut = {c4 * Tj.y, -c4 * Tj.x, R(0)};
        sut = {c4 * Ti.y, -c4 * Ti.x, R(0)};
      }
    }
  }
  if (BLK & kUF) {
    const R s = guf * (dx * Fj.x + dy * Fj.y + dz * Fj.z);
    a.uf.x += bij * (fuf * Fj.x + s * dx + uf.x);
    a.uf.y += bij * (fuf * Fj.y + s * dy + uf.y);
    a.uf.z += bij * (fuf * Fj.z + s * dz + uf.z);
    const R ss = guf * (dx * Fi.x + dy * Fi.y + dz * Fi.z);
    su.x += bij * (fuf * Fi.x + ss * dx + suf.x);
    su.y += bij * (fuf * Fi.y + ss * dy + suf.y);
    su.z += bij * (fuf * Fi.z + ss * dz + suf.z);
  }
  if (BLK & kWT) {
    const R s = gwt * (dx * Tj.x + dy * Tj.y + dz * Tj.z);
    a.wt.x += bij * (fwt * Tj.x + s * dx + wt.x);
    a.wt.y += bij * (fwt * Tj.y + s * dy + wt.y);
    a.wt.z += bij * (fwt * Tj.z + s * dz + wt.z);
    const R ss = gwt * (dx * Ti.x + dy * Ti.y + dz * Ti.z);
    sw.x += bij * (fwt * Ti.x + ss * dx + swt.x);
    sw.y += bij * (fwt * Ti.y + ss * dy + swt.y);
    sw.z += bij * (fwt * Ti.z + ss * dz + swt.z);
  }
  if (BLK & kUT) {
    a.ut.x += bij * (c * (Tj.y * dz - Tj.z * dy) + ut.x);
    a.ut.y += bij * (c * (Tj.z * dx - Tj.x * dz) + ut.y);
    a.ut.z += bij * (c * (Tj.x * dy - Tj.y * dx) + ut.z);
    const R b2 = kut * bij;
    su.x += b2 * (sut.x - c * (Ti.y * dz - Ti.z * dy));
    su.y += b2 * (sut.y - c * (Ti.z * dx - Ti.x * dz));
    su.z += b2 * (sut.z - c * (Ti.x * dy - Ti.y * dx));
  }
  if (BLK & kWF) {
    a.wf.x += bij * (c * (Fj.y * dz - Fj.z * dy) + wf.x);
    a.wf.y += bij * (c * (Fj.z * dx - Fj.x * dz) + wf.y);
    a.wf.z += bij * (c * (Fj.x * dy - Fj.y * dx) + wf.z);
    const R b2 = kwf * bij;
    sw.x += b2 * (swf.x - c * (Fi.y * dz - Fi.z * dy));
    sw.y += b2 * (swf.y - c * (Fi.z * dx - Fi.x * dz));
    sw.z += b2 * (swf.z - c * (Fi.x * dy - Fi.y * dx));
  }
}

}  // namespace lm
