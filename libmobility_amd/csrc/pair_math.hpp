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
// FP64: hardware estimate refined by two Newton steps y (2 - x y) (relative
// error e0^4 for an estimate error e0).  Callers pass finite x > 0.
__host__ __device__ inline double lm_rcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double y = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, y, 1.0);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-x, y, 1.0);
  return __builtin_fma(y, e, y);
#else
  return 1.0 / x;
#endif
}
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

// Open-boundary UF coefficients times 4/3, so that the far field costs one
// operation less: fuf = 1/r + (2/3)/r^3, guf = 1/r^3 - 2/r^5 (4/3 of
// rpy_far's); the caller scales its sums by 3/4.
template <class R>
__host__ __device__ inline void rpy_far_uf34(R ir, RpyCoef<R>& k) {
  const R ir2 = ir * ir;
  const R ir3 = ir * ir2;
  k.fuf = lm_fma(R(2.0 / 3.0), ir3, ir);
  k.guf = lm_fma(R(-2) * ir2, ir3, ir3);
}

// rpy_near_select for the 4/3-scaled coefficients
template <class R>
__host__ __device__ inline void rpy_near_select_uf34(R r2, R ir, RpyCoef<R>& k) {
  RpyCoef<R> n = k;
  rpy_near<R, kUF>(r2, ir, n);
  const bool is_near = r2 < R(4);
  k.fuf = is_near ? n.fuf * R(4.0 / 3.0) : k.fuf;
  k.guf = is_near ? n.guf * R(4.0 / 3.0) : k.guf;
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
  if (BLK & (kUT | kWF)) {  // v x (c d)
    const R cdx = c * dx, cdy = c * dy, cdz = c * dz;
    if (BLK & kUT) {  // c (T x d)
      a.ut.x = lm_fma(ty, cdz, lm_fma(-tz, cdy, a.ut.x));
      a.ut.y = lm_fma(tz, cdx, lm_fma(-tx, cdz, a.ut.y));
      a.ut.z = lm_fma(tx, cdy, lm_fma(-ty, cdx, a.ut.z));
    }
    if (BLK & kWF) {  // c (F x d)
      a.wf.x = lm_fma(fy, cdz, lm_fma(-fz, cdy, a.wf.x));
      a.wf.y = lm_fma(fz, cdx, lm_fma(-fx, cdz, a.wf.y));
      a.wf.z = lm_fma(fx, cdy, lm_fma(-fy, cdx, a.wf.z));
    }
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

// The bottom wall's UF image coefficients f1..f5 (their forms are stated in
// the comment on the bottom wall further down), each times the damping b, in
// Horner form in u = 1/|R|^2; bR = b/|R|, hw = h_j/R_z, ez = R_z/|R|.
template <class R>
struct WallUF {
  R F1, F2, F3, F4, F5;
};

template <class R>
__host__ __device__ inline WallUF<R> wall_uf_coefs(R hw, R ez, R ez2, R u, R bR) {
  WallUF<R> w;
  const R hq = hw * (R(1) - hw) * ez2;
  w.F1 = bR * lm_fma(u, lm_fma(u, lm_fma(R(-2.5), ez2, R(0.5)), lm_fma(R(1.5), ez2, R(-0.5))),
                     lm_fma(R(-1.5), hq, R(-0.75)));
  w.F2 = bR * lm_fma(u, lm_fma(u, lm_fma(R(17.5), ez2, R(-2.5)), lm_fma(R(-7.5), ez2, R(1.5))),
                     lm_fma(R(4.5), hq, R(-0.75)));
  const R ezbR = ez * bR;
  const R c30 = R(1.5) * hw * lm_fma(R(-6) * (R(1) - hw), ez2, R(1));
  w.F3 = ezbR * lm_fma(u, lm_fma(u, lm_fma(R(-35), ez2, R(10)), lm_fma(R(15), ez2, R(-3))), c30);
  w.F4 = ezbR * lm_fma(R(-5) * u, u, R(1.5) * hw);
  w.F5 = bR * lm_fma(u, lm_fma(u, lm_fma(R(15), ez2, R(-2)), R(-3) * ez2), R(-3) * hw * hw * ez2);
  return w;
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
    const R ex = dx * iR, ey = dy * iR, ez = Rz * iR;
    const WallUF<R> f = wall_uf_coefs<R>(hj * lm_rcp(Rz), ez, ez * ez, iR * iR, b * iR);
    k.a0 = lm_fma(b, fuf, f.F1);
    k.ex = ex;
    k.ey = ey;
    k.ez = ez;
    k.gx = f.F2 * ex;
    k.gy = f.F2 * ey;
    k.gz = lm_fma(f.F2, ez, f.F4);
    k.hx = f.F3 * ex;
    k.hy = f.F3 * ey;
    k.hz = lm_fma(f.F3, ez, f.F5);
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

// ---- bottom wall, and one pair in both directions -------------------------
// Image system of the bottom wall (Swan & Brady, hydrodynamicKernels.cuh:187-359),
// damped by b_i b_j: R = (d_x, d_y, h_i + h_j), e = R/|R|, u = 1/|R|^2,
// q = (e_y, -e_x, 0), h = h_j/R_z (h_i/R_z where marked):
//   UF:  M v = f1 v + (f2 (e.v) + f3 v_z) e + z^ (f4 (e.v) + f5 v_z), with t1 = 5 e_z^2 - 1,
//        hq = h (1 - h) e_z^2 and
//        f1 = -(3 (1 + 2 hq) + 2 (1 - 3 e_z^2) u + 2 t1 u^2) / (4 |R|)
//        f2 = -(3 (1 - 6 hq) + 6 t1 u + 10 (1 - 7 e_z^2) u^2) / (4 |R|)
//        f3 = e_z (3 h (1 - 6 (1 - h) e_z^2) + 6 t1 u + 10 (2 - 7 e_z^2) u^2) / (2 |R|)
//        f4 = e_z (3 h - 10 u^2) / (2 |R|)
//        f5 = -(3 h^2 e_z^2 + 3 e_z^2 u + (2 - 15 e_z^2) u^2) / |R|
//   WT:  M v = g1 v + g2 (e.v) e + g3 (e.v) z^ + g4 (q.v) q, with
//        g1 = (1 - 6 e_z^2) u / (2 |R|), g2 = -3 u / (2 |R|), g3 = 3 e_z u / |R|, g4 = 3 u / |R|
//   WF:  B v = q (k3 (e.v) + (k1 + k2) v_z) + kz (z^ x v) - k1 (q.v) z^
//   UT:  C v = k3 (q.v) e - kz (z^ x v) + k1 v_z q + gam (q.v) z^ (h = h_i),
//        gam = -(k1 + k2 + 2 k3 e_z), with k1 = u, k2 = 2 u (6 h e_z^2 + (1 - 10 e_z^2) u),
//        k3 = -2 e_z u (3 h - 5 u), kz = e_z u (1 - 2 h + 2 u).
// Self pairs (r = 0, h_i = h_j) have their own forms (B1-B3, :190-202, :256-261,
// :290-294, :327-331).
//
// The grand mobility matrix is symmetric: blockwise M_ji = M_ij^T, with UF and
// WT transposing into themselves and the cross blocks swapping,
// M_UT(ji) = M_WF(ij)^T and M_WF(ji) = M_UT(ij)^T.  So the coefficients of the
// pair (i <- j) also give what j receives from i:
//   UF^T v = f1 v + (f2 (e.v) + f4 v_z) e + z^ (f3 (e.v) + f5 v_z)   (f3, f4 swap)
//   WT^T v = g1 v + (g2 (e.v) + g3 v_z) e + g4 (q.v) q
//   B^T v = (k3 e + (k1 + k2) z^) (q.v) - kz (z^ x v) - k1 v_z q
//   C^T v = (k3 (e.v) + gam v_z) q + kz (z^ x v) + k1 (q.v) z^
// and the open parts: UF, WT symmetric; c (v x d) transposes to -c (v x d).
// The target accumulates M_ij [F_j; T_j] into a; the source sums are kept in
// units of t0 (su: UF part + kut * UT part, kut = rt0/t0) and r0 (sw: WT part
// + kwf * WF part, kwf = rt0/r0).

// Open boundaries: both directions of every block from the RPY coefficients.
template <class R, int BLK>
__host__ __device__ inline void open_both(R dx, R dy, R dz, R fuf, R guf, R fwt, R gwt, R c,
                                          const Vec3<R>& Fj, const Vec3<R>& Tj,
                                          const Vec3<R>& Fi, const Vec3<R>& Ti, R kut, R kwf,
                                          PairAcc<R>& a, Vec3<R>& su, Vec3<R>& sw) {
  if (BLK & kUF) {
    const R s = guf * (dx * Fj.x + dy * Fj.y + dz * Fj.z);
    a.uf.x = lm_fma(s, dx, lm_fma(fuf, Fj.x, a.uf.x));
    a.uf.y = lm_fma(s, dy, lm_fma(fuf, Fj.y, a.uf.y));
    a.uf.z = lm_fma(s, dz, lm_fma(fuf, Fj.z, a.uf.z));
    const R ss = guf * (dx * Fi.x + dy * Fi.y + dz * Fi.z);
    su.x = lm_fma(ss, dx, lm_fma(fuf, Fi.x, su.x));
    su.y = lm_fma(ss, dy, lm_fma(fuf, Fi.y, su.y));
    su.z = lm_fma(ss, dz, lm_fma(fuf, Fi.z, su.z));
  }
  if (BLK & kWT) {
    const R s = gwt * (dx * Tj.x + dy * Tj.y + dz * Tj.z);
    a.wt.x = lm_fma(s, dx, lm_fma(fwt, Tj.x, a.wt.x));
    a.wt.y = lm_fma(s, dy, lm_fma(fwt, Tj.y, a.wt.y));
    a.wt.z = lm_fma(s, dz, lm_fma(fwt, Tj.z, a.wt.z));
    const R ss = gwt * (dx * Ti.x + dy * Ti.y + dz * Ti.z);
    sw.x = lm_fma(ss, dx, lm_fma(fwt, Ti.x, sw.x));
    sw.y = lm_fma(ss, dy, lm_fma(fwt, Ti.y, sw.y));
    sw.z = lm_fma(ss, dz, lm_fma(fwt, Ti.z, sw.z));
  }
  // cross blocks: v x (c d), with the source side's factors -kut, -kwf on
  // the target's own torque and force (fixed over the source loop)
  const R cdx = c * dx, cdy = c * dy, cdz = c * dz;
  if (BLK & kUT) {  // c (T_j x d) to i; -c (T_i x d) to j, in units of t0
    a.ut.x = lm_fma(Tj.y, cdz, lm_fma(-Tj.z, cdy, a.ut.x));
    a.ut.y = lm_fma(Tj.z, cdx, lm_fma(-Tj.x, cdz, a.ut.y));
    a.ut.z = lm_fma(Tj.x, cdy, lm_fma(-Tj.y, cdx, a.ut.z));
    const R kx = -kut * Ti.x, ky = -kut * Ti.y, kz = -kut * Ti.z;
    su.x = lm_fma(ky, cdz, lm_fma(-kz, cdy, su.x));
    su.y = lm_fma(kz, cdx, lm_fma(-kx, cdz, su.y));
    su.z = lm_fma(kx, cdy, lm_fma(-ky, cdx, su.z));
  }
  if (BLK & kWF) {  // c (F_j x d) to i; -c (F_i x d) to j, in units of r0
    a.wf.x = lm_fma(Fj.y, cdz, lm_fma(-Fj.z, cdy, a.wf.x));
    a.wf.y = lm_fma(Fj.z, cdx, lm_fma(-Fj.x, cdz, a.wf.y));
    a.wf.z = lm_fma(Fj.x, cdy, lm_fma(-Fj.y, cdx, a.wf.z));
    const R kx = -kwf * Fi.x, ky = -kwf * Fi.y, kz = -kwf * Fi.z;
    sw.x = lm_fma(ky, cdz, lm_fma(-kz, cdy, sw.x));
    sw.y = lm_fma(kz, cdx, lm_fma(-kx, cdz, sw.y));
    sw.z = lm_fma(kx, cdy, lm_fma(-ky, cdx, sw.z));
  }
}

// Bottom wall, r > 0: the open part and the image terms of each block merged
// into one set of coefficients per pair, with the damping b = b_i b_j and the
// unit factors kut, kwf folded in, applied to i and (kToJ) to j.  e's in-plane
// part is d's divided by |R|; for a vector v (F or T of the other particle)
// c_v = d_y v_x - d_x v_y = |R| (q.v).
//   UF:  x, y: A v + (G (d.v) + Q/|R|) d,  z: (A + F5) v_z + G (d.v) d_z + Q e_z + F4 (e.v)
//        with A = b (fuf + f1), G = b guf, Fk = b fk, Q = F2 (e.v) + F3 v_z; f3 and f4
//        swap for j.
//   WT:  x: Aw v_x + (Gw (d.v) + G2 (e.v)/|R|) d_x + G4 c_v d_y,  y: ... - G4 c_v d_x,
//        z: Aw v_z + Gw (d.v) d_z + b (g2 e_z + g3) (e.v); for j, G2 (e.v) + G3 v_z
//        takes the place of G2 (e.v) and of the z coefficient.
//   WF, UT and their transposes: x, y from (v_y, -v_x) and (d_x, d_y), z from c_v.
// The polynomials in u are evaluated in Horner form.
template <class R, int BLK, bool kToJ = true>
__host__ __device__ inline void wall_image_both(R dx, R dy, R dz, R hi, R hj, R b, R fuf, R guf,
                                                R fwt, R gwt, R c, const Vec3<R>& Fj,
                                                const Vec3<R>& Tj, const Vec3<R>& Fi,
                                                const Vec3<R>& Ti, R kut, R kwf, PairAcc<R>& a,
                                                Vec3<R>& su, Vec3<R>& sw) {
  constexpr bool kCross = (BLK & (kUT | kWF)) != 0;
  const R Rz = hi + hj;
  const R iR = lm_rsqrt(dx * dx + dy * dy + Rz * Rz);
  const R u = iR * iR;
  const R ez = Rz * iR;
  const R ez2 = ez * ez;
  const R iRz = lm_rcp(Rz);
  const R hw = hj * iRz;  // h_j / (h_i + h_j)
  const R bR = b * iR;
  const R bR3 = bR * u;
  // projections of the source's (j's) force and torque, and of the target's
  const R pF = dx * Fj.x + dy * Fj.y, pFi = kToJ ? dx * Fi.x + dy * Fi.y : R(0);
  const R eF = lm_fma(pF, iR, ez * Fj.z), eFi = kToJ ? lm_fma(pFi, iR, ez * Fi.z) : R(0);
  const R pT = dx * Tj.x + dy * Tj.y, pTi = kToJ ? dx * Ti.x + dy * Ti.y : R(0);
  const R cT = dy * Tj.x - dx * Tj.y, cTi = kToJ ? dy * Ti.x - dx * Ti.y : R(0);
  if (BLK & kUF) {
    const WallUF<R> f = wall_uf_coefs<R>(hw, ez, ez2, u, bR);
    const R F2 = f.F2, F3 = f.F3, F4 = f.F4, F5 = f.F5;
    const R A = lm_fma(b, fuf, f.F1);
    const R A5 = A + F5;
    const R G = b * guf;
    // i from F_j
    const R GdF = G * lm_fma(dz, Fj.z, pF);
    const R Q = lm_fma(F2, eF, F3 * Fj.z);
    const R S = lm_fma(Q, iR, GdF);
    a.uf.x = lm_fma(S, dx, lm_fma(A, Fj.x, a.uf.x));
    a.uf.y = lm_fma(S, dy, lm_fma(A, Fj.y, a.uf.y));
    a.uf.z = lm_fma(F4, eF, lm_fma(Q, ez, lm_fma(GdF, dz, lm_fma(A5, Fj.z, a.uf.z))));
    if (kToJ) {  // j from F_i (f3 and f4 swap)
      const R GdFi = G * lm_fma(dz, Fi.z, pFi);
      const R Qs = lm_fma(F2, eFi, F4 * Fi.z);
      const R Ss = lm_fma(Qs, iR, GdFi);
      su.x = lm_fma(Ss, dx, lm_fma(A, Fi.x, su.x));
      su.y = lm_fma(Ss, dy, lm_fma(A, Fi.y, su.y));
      su.z = lm_fma(F3, eFi, lm_fma(Qs, ez, lm_fma(GdFi, dz, lm_fma(A5, Fi.z, su.z))));
    }
  }
  if (BLK & kWT) {
    const R Aw = lm_fma(b, fwt, bR3 * lm_fma(R(-3), ez2, R(0.5)));  // b (fwt + g1)
    const R Gw = b * gwt;
    const R G2 = R(-1.5) * bR3;
    const R G2iR = G2 * iR;
    const R G3 = R(3) * bR3 * ez;
    const R G23 = R(1.5) * bR3 * ez;  // b (g2 e_z + g3)
    const R G4 = R(3) * bR3 * u;
    // i from T_j
    const R eT = lm_fma(pT, iR, ez * Tj.z);
    const R GdT = Gw * lm_fma(dz, Tj.z, pT);
    const R Sw = lm_fma(G2iR, eT, GdT);
    const R P = G4 * cT;
    a.wt.x = lm_fma(P, dy, lm_fma(Sw, dx, lm_fma(Aw, Tj.x, a.wt.x)));
    a.wt.y = lm_fma(-P, dx, lm_fma(Sw, dy, lm_fma(Aw, Tj.y, a.wt.y)));
    a.wt.z = lm_fma(G23, eT, lm_fma(GdT, dz, lm_fma(Aw, Tj.z, a.wt.z)));
    if (kToJ) {  // j from T_i
      const R eTi = lm_fma(pTi, iR, ez * Ti.z);
      const R SS = lm_fma(G2, eTi, G3 * Ti.z);
      const R GdTi = Gw * lm_fma(dz, Ti.z, pTi);
      const R Ss = lm_fma(SS, iR, GdTi);
      const R Ps = G4 * cTi;
      sw.x = lm_fma(Ps, dy, lm_fma(Ss, dx, lm_fma(Aw, Ti.x, sw.x)));
      sw.y = lm_fma(-Ps, dx, lm_fma(Ss, dy, lm_fma(Aw, Ti.y, sw.y)));
      sw.z = lm_fma(SS, ez, lm_fma(GdTi, dz, lm_fma(Aw, Ti.z, sw.z)));
    }
  }
  if (kCross) {
    // k1 = u; k2, k3, kz with h = h_j (w: WF to i, UT to j) and h = h_i (u: UT to i, WF to j)
    const R hu = hi * iRz;
    const R ezu = ez * u;
    const R t20 = lm_fma(R(-20), ez2, R(2)) * u;
    const R k2w = u * lm_fma(R(12) * hw, ez2, t20);
    const R k3w = ezu * lm_fma(R(10), u, R(-6) * hw);
    const R kzw = ezu * lm_fma(R(2), u - hw, R(1));
    const R k2u = u * lm_fma(R(12) * hu, ez2, t20);
    const R k3u = ezu * lm_fma(R(10), u, R(-6) * hu);
    const R kzu = ezu * lm_fma(R(2), u - hu, R(1));
    const R bc = b * c;
    const R Cz = bc - bR3;                   // b (c - k1/R)
    const R Cw = lm_fma(bc, dz, -b * kzw);  // b (c d_z - kz_w)
    const R Cu = lm_fma(bc, dz, b * kzu);   // b (c d_z + kz_u)
    if (BLK & kWF) {
      // i from F_j: x: F_y Cw + d_y Wv, y: -F_x Cw - d_x Wv, z: Cz (d_y F_x - d_x F_y)
      const R K3w = bR * k3w;
      const R KCw = lm_fma(bR, u + k2w, -bc);
      const R Wv = lm_fma(K3w, eF, KCw * Fj.z);
      a.wf.x = lm_fma(dy, Wv, lm_fma(Fj.y, Cw, a.wf.x));
      a.wf.y = lm_fma(-dx, Wv, lm_fma(-Fj.x, Cw, a.wf.y));
      a.wf.z = lm_fma(Cz, dy * Fj.x - dx * Fj.y, a.wf.z);
      if (kToJ) {
        // j from F_i (units of r0): x: -F_y Cf + d_y Vs, y: F_x Cf - d_x Vs,
        // z: -kwf Cz (d_y F_x - d_x F_y)
        const R Cf = kwf * Cu;
        const R gam = -lm_fma(R(2) * k3u, ez, u + k2u);
        const R Vs = kwf * lm_fma(bR * k3u, eFi, lm_fma(bR, gam, bc) * Fi.z);
        sw.x = lm_fma(dy, Vs, lm_fma(-Fi.y, Cf, sw.x));
        sw.y = lm_fma(-dx, Vs, lm_fma(Fi.x, Cf, sw.y));
        sw.z = lm_fma(-kwf * Cz, dy * Fi.x - dx * Fi.y, sw.z);
      }
    }
    if (BLK & kUT) {
      // i from T_j: x: T_y Cu + d_x m + d_y n, y: -T_x Cu + d_y m - d_x n, z: Zu (d_y T_x - d_x T_y)
      const R m = b * u * k3u * cT;
      const R n = -Cz * Tj.z;
      const R Zu = bc - bR * lm_fma(k3u, ez, u + k2u);
      a.ut.x = lm_fma(dy, n, lm_fma(dx, m, lm_fma(Tj.y, Cu, a.ut.x)));
      a.ut.y = lm_fma(-dx, n, lm_fma(dy, m, lm_fma(-Tj.x, Cu, a.ut.y)));
      a.ut.z = lm_fma(Zu, cT, a.ut.z);
      if (kToJ) {
        // j from T_i (units of t0): x: T_y Cs + d_x ms + d_y ns, y: -T_x Cs + d_y ms - d_x ns
        const R Cs = -kut * Cw;
        const R ms = kut * b * u * k3w * cTi;
        const R ns = kut * Cz * Ti.z;
        const R Zs = kut * (bR * lm_fma(k3w, ez, u + k2w) - bc);
        su.x = lm_fma(dy, ns, lm_fma(dx, ms, lm_fma(Ti.y, Cs, su.x)));
        su.y = lm_fma(-dx, ns, lm_fma(dy, ms, lm_fma(-Ti.x, Cs, su.y)));
        su.z = lm_fma(Zs, cTi, su.z);
      }
    }
  }
}

// Bottom wall, self pair (r = 0, h_i = h_j): Swan & Brady B1-B3; each block
// equals its transpose's counterpart, so j (kToJ) gets the same forms.  With
// d = 0 the open part is fuf v (UF), fwt v (WT) and nothing for the cross blocks.
template <class R, int BLK, bool kToJ = true>
__host__ __device__ inline void wall_self_both(R hj, R b, R fuf, R fwt, const Vec3<R>& Fj,
                                               const Vec3<R>& Tj, const Vec3<R>& Fi,
                                               const Vec3<R>& Ti, R kut, R kwf, PairAcc<R>& a,
                                               Vec3<R>& su, Vec3<R>& sw) {
  const R ih = R(1) / hj;
  const R ih2 = ih * ih;
  const R ih3 = ih2 * ih;
  if (BLK & kUF) {
    const R ih5 = ih3 * ih2;
    const R cp = b * (fuf + R(-0.0625) * (R(9) * ih - R(2) * ih3 + ih5));
    const R cz = b * (fuf + R(-0.125) * (R(9) * ih - R(4) * ih3 + ih5));
    a.uf.x += cp * Fj.x;
    a.uf.y += cp * Fj.y;
    a.uf.z += cz * Fj.z;
    if (kToJ) {
      su.x += cp * Fi.x;
      su.y += cp * Fi.y;
      su.z += cz * Fi.z;
    }
  }
  if (BLK & kWT) {
    const R cp = b * (fwt + R(-0.3125) * ih3);
    const R cz = b * (fwt + R(-0.125) * ih3);
    a.wt.x += cp * Tj.x;
    a.wt.y += cp * Tj.y;
    a.wt.z += cz * Tj.z;
    if (kToJ) {
      sw.x += cp * Ti.x;
      sw.y += cp * Ti.y;
      sw.z += cz * Ti.z;
    }
  }
  const R c4 = b * R(0.125) * ih3 * ih;
  if (BLK & kWF) {  // (-c4 v_y, c4 v_x, 0)
    a.wf.x -= c4 * Fj.y;
    a.wf.y += c4 * Fj.x;
    if (kToJ) {
      sw.x -= kwf * c4 * Fi.y;
      sw.y += kwf * c4 * Fi.x;
    }
  }
  if (BLK & kUT) {  // (c4 v_y, -c4 v_x, 0)
    a.ut.x += c4 * Tj.y;
    a.ut.y -= c4 * Tj.x;
    if (kToJ) {
      su.x += kut * c4 * Ti.y;
      su.y -= kut * c4 * Ti.x;
    }
  }
}

// Bottom-wall contribution of source j to target i (open RPY + image terms,
// damped by b_i b_j).  (xi, yi, hi, bi) and (xj, yj, hj, bj) are position
// records as described at the top of this file.
template <class R, int BLK>
__host__ __device__ inline void wall_pair(R xi, R yi, R hi, R bi, R xj, R yj, R hj, R bj, R fx,
                                          R fy, R fz, R tx, R ty, R tz, PairAcc<R>& a) {
  const R dx = xi - xj, dy = yi - yj, dz = hi - hj;
  const R r2 = dx * dx + dy * dy + dz * dz;
  R fuf, guf, fwt, gwt, c;
  rpy_coeffs<R, BLK>(r2, fuf, guf, fwt, gwt, c);
  const Vec3<R> Fj = {fx, fy, fz}, Tj = {tx, ty, tz}, none = {R(0), R(0), R(0)};
  Vec3<R> su = none, sw = none;  // not written (target only)
  if (r2 != R(0)) {
    wall_image_both<R, BLK, false>(dx, dy, dz, hi, hj, bi * bj, fuf, guf, fwt, gwt, c, Fj, Tj,
                                   none, none, R(0), R(0), a, su, sw);
  } else {
    lm_rare_path();
    wall_self_both<R, BLK, false>(hj, bi * bj, fuf, fwt, Fj, Tj, none, none, R(0), R(0), a, su,
                                  sw);
  }
}

template <class R, bool WALL, int BLK>
__host__ __device__ inline void pair_both(R xi, R yi, R hi, R bi, R xj, R yj, R hj, R bj,
                                          const Vec3<R>& Fj, const Vec3<R>& Tj,
                                          const Vec3<R>& Fi, const Vec3<R>& Ti, R kut, R kwf,
                                          PairAcc<R>& a, Vec3<R>& su, Vec3<R>& sw) {
  constexpr bool kCross = (BLK & (kUT | kWF)) != 0;
  const R dx = xi - xj, dy = yi - yj, dz = hi - hj;
  const R r2 = dx * dx + dy * dy + dz * dz;
  R fuf, guf, fwt, gwt, c;
  rpy_coeffs<R, BLK | (kCross ? (kUT | kWF) : 0)>(r2, fuf, guf, fwt, gwt, c);
  if (!WALL) {
    open_both<R, BLK>(dx, dy, dz, fuf, guf, fwt, gwt, c, Fj, Tj, Fi, Ti, kut, kwf, a, su, sw);
  } else if (r2 != R(0)) {
    wall_image_both<R, BLK>(dx, dy, dz, hi, hj, bi * bj, fuf, guf, fwt, gwt, c, Fj, Tj, Fi, Ti,
                            kut, kwf, a, su, sw);
  } else {
    lm_rare_path();
    wall_self_both<R, BLK>(hj, bi * bj, fuf, fwt, Fj, Tj, Fi, Ti, kut, kwf, a, su, sw);
  }
}

}  // namespace lm
