/*
 * TEST INFRASTRUCTURE ONLY.  CPU oracle for the NBody mobility operator.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load this file's library; the product path never does.
 *
 * It restates the reference pair arithmetic as explicit 3x3 matrices (the
 * reference applies the same coefficients to vectors).  Each function cites
 * the reference lines it follows (paths relative to the reference root).
 *
 *   solvers/NBody/extra/hydrodynamicKernels.cuh
 *     RPY_UF  :41-57      RPY_WT :63-86     RPY_UT/RPY_WF :89-118
 *     OpenBoundary::dotProduct*           :121-174
 *     BottomWall::wallCorrection_UF       :187-249
 *     BottomWall::wallCorrection_WT       :253-282
 *     BottomWall::wallCorrection_UT       :287-319
 *     BottomWall::wallCorrection_WF       :324-359
 *     BottomWall::dotProduct + blocks     :362-441
 *   solvers/NBody/extra/NbodyRPY.cu:254-287  (which blocks are active)
 *   solvers/NBody/mobility.h:99-106          (t0, rt0, r0)
 *
 * Pinning: tests/test_oracle_golden.py checks every function here against
 * the reference's golden matrices (free-space pair 12x12, bottom-wall self
 * 6x6 and pair 12x12), the closed-form self mobilities, symmetry, positive
 * definiteness, branch continuity at r = 2a and z = a, and the independent
 * Python scalar restatement in oracle/rpy_scalar.py.
 *
 * Units: every geometric quantity inside the block functions is scaled by
 * the hydrodynamic radius a (r~ = r/a), exactly as the reference does with
 * invrh.  The returned blocks are dimensionless; callers multiply by t0, rt0
 * and r0.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double m3[3][3];

static void m3_zero(m3 m) { memset(m, 0, sizeof(m3)); }

/* f and g/r^2 of the translational RPY tensor, scaled units.
 * hydrodynamicKernels.cuh:41-57 (far field r>=2, overlap branch r<2). */
static void rpy_uf_coeffs(double r, double *f, double *g2) {
  if (r >= 2.0) {
    double ir = 1.0 / r, ir2 = ir * ir;
    *f = (0.75 + 0.5 * ir2) * ir;
    *g2 = (0.75 - 1.5 * ir2) * ir * ir2;
  } else {
    *f = 1.0 - 0.28125 * r;
    *g2 = (r > 0.0) ? 0.09375 / r : 0.0;
  }
}

/* Rotational RPY coefficients, hydrodynamicKernels.cuh:63-86. */
static void rpy_wt_coeffs(double r, double *f, double *g2) {
  if (r >= 2.0) {
    double ir = 1.0 / r, ir2 = ir * ir, ir3 = ir * ir2;
    *f = -0.5 * ir3;
    *g2 = 1.5 * ir2 * ir3;
  } else {
    *f = 1.0 - 0.84375 * r + 0.078125 * r * r * r;
    *g2 = (r > 0.0) ? (0.28125 / r - 0.046875 * r) : 0.0;
  }
}

/* Coupling coefficient c of U = c (T x r~), W = c (F x r~).
 * hydrodynamicKernels.cuh:89-118 together with :161-173. */
static double rpy_rt_coeff(double r) {
  if (r >= 2.0) return 1.0 / (r * r * r);
  return 0.5 * (1.0 - 0.375 * r);
}

/* f I + g2 d d^T (g term dropped at r = 0), hydrodynamicKernels.cuh:143-159 */
static void open_sym_block(const double d[3], double r, double f, double g2, m3 A) {
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      A[a][b] = (a == b ? f : 0.0) + (r > 0.0 ? g2 * d[a] * d[b] : 0.0);
}

/* Matrix of v -> c (v x d): (v x d)_x = v_y d_z - v_z d_y, ...
 * Same as the (M_xy, M_xz, M_yz) application at hydrodynamicKernels.cuh:161-173. */
static void open_cross_block(const double d[3], double c, m3 A) {
  A[0][0] = 0.0;       A[0][1] = c * d[2];  A[0][2] = -c * d[1];
  A[1][0] = -c * d[2]; A[1][1] = 0.0;       A[1][2] = c * d[0];
  A[2][0] = c * d[1];  A[2][1] = -c * d[0]; A[2][2] = 0.0;
}

/* Wall correction to UF, hydrodynamicKernels.cuh:187-249.
 * R = (r_ij + 2 h_j z)/a, h = h_j/a. */
static void wall_uf(const double R[3], int self, double h, m3 C) {
  m3_zero(C);
  if (self) {
    double iz = 1.0 / h, iz3 = iz * iz * iz, iz5 = iz3 * iz * iz;
    double par = -(9.0 * iz - 2.0 * iz3 + iz5) / 16.0;
    double per = -(9.0 * iz - 4.0 * iz3 + iz5) / 8.0;
    C[0][0] = par; C[1][1] = par; C[2][2] = per;
    return;
  }
  double hh = h / R[2];
  double iR = 1.0 / sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
  double e[3] = {R[0] * iR, R[1] * iR, R[2] * iR};
  double iR3 = iR * iR * iR, iR5 = iR3 * iR * iR;
  double ez2 = e[2] * e[2];
  double f1 = -(3.0 * (1.0 + 2.0 * hh * (1.0 - hh) * ez2) * iR + 2.0 * (1.0 - 3.0 * ez2) * iR3 -
                2.0 * (1.0 - 5.0 * ez2) * iR5) / 4.0;
  double f2 = -(3.0 * (1.0 - 6.0 * hh * (1.0 - hh) * ez2) * iR - 6.0 * (1.0 - 5.0 * ez2) * iR3 +
                10.0 * (1.0 - 7.0 * ez2) * iR5) / 4.0;
  double f3 = e[2] * (3.0 * hh * (1.0 - 6.0 * (1.0 - hh) * ez2) * iR - 6.0 * (1.0 - 5.0 * ez2) * iR3 +
                      10.0 * (2.0 - 7.0 * ez2) * iR5) / 2.0;
  double f4 = e[2] * (3.0 * hh * iR - 10.0 * iR5) / 2.0;
  double f5 = -(3.0 * hh * hh * ez2 * iR + 3.0 * ez2 * iR3 + (2.0 - 15.0 * ez2) * iR5);
  C[0][0] = f1 + f2 * e[0] * e[0];
  C[0][1] = f2 * e[0] * e[1];
  C[0][2] = f2 * e[0] * e[2] + f3 * e[0];
  C[1][0] = f2 * e[1] * e[0];
  C[1][1] = f1 + f2 * e[1] * e[1];
  C[1][2] = f2 * e[1] * e[2] + f3 * e[1];
  C[2][0] = f2 * e[2] * e[0] + f4 * e[0];
  C[2][1] = f2 * e[2] * e[1] + f4 * e[1];
  C[2][2] = f1 + f2 * e[2] * e[2] + f3 * e[2] + f4 * e[2] + f5;
}

/* Wall correction to WT, hydrodynamicKernels.cuh:253-282 (8 pi eta a^3 units). */
static void wall_wt(const double R[3], int self, double h, m3 C) {
  m3_zero(C);
  if (self) {
    double iz = 1.0 / h, iz3 = iz * iz * iz;
    C[0][0] = -iz3 * 0.3125; C[1][1] = -iz3 * 0.3125; C[2][2] = -iz3 * 0.125;
    return;
  }
  double iR = 1.0 / sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
  double iR3 = iR * iR * iR;
  double e[3] = {R[0] * iR, R[1] * iR, R[2] * iR};
  double f1 = (1.0 - 6.0 * e[2] * e[2]) * iR3 * 0.5;
  double f2 = -iR3 * 1.5;
  double f3 = 3.0 * iR3 * e[2];
  double f4 = 3.0 * iR3;
  C[0][0] = f1 + f2 * e[0] * e[0] + f4 * e[1] * e[1];
  C[0][1] = (f2 - f4) * e[0] * e[1];
  C[0][2] = f2 * e[0] * e[2];
  C[1][0] = (f2 - f4) * e[0] * e[1];
  C[1][1] = f1 + f2 * e[1] * e[1] + f4 * e[0] * e[0];
  C[1][2] = f2 * e[1] * e[2];
  C[2][0] = f2 * e[2] * e[0] + f3 * e[0];
  C[2][1] = f2 * e[2] * e[1] + f3 * e[1];
  C[2][2] = f1 + f2 * e[2] * e[2] + f3 * e[2];
}

/* The four scalar factors shared by the UT and WF wall corrections,
 * hydrodynamicKernels.cuh:296-307 and :333-345. */
static void wall_rt_factors(const double R[3], double h, double e[3], double *f1, double *f2,
                            double *f3, double *f4) {
  double hh = h / R[2];
  double iR = 1.0 / sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
  double iR2 = iR * iR, iR4 = iR2 * iR2;
  e[0] = R[0] * iR; e[1] = R[1] * iR; e[2] = R[2] * iR;
  *f1 = iR2;
  *f2 = (6.0 * hh * e[2] * e[2] * iR2 + (1.0 - 10.0 * e[2] * e[2]) * iR4) * 2.0;
  *f3 = -e[2] * (3.0 * hh * iR2 - 5.0 * iR4) * 2.0;
  *f4 = -e[2] * (hh * iR2 - iR4) * 2.0;
}

/* Wall correction to WF (C3 of Swan & Brady), hydrodynamicKernels.cuh:324-359.
 * Entries are the negated coefficients of the reference's "-=" updates. */
static void wall_wf(const double R[3], int self, double h, m3 C) {
  m3_zero(C);
  if (self) {
    double iz = 1.0 / h, iz4 = iz * iz * iz * iz;
    C[0][1] = -iz4 * 0.125;
    C[1][0] = iz4 * 0.125;
    return;
  }
  double e[3], f1, f2, f3, f4;
  wall_rt_factors(R, h, e, &f1, &f2, &f3, &f4);
  C[0][0] = -(-f3 * e[0] * e[1]);
  C[0][1] = -(f1 * e[2] - f3 * e[1] * e[1] + f4);
  C[0][2] = -(-f1 * e[1] - f2 * e[1] - f3 * e[1] * e[2]);
  C[1][0] = -(-f1 * e[2] + f3 * e[0] * e[0] - f4);
  C[1][1] = -(f3 * e[0] * e[1]);
  C[1][2] = -(f1 * e[0] + f2 * e[0] + f3 * e[0] * e[2]);
  C[2][0] = -(f1 * e[1]);
  C[2][1] = -(-f1 * e[0]);
  C[2][2] = 0.0;
}

/* Wall correction to UT (C3^T), hydrodynamicKernels.cuh:287-319.  The caller
 * passes R = (-r_ij + 2 h_i z)/a and h = h_i/a (:422-430). */
static void wall_ut(const double R[3], int self, double h, m3 C) {
  m3_zero(C);
  if (self) {
    double iz = 1.0 / h, iz4 = iz * iz * iz * iz;
    C[0][1] = iz4 * 0.125;
    C[1][0] = -iz4 * 0.125;
    return;
  }
  double e[3], f1, f2, f3, f4;
  wall_rt_factors(R, h, e, &f1, &f2, &f3, &f4);
  C[0][0] = -(-f3 * e[0] * e[1]);
  C[0][1] = -(-f1 * e[2] + f3 * e[0] * e[0] - f4);
  C[0][2] = -(f1 * e[1]);
  C[1][0] = -(f1 * e[2] - f3 * e[1] * e[1] + f4);
  C[1][1] = -(f3 * e[0] * e[1]);
  C[1][2] = -(-f1 * e[0]);
  C[2][0] = -(-f1 * e[1] - f2 * e[1] - f3 * e[1] * e[2]);
  C[2][1] = -(f1 * e[0] + f2 * e[0] + f3 * e[0] * e[2]);
  C[2][2] = 0.0;
}

/*
 * 6x6 coupling block between target i and source j, dimensionless.
 * Output layout blk[row][col] with rows (U_x,U_y,U_z,W_x,W_y,W_z) of i and
 * columns (F_x,F_y,F_z,T_x,T_y,T_z) of j; the caller applies t0 (UF),
 * rt0 (UT, WF) and r0 (WT).
 *   wall = 0: OpenBoundary::dotProduct, hydrodynamicKernels.cuh:124-140
 *   wall = 1: BottomWall::dotProduct,   hydrodynamicKernels.cuh:362-441
 * pi, pj are in the caller's units; a is the hydrodynamic radius.
 */
void oracle_pair_block(int wall, const double *pi_in, const double *pj_in, double a,
                       double *blk /* 36 */) {
  double pi[3] = {pi_in[0], pi_in[1], pi_in[2]};
  double pj[3] = {pj_in[0], pj_in[1], pj_in[2]};
  double bij = 1.0;
  if (wall) {
    /* damping, Usabiaga et al. 2017 App. 1, hydrodynamicKernels.cuh:368-375 */
    double bi = fmax(fmin(pi[2] / a, 1.0), 0.0);
    double bj = fmax(fmin(pj[2] / a, 1.0), 0.0);
    bij = bi * bj;
    pi[2] = fmax(pi[2], a);
    pj[2] = fmax(pj[2], a);
  }
  double d[3] = {(pi[0] - pj[0]) / a, (pi[1] - pj[1]) / a, (pi[2] - pj[2]) / a};
  double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  double f, g2;
  m3 UF, WT, UT, WF;
  rpy_uf_coeffs(r, &f, &g2);
  open_sym_block(d, r, f, g2, UF);
  rpy_wt_coeffs(r, &f, &g2);
  open_sym_block(d, r, f, g2, WT);
  double c = rpy_rt_coeff(r);
  open_cross_block(d, c, UT);
  open_cross_block(d, c, WF);
  if (wall) {
    int self = (r == 0.0);
    double hi = pi[2] / a, hj = pj[2] / a;
    double R[3] = {d[0], d[1], d[2] + 2.0 * hj};     /* :399, :411, :438 */
    double Rt[3] = {-d[0], -d[1], -d[2] + 2.0 * hi}; /* :427-428 */
    m3 C;
    wall_uf(R, self, hj, C);
    for (int x = 0; x < 3; x++) for (int y = 0; y < 3; y++) UF[x][y] += C[x][y];
    wall_wt(R, self, hj, C);
    for (int x = 0; x < 3; x++) for (int y = 0; y < 3; y++) WT[x][y] += C[x][y];
    wall_ut(Rt, self, hi, C);
    for (int x = 0; x < 3; x++) for (int y = 0; y < 3; y++) UT[x][y] += C[x][y];
    wall_wf(R, self, hj, C);
    for (int x = 0; x < 3; x++) for (int y = 0; y < 3; y++) WF[x][y] += C[x][y];
  }
  for (int x = 0; x < 3; x++)
    for (int y = 0; y < 3; y++) {
      blk[x * 6 + y] = bij * UF[x][y];
      blk[x * 6 + 3 + y] = bij * UT[x][y];
      blk[(3 + x) * 6 + y] = bij * WF[x][y];
      blk[(3 + x) * 6 + 3 + y] = bij * WT[x][y];
    }
}

/* Mobility constants, mobility.h:99-106. */
static void mobility_constants(double eta, double a, double *t0, double *rt0, double *r0) {
  *t0 = 1.0 / (6.0 * M_PI * eta * a);
  *rt0 = 1.0 / (8.0 * M_PI * eta * a * a);
  *r0 = 1.0 / (8.0 * M_PI * eta * a * a * a);
}

/* Batch id of particle i; batches are contiguous ranges given by offsets
 * (sys_off[0]=0 ... sys_off[nsys]=N).  mobility.h:154-161 / NbodyRPY.cu:36,77-78
 * use equal sizes; the ragged form reduces to that when all sizes match. */
static int64_t sys_of(const int64_t *off, int64_t nsys, int64_t i) {
  int64_t lo = 0, hi = nsys;
  while (hi - lo > 1) {
    int64_t mid = (lo + hi) / 2;
    if (off[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

/*
 * Dense grand mobility matrix in the layout of the reference's tests
 * (tests/utils.py:62-84): rows/cols [U_1..U_N, W_1..W_N] when angular,
 * [U_1..U_N] otherwise.  M is (dim x dim) row-major, dim = 3N or 6N.
 * wall_flags[s] selects the kernel per system s.
 */
void oracle_dense_matrix(int64_t N, const double *pos, const int64_t *sys_off, int64_t nsys,
                         const int *wall_flags, double eta, double a, int angular, double *M) {
  double t0, rt0, r0;
  mobility_constants(eta, a, &t0, &rt0, &r0);
  int64_t dim = angular ? 6 * N : 3 * N;
  memset(M, 0, sizeof(double) * dim * dim);
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t i = 0; i < N; i++) {
    int64_t s = sys_of(sys_off, nsys, i);
    double blk[36];
    for (int64_t j = sys_off[s]; j < sys_off[s + 1]; j++) {
      oracle_pair_block(wall_flags[s], pos + 3 * i, pos + 3 * j, a, blk);
      for (int x = 0; x < 3; x++)
        for (int y = 0; y < 3; y++) {
          M[(3 * i + x) * dim + 3 * j + y] = t0 * blk[x * 6 + y];
          if (angular) {
            M[(3 * i + x) * dim + 3 * N + 3 * j + y] = rt0 * blk[x * 6 + 3 + y];
            M[(3 * N + 3 * i + x) * dim + 3 * j + y] = rt0 * blk[(3 + x) * 6 + y];
            M[(3 * N + 3 * i + x) * dim + 3 * N + 3 * j + y] = r0 * blk[(3 + x) * 6 + 3 + y];
          }
        }
    }
  }
}

/* Rows `rows[0..nrows)` (all rows when rows == NULL, nrows = N) of the
 * matrix-free product below; U/W receive 3*nrows values. */
void oracle_mdot_rows(int64_t N, const double *pos, const int64_t *sys_off, int64_t nsys,
                      const int *wall_flags, double eta, double a, int angular, const double *F,
                      const double *T, int64_t nrows, const int64_t *rows, double *U, double *W) {
  double t0, rt0, r0;
  mobility_constants(eta, a, &t0, &rt0, &r0);
  int useT = (T != 0) && angular;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t q = 0; q < nrows; q++) {
    const int64_t i = rows ? rows[q] : q;
    int64_t s = sys_of(sys_off, nsys, i);
    double u[3] = {0, 0, 0}, w[3] = {0, 0, 0};
    double blk[36];
    for (int64_t j = sys_off[s]; j < sys_off[s + 1]; j++) {
      oracle_pair_block(wall_flags[s], pos + 3 * i, pos + 3 * j, a, blk);
      for (int x = 0; x < 3; x++)
        for (int y = 0; y < 3; y++) {
          if (F) {
            u[x] += t0 * blk[x * 6 + y] * F[3 * j + y];
            if (angular) w[x] += rt0 * blk[(3 + x) * 6 + y] * F[3 * j + y];
          }
          if (useT) {
            u[x] += rt0 * blk[x * 6 + 3 + y] * T[3 * j + y];
            w[x] += r0 * blk[(3 + x) * 6 + 3 + y] * T[3 * j + y];
          }
        }
    }
    for (int x = 0; x < 3; x++) {
      U[3 * q + x] = u[x];
      if (angular && W) W[3 * q + x] = w[x];
    }
  }
}

/*
 * Matrix-free O(N^2) product, FP64, OpenMP over targets.  F or T may be NULL
 * (absent).  Block selection follows NbodyRPY.cu:264-275: UF iff F, WF iff F
 * and angular, UT and WT iff T and angular.  U (3N) is overwritten; W (3N) is
 * overwritten when angular (NbodyRPY.cu:91-96).
 */
void oracle_mdot(int64_t N, const double *pos, const int64_t *sys_off, int64_t nsys,
                 const int *wall_flags, double eta, double a, int angular, const double *F,
                 const double *T, double *U, double *W) {
  oracle_mdot_rows(N, pos, sys_off, nsys, wall_flags, eta, a, angular, F, T, N, 0, U, W);
}

/* Open-boundary UF-only product specialised to the far/overlap branches
 * (same arithmetic as oracle_mdot with angular = 0, no wall), used for the
 * CPU baseline at large N where the generic 6x6 path would waste time. */
void oracle_mdot_open_uf_rows(int64_t N, const double *pos, double eta, double a,
                              const double *F, int64_t row0, int64_t nrows, double *U) {
  double t0 = 1.0 / (6.0 * M_PI * eta * a);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = row0; i < row0 + nrows; i++) {
    double u0 = 0, u1 = 0, u2 = 0;
    for (int64_t j = 0; j < N; j++) {
      double d[3] = {(pos[3 * i] - pos[3 * j]) / a, (pos[3 * i + 1] - pos[3 * j + 1]) / a,
                     (pos[3 * i + 2] - pos[3 * j + 2]) / a};
      double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      double f, g2;
      rpy_uf_coeffs(r, &f, &g2);
      const double *fj = F + 3 * j;
      double gv = (r > 0.0) ? g2 * (d[0] * fj[0] + d[1] * fj[1] + d[2] * fj[2]) : 0.0;
      u0 += f * fj[0] + gv * d[0];
      u1 += f * fj[1] + gv * d[1];
      u2 += f * fj[2] + gv * d[2];
    }
    U[3 * i] = t0 * u0;
    U[3 * i + 1] = t0 * u1;
    U[3 * i + 2] = t0 * u2;
  }
}

void oracle_mdot_open_uf(int64_t N, const double *pos, double eta, double a, const double *F,
                         double *U) {
  oracle_mdot_open_uf_rows(N, pos, eta, a, F, 0, N, U);
}

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
