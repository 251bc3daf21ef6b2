"""TEST INFRASTRUCTURE ONLY -- CPU oracle for the NBody + Lanczos path.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
import this package; it is the checker, never the thing measured or shipped.

Contents
  * ctypes front-end of ``rpy_oracle.c`` (pair blocks, dense grand mobility
    matrix, matrix-free FP64 product), built by ``oracle/Makefile``.
  * ``LanczosOracle``: numpy restatement of the reference Lanczos solver,
    include/lanczos/LanczosAlgorithm.cu:26-232 (plain three-term recurrence,
    breakdown rule, error estimate, adaptive check schedule, 200-step limit).
  * ``dense_sqrt``: FP64 eigen square root (numpy ``eigh``), the method the
    reference tests use (tests/test_fluctuation_dissipation.py:32-36).
  * ``rfd_drift``: random finite differences,
    include/MobilityInterface/random_finite_differences.h:30-63.
  * ``deterministic_divM``: central-difference divergence,
    tests/test_thermal_drift.py:179-207.

Pinning status (each checked in tests/test_oracle_*.py):
  pair blocks        golden npz fixtures + closed forms + continuity + SPD
  dense / mdot       equal to each other and to the blocks (brute force)
  LanczosOracle      converges to dense_sqrt on SPD matrices; tridiagonal
                     step against scipy.linalg.eigh_tridiagonal
  dense_sqrt         (dense_sqrt)^2 == M
  deterministic_divM closed-form d/dz of the Swan-Brady self mobility above
                     the wall and of the damped self mobility below one
                     radius; zero divergence of open RPY (test_oracle_drift.py)
  rfd_drift          the same closed form for one particle; the sum over unit
                     displacements equals deterministic_divM
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "librpy_oracle.so")
_lib = None


def build():
    """Compile the C oracle (make, gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    dp = ctypes.POINTER(ctypes.c_double)
    i64p = ctypes.POINTER(ctypes.c_int64)
    ip = ctypes.POINTER(ctypes.c_int)
    lib.oracle_pair_block.argtypes = [ctypes.c_int, dp, dp, ctypes.c_double, dp]
    lib.oracle_pair_block.restype = None
    lib.oracle_dense_matrix.argtypes = [ctypes.c_int64, dp, i64p, ctypes.c_int64, ip,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_int, dp]
    lib.oracle_dense_matrix.restype = None
    lib.oracle_mdot.argtypes = [ctypes.c_int64, dp, i64p, ctypes.c_int64, ip, ctypes.c_double,
                                ctypes.c_double, ctypes.c_int, dp, dp, dp, dp]
    lib.oracle_mdot.restype = None
    lib.oracle_mdot_rows.argtypes = [ctypes.c_int64, dp, i64p, ctypes.c_int64, ip, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int, dp, dp, ctypes.c_int64, i64p,
                                     dp, dp]
    lib.oracle_mdot_rows.restype = None
    lib.oracle_mdot_open_uf.argtypes = [ctypes.c_int64, dp, ctypes.c_double, ctypes.c_double,
                                        dp, dp]
    lib.oracle_mdot_open_uf.restype = None
    lib.oracle_mdot_open_uf_rows.argtypes = [ctypes.c_int64, dp, ctypes.c_double, ctypes.c_double,
                                             dp, ctypes.c_int64, ctypes.c_int64, dp]
    lib.oracle_mdot_open_uf_rows.restype = None
    lib.oracle_num_threads.restype = ctypes.c_int
    _lib = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def num_threads():
    return int(_load().oracle_num_threads())


def mobility_constants(eta, a):
    """t0, rt0, r0 of solvers/NBody/mobility.h:99-106."""
    t0 = 1.0 / (6 * np.pi * eta * a)
    rt0 = 1.0 / (8 * np.pi * eta * a * a)
    r0 = 1.0 / (8 * np.pi * eta * a ** 3)
    return t0, rt0, r0


def pair_block(wall, pi, pj, a):
    """Dimensionless 6x6 coupling block (rows U,W of i; cols F,T of j)."""
    pi = np.ascontiguousarray(pi, dtype=np.float64)
    pj = np.ascontiguousarray(pj, dtype=np.float64)
    out = np.zeros(36)
    _load().oracle_pair_block(int(wall), _dptr(pi), _dptr(pj), float(a), _dptr(out))
    return out.reshape(6, 6)


def _systems(N, sys_sizes, walls):
    if sys_sizes is None:
        sys_sizes = [N]
    sizes = np.asarray(sys_sizes, dtype=np.int64)
    off = np.zeros(len(sizes) + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    assert off[-1] == N
    if np.isscalar(walls) or isinstance(walls, bool):
        walls = [int(walls)] * len(sizes)
    wf = np.ascontiguousarray(walls, dtype=np.int32)
    return off, wf


def dense_matrix(pos, eta, a, angular=False, wall=False, sys_sizes=None):
    """Dense grand mobility matrix, layout of tests/utils.py:62-84.

    ``wall`` may be a bool or a per-system list; positions are relative to
    the wall at z = 0 (the caller subtracts wallHeight, mobility.h:124-136).
    """
    pos = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1))
    N = pos.size // 3
    off, wf = _systems(N, sys_sizes, wall)
    dim = 6 * N if angular else 3 * N
    M = np.zeros((dim, dim))
    _load().oracle_dense_matrix(N, _dptr(pos), off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                len(off) - 1, wf.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                float(eta), float(a), int(angular), _dptr(M))
    return M


def mdot(pos, eta, a, F=None, T=None, angular=False, wall=False, sys_sizes=None):
    """Matrix-free FP64 product; returns (U, W or None) as flat 3N arrays."""
    pos = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1))
    N = pos.size // 3
    off, wf = _systems(N, sys_sizes, wall)
    Fp = None if F is None else np.ascontiguousarray(np.asarray(F, np.float64).reshape(-1))
    Tp = None if T is None else np.ascontiguousarray(np.asarray(T, np.float64).reshape(-1))
    U = np.zeros(3 * N)
    W = np.zeros(3 * N) if angular else None
    _load().oracle_mdot(N, _dptr(pos), off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                        len(off) - 1, wf.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                        float(eta), float(a), int(angular),
                        _dptr(Fp) if Fp is not None else None,
                        _dptr(Tp) if Tp is not None else None,
                        _dptr(U), _dptr(W) if W is not None else None)
    return U, W


def mdot_rows(pos, eta, a, rows, F=None, T=None, angular=False, wall=False, sys_sizes=None):
    """Selected target rows of the matrix-free FP64 product: (U, W or None),
    each of shape (len(rows), 3)."""
    pos = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1))
    N = pos.size // 3
    off, wf = _systems(N, sys_sizes, wall)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    Fp = None if F is None else np.ascontiguousarray(np.asarray(F, np.float64).reshape(-1))
    Tp = None if T is None else np.ascontiguousarray(np.asarray(T, np.float64).reshape(-1))
    U = np.zeros(3 * rows.size)
    W = np.zeros(3 * rows.size) if angular else None
    _load().oracle_mdot_rows(N, _dptr(pos), off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                             len(off) - 1, wf.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                             float(eta), float(a), int(angular),
                             _dptr(Fp) if Fp is not None else None,
                             _dptr(Tp) if Tp is not None else None, rows.size,
                             rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                             _dptr(U), _dptr(W) if W is not None else None)
    return U.reshape(-1, 3), (W.reshape(-1, 3) if W is not None else None)


def mdot_open_uf(pos, eta, a, F):
    """Open-boundary UF product (CPU baseline kernel)."""
    pos = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1))
    F = np.ascontiguousarray(np.asarray(F, dtype=np.float64).reshape(-1))
    N = pos.size // 3
    U = np.zeros(3 * N)
    _load().oracle_mdot_open_uf(N, _dptr(pos), float(eta), float(a), _dptr(F), _dptr(U))
    return U


def mdot_open_uf_rows(pos, eta, a, F, row0, nrows):
    """Rows [row0, row0 + nrows) of the open-boundary UF product (CPU baseline
    sample); returns the 3*nrows velocity entries."""
    pos = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1))
    F = np.ascontiguousarray(np.asarray(F, dtype=np.float64).reshape(-1))
    N = pos.size // 3
    U = np.zeros(3 * N)
    _load().oracle_mdot_open_uf_rows(N, _dptr(pos), float(eta), float(a), _dptr(F), int(row0),
                                     int(nrows), _dptr(U))
    return U[3 * row0: 3 * (row0 + nrows)]


def dense_sqrt(M):
    """Symmetric FP64 square root V sqrt(L) V^T (numpy eigh)."""
    lam, V = np.linalg.eigh(M)
    return (V * np.sqrt(np.clip(lam, 0.0, None))) @ V.T


class LanczosError(RuntimeError):
    pass


class LanczosOracle:
    """Restatement of lanczos::Solver (include/lanczos/LanczosAlgorithm.cu).

    State that persists between ``run`` calls, as in the reference: the
    adaptive ``check_convergence_steps`` (LanczosAlgorithm.h:48, cu:171) and
    the previous result ``oldBz`` (cu:175, kept by ``resize``).
    """

    def __init__(self, iteration_hard_limit=200):
        self.check_convergence_steps = 3
        self.iteration_hard_limit = iteration_hard_limit
        self.oldBz = np.zeros(0)
        self.last_iterations = 0
        self.trace = None  # optional list receiving (iteration, error) of every check

    @staticmethod
    def tridiagonal_sqrt_e1(hdiag, hsup):
        """T^{1/2} e1 via the eigen-decomposition (cu:48-85)."""
        m = len(hdiag)
        T = np.diag(hdiag) + np.diag(hsup[: m - 1], 1) + np.diag(hsup[: m - 1], -1)
        lam, P = np.linalg.eigh(T)
        with np.errstate(invalid="ignore"):
            h = np.sqrt(lam) * P[0, :]
        return P @ h

    def run(self, dot, z, tolerance, callback=None):
        z = np.asarray(z, dtype=np.float64)
        n = z.size
        # oldBz.resize(N + 1): keeps the previous contents, pads with zeros (cu:175)
        old = np.zeros(n + 1)
        k = min(self.oldBz.size, n + 1)
        old[:k] = self.oldBz[:k]
        self.oldBz = old
        normz = np.linalg.norm(z)
        V = [z / normz]                                        # cu:110-118
        hdiag, hsup = [], []
        check = min(self.check_convergence_steps, self.iteration_hard_limit - 2)
        for i in range(self.iteration_hard_limit):
            # nextIteration, cu:120-153
            w = np.asarray(dot(V[i]), dtype=np.float64).copy()
            if i > 0:
                w -= hsup[i - 1] * V[i - 1]
            alpha = float(w @ V[i])
            hdiag.append(alpha)
            w -= alpha * V[i]
            beta = float(np.linalg.norm(w))
            if beta < 1e-3 * alpha / normz:
                beta = 0.0
            hsup.append(beta)
            if beta > 0.0:
                w = w / beta
            else:
                w = np.zeros(n)
                w[0] = 1.0
            V.append(w)
            if i >= check or callback is not None:
                s = self.tridiagonal_sqrt_e1(np.array(hdiag), np.array(hsup))
                Bz = normz * (np.array(V[: i + 1]).T @ s)     # cu:158-167
                if i > 0:
                    prev = self.oldBz[:n]
                    with np.errstate(divide="ignore", invalid="ignore"):
                        err = abs(np.linalg.norm(Bz - prev) / np.linalg.norm(prev))  # cu:202-222
                    if np.isnan(err):
                        raise LanczosError("[Lanczos] Unknown error (found NaN in result guess)")
                    if self.trace is not None:
                        self.trace.append((i, err))
                    if callback is not None:
                        callback(i, err)
                    if err <= tolerance:
                        if i - 2 > self.check_convergence_steps:    # cu:224-232
                            self.check_convergence_steps += 1
                        else:
                            self.check_convergence_steps = max(1, self.check_convergence_steps - 2)
                        self.last_iterations = i
                        # computeError left oldBz = oldBz - Bz (cu:211); it is kept
                        self.oldBz[:n] = prev - Bz
                        return Bz, i
                self.oldBz[:n] = Bz
        raise LanczosError("[Lanczos] Could not converge")


def rfd_drift(pos, W, eta, a, delta, prefactor=1.0, angular=False, wall=True, sys_sizes=None,
              rows=None):
    """(prefactor/delta) [M(q + delta/2 W) W - M(q - delta/2 W) W] with W used
    as both displacement and force (random_finite_differences.h:46-62); the
    step delta is already multiplied by a (mobility.h:224).  With ``rows``,
    only those target particles (every source moves), shaped (len(rows), 3)."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1)
    W = np.asarray(W, dtype=np.float64).reshape(-1)
    kw = dict(F=W, angular=angular, wall=wall, sys_sizes=sys_sizes)
    if rows is None:
        Up, Wp = mdot(pos + 0.5 * delta * W, eta, a, **kw)
        Um, Wm = mdot(pos - 0.5 * delta * W, eta, a, **kw)
    else:
        Up, Wp = mdot_rows(pos + 0.5 * delta * W, eta, a, rows, **kw)
        Um, Wm = mdot_rows(pos - 0.5 * delta * W, eta, a, rows, **kw)
    lin = (prefactor / delta) * (Up - Um)
    ang = (prefactor / delta) * (Wp - Wm) if angular else None
    return lin, ang


def deterministic_divM(pos, eta, a, delta, angular=False, wall=True):
    """Central-difference divergence of M, tests/test_thermal_drift.py:179-207."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1)
    N = pos.size // 3
    size = 6 * N if angular else 3 * N
    div = np.zeros(size)
    for j in range(3 * N):
        pp = pos.copy()
        pm = pos.copy()
        pp[j] += 0.5 * delta
        pm[j] -= 0.5 * delta
        Mp = dense_matrix(pp, eta, a, angular=angular, wall=wall)
        Mm = dense_matrix(pm, eta, a, angular=angular, wall=wall)
        div += Mp[:, j] - Mm[:, j]
    return div / delta
