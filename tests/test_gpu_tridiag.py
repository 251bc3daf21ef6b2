"""GPU checks of the Lanczos tridiagonal square root across condition numbers.

The kernel (lanczos_kernels.hip, tridiag_sqrt_body) replaces host
LAPACKE ?steqr + cblas ?gemv (include/lanczos/LanczosAlgorithm.cu:48-85).
lm_selftest_tridiag_sqrt runs the same device code on a given T:
  * against the numpy restatement of the kernel (tests/tridiag_rule.py), whose
    range is pinned on the CPU against a 30-digit reference (1e-13 below kappa
    1e16: the two differ only in the order of the node sum and in exp/sinh
    rounding; looser where the bracket sits on the 1e-16 eigenvalue floor);
  * against the 30-digit reference directly, kappa 1e2 .. 1e16 (<= 1e-12)
    and 1e18 beyond the eigenvalue floor (<= 1e-7);
  * sizes 1 and 1024 (the iteration limit), and NaN for a negative
    eigenvalue (which the solver reports as the reference does).
End to end, damped wall particles (z down to 0.005a, hydrodynamicKernels.cuh
:362-390) make M ill-conditioned; sqrtMdotW_with_noise is compared with the
dense FP64 eigen square root as the other Lanczos tests are.
"""
import ctypes

import numpy as np
import pytest

import oracle
from tests import tridiag_rule
from tests.test_tridiag_rule import _mp_sqrt_e1, _T

pytestmark = pytest.mark.gpu


def _device(a, b):
    from libmobility_amd import _ffi
    lib = _ffi.load()
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b if len(b) else [0.0], dtype=np.float64)
    out = np.zeros(a.size)
    dp = ctypes.POINTER(ctypes.c_double)
    _ffi.check(lib.lm_selftest_tridiag_sqrt(a.ctypes.data_as(dp), b.ctypes.data_as(dp), a.size,
                                            out.ctypes.data_as(dp)))
    return out


def _rel(x, y):
    return np.linalg.norm(x - y) / np.linalg.norm(y)


@pytest.mark.parametrize("kappa,bound", [(1e2, 1e-12), (1e6, 1e-12), (1e9, 1e-12),
                                         (1e12, 1e-12), (1e14, 1e-12), (1e16, 3e-12),
                                         (1e18, 1e-7)])
def test_device_tridiagonal_square_root_range(kappa, bound):
    a, b = _T(kappa)
    got = _device(a, b)
    # at 1e16 and beyond, the bracket sits on the 1e-16 floor, which rounds
    # differently on the two sides
    agree = 1e-13 if kappa < 1e16 else (1e-12 if kappa == 1e16 else 1e-9)
    assert _rel(got, tridiag_rule.sqrt_e1(a, b)) <= agree
    assert _rel(got, _mp_sqrt_e1(a, b)) <= bound


def test_device_tridiagonal_square_root_sizes_and_negative_spectrum():
    assert np.allclose(_device([9.0], []), [3.0], rtol=1e-14)
    rng = np.random.default_rng(4)
    m = 1024
    a = rng.uniform(2.0, 5.0, m)
    b = rng.uniform(-0.9, 0.9, m - 1)  # diagonally dominant: SPD
    got = _device(a, b)
    assert _rel(got, tridiag_rule.sqrt_e1(a, b)) <= 1e-13
    a2, b2 = _T(1e4, seed=2, m=16)
    assert np.all(np.isnan(_device(a2 - 2 * 3.7, b2)))


@pytest.mark.parametrize("angular", [False, True])
def test_damped_wall_particles_square_root(angular):
    """Particles between 0.005a and 0.1a above the wall next to ordinary ones:
    the damping b = z/a shrinks their self mobility by up to 4e4 (kappa of M
    1.4e6; the reference algorithm needs 88 and 120 iterations at tol 1e-6)."""
    rng = np.random.default_rng(12)
    N, a, eta, tol = 60, 1.0, 1 / (6 * np.pi), 1e-6
    pos = rng.uniform(0, 12.0, (N, 3))
    pos[:, 2] = rng.uniform(1.2, 5.0, N)
    pos[:12, 2] = np.geomspace(0.005, 0.1, 12)
    from libmobility_amd import NBody
    s = NBody("open", "open", "single_wall")
    s.setParameters(wallHeight=0.0)
    s.initialize(viscosity=eta, hydrodynamicRadius=a, includeAngular=angular, tolerance=tol)
    s.setPositions(pos)
    M = oracle.dense_matrix(pos, eta, a, angular=angular, wall=True)
    lam = np.linalg.eigvalsh(M)
    assert lam.max() / lam.min() > 1e4
    exact = oracle.dense_sqrt(M)
    lz = oracle.LanczosOracle()
    for call in range(2):
        W = rng.normal(size=M.shape[0])
        u, w = s.sqrtMdotW_with_noise(W)
        got = u.ravel() if not angular else np.concatenate([u.ravel(), w.ravel()])
        ref, _ = lz.run(lambda v: M @ v, W, tol)
        ex = exact @ W
        err_ref = np.linalg.norm(ref - ex)
        assert np.linalg.norm(got - ex) <= 3 * err_ref + 5 * tol * np.linalg.norm(ex)
