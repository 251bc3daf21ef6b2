"""Pins the oracle's Lanczos restatement and dense square root (CPU only)."""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle


def _spd(n, kappa, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    lam = np.exp(rng.uniform(0, np.log(kappa), n))
    return (Q * lam) @ Q.T


def test_dense_sqrt_squares_back():
    A = _spd(40, 1e4, 0)
    B = oracle.dense_sqrt(A)
    assert np.allclose(B @ B, A, rtol=1e-10, atol=1e-10 * np.abs(A).max())
    assert np.allclose(B, B.T, atol=1e-12 * np.abs(B).max())


def test_tridiagonal_sqrt_against_scipy():
    rng = np.random.default_rng(1)
    for m in (1, 2, 7, 40):
        a = rng.uniform(1, 3, m)
        b = rng.uniform(0.1, 0.5, m)
        lam, P = sl.eigh_tridiagonal(a, b[: m - 1])
        ref = P @ (np.sqrt(lam) * P[0])
        got = oracle.LanczosOracle.tridiagonal_sqrt_e1(a, b)
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("kappa", [1.0, 30.0, 1e3])
def test_lanczos_converges_to_sqrt(kappa):
    A = _spd(120, kappa, 2)
    B = oracle.dense_sqrt(A)
    lz = oracle.LanczosOracle()
    rng = np.random.default_rng(3)
    for call in range(3):
        z = rng.normal(size=120)
        y, it = lz.run(lambda v: A @ v, z, 1e-8)
        assert np.linalg.norm(y - B @ z) <= 1e-6 * np.linalg.norm(B @ z)
        assert 0 < it < 200


def test_lanczos_state_and_schedule():
    """Adaptive check schedule of LanczosAlgorithm.cu:224-232."""
    A = _spd(60, 50.0, 4)
    lz = oracle.LanczosOracle()
    z = np.random.default_rng(5).normal(size=60)
    _, it = lz.run(lambda v: A @ v, z, 1e-6)
    assert lz.check_convergence_steps == (4 if it - 2 > 3 else 1)


def test_lanczos_not_spd_raises():
    A = -np.eye(5)
    with pytest.raises(oracle.LanczosError, match="NaN"):
        oracle.LanczosOracle().run(lambda v: A @ v, np.ones(5), 1e-6)


def test_lanczos_nonconvergence_raises():
    A = _spd(300, 1e8, 6)
    lz = oracle.LanczosOracle(iteration_hard_limit=10)
    with pytest.raises(oracle.LanczosError, match="Could not converge"):
        lz.run(lambda v: A @ v, np.ones(300), 1e-12)


def test_stopping_rule_is_rounding_sensitive():
    """Why GPU/oracle iteration counts are compared within a window: the
    reference's error estimate ||y_i - y_{i-1}|| / ||y_{i-1}|| near the
    tolerance changes by O(10%) under a 1e-16 relative perturbation of M v,
    and a mere reordering of the matrix-vector sum shifts the converged step."""
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, 16.0, (200, 3))
    M = oracle.dense_matrix(pos, 1 / (6 * np.pi), 1.0)
    W = np.random.default_rng(7).normal(size=600)
    runs = {}
    for name, dot in (("plain", lambda v: M @ v),
                      ("reordered", lambda v: M[:, ::-1] @ v[::-1])):
        lz = oracle.LanczosOracle()
        lz.trace = []
        y, it = lz.run(dot, W, 1e-6)
        runs[name] = (y, it, lz.trace)
    (y0, it0, tr0), (y1, it1, tr1) = runs["plain"], runs["reordered"]
    assert it0 != it1 or abs(tr0[-1][1] - tr1[-1][1]) > 1e-2 * tr0[-1][1]
    exact = oracle.dense_sqrt(M) @ W
    for y in (y0, y1):
        assert np.linalg.norm(y - exact) <= 20e-6 * np.linalg.norm(exact)
