"""Range of the device tridiagonal square root, restated in numpy
(tests/tridiag_rule.py), against references that do not share its method:

  * a 30-digit eigen decomposition (mpmath) of the same T, for condition
    numbers 1e2 .. 1e16 (<= 1e-12 relative) and 1e18, beyond the 1e-16
    eigenvalue floor (<= 1e-7: eigenvalues below the floor count as lying on
    it, which moves their square roots by at most 1e-8 of the largest);
  * scipy.linalg.eigh_tridiagonal (LAPACK), whose own error grows with the
    condition number (6e-11 against the 30-digit reference at 1e16), up to
    1e12 (<= 1e-13 .. 3e-12);
  * the scalar rule: lam * q(lam) = sqrt(lam) over the whole bracketed
    spectrum (<= 2e-13 up to kappa 1e16).

T is the tridiagonal of a full-length Lanczos run (with reorthogonalisation)
on a diagonal matrix of prescribed spectrum, so it has exactly that spectrum:
half log-spaced across [1, kappa] (isolated small eigenvalues, as damped wall
particles give) and half clustered in the top decade.
"""
import mpmath as mp
import numpy as np
import pytest
import scipy.linalg as sl

from tests import tridiag_rule


def _lanczos_T(lam, seed):
    rng = np.random.default_rng(seed)
    n = lam.size
    v = rng.normal(size=n)
    V = [v / np.linalg.norm(v)]
    a, b = [], []
    for i in range(n):
        w = lam * V[-1]
        if i > 0:
            w -= b[-1] * V[-2]
        a.append(w @ V[-1])
        w -= a[-1] * V[-1]
        for u in V:
            w -= (w @ u) * u
        if i < n - 1:
            b.append(np.linalg.norm(w))
            V.append(w / b[-1])
    return np.array(a), np.array(b)


def _T(kappa, seed=0, m=32):
    rng = np.random.default_rng(seed)
    lam = np.concatenate([np.logspace(0, np.log10(kappa), m // 2),
                          rng.uniform(kappa / 10, kappa, m - m // 2)]) * 3.7
    return _lanczos_T(lam, seed)


def _mp_sqrt_e1(a, b):
    mp.mp.dps = 30
    m = len(a)
    T = mp.matrix(m, m)
    for i in range(m):
        T[i, i] = mp.mpf(a[i])
        if i < m - 1:
            T[i, i + 1] = T[i + 1, i] = mp.mpf(b[i])
    E, Q = mp.eigsy(T)
    y = [mp.mpf(0)] * m
    for k in range(m):
        c = mp.sqrt(max(E[k], 0)) * Q[0, k]
        for i in range(m):
            y[i] += c * Q[i, k]
    return np.array([float(v) for v in y])


def _rel(x, y):
    return np.linalg.norm(x - y) / np.linalg.norm(y)


@pytest.mark.parametrize("kappa,bound", [(1e2, 1e-13), (1e6, 1e-13), (1e9, 1e-13),
                                         (1e12, 1e-12), (1e14, 1e-12), (1e16, 3e-12),
                                         (1e18, 1e-7)])
def test_device_rule_against_30_digit_reference(kappa, bound):
    a, b = _T(kappa)
    assert _rel(tridiag_rule.sqrt_e1(a, b), _mp_sqrt_e1(a, b)) <= bound


@pytest.mark.parametrize("kappa,bound", [(1.0, 1e-13), (1e3, 1e-13), (1e8, 1e-12), (1e10, 1e-12),
                                         (1e12, 3e-12)])
def test_device_rule_against_eigh_tridiagonal(kappa, bound):
    a, b = _T(kappa, seed=3, m=48)
    lam, V = sl.eigh_tridiagonal(a, b)
    ref = V @ (np.sqrt(np.clip(lam, 0, None)) * V[0])
    assert _rel(tridiag_rule.sqrt_e1(a, b), ref) <= bound


@pytest.mark.parametrize("kappa", [1e2, 1e9, 1e12, 1e16])
def test_bracket_and_scalar_quadrature(kappa):
    a, b = _T(kappa, seed=1)
    lam = sl.eigh_tridiagonal(a, b, eigvals_only=True)
    lmin, lmax = tridiag_rule.bounds(a, b)
    # the bracket is a lower bound within 7.5% of the smallest eigenvalue, or
    # the floor 1e-16 lmax above it
    floor = 1e-16 * lmax
    assert lmin <= max(lam.min(), floor) * (1 + 1e-12)
    assert lmin >= lam.min() / 1.0747 * (1 - 1e-9)
    assert lmax >= lam.max()
    sig, w = tridiag_rule.nodes(lmin, lmax)
    x = np.logspace(np.log10(lmin), np.log10(lmax), 1500)
    q = (w[None, :] / (x[:, None] + sig[None, :])).sum(1)
    assert np.max(np.abs(x * q - np.sqrt(x)) / np.sqrt(x)) <= 2e-13


def test_one_by_one_and_negative_spectrum():
    assert np.allclose(tridiag_rule.sqrt_e1([4.0], []), [2.0], rtol=1e-14)
    a, b = _T(1e4, seed=2, m=16)
    with np.errstate(all="ignore"):
        lam = sl.eigh_tridiagonal(a - 2 * 3.7, b, eigvals_only=True)
    assert lam.min() < 0  # the kernel answers NaN there; the rule needs SPD
