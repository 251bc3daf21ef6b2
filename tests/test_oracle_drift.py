"""Pins of the oracle's thermal-drift functions against closed forms.

``oracle.deterministic_divM`` restates the central-difference divergence of
the reference's tests (tests/test_thermal_drift.py:179-207) and
``oracle.rfd_drift`` the random finite difference of
include/MobilityInterface/random_finite_differences.h:46-62.  Neither is
checked here against another finite difference of the same matrix: the
expected values come from the mathematics.

* One particle at height h >= a above the wall: the self mobility is the
  Swan & Brady (2007, eq. B1) series, M_zz = t0 (1 - 9/(8h) + 1/(2h^3) -
  1/(8h^5)) with h = z/a, M_xx = M_yy independent of z.  The divergence is
  d M_zz / dz = (t0/a) (9/(8h^2) - 3/(2h^4) + 5/(8h^6)) along z, zero along
  x and y, and zero for the angular rows (the self W-F coupling only has x-y
  entries, hydrodynamicKernels.cuh:287-359, and none depends on x or y).
* Below one radius the reference damps every block by b_i b_j with
  b = z/a and clamps the height to a (hydrodynamicKernels.cuh:362-390), so
  M_zz = (z/a)^2 t0 (1 - 9/8 + 1/2 - 1/8) = t0 z^2 / (4 a^2) and the drift is
  t0 z / (2 a^2).
* Open boundaries: the RPY tensor f(r) I + g(r) r^r^ is divergence-free in
  both branches (f' + g' + 2g/r = 0 for r >= 2a and for r < 2a), so the
  divergence of the open grand mobility matrix vanishes.
"""
import numpy as np
import pytest

import oracle


def _swan_brady_drift(h):
    return 9.0 / (8 * h**2) - 3.0 / (2 * h**4) + 5.0 / (8 * h**6)


@pytest.mark.parametrize("a", [1.0, 1.3])
@pytest.mark.parametrize("h", [1.2, 1.7, 2.5, 4.0, 7.3])
@pytest.mark.parametrize("angular", [False, True])
def test_deterministic_divM_one_particle_above_wall(a, h, angular):
    eta = 0.7
    t0 = 1.0 / (6 * np.pi * eta * a)
    pos = np.array([0.4, -1.1, h * a])
    delta = 1e-4 * a
    div = oracle.deterministic_divM(pos, eta, a, delta, angular=angular, wall=True)
    assert div.shape == ((6,) if angular else (3,))
    expect = t0 / a * _swan_brady_drift(h)
    assert abs(div[2] - expect) <= 1e-7 * abs(expect)
    assert np.all(np.abs(div[:2]) <= 1e-9 * abs(expect))
    if angular:
        assert np.all(np.abs(div[3:]) <= 1e-9 * abs(expect))


@pytest.mark.parametrize("zr", [0.3, 0.6, 0.85])
def test_deterministic_divM_one_particle_damped(zr):
    eta, a = 1.0 / (6 * np.pi), 1.4
    t0 = 1.0 / (6 * np.pi * eta * a)
    pos = np.array([2.0, 3.0, zr * a])
    div = oracle.deterministic_divM(pos, eta, a, 1e-4 * a, wall=True)
    expect = t0 * zr * a / (2 * a**2)
    assert abs(div[2] - expect) <= 1e-7 * abs(expect)
    assert np.all(np.abs(div[:2]) <= 1e-9 * abs(expect))


def test_deterministic_divM_open_is_divergence_free():
    """Far and overlapping pairs: the open RPY divergence is zero."""
    rng = np.random.default_rng(3)
    eta, a = 0.9, 1.1
    pos = rng.uniform(0, 4.0, (6, 3))  # several pairs closer than 2a
    d = np.linalg.norm(pos[:, None] - pos[None], axis=-1)
    assert np.sum((d > 0) & (d < 2 * a)) >= 2 and np.sum(d > 2 * a) >= 2
    div = oracle.deterministic_divM(pos, eta, a, 1e-4 * a, angular=True, wall=False)
    t0 = 1.0 / (6 * np.pi * eta * a)
    assert np.max(np.abs(div)) <= 1e-7 * t0 / a


def test_deterministic_divM_pair_near_wall_matches_the_sum_of_self_terms_far_apart():
    """Two particles 10^4 radii apart: each one's drift is its own closed form."""
    eta, a = 1.0 / (6 * np.pi), 1.0
    pos = np.array([[0.0, 0.0, 1.5], [1e4, 0.0, 3.0]])
    div = oracle.deterministic_divM(pos, eta, a, 1e-4, wall=True).reshape(2, 3)
    for k, h in enumerate((1.5, 3.0)):
        assert abs(div[k, 2] - _swan_brady_drift(h)) <= 1e-6 * _swan_brady_drift(h)


def test_rfd_drift_one_particle_closed_form():
    """One random finite difference with W = e_z gives (M_zz(z+d/2) -
    M_zz(z-d/2)) / d, the closed-form derivative; W = e_x gives no z drift."""
    eta, a, h = 0.5, 1.2, 2.2
    t0 = 1.0 / (6 * np.pi * eta * a)
    pos = np.array([1.0, 1.0, h * a])
    lin, ang = oracle.rfd_drift(pos, np.array([0.0, 0.0, 1.0]), eta, a, 1e-4 * a, angular=True)
    expect = t0 / a * _swan_brady_drift(h)
    assert abs(lin[2] - expect) <= 1e-7 * expect
    assert np.all(np.abs(lin[:2]) <= 1e-9 * expect) and np.all(np.abs(ang) <= 1e-9 * expect)
    lin, _ = oracle.rfd_drift(pos, np.array([1.0, 0.0, 0.0]), eta, a, 1e-4 * a)
    assert np.all(np.abs(lin) <= 1e-9 * expect)


def test_rfd_drift_prefactor_and_unit_vectors_sum_to_the_divergence():
    """sum_j RFD(W = e_j) equals the central-difference divergence: the RFD
    with a unit displacement is one column difference of M (matrix-free
    product here, dense columns there)."""
    rng = np.random.default_rng(8)
    eta, a = 0.8, 1.0
    pos = rng.uniform(0, 5, (4, 3))
    pos[:, 2] = rng.uniform(0.6, 4.0, 4)
    delta = 1e-4
    total = np.zeros(12)
    for j in range(12):
        e = np.zeros(12)
        e[j] = 1.0
        lin, _ = oracle.rfd_drift(pos, e, eta, a, delta)
        total += lin
    ref = oracle.deterministic_divM(pos, eta, a, delta, wall=True)
    assert np.allclose(total, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())
    lin2, _ = oracle.rfd_drift(pos, e, eta, a, delta, prefactor=2.5)
    lin1, _ = oracle.rfd_drift(pos, e, eta, a, delta)
    assert np.allclose(lin2, 2.5 * lin1, rtol=1e-14, atol=0)
