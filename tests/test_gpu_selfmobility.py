"""SelfMobility (solvers/SelfMobility/mobility.h:17-111) on the GPU: the
diagonal mobility in closed form (test_mobility_matrix.py's self-mobility
check), the statistics of its square root (test_fluctuation_dissipation.py),
the zero drift of the base class (MobilityInterface.h:211-216) and the
Langevin composition (MobilityInterface.h:191-206).  The noise comes from
this library's Philox generator, not the reference's std::normal_distribution
stream, so noise values are compared through their statistics only."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ETA, A = 0.7, 1.3
LIN = 1.0 / (6 * np.pi * ETA * A)
ANG = 1.0 / (8 * np.pi * ETA * A ** 3)


def _solver(angular=False, seed=11):
    from libmobility_amd import SelfMobility
    s = SelfMobility("open", "open", "open")
    s.setParameters(parameter=5.0)
    s.initialize(viscosity=ETA, hydrodynamicRadius=A, includeAngular=angular, seed=seed)
    return s


@pytest.mark.parametrize("framework", ["numpy", "torch_gpu"])
def test_mdot_is_the_diagonal_mobility(framework):
    rng = np.random.default_rng(1)
    N = 37
    pos, F, T = rng.normal(size=(N, 3)), rng.normal(size=(N, 3)), rng.normal(size=(N, 3))
    s = _solver(angular=True)
    if framework == "torch_gpu":
        import torch
        pos, F, T = (torch.from_numpy(x).to("cuda:0") for x in (pos, F, T))
    s.setPositions(pos)
    U, W = s.Mdot(F, T)
    if framework == "torch_gpu":
        U, W, F, T = (x.cpu().numpy() for x in (U, W, F, T))
    assert np.allclose(U, LIN * F, rtol=1e-15, atol=0)
    assert np.allclose(W, ANG * T, rtol=1e-15, atol=0)
    U, W = s.Mdot(forces=F)  # only forces: zero angular velocity
    assert np.allclose(U, LIN * F, rtol=1e-15, atol=0) and np.all(W == 0)


def test_dense_matrix_is_diagonal():
    """Columns of the 6N x 6N matrix (tests/utils.py:62-84)."""
    N = 4
    s = _solver(angular=True)
    s.setPositions(np.random.default_rng(2).normal(size=(N, 3)))
    M = np.zeros((6 * N, 6 * N))
    for c in range(6 * N):
        e = np.zeros(6 * N)
        e[c] = 1.0
        u, w = s.Mdot(e[:3 * N], e[3 * N:])
        M[:, c] = np.concatenate([u.ravel(), w.ravel()])
    ref = np.diag(np.concatenate([np.full(3 * N, LIN), np.full(3 * N, ANG)]))
    assert np.abs(M - ref).max() <= 1e-15 * LIN


def test_sqrt_statistics_match_the_mobility():
    """<(M^1/2 W)(M^1/2 W)^T> = M: per-entry variance, zero mean and no
    correlation between entries, over many independent draws."""
    N = 6
    s = _solver(angular=True, seed=123)
    s.setPositions(np.zeros((N, 3)))
    draws = []
    for _ in range(3000):
        u, w = s.sqrtMdotW()
        draws.append(np.concatenate([u.ravel(), w.ravel()]))
    X = np.array(draws)  # (samples, 6N)
    C = X.T @ X / X.shape[0]
    ref = np.concatenate([np.full(3 * N, LIN), np.full(3 * N, ANG)])
    # 3000 samples: relative standard error sqrt(2/3000) = 2.6 %
    assert np.all(np.abs(np.diag(C) / ref - 1) < 0.12)
    off = C / np.sqrt(np.outer(np.diag(C), np.diag(C)))
    np.fill_diagonal(off, 0)
    assert np.abs(off).max() < 0.1
    assert np.abs(X.mean(axis=0) / np.sqrt(ref)).max() < 0.1


def test_noise_is_reproducible_per_seed_and_scales_with_the_prefactor():
    pos = np.zeros((5, 3))
    a, b, c = _solver(seed=7), _solver(seed=7), _solver(seed=8)
    for s in (a, b, c):
        s.setPositions(pos)
    ua, _ = a.sqrtMdotW(prefactor=1.0)
    ub, _ = b.sqrtMdotW(prefactor=2.5)
    uc, _ = c.sqrtMdotW(prefactor=1.0)
    assert np.allclose(ub, 2.5 * ua, rtol=1e-14, atol=0)
    assert not np.allclose(ua, uc)
    ua2, _ = a.sqrtMdotW()
    assert not np.allclose(ua, ua2)  # new numbers each call
    assert np.all(a.sqrtMdotW(prefactor=0.0)[0] == 0)


def test_divm_is_zero_and_langevin_composes():
    rng = np.random.default_rng(3)
    N = 8
    F, T = rng.normal(size=(N, 3)), rng.normal(size=(N, 3))
    a, b = _solver(angular=True, seed=5), _solver(angular=True, seed=5)
    for s in (a, b):
        s.setPositions(rng.normal(size=(N, 3)))
    d, dw = a.divM()
    assert np.all(d == 0) and np.all(dw == 0)
    dt, kbt = 0.3, 1.7
    v, w = a.LangevinVelocities(dt, kbt, F, T)
    u0, w0 = b.Mdot(F, T)
    nu, nw = b.sqrtMdotW(prefactor=np.sqrt(2 * kbt / dt))
    assert np.allclose(v, u0 + nu, rtol=1e-13, atol=1e-15)
    assert np.allclose(w, w0 + nw, rtol=1e-13, atol=1e-15)
    h, hw = b.hydrodynamicVelocities(F, T)
    assert h.shape == (N, 3) and hw.shape == (N, 3)


def test_libmobility_package_surface():
    """solvers/__init__.py:1-6: the four classes import from libMobility."""
    import libMobility
    from libMobility import DPStokes, NBody, PSE, SelfMobility
    assert set(libMobility.__all__) == {"PSE", "SelfMobility", "NBody", "DPStokes"}
    s = SelfMobility("open", "open", "open")
    s.setParameters(parameter=5.0)
    s.initialize(viscosity=1.0, hydrodynamicRadius=1.0)
    s.setPositions(np.zeros((2, 3)))
    u, _ = s.Mdot(np.ones((2, 3)))
    assert np.allclose(u, 1 / (6 * np.pi))
    n = NBody("open", "open", "open")
    n.setParameters()
    with pytest.raises(RuntimeError):
        PSE("periodic", "periodic", "periodic")
    with pytest.raises(RuntimeError):
        DPStokes("periodic", "periodic", "open")
