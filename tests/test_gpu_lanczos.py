"""GPU parity of sqrtMdotW (Lanczos) and divM (random finite differences).

Checks, through the C ABI:
  * explicit-W parity with the oracle's restatement of lanczos::Solver
    (converged iteration within a small window, results within a few x tol), including the
    persistent state between calls (adaptive check schedule, previous
    estimate), LanczosAlgorithm.cu:171-232;
  * agreement with the dense FP64 eigen square root applied to the same W,
    bounded by the error of the oracle's restated algorithm on that W (the
    stopping rule measures successive change, not the error);
  * fluctuation-dissipation: sample covariance of sqrtMdotW -> M
    (test_fluctuation_dissipation.py:19-112 in spirit);
  * prefactor linearity with equal seeds (test_interface.py:384-418);
  * RFD divM mean vs the deterministic divergence (test_thermal_drift.py).
"""
import numpy as np
import pytest

import oracle
from tests.test_gpu_mdot import _solver, compute_M

pytestmark = pytest.mark.gpu


def _open_cloud(N, L, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, L, (N, 3))


@pytest.mark.parametrize("wall,angular", [(False, False), (False, True), (True, False), (True, True)])
def test_lanczos_explicit_noise_matches_oracle(wall, angular):
    rng = np.random.default_rng(7)
    N = 200
    a, eta, tol = 1.0, 1 / (6 * np.pi), 1e-6
    pos = _open_cloud(N, 16.0, 3)
    if wall:
        pos[:, 2] = rng.uniform(1.2, 8.0, N)
    from libmobility_amd import NBody
    s = NBody("open", "open", "single_wall" if wall else "open")
    s.setParameters(wallHeight=0.0) if wall else s.setParameters()
    s.initialize(viscosity=eta, hydrodynamicRadius=a, includeAngular=angular, tolerance=tol)
    s.setPositions(pos)
    M = oracle.dense_matrix(pos, eta, a, angular=angular, wall=wall)
    B = oracle.dense_sqrt(M)
    lz = oracle.LanczosOracle()
    n = M.shape[0]
    for call in range(4):  # state carried between calls
        W = rng.normal(size=n)
        u, w = s.sqrtMdotW_with_noise(W, prefactor=1.0)
        got = u.ravel() if not angular else np.concatenate([u.ravel(), w.ravel()])
        ref, it = lz.run(lambda v: M @ v, W, tol)
        # The stopping rule of the reference (no reorthogonalisation) reacts to
        # rounding at the 1e-16 level (tests/test_oracle_lanczos.py::
        # test_stopping_rule_is_rounding_sensitive), so the converged iteration
        # may differ by a few steps; the result must still meet the tolerance.
        assert abs(s.lastLanczosIterations() - it) <= 6
        assert np.linalg.norm(got - ref) <= 5 * tol * np.linalg.norm(ref)
        exact = B @ W
        err_ref = np.linalg.norm(ref - exact)
        assert np.linalg.norm(got - exact) <= 3 * err_ref + 5 * tol * np.linalg.norm(exact)


def test_lanczos_c1_fp64_against_eigh():
    """C1: N=1024 open, cube 24a, eta 1/(6 pi), tol 1e-6."""
    rng = np.random.default_rng(0)
    pos = rng.uniform(0, 24.0, (1024, 3))
    eta = 1 / (6 * np.pi)
    s = _solver(False, 1.0, eta)
    s.initialize(viscosity=eta, hydrodynamicRadius=1.0, tolerance=1e-6)
    s.setPositions(pos)
    W = rng.normal(size=3072)
    u, _ = s.sqrtMdotW_with_noise(W)
    M = oracle.dense_matrix(pos, eta, 1.0)
    exact = oracle.dense_sqrt(M) @ W
    err = np.linalg.norm(u.ravel() - exact) / np.linalg.norm(exact)
    # The stopping rule bounds the change between successive estimates, not
    # the error; the bound is the error of the restated reference algorithm
    # on the same W.
    ref, _ = oracle.LanczosOracle().run(lambda v: M @ v, W, 1e-6)
    err_ref = np.linalg.norm(ref - exact) / np.linalg.norm(exact)
    assert err <= 3 * err_ref + 5e-6, (err, err_ref)


def test_prefactor_linearity_same_seed():
    pos = _open_cloud(10, 6.0, 1)
    outs = []
    for pref in (1.0, 2.45):
        s = _solver(True, 1.0, 1.0, angular=True)
        s.initialize(viscosity=1.0, hydrodynamicRadius=1.0, includeAngular=True, seed=1234)
        s.setPositions(pos + np.array([0, 0, 1.0]))
        outs.append((s.sqrtMdotW(prefactor=pref), s.divM(prefactor=pref)))
    (sq1, dv1), (sq2, dv2) = outs
    for x1, x2 in zip(sq1 + dv1, sq2 + dv2):
        assert np.allclose(x2, 2.45 * x1, atol=1e-12, rtol=1e-10)


def test_fluctuation_dissipation_covariance():
    rng = np.random.default_rng(2)
    N = 4
    pos = rng.uniform(0, 4.0, (N, 3))
    pos[:, 2] += 1.5
    s = _solver(True, 1.0, 1.0, angular=True)
    s.initialize(viscosity=1.0, hydrodynamicRadius=1.0, includeAngular=True, tolerance=1e-8,
                 seed=99)
    s.setPositions(pos)
    M = compute_M(s, N, True)
    ns = 4000
    X = np.empty((ns, 6 * N))
    for k in range(ns):
        u, w = s.sqrtMdotW()
        X[k] = np.concatenate([u.ravel(), w.ravel()])
    # whitened samples M^{-1/2} x must be i.i.d. N(0,1)
    lam, V = np.linalg.eigh(M)
    Z = (X @ V) / np.sqrt(lam)
    mean = Z.mean(axis=0)
    cov = np.cov(Z.T)
    assert np.abs(mean).max() < 5.0 / np.sqrt(ns) * 1.3
    assert np.abs(cov - np.eye(6 * N)).max() < 6.0 * np.sqrt(2.0 / ns)


def test_rfd_divm_matches_deterministic_divergence():
    rng = np.random.default_rng(4)
    N = 3
    a = 1.0
    pos = rng.uniform(0, 3.0, (N, 3))
    pos[:, 2] = rng.uniform(1.1, 2.5, N)
    delta = 1e-3
    s = _solver(True, a, 1.0, angular=True, delta=delta)
    s.initialize(viscosity=1.0, hydrodynamicRadius=a, includeAngular=True, seed=5)
    s.setPositions(pos)
    det = oracle.deterministic_divM(pos, 1.0, a, delta * a, angular=True, wall=True)
    ns = 20000
    acc = np.zeros(6 * N)
    acc2 = np.zeros(6 * N)
    for _ in range(ns):
        u, w = s.divM()
        x = np.concatenate([u.ravel(), w.ravel()])
        acc += x
        acc2 += x * x
    mean = acc / ns
    sem = np.sqrt(np.maximum(acc2 / ns - mean ** 2, 0) / ns)
    assert np.all(np.abs(mean - det) <= 6 * sem + 1e-3 * np.abs(det) + 1e-9)


def test_divm_open_is_zero_and_positions_unchanged():
    pos = _open_cloud(10, 8.0, 3)
    s = _solver(False, 1.0, 1.0)
    s.setPositions(pos)
    u, _ = s.divM()
    assert np.all(u == 0)
    sw = _solver(True, 1.0, 1.0)
    sw.setPositions(pos + np.array([0, 0, 1.5]))
    F = np.random.default_rng(1).normal(size=pos.shape)
    m1, _ = sw.Mdot(F)
    sw.divM()
    m2, _ = sw.Mdot(F)
    assert np.array_equal(m1, m2)  # test_thermal_drift.py:45-70
    d1, _ = sw.divM()
    d2, _ = sw.divM()
    assert np.any(np.abs(d1 - d2) > 1e-5)  # test_thermal_drift.py:73-102


def test_langevin_equals_mdot_at_zero_temperature():
    # test_interface.py:334-355
    pos = _open_cloud(5, 5.0, 8) + np.array([0, 0, 1.2])
    rng = np.random.default_rng(0)
    F = rng.normal(size=pos.shape)
    T = rng.normal(size=pos.shape)
    for wall in (False, True):
        s = _solver(wall, 1.0, 1.0, angular=True)
        s.setPositions(pos)
        mf, mt = s.Mdot(F, T)
        lf, lt = s.LangevinVelocities(1.0, 0.0, F, T)
        assert np.allclose(mf, lf, atol=1e-14) and np.allclose(mt, lt, atol=1e-14)


def test_hydrodynamic_velocities_decomposition():
    """hydrodynamicVelocities = Mdot + sqrtMdotW + divM with unit prefactors (same seeds)."""
    pos = _open_cloud(20, 8.0, 9) + np.array([0, 0, 1.3])
    F = np.random.default_rng(3).normal(size=pos.shape)
    s1 = _solver(True, 1.0, 1.0)
    s1.initialize(viscosity=1.0, hydrodynamicRadius=1.0, seed=42)
    s1.setPositions(pos)
    total, _ = s1.hydrodynamicVelocities(F)
    s2 = _solver(True, 1.0, 1.0)
    s2.initialize(viscosity=1.0, hydrodynamicRadius=1.0, seed=42)
    s2.setPositions(pos)
    m, _ = s2.Mdot(F)
    q, _ = s2.sqrtMdotW(1.0)
    d, _ = s2.divM(1.0)
    assert np.allclose(total, m + q + d, atol=1e-12, rtol=1e-12)


def test_iteration_limit_beyond_default():
    """lanczos::Solver::iterationHardLimit (LanczosAlgorithm.h:49) raised to
    300: a tolerance no estimate can meet runs every iteration, with the
    estimate and error computed at each one (callback set, cu:183), and ends in
    "[Lanczos] Could not converge" (cu:200) as the restated solver does.  The
    first errors follow the oracle's; the later ones stay at rounding level."""
    rng = np.random.default_rng(3)
    N = 200
    pos = rng.uniform(0, 16.0, (N, 3))
    eta = 1 / (6 * np.pi)
    W = np.random.default_rng(5).normal(size=3 * N)
    M = oracle.dense_matrix(pos, eta, 1.0)
    errs_ref = []
    with pytest.raises(oracle.LanczosError, match="Could not converge"):
        oracle.LanczosOracle(iteration_hard_limit=300).run(
            lambda v: M @ v, W, -1.0, callback=lambda i, e: errs_ref.append(e))
    from libmobility_amd import NBody
    errs = []
    s = NBody("open", "open", "open")
    s.setParameters()
    s.initialize(viscosity=eta, hydrodynamicRadius=1.0, tolerance=-1.0,
                 lanczos_callback=lambda i, e: errs.append((i, e)))
    s.setLanczosIterationLimit(300)
    s.setPositions(pos)
    with pytest.raises(RuntimeError, match="Could not converge"):
        s.sqrtMdotW_with_noise(W)
    assert [i for i, _ in errs] == list(range(1, 300))
    e = np.array([x for _, x in errs])
    assert np.all(np.isfinite(e))
    assert np.allclose(e[:10], errs_ref[:10], rtol=1e-6)
    assert e[250:].max() < 1e-12
    # limits outside [3, 1024] are refused; the default stays 200
    with pytest.raises(RuntimeError, match="iteration limit"):
        s.setLanczosIterationLimit(1025)
    calls = []
    s2 = NBody("open", "open", "open")
    s2.setParameters()
    s2.initialize(viscosity=eta, hydrodynamicRadius=1.0, tolerance=-1.0,
                  lanczos_callback=lambda i, e: calls.append(i))
    s2.setPositions(pos)
    with pytest.raises(RuntimeError, match="Could not converge"):
        s2.sqrtMdotW_with_noise(W)
    assert calls == list(range(1, 200))


def _reorth_lanczos(M, z, tol, check=3, limit=400):
    """numpy Lanczos with the reference's recurrence, breakdown rule and
    stopping rule (LanczosAlgorithm.cu:120-222), plus full reorthogonalisation
    of each new basis vector (classical Gram-Schmidt, FP64) -- the extension
    lm_nbody_set_lanczos_reorthogonalization adds.  Pinned by the exact
    square root in the tests below."""
    n = z.size
    normz = np.linalg.norm(z)
    V = np.zeros((limit + 1, n))
    V[0] = z / normz
    a, b = [], []
    old = np.zeros(n)
    for i in range(limit):
        w = M @ V[i]
        if i > 0:
            w -= b[-1] * V[i - 1]
        al = w @ V[i]
        w -= al * V[i]
        w -= V[:i + 1].T @ (V[:i + 1] @ w)
        be = np.linalg.norm(w)
        if be < 1e-3 * al / normz:
            be = 0.0
        a.append(al)
        b.append(be)
        V[i + 1] = w / be if be > 0 else np.eye(n)[0]
        if i >= check:
            T = np.diag(a) + np.diag(b[:-1], 1) + np.diag(b[:-1], -1)
            lam, P = np.linalg.eigh(T)
            y = normz * (V[:i + 1].T @ (P @ (np.sqrt(np.clip(lam, 0, None)) * P[0])))
            if i > check and np.linalg.norm(old - y) / np.linalg.norm(old) <= tol:
                return y, i
            old = y
    raise RuntimeError("no convergence")


@pytest.mark.parametrize("wall,N,L", [(False, 1024, 24.0), (True, 600, 20.0)])
def test_reorthogonalised_lanczos_against_eigh(wall, N, L):
    """With full FP64 reorthogonalisation the square root follows the exact
    one: on the C1 cloud (overlapping spheres, where the plain recurrence's
    stop moves with rounding and its error against eigh is 1.2-3e-5 at tol
    1e-6) the error is below 1e-5, and the converged iteration is within one
    of the numpy recurrence with the same reorthogonalisation."""
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, L, (N, 3))
    if wall:
        pos[:, 2] = rng.uniform(1.2, 8.0, N)
    eta, tol = 1 / (6 * np.pi), 1e-6
    s = _solver(wall, 1.0, eta)
    s.initialize(viscosity=eta, hydrodynamicRadius=1.0, tolerance=tol)
    s.setLanczosReorthogonalization(True)
    s.setPositions(pos)
    M = oracle.dense_matrix(pos, eta, 1.0, wall=wall)
    B = oracle.dense_sqrt(M)
    check = 3  # the adaptive schedule of LanczosAlgorithm.cu:171, 224-232
    for call in range(2):
        W = rng.normal(size=3 * N)
        u, _ = s.sqrtMdotW_with_noise(W)
        exact = B @ W
        err = np.linalg.norm(u.ravel() - exact) / np.linalg.norm(exact)
        ref, it = _reorth_lanczos(M, W, tol, check=check)
        check = check + 1 if it - 2 > check else max(1, check - 2)
        assert err <= 1e-5, err
        assert abs(s.lastLanczosIterations() - it) <= 1, (s.lastLanczosIterations(), it)
        assert np.linalg.norm(u.ravel() - ref) <= 5 * tol * np.linalg.norm(exact)


def test_reorthogonalisation_makes_fp32_and_fp64_pairs_stop_together():
    """Without reorthogonalisation the FP32-pair operator loses orthogonality
    faster and stops later than FP64 (C1 cloud: 79 vs 70 iterations in numpy);
    with it both stop within one iteration of each other and of the exact
    rate."""
    rng = np.random.default_rng(5)
    pos = rng.uniform(0, 24.0, (1024, 3))
    eta = 1 / (6 * np.pi)
    W = rng.normal(size=3072)
    its = {}
    for prec in ("fp64", "fp32"):
        s = _solver(False, 1.0, eta, precision=prec)
        s.initialize(viscosity=eta, hydrodynamicRadius=1.0, tolerance=1e-6)
        s.setLanczosReorthogonalization(True)
        s.setPositions(pos)
        s.sqrtMdotW_with_noise(W)
        its[prec] = s.lastLanczosIterations()
    assert abs(its["fp32"] - its["fp64"]) <= 1, its
