"""Ragged ensembles of independent systems, open and single-wall mixed, in one
(persistent) launch: parity with the oracle's block-diagonal mobility and with
per-system solves."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _ensemble(seed=0, sizes=(64, 300, 17, 1000, 5, 129, 700), wall_height=0.0):
    rng = np.random.default_rng(seed)
    kinds = [s % 2 for s in range(len(sizes))]
    pos = []
    for sz, k in zip(sizes, kinds):
        L = (sz * 4 * np.pi / 3 / 0.15) ** (1 / 3)
        p = rng.uniform(0, L, (sz, 3))
        if k:
            p[:, 2] = rng.uniform(0.6, 8.0, sz) + wall_height
        pos.append(p)
    return list(sizes), kinds, np.concatenate(pos)


def _solver(sizes, kinds, pos, angular=False, precision="fp64", persistent=-1, wall_height=0.0,
            tol=1e-6):
    from libmobility_amd import NBody
    s = NBody("open", "open", "single_wall")
    s.setParameters(wallHeight=wall_height)
    s.setPairPrecision(precision)
    s.initialize(viscosity=0.9, hydrodynamicRadius=1.1, includeAngular=angular, tolerance=tol, seed=3)
    s.setSystems(sizes, ["single_wall" if k else "open" for k in kinds])
    s.setPersistent(persistent)
    s.setPositions(pos)
    return s


@pytest.mark.parametrize("angular", [False, True])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_mixed_ensemble_mdot_vs_oracle(angular, precision, tol):
    sizes, kinds, pos = _ensemble(wall_height=2.5)
    rng = np.random.default_rng(1)
    F = rng.normal(size=pos.shape)
    T = rng.normal(size=pos.shape) if angular else None
    s = _solver(sizes, kinds, pos, angular, precision, wall_height=2.5)
    U, W = s.Mdot(F, T)
    shifted = pos.copy()
    off = np.concatenate([[0], np.cumsum(sizes)])
    for k, a, b in zip(kinds, off[:-1], off[1:]):
        if k:
            shifted[a:b, 2] -= 2.5
    Uo, Wo = oracle.mdot(shifted, 0.9, 1.1, F=F, T=T, angular=angular, wall=kinds, sys_sizes=sizes)
    assert np.linalg.norm(U.ravel() - Uo) <= tol * np.linalg.norm(Uo)
    if angular:
        assert np.linalg.norm(W.ravel() - Wo) <= tol * np.linalg.norm(Wo)


def test_persistent_and_plain_launch_agree_bitwise():
    sizes, kinds, pos = _ensemble(seed=4)
    F = np.random.default_rng(2).normal(size=pos.shape)
    U1, _ = _solver(sizes, kinds, pos, persistent=1).Mdot(F)
    U0, _ = _solver(sizes, kinds, pos, persistent=0).Mdot(F)
    assert np.array_equal(U0, U1)


def test_ensemble_sqrt_and_drift():
    sizes, kinds, pos = _ensemble(seed=7, sizes=(40, 25, 60, 33))
    s = _solver(sizes, kinds, pos, tol=1e-8)
    M = oracle.dense_matrix(pos, 0.9, 1.1, wall=kinds, sys_sizes=sizes)
    W = np.random.default_rng(5).normal(size=M.shape[0])
    u, _ = s.sqrtMdotW_with_noise(W)
    exact = oracle.dense_sqrt(M) @ W
    ref, _ = oracle.LanczosOracle().run(lambda v: M @ v, W, 1e-8)
    err_ref = np.linalg.norm(ref - exact)
    assert np.linalg.norm(u.ravel() - exact) <= 3 * err_ref + 1e-7 * np.linalg.norm(exact)
    # per-system solves of the same noise give the same square root
    off = np.concatenate([[0], np.cumsum(sizes)])
    for k, a, b in zip(kinds, off[:-1], off[1:]):
        Ms = M[3 * a:3 * b, 3 * a:3 * b]
        ys = oracle.dense_sqrt(Ms) @ W[3 * a:3 * b]
        assert np.linalg.norm(u.ravel()[3 * a:3 * b] - ys) <= 1e-5 * np.linalg.norm(ys)
    # thermal drift: exactly zero on the open systems, not on the wall ones
    d, _ = s.divM()
    d = d.reshape(-1, 3)
    for k, a, b in zip(kinds, off[:-1], off[1:]):
        if k:
            assert np.any(d[a:b] != 0)
        else:
            assert np.all(d[a:b] == 0)


def _single(pos, kind, angular, tol):
    from libmobility_amd import NBody
    s = NBody("open", "open", "single_wall" if kind else "open")
    s.setParameters(wallHeight=0.0) if kind else s.setParameters()
    s.initialize(viscosity=0.9, hydrodynamicRadius=1.1, includeAngular=angular, tolerance=tol, seed=3)
    s.setPositions(pos)
    return s


def _split(v, off, N, angular):
    """Per-system vectors [forces_s, torques_s] of a global [forces, torques] vector."""
    v = np.asarray(v).ravel()
    out = []
    for a, b in zip(off[:-1], off[1:]):
        part = v[3 * a:3 * b]
        if angular:
            part = np.concatenate([part, v[3 * N + 3 * a:3 * N + 3 * b]])
        out.append(part)
    return out


@pytest.mark.parametrize("angular", [False, True])
def test_per_system_lanczos_matches_separate_solvers(angular):
    """One recurrence per system (LanczosAlgorithm.cu:171-232 on each diagonal
    block) against separate per-system solvers and the oracle's restatement,
    over three calls so the per-system check schedule and previous estimate
    carry over as in separate solvers."""
    sizes, kinds, pos = _ensemble(seed=11, sizes=(40, 25, 60, 33, 7))
    N = pos.shape[0]
    off = np.concatenate([[0], np.cumsum(sizes)])
    tol = 1e-6
    ens = _solver(sizes, kinds, pos, angular=angular, tol=tol)
    ens.setLanczosPerSystem(True)
    singles = [_single(pos[a:b], k, angular, tol) for a, b, k in zip(off[:-1], off[1:], kinds)]
    oracles = [oracle.LanczosOracle() for _ in sizes]
    Ms = [oracle.dense_matrix(pos[a:b], 0.9, 1.1, angular=angular, wall=bool(k))
          for a, b, k in zip(off[:-1], off[1:], kinds)]
    n = (6 if angular else 3) * N
    rng = np.random.default_rng(4)
    for call in range(3):
        W = rng.normal(size=n)
        u, w = ens.sqrtMdotW_with_noise(W)
        got = _split(np.concatenate([u.ravel(), w.ravel()]) if angular else u, off, N, angular)
        its = ens.lastLanczosIterationsPerSystem()
        assert its.shape == (len(sizes),)
        for s, (Ws, g) in enumerate(zip(_split(W, off, N, angular), got)):
            us, ws = singles[s].sqrtMdotW_with_noise(Ws)
            one = np.concatenate([us.ravel(), ws.ravel()]) if angular else us.ravel()
            ref, it_ref = oracles[s].run(lambda v, M=Ms[s]: M @ v, Ws, tol)
            # the stopping rule reacts to rounding (see test_gpu_lanczos.py)
            assert abs(its[s] - singles[s].lastLanczosIterations()) <= 6, (call, s)
            assert abs(its[s] - it_ref) <= 6, (call, s)
            assert np.linalg.norm(g - one) <= 10 * tol * np.linalg.norm(one), (call, s)
            exact = oracle.dense_sqrt(Ms[s]) @ Ws
            err_ref = np.linalg.norm(ref - exact)
            assert np.linalg.norm(g - exact) <= 3 * err_ref + 5 * tol * np.linalg.norm(exact), (call, s)
    assert ens.lastLanczosIterations() == max(its)


def _rsa_ensemble(seed, sizes, a=1.1, phi=0.15):
    """Non-overlapping spheres (volume fraction phi) per system; wall systems in
    a slab 1.2 <= z < 9.2 above the wall at z = 0.  Overlapping random clouds
    make M so ill-conditioned that the Lanczos stopping step reacts to
    rounding; these systems converge in 30-45 steps whatever the rounding."""
    from libmobility_amd.workloads import rsa_box, rsa_box_side
    kinds = [s % 2 for s in range(len(sizes))]
    pos = []
    for k, (sz, kind) in enumerate(zip(sizes, kinds)):
        if kind:
            lxy = np.sqrt(sz * 4 * np.pi * a ** 3 / 3 / phi / 8.0)
            pos.append(rsa_box(sz, (lxy, lxy, 8.0), seed + k, a=a, origin=(0.0, 0.0, 1.2)))
        else:
            L = rsa_box_side(sz, phi, a)
            pos.append(rsa_box(sz, (L, L, L), seed + k, a=a))
    return list(sizes), kinds, np.concatenate(pos)


@pytest.mark.parametrize("angular,persistent", [(False, 1), (False, 0), (True, 1)])
def test_per_system_lanczos_large_systems_match_separate_solvers(angular, persistent):
    """Systems of several source chunks: once a system converges the pair work
    list is rebuilt for the systems still iterating (finer tiles, fewer
    chunks) and the chunk sum covers only their particles; every system still
    equals its own solver (LanczosAlgorithm.cu:171-232)."""
    sizes, kinds, pos = _rsa_ensemble(seed=21, sizes=(2100, 40, 900, 1300, 5))
    N = pos.shape[0]
    off = np.concatenate([[0], np.cumsum(sizes)])
    tol = 1e-6
    ens = _solver(sizes, kinds, pos, angular=angular, tol=tol, persistent=persistent)
    ens.setLanczosPerSystem(True)
    twin = _solver(sizes, kinds, pos, angular=angular, tol=tol, persistent=persistent)
    twin.setLanczosPerSystem(True)
    singles = [_single(pos[a:b], k, angular, tol) for a, b, k in zip(off[:-1], off[1:], kinds)]
    n = (6 if angular else 3) * N
    rng = np.random.default_rng(5)
    for call in range(2):
        W = rng.normal(size=n)
        u, w = ens.sqrtMdotW_with_noise(W)
        u2, w2 = twin.sqrtMdotW_with_noise(W)
        assert np.array_equal(u, u2)  # the rebuilt work lists are deterministic
        got = _split(np.concatenate([u.ravel(), w.ravel()]) if angular else u, off, N, angular)
        its = ens.lastLanczosIterationsPerSystem()
        assert len(set(its.tolist())) > 1  # systems left the recurrence at different steps
        for s, (Ws, g) in enumerate(zip(_split(W, off, N, angular), got)):
            us, ws = singles[s].sqrtMdotW_with_noise(Ws)
            one = np.concatenate([us.ravel(), ws.ravel()]) if angular else us.ravel()
            assert abs(its[s] - singles[s].lastLanczosIterations()) <= 6, (call, s)
            assert np.linalg.norm(g - one) <= 10 * tol * np.linalg.norm(one), (call, s)


def test_per_system_lanczos_with_preconditioner_is_a_square_root():
    from libmobility_amd.precond import LearnedPreconditioner
    sizes, kinds, pos = _ensemble(seed=12, sizes=(20, 14, 9))
    s = _solver(sizes, kinds, pos, tol=1e-9)
    s.setLanczosPerSystem(True)
    s.setPreconditioner(LearnedPreconditioner(seed=9))
    n = 3 * pos.shape[0]
    B = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        B[:, k] = s.sqrtMdotW_with_noise(e)[0].ravel()
    M = oracle.dense_matrix(pos, 0.9, 1.1, wall=kinds, sys_sizes=sizes)
    assert np.linalg.norm(B @ B.T - M) <= 1e-5 * np.linalg.norm(M)


def test_per_system_lanczos_limit_and_callback():
    sizes, kinds, pos = _ensemble(seed=13, sizes=(30, 12))
    from libmobility_amd import NBody
    seen = []
    s = NBody("open", "open", "single_wall")
    s.setParameters(wallHeight=0.0)
    s.initialize(viscosity=0.9, hydrodynamicRadius=1.1, tolerance=-1.0,
                 lanczos_callback=lambda i, e: seen.append((i, e)))
    s.setSystems(sizes, ["single_wall" if k else "open" for k in kinds])
    s.setLanczosPerSystem(True)
    s.setLanczosIterationLimit(40)
    s.setPositions(pos)
    with pytest.raises(RuntimeError, match="Could not converge"):
        s.sqrtMdotW_with_noise(np.random.default_rng(0).normal(size=3 * pos.shape[0]))
    assert [i for i, _ in seen] == list(range(1, 40))
    assert all(np.isfinite(e) for _, e in seen)
