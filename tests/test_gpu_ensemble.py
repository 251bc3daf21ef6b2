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
