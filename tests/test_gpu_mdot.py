"""GPU parity of Mdot against the CPU oracle and the reference's golden fixtures.

Mirrors tests/test_mobility_matrix.py and tests/test_wall_mobility.py of the
reference (NBody rows), run through the C ABI (libmobility_amd.NBody).
Tolerances: FP64 pair mode 1e-12 relative to the largest entry (golden wall
fixtures: 100 eps as in test_wall_mobility.py:35); FP32 pair mode (FP32 pair
terms, FP64 accumulation) 1e-6 relative in norm.
"""
import os

import numpy as np
import pytest

import oracle
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

ALGOS = ["naive", "block", "fast", "advise"]


def _solver(wall, a, eta=1.0, angular=False, algorithm="advise", precision="fp64",
            wallHeight=0.0, **kw):
    from libmobility_amd import NBody
    s = NBody("open", "open", "single_wall" if wall else "open")
    if wall:
        s.setParameters(algorithm=algorithm, wallHeight=wallHeight, **kw)
    else:
        s.setParameters(algorithm=algorithm, **kw)
    s.setPairPrecision(precision)
    s.initialize(viscosity=eta, hydrodynamicRadius=a, includeAngular=angular)
    return s


def compute_M(solver, n, angular):
    """Dense matrix column by column, tests/utils.py:62-84."""
    size = 6 * n if angular else 3 * n
    M = np.zeros((size, size))
    I = np.identity(size)
    for i in range(size):
        f = I[: 3 * n, i].copy()
        if angular:
            t = I[3 * n:, i].copy()
            u, w = solver.Mdot(f, t)
            M[:, i] = np.concatenate((u.reshape(3 * n), w.reshape(3 * n)))
        else:
            M[:, i] = solver.Mdot(f)[0].reshape(3 * n)
    return M


@pytest.mark.parametrize("algorithm", ALGOS)
def test_self_mobility_closed_form(algorithm):
    # test_mobility_matrix.py:97-149
    a, eta = 0.9123, 1.123
    s = _solver(False, a, eta, angular=True, algorithm=algorithm)
    s.setPositions(np.zeros((1, 3)))
    f = np.ones(3)
    t = 2 * np.ones(3)
    u, w = s.Mdot(f, t)
    m0 = 1 / (6 * np.pi * eta * a)
    t0 = 1 / (8 * np.pi * eta * a ** 3)
    assert np.allclose(u, m0 * f, rtol=0, atol=1e-13)
    assert np.allclose(w, t0 * t, rtol=0, atol=1e-13)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("algorithm", ALGOS)
def test_pair_freespace_golden(algorithm, precision):
    # test_mobility_matrix.py:186-214 + pair_mobility_nbody_freespace.npz
    ref = np.load(os.path.join(GOLDEN, "pair_mobility_nbody_freespace.npz"))
    s = None
    tol = 1e-12 if precision == "fp64" else 2e-6
    for i in range(len(ref["r_vecs"])):
        a = ref["a"].ravel()[i]
        eta = ref["eta"].ravel()[i]
        s = _solver(False, a, eta, angular=True, algorithm=algorithm, precision=precision)
        s.setPositions(ref["r_vecs"][i])
        M = compute_M(s, 2, True)
        scale = np.abs(ref["M"][i]).max()
        assert np.abs(M - ref["M"][i]).max() <= tol * scale


def _wall_norm(nP):
    eta = 1 / 4 / np.sqrt(np.pi)
    n = 3 * nP
    norm = np.zeros((2 * n, 2 * n))
    norm[:n, :n] = 1 / (6 * np.pi * eta)
    norm[n:, n:] = 1 / (8 * np.pi * eta)
    norm[n:, :n] = 1 / (6 * np.pi * eta)
    norm[:n, n:] = 1 / (6 * np.pi * eta)
    return eta, norm


@pytest.mark.parametrize("wallHeight", [0, 5.4, -10])
def test_wall_self_golden(wallHeight):
    # test_wall_mobility.py:138-197 with self_mobility_bw_ref_noimg.npz
    ref = np.load(os.path.join(GOLDEN, "self_mobility_bw_ref_noimg.npz"))
    eta, norm = _wall_norm(1)
    s = _solver(True, 1.0, eta, angular=True, wallHeight=wallHeight)
    tol = 100 * np.finfo(np.float64).eps
    for k, h in enumerate(ref["heights"].ravel()):
        s.setPositions(np.array([[38.4, 38.4, h + wallHeight]]))
        M = compute_M(s, 1, True) / norm
        assert np.allclose(M, ref["M"][k], atol=tol, rtol=tol)


@pytest.mark.parametrize("offset", ["x", "y"])
@pytest.mark.parametrize("wallHeight", [0, 5.4, -10])
def test_wall_pair_golden(offset, wallHeight):
    # test_wall_mobility.py:200-276 with pair_mobility_bw_ref_noimg_offset_{x,y}.npz
    ref = np.load(os.path.join(GOLDEN, f"pair_mobility_bw_ref_noimg_offset_{offset}.npz"))
    eta, norm = _wall_norm(2)
    s = _solver(True, 1.0, eta, angular=True, wallHeight=wallHeight)
    tol = 100 * np.finfo(np.float64).eps
    ov = np.array([1.0, 0, 0]) if offset == "x" else np.array([0, 1.0, 0])
    for i, sep in enumerate([3.0, 4.0, 8.0]):
        for k, h in enumerate(np.asarray(ref["heights"]).ravel()):
            c = np.array([38.4, 38.4, h + wallHeight])
            s.setPositions(np.array([c + sep * ov / 2, c - sep * ov / 2]))
            M = compute_M(s, 2, True) / norm
            assert np.allclose(M, ref["M"][i, k], atol=tol, rtol=tol)


def _c1_inputs(N=1024, seed=0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, 24.0, (N, 3))
    F = rng.normal(size=(N, 3))
    return pos, F


@pytest.mark.parametrize("algorithm", ALGOS)
def test_open_mdot_c1_fp64(algorithm):
    """N=1024 open RPY, cube 24a (overlaps kept), eta = 1/(6 pi): rel. error <= 1e-12."""
    pos, F = _c1_inputs()
    eta = 1 / (6 * np.pi)
    s = _solver(False, 1.0, eta, algorithm=algorithm)
    s.setPositions(pos)
    U, W = s.Mdot(F)
    assert W is None and U.shape == F.shape
    M = oracle.dense_matrix(pos, eta, 1.0)
    ref = M @ F.ravel()
    err = np.linalg.norm(U.ravel() - ref) / np.linalg.norm(ref)
    assert err <= 1e-12, err
    assert np.abs(U.ravel() - ref).max() <= 1e-12 * np.abs(ref).max()


def test_open_mdot_c1_fp32():
    pos, F = _c1_inputs()
    eta = 1 / (6 * np.pi)
    s = _solver(False, 1.0, eta, precision="fp32")
    s.setPositions(pos)
    U, _ = s.Mdot(F)
    ref = oracle.dense_matrix(pos, eta, 1.0) @ F.ravel()
    err = np.linalg.norm(U.ravel() - ref) / np.linalg.norm(ref)
    assert err <= 1e-6, err


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
@pytest.mark.parametrize("wall", [False, True])
@pytest.mark.parametrize("which", ["F", "T", "FT"])
def test_mdot_angular_vs_oracle(precision, tol, wall, which):
    rng = np.random.default_rng(11)
    N = 700
    a = 1.3
    pos = rng.uniform(0, 20.0, (N, 3))
    if wall:
        pos[:, 2] = rng.uniform(0.3 * a, 8.0, N)  # includes damped heights z < a
    F = rng.normal(size=(N, 3)) if "F" in which else None
    T = rng.normal(size=(N, 3)) if "T" in which else None
    s = _solver(wall, a, 0.7, angular=True, precision=precision)
    s.setPositions(pos)
    U, W = s.Mdot(F, T)
    Uo, Wo = oracle.mdot(pos, 0.7, a, F=F, T=T, angular=True, wall=wall)
    for got, ref in ((U, Uo), (W, Wo)):
        err = np.linalg.norm(got.ravel() - ref) / np.linalg.norm(ref)
        assert err <= tol, err


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
@pytest.mark.parametrize("sym", [0, 1])
def test_mdot_wall_far_above_and_spread_out(precision, tol, sym):
    """Heights up to 300a and separations up to 400a: the image terms are then
    small against the open part (f1..f5 of the merged wall coefficients in
    their Horner form), in the ordered and the symmetric evaluation."""
    rng = np.random.default_rng(17)
    N = 600
    a = 1.0
    pos = rng.uniform(0, 400.0, (N, 3))
    pos[:, 2] = np.exp(rng.uniform(np.log(0.5), np.log(300.0), N))
    F = rng.normal(size=(N, 3))
    T = rng.normal(size=(N, 3))
    s = _solver(True, a, 0.7, angular=True, precision=precision)
    s.setPairSymmetric(sym)
    s.setPositions(pos)
    U, W = s.Mdot(F, T)
    assert s.lastPairSymmetric() == bool(sym)
    Uo, Wo = oracle.mdot(pos, 0.7, a, F=F, T=T, angular=True, wall=True)
    for got, ref in ((U, Uo), (W, Wo)):
        err = np.linalg.norm(got.ravel() - ref) / np.linalg.norm(ref)
        assert err <= tol, err


def test_batched_equals_per_system():
    rng = np.random.default_rng(5)
    nb, npb = 6, 97
    pos = rng.uniform(0, 10, (nb * npb, 3))
    F = rng.normal(size=(nb * npb, 3))
    s = _solver(False, 1.0, 1.0, Nbatch=nb, NperBatch=npb)
    s.setPositions(pos)
    U, _ = s.Mdot(F)
    Uo, _ = oracle.mdot(pos, 1.0, 1.0, F=F, sys_sizes=[npb] * nb)
    assert np.abs(U.ravel() - Uo).max() <= 1e-12 * np.abs(Uo).max()
    for b in range(nb):
        sl = slice(b * npb, (b + 1) * npb)
        one = _solver(False, 1.0, 1.0)
        one.setPositions(pos[sl])
        Ub, _ = one.Mdot(F[sl])
        assert np.abs(U[sl] - Ub).max() <= 1e-12 * np.abs(Ub).max()


def test_ragged_systems_mixed_sizes():
    rng = np.random.default_rng(9)
    sizes = [1, 64, 300, 5, 1030]
    N = sum(sizes)
    pos = rng.uniform(0, 12, (N, 3))
    pos[:, 2] += 1.0
    F = rng.normal(size=(N, 3))
    T = rng.normal(size=(N, 3))
    s = _solver(True, 1.0, 1.0, angular=True)
    s.setSystems(sizes)
    s.setPositions(pos)
    U, W = s.Mdot(F, T)
    Uo, Wo = oracle.mdot(pos, 1.0, 1.0, F=F, T=T, angular=True, wall=True, sys_sizes=sizes)
    assert np.abs(U.ravel() - Uo).max() <= 1e-12 * np.abs(Uo).max()
    assert np.abs(W.ravel() - Wo).max() <= 1e-12 * np.abs(Wo).max()


@pytest.mark.parametrize("wall", [False, True])
@pytest.mark.parametrize("n", [1, 2, 3, 10])
def test_symmetry_and_spd(wall, n):
    # test_mobility_matrix.py:8-64 and test_fluctuation_dissipation.py:115-138
    rng = np.random.default_rng(n)
    for a in (1.0, 0.95, 1.12):
        s = _solver(wall, a, 1.0, angular=True)
        pos = rng.random((n, 3)) * 10 - np.array([5, 5, 0])
        s.setPositions(pos)
        M = compute_M(s, n, True)
        assert np.abs(M - M.T).max() <= 1e-12 * np.abs(M).max()
        assert np.all(np.linalg.eigvalsh(0.5 * (M + M.T)) > 0)
        Mo = oracle.dense_matrix(pos, 1.0, a, angular=True, wall=wall)
        assert np.abs(M - Mo).max() <= 1e-12 * np.abs(Mo).max()


def test_wall_height_shift_and_below_wall_error():
    # test_interface.py:421-430
    from libmobility_amd import NBody
    s = NBody("open", "open", "single_wall")
    s.setParameters(wallHeight=1.0)
    s.initialize(viscosity=1.0, hydrodynamicRadius=1.0)
    with pytest.raises(RuntimeError, match="fallen below the wall"):
        s.setPositions(np.array([1.0, 2.0, 0.5]))
