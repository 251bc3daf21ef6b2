"""Several vectors per pass over the pairs (extension): the multi-vector
product equals separate NBody::Mdot calls (mobility.h:166-182) and the oracle;
the multi-vector square root equals separate lanczos::Solver runs
(LanczosAlgorithm.cu:171-232) on each vector, draws the noise successive
sqrtMdotW calls would draw, and with the learned preconditioner is a square
root of M."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _solver(wall, precision="fp64", tol=1e-6, seed=5, sizes=None, kinds=None, per_system=False):
    from libmobility_amd import NBody
    s = NBody("open", "open", "single_wall" if wall else "open")
    s.setParameters(wallHeight=0.0) if wall else s.setParameters()
    s.setPairPrecision(precision)
    s.initialize(viscosity=0.8, hydrodynamicRadius=1.1, tolerance=tol, seed=seed)
    if sizes is not None:
        s.setSystems(sizes, kinds)
        s.setLanczosPerSystem(per_system)
    return s


def _cloud(N, wall, seed):
    rng = np.random.default_rng(seed)
    L = (N * 4 * np.pi / 3 / 0.15) ** (1 / 3)
    p = rng.uniform(0, L, (N, 3))
    if wall:
        p[:, 2] = rng.uniform(0.5, 8.0, N)  # includes z < a
    return p


CASES = [("open", False, None), ("wall", True, None), ("mixed", True, [40, 25, 60])]


@pytest.mark.parametrize("name,wall,sizes", CASES)
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32", 3e-6)])
@pytest.mark.parametrize("nrhs", [1, 2, 3, 4, 5, 7])
def test_mdot_multi_equals_single_products(name, wall, sizes, precision, tol, nrhs):
    N = 125 if sizes is None else sum(sizes)
    pos = _cloud(N, wall, 1)
    kinds = None if sizes is None else [k % 2 for k in range(len(sizes))]
    s = _solver(wall, precision, sizes=sizes, kinds=kinds)
    s.setPositions(pos)
    F = np.random.default_rng(2).normal(size=(nrhs, N, 3))
    U = s.MdotMulti(F)
    assert U.shape == F.shape
    for r in range(nrhs):
        u1, _ = s.Mdot(F[r])
        assert np.linalg.norm(U[r] - u1) <= tol * np.linalg.norm(u1), r
        wk = kinds if kinds is not None else wall
        uo, _ = oracle.mdot(pos, 0.8, 1.1, F=F[r], wall=wk, sys_sizes=sizes)
        assert np.linalg.norm(U[r].ravel() - uo) <= max(tol, 1e-12) * np.linalg.norm(uo), r


@pytest.mark.parametrize("wall", [False, True])
def test_mdot_multi_with_source_chunks(wall):
    """Large enough for several source chunks (partial slabs per vector)."""
    N = 3000
    pos = _cloud(N, wall, 3)
    s = _solver(wall)
    s.setPositions(pos)
    F = np.random.default_rng(4).normal(size=(6, 3 * N))
    U = s.MdotMulti(F)
    for r in range(6):
        u1, _ = s.Mdot(F[r])
        assert np.linalg.norm(U[r] - u1.ravel()) <= 1e-13 * np.linalg.norm(u1)


@pytest.mark.parametrize("wall", [False, True])
def test_sqrt_multi_equals_separate_solves(wall):
    N = 80
    pos = _cloud(N, wall, 5)
    tol = 1e-7
    s = _solver(wall, tol=tol)
    s.setPositions(pos)
    nrhs = 5
    W = np.random.default_rng(6).normal(size=(nrhs, 3 * N))
    Y = s.sqrtMdotWMulti_with_noise(W)
    its = s.lastLanczosIterationsMulti(nrhs)
    M = oracle.dense_matrix(pos, 0.8, 1.1, wall=wall)
    B = oracle.dense_sqrt(M)
    for r in range(nrhs):
        single = _solver(wall, tol=tol)
        single.setPositions(pos)
        y1, _ = single.sqrtMdotW_with_noise(W[r])
        ref, it_ref = oracle.LanczosOracle().run(lambda v: M @ v, W[r], tol)
        assert abs(its[r] - single.lastLanczosIterations()) <= 6
        assert abs(its[r] - it_ref) <= 6
        assert np.linalg.norm(Y[r] - y1.ravel()) <= 10 * tol * np.linalg.norm(y1)
        exact = B @ W[r]
        err_ref = np.linalg.norm(ref - exact)
        assert np.linalg.norm(Y[r] - exact) <= 3 * err_ref + 5 * tol * np.linalg.norm(exact)


def test_sqrt_multi_draws_the_noise_of_successive_calls():
    N = 60
    pos = _cloud(N, False, 7)
    a = _solver(False, tol=1e-9, seed=2024)
    b = _solver(False, tol=1e-9, seed=2024)
    a.setPositions(pos)
    b.setPositions(pos)
    Y = a.sqrtMdotWMulti(3)
    assert Y.shape == (3, N, 3)
    for r in range(3):
        y, _ = b.sqrtMdotW()
        assert np.linalg.norm(Y[r] - y) <= 1e-6 * np.linalg.norm(y)


def test_sqrt_multi_per_system_ensemble():
    sizes = [30, 12, 45]
    kinds = [0, 1, 0]
    N = sum(sizes)
    pos = _cloud(N, True, 8)
    tol = 1e-7
    s = _solver(True, tol=tol, sizes=sizes, kinds=kinds, per_system=True)
    s.setPositions(pos)
    W = np.random.default_rng(9).normal(size=(3, 3 * N))
    Y = s.sqrtMdotWMulti_with_noise(W)
    off = np.concatenate([[0], np.cumsum(sizes)])
    for r in range(3):
        for k, a, b in zip(kinds, off[:-1], off[1:]):
            Ms = oracle.dense_matrix(pos[a:b], 0.8, 1.1, wall=bool(k))
            ex = oracle.dense_sqrt(Ms) @ W[r, 3 * a:3 * b]
            assert np.linalg.norm(Y[r, 3 * a:3 * b] - ex) <= 1e-5 * np.linalg.norm(ex)


def test_sqrt_multi_with_preconditioner_is_a_square_root():
    from libmobility_amd.precond import LearnedPreconditioner
    N = 20
    pos = _cloud(N, False, 10)
    s = _solver(False, tol=1e-9)
    s.setPositions(pos)
    s.setPreconditioner(LearnedPreconditioner(seed=4))
    n = 3 * N
    B = s.sqrtMdotWMulti_with_noise(np.eye(n)).T  # column k = B e_k
    M = oracle.dense_matrix(pos, 0.8, 1.1)
    assert np.linalg.norm(B @ B.T - M) <= 1e-5 * np.linalg.norm(M)


def test_multi_refuses_angular_solvers_and_bad_shapes():
    from libmobility_amd import NBody
    s = NBody("open", "open", "open")
    s.setParameters()
    s.initialize(viscosity=1.0, hydrodynamicRadius=1.0, includeAngular=True)
    s.setPositions(_cloud(5, False, 11))
    with pytest.raises(RuntimeError, match="translational"):
        s.MdotMulti(np.ones((2, 15)))
    t = _solver(False)
    t.setPositions(_cloud(5, False, 11))
    with pytest.raises(RuntimeError):
        t.MdotMulti(np.ones(15))  # a single vector is not a stack
    with pytest.raises(RuntimeError):
        t.MdotMulti(np.ones((2, 14)))
