"""GPU checks of the coarse level of the sqrtMdotW preconditioner (no
reference counterpart; the oracle is oracle/coarse.py, pinned in
tests/test_oracle_coarse.py).

  * the cells of a build are the oracle's cells for the same positions;
  * L L^T equals the oracle's Q^T M Q (FP32 pair terms: 1e-5 relative) and
    L^{-1} L = I to FP64 rounding, also for a block large enough for the
    two-by-two-tile trailing updates;
  * the result is a square root of M: ||B B^T - M|| / ||M|| <= 1e-5 from
    dense columns, open / wall / wall with torques, with and without the
    learned blocks;
  * sqrtMdotW equals the oracle's two-level Lanczos run with the build's L
    on the same W;
  * fewer Lanczos iterations than plain Lanczos on an RSA packing;
  * the multi-vector square root agrees with the single-vector one;
  * ensembles: per-system cells and factors (the one-workgroup batched path
    and the per-system launches), B B^T = M over a mixed open / wall
    ensemble, per-system square roots against the oracle;
  * errors: negative cell targets, queries before a build or of a missing
    system.
"""
import numpy as np
import pytest

import oracle
from oracle import coarse

pytestmark = pytest.mark.gpu

ETA = 1.0 / (6 * np.pi)


def _solver(pos, wall=False, angular=False, tol=1e-6, ppc=8):
    from libmobility_amd import NBody
    s = NBody("open", "open", "single_wall" if wall else "open")
    if wall:
        s.setParameters(wallHeight=0.0)
    else:
        s.setParameters()
    s.initialize(viscosity=ETA, hydrodynamicRadius=1.0, includeAngular=angular, tolerance=tol)
    s.setPositions(pos)
    s.setCoarseLevel(ppc)
    return s


def _wall_positions(N, seed):
    rng = np.random.default_rng(seed)
    L = (N / 0.05 / 8.0) ** 0.5
    return np.column_stack([rng.uniform(0, L, N), rng.uniform(0, L, N), rng.uniform(1.5, 9.5, N)])


def _records_xyz(pos, wall):
    """Positions as the solver's records hold them (a = 1): wall heights are
    max(z, a)."""
    xyz = np.array(pos, dtype=np.float64).reshape(-1, 3).copy()
    if wall:
        xyz[:, 2] = np.maximum(xyz[:, 2], 1.0)
    return xyz


@pytest.mark.parametrize("wall", [False, True])
@pytest.mark.parametrize("ppc", [24, 100])
def test_cells_and_factors_match_the_oracle(wall, ppc):
    """ppc 100: cells of more than 64 particles, which the coarse matrix
    kernel takes in several 64 x 64 chunks."""
    from libmobility_amd.workloads import rsa_cube
    N = 1500
    pos = _wall_positions(N, 3) if wall else rsa_cube(N, 0.2, seed=3)[0]
    s = _solver(pos, wall=wall, ppc=ppc)
    s.sqrtMdotW_with_noise(np.random.default_rng(0).normal(size=3 * N))
    cell_of, grid, C = s.coarseCells()
    ref, g, Cr = coarse.coarse_cells(_records_xyz(pos, wall), ppc)
    assert C == Cr and list(grid) == list(g)
    assert np.array_equal(cell_of, ref)
    L, X = s.coarseFactors()
    M = oracle.dense_matrix(pos, ETA, 1.0, wall=wall) / (1.0 / (6 * np.pi * ETA))
    Mc = coarse.coarse_matrix(M, cell_of, C)
    assert np.allclose(L, np.tril(L))
    rel = np.linalg.norm(L @ L.T - Mc) / np.linalg.norm(Mc)
    if ppc == 100:
        assert np.bincount(cell_of).max() > 64
    print(f"[coarse] wall={wall} ppc={ppc} C={C} |LL^T - Mc|/|Mc| = {rel:.2e}")
    assert rel <= 1e-5
    assert np.abs(X @ L - np.eye(3 * C)).max() <= 1e-11
    assert s.coarseBuildMs() > 0.0


@pytest.mark.parametrize("case", ["open", "open+net", "wall", "wall+angular", "wall+angular+net"])
def test_two_level_result_is_a_square_root(case):
    from libmobility_amd.precond import LearnedPreconditioner
    from libmobility_amd.workloads import rsa_cube
    wall = case.startswith("wall")
    angular = "angular" in case
    N = 36
    pos = _wall_positions(N, 5) if wall else rsa_cube(N, 0.2, seed=4)[0]
    s = _solver(pos, wall=wall, angular=angular, tol=1e-8, ppc=4)
    if "net" in case:
        net = LearnedPreconditioner(seed=9)
        s.setPreconditioner(net)
    n = (6 if angular else 3) * N
    B = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        u, w = s.sqrtMdotW_with_noise(e)
        B[:, k] = np.concatenate([u.ravel(), w.ravel()]) if angular else u.ravel()
    M = oracle.dense_matrix(pos, ETA, 1.0, angular=angular, wall=wall)
    rel = np.linalg.norm(B @ B.T - M) / np.linalg.norm(M)
    print(f"[coarse] {case}: cells {s.coarseCells()[2]}, |BB^T - M|/|M| = {rel:.2e}")
    assert rel <= 1e-5


def test_sqrt_matches_the_oracle_two_level_lanczos():
    from libmobility_amd.workloads import rsa_cube
    N = 600
    pos, _ = rsa_cube(N, 0.2, seed=8)
    s = _solver(pos, tol=1e-8, ppc=16)
    W = np.random.default_rng(5).normal(size=3 * N)
    u, _ = s.sqrtMdotW_with_noise(W)
    cell_of, _, C = s.coarseCells()
    L, X = s.coarseFactors()
    M = oracle.dense_matrix(pos, ETA, 1.0)
    ref, it = coarse.TwoLevel(coarse.coarse_Q(cell_of, C), L, X).sqrt(M, W, 1e-8)
    rel = np.linalg.norm(u.ravel() - ref) / np.linalg.norm(ref)
    print(f"[coarse] N={N}: GPU {s.lastLanczosIterations()} / oracle {it} iterations, rel {rel:.2e}")
    assert rel <= 1e-6
    assert abs(s.lastLanczosIterations() - it) <= 2


def test_coarse_level_lowers_iterations():
    from libmobility_amd.workloads import rsa_cube
    N = 8192
    pos, _ = rsa_cube(N, 0.2, seed=12)
    W = np.random.default_rng(6).normal(size=3 * N)
    s = _solver(pos, ppc=0)
    s.setPairPrecision("fp32")
    s.sqrtMdotW_with_noise(W)
    plain = s.lastLanczosIterations()
    s.setCoarseLevel(32)
    s.sqrtMdotW_with_noise(W)
    two = s.lastLanczosIterations()
    print(f"[coarse] N={N}: plain {plain}, coarse(32) {two} iterations")
    assert two <= 0.6 * plain


def test_multi_vector_sqrt_with_coarse_level():
    from libmobility_amd.workloads import rsa_cube
    N = 900
    pos, _ = rsa_cube(N, 0.2, seed=13)
    s = _solver(pos, tol=1e-8, ppc=16)
    rng = np.random.default_rng(7)
    Ws = rng.normal(size=(3, 3 * N))
    U = s.sqrtMdotWMulti_with_noise(Ws)
    for r in range(3):
        u, _ = s.sqrtMdotW_with_noise(Ws[r])
        assert np.linalg.norm(U[r].ravel() - u.ravel()) <= 1e-6 * np.linalg.norm(u)


def test_coarse_level_errors():
    from libmobility_amd.workloads import rsa_cube
    pos, _ = rsa_cube(200, 0.2, seed=14)
    s = _solver(pos, ppc=0)
    with pytest.raises(RuntimeError, match="particles per cell"):
        s.setCoarseLevel(-1)
    with pytest.raises(RuntimeError, match="no coarse level"):
        s.coarseCells()
    with pytest.raises(RuntimeError, match="no such system"):
        s.setCoarseLevel(8)
        s.sqrtMdotW()
        s.coarseFactors(5)


@pytest.mark.parametrize("case", ["one", "plane", "fewer_than_a_cell", "wall_below_a"])
def test_coarse_level_edge_cases(case):
    """Degenerate cell grids: one particle (one cell), coplanar particles (an
    axis of zero extent), fewer particles than one cell target, and wall
    particles below z = a (cells use the clamped heights).  The result is a
    square root of M (dense columns)."""
    rng = np.random.default_rng(21)
    wall = case == "wall_below_a"
    if case == "one":
        pos = np.array([[0.3, -1.2, 2.0]])
    elif case == "plane":
        pos = np.column_stack([rng.uniform(0, 12, 20), rng.uniform(0, 9, 20), np.full(20, 4.0)])
    elif case == "fewer_than_a_cell":
        pos = rng.uniform(0, 10, size=(12, 3))
    else:
        pos = np.column_stack([rng.uniform(0, 10, 16), rng.uniform(0, 10, 16),
                               rng.uniform(0.2, 3.0, 16)])
    N = pos.shape[0]
    s = _solver(pos, wall=wall, tol=1e-10, ppc=6 if case != "fewer_than_a_cell" else 64)
    n = 3 * N
    B = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        B[:, k] = s.sqrtMdotW_with_noise(e)[0].ravel()
    cell_of, grid, C = s.coarseCells()
    ref, g, Cr = coarse.coarse_cells(_records_xyz(pos, wall), 6 if case != "fewer_than_a_cell" else 64)
    assert np.array_equal(cell_of, ref) and C == Cr
    M = oracle.dense_matrix(pos, ETA, 1.0, wall=wall)
    rel = np.linalg.norm(B @ B.T - M) / np.linalg.norm(M)
    print(f"[coarse] {case}: N={N} cells {C} grid {list(grid)} |BB^T - M|/|M| = {rel:.2e}")
    assert rel <= 1e-7


def test_refresh_interval_keeps_an_exact_square_root():
    """With a refresh interval, the preconditioner built for earlier positions
    is kept: the result is still a square root of the current M (dense
    columns), the cells are those of the earlier positions, and the build
    comes back after `interval` setPositions calls or when N changes."""
    from libmobility_amd.precond import LearnedPreconditioner
    from libmobility_amd.workloads import rsa_cube
    N = 36
    pos0, _ = rsa_cube(N, 0.2, seed=31)
    pos1 = pos0[::-1] * np.array([1.0, 0.7, 1.3]) + 2.0  # other cells for the same particles
    s = _solver(pos0, tol=1e-8, ppc=4)
    net = LearnedPreconditioner(seed=9)
    s.setPreconditioner(net)
    s.setPreconditionerRefresh(3)
    with pytest.raises(RuntimeError, match="refresh interval"):
        s.setPreconditionerRefresh(0)
    s.sqrtMdotW_with_noise(np.ones(3 * N))
    assert s.coarseBuilds() == 1
    cells0 = s.coarseCells()[0].copy()
    s.setPositions(pos1)
    n = 3 * N
    B = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        B[:, k] = s.sqrtMdotW_with_noise(e)[0].ravel()
    assert s.coarseBuilds() == 1
    assert np.array_equal(s.coarseCells()[0], cells0)
    M1 = oracle.dense_matrix(pos1, ETA, 1.0)
    rel = np.linalg.norm(B @ B.T - M1) / np.linalg.norm(M1)
    print(f"[coarse] stale preconditioner: |BB^T - M|/|M| = {rel:.2e}")
    assert rel <= 1e-5
    ref1, _, _ = coarse.coarse_cells(pos1, 4)
    assert not np.array_equal(ref1, cells0)
    s.setPositions(pos1)
    s.sqrtMdotW_with_noise(np.ones(3 * N))
    assert s.coarseBuilds() == 1
    s.setPositions(pos1)
    s.sqrtMdotW_with_noise(np.ones(3 * N))
    assert s.coarseBuilds() == 2
    assert np.array_equal(s.coarseCells()[0], ref1)
    pos2, _ = rsa_cube(N + 4, 0.2, seed=32)
    s.setPositions(pos2)
    s.sqrtMdotW_with_noise(np.ones(3 * (N + 4)))
    assert s.coarseBuilds() == 3


def test_two_level_fluctuation_dissipation():
    """Samples of the two-level square root (learned blocks + coarse level,
    fresh noise each call) have covariance M: whitened by the eigenvectors of
    the oracle's dense M, the sample covariance is the identity within its
    sampling error."""
    from libmobility_amd.precond import LearnedPreconditioner
    from libmobility_amd.workloads import rsa_cube
    N = 20
    pos, _ = rsa_cube(N, 0.2, seed=41)
    s = _solver(pos, tol=1e-8, ppc=4)
    s.initialize(viscosity=ETA, hydrodynamicRadius=1.0, tolerance=1e-8, seed=91)
    s.setPositions(pos)
    s.setCoarseLevel(4)
    net = LearnedPreconditioner(seed=5)
    s.setPreconditioner(net)
    M = oracle.dense_matrix(pos, ETA, 1.0)
    ns = 3000
    X = np.array([s.sqrtMdotW()[0].ravel() for _ in range(ns)])
    assert s.coarseCells()[2] > 1
    lam, V = np.linalg.eigh(M)
    Z = (X @ V) / np.sqrt(lam)
    cov = np.cov(Z.T)
    err = np.abs(cov - np.eye(3 * N)).max()
    print(f"[coarse] FDT: max |cov - I| = {err:.3f} (bound {6.0 * np.sqrt(2.0 / ns):.3f})")
    assert err < 6.0 * np.sqrt(2.0 / ns)


def _ensemble(sizes, kinds, seed):
    from libmobility_amd.workloads import rsa_cube
    rng = np.random.default_rng(seed)
    pos = []
    for q, (sz, k) in enumerate(zip(sizes, kinds)):
        if sz == 0:
            continue
        if k:
            L = (sz / 0.05 / 5.0) ** 0.5
            pos.append(np.column_stack([rng.uniform(0, L, sz), rng.uniform(0, L, sz),
                                        rng.uniform(1.5, 6.5, sz)]))
        else:
            pos.append(rsa_cube(sz, 0.2, seed=seed * 100 + q)[0])
    return np.concatenate(pos)


def _ens_solver(sizes, kinds, pos, tol, ppc, per_system=False):
    from libmobility_amd import NBody
    s = NBody("open", "open", "single_wall")
    s.setParameters(wallHeight=0.0)
    s.initialize(viscosity=ETA, hydrodynamicRadius=1.0, tolerance=tol, seed=7)
    s.setSystems(sizes, ["single_wall" if k else "open" for k in kinds])
    s.setPositions(pos)
    s.setCoarseLevel(ppc)
    s.setLanczosPerSystem(per_system)
    return s


@pytest.mark.parametrize("path", ["batched", "per_system_launches"])
def test_ensemble_cells_and_factors_match_the_oracle(path):
    """Every system gets its own cells and coarse block: the cells are the
    oracle's rule applied to the system alone, L_s L_s^T is the oracle's
    Q_s^T M_s Q_s and X_s L_s = I.  Nine multi-tile blocks take the batched
    one-workgroup path, two take the per-system launches."""
    if path == "batched":
        sizes = [70, 95, 0, 120, 60, 88, 101, 77, 64, 110]
        kinds = [0, 1, 0, 1, 0, 0, 1, 0, 1, 0]
        ppc = 2
    else:
        sizes = [300, 0, 220]
        kinds = [1, 0, 0]
        ppc = 2
    pos = _ensemble(sizes, kinds, 3)
    s = _ens_solver(sizes, kinds, pos, 1e-6, ppc)
    s.sqrtMdotW_with_noise(np.random.default_rng(1).normal(size=pos.size))
    cell_of, _, C = s.coarseCells()
    off = np.concatenate([[0], np.cumsum(sizes)])
    c0 = 0
    for q, (a, b) in enumerate(zip(off[:-1], off[1:])):
        if a == b:
            assert s.coarseFactors(q)[0].shape == (0, 0)
            continue
        ps = pos[a:b]
        ref, _, Cq = coarse.coarse_cells(_records_xyz(ps, kinds[q]), ppc)
        assert np.array_equal(cell_of[a:b], ref + c0), q
        L, X = s.coarseFactors(q)
        assert L.shape == (3 * Cq, 3 * Cq)
        M = oracle.dense_matrix(ps, ETA, 1.0, wall=bool(kinds[q]))
        Mc = coarse.coarse_matrix(M, ref, Cq)
        rel = np.linalg.norm(L @ L.T - Mc) / np.linalg.norm(Mc)
        assert rel <= 1e-5, (q, rel)
        assert np.abs(X @ L - np.eye(3 * Cq)).max() <= 1e-10, q
        c0 += Cq
    assert c0 == C


@pytest.mark.parametrize("per_system", [False, True])
def test_ensemble_two_level_is_a_square_root(per_system):
    sizes, kinds = [14, 0, 12, 10, 9, 11, 8, 13, 10, 12], [0, 0, 1, 1, 0, 1, 0, 1, 0, 0]
    pos = _ensemble(sizes, kinds, 5)
    s = _ens_solver(sizes, kinds, pos, 1e-9, 3, per_system=per_system)
    n = pos.size
    B = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        B[:, k] = s.sqrtMdotW_with_noise(e)[0].ravel()
    M = oracle.dense_matrix(pos, ETA, 1.0, wall=kinds, sys_sizes=sizes)
    rel = np.linalg.norm(B @ B.T - M) / np.linalg.norm(M)
    print(f"[coarse] ensemble (per system {per_system}): cells {s.coarseCells()[2]}, "
          f"|BB^T - M|/|M| = {rel:.2e}")
    assert rel <= 1e-6


def test_ensemble_per_system_sqrt_matches_the_oracle():
    """Per-system recurrences with the coarse level: each system's result is
    the oracle's two-level Lanczos on that system with the build's L_s."""
    sizes, kinds = [300, 0, 260, 200], [0, 0, 1, 0]
    pos = _ensemble(sizes, kinds, 9)
    s = _ens_solver(sizes, kinds, pos, 1e-8, 12, per_system=True)
    W = np.random.default_rng(3).normal(size=pos.size)
    u, _ = s.sqrtMdotW_with_noise(W)
    u = u.ravel()
    cell_of, _, _ = s.coarseCells()
    off = np.concatenate([[0], np.cumsum(sizes)])
    c0 = 0
    for q, (a, b) in enumerate(zip(off[:-1], off[1:])):
        if a == b:
            continue
        L, X = s.coarseFactors(q)
        Cq = L.shape[0] // 3
        M = oracle.dense_matrix(pos[a:b], ETA, 1.0, wall=bool(kinds[q]))
        Q = coarse.coarse_Q(cell_of[a:b] - c0, Cq)
        ref, _ = coarse.TwoLevel(Q, L, X).sqrt(M, W[3 * a:3 * b], 1e-8)
        rel = np.linalg.norm(u[3 * a:3 * b] - ref) / np.linalg.norm(ref)
        print(f"[coarse] system {q}: cells {Cq}, rel {rel:.2e}")
        assert rel <= 1e-6, (q, rel)
        c0 += Cq


def test_sqrt_matches_the_oracle_with_long_coarse_rows():
    """A coarse block wider than one wave's matvec rows (3C > 2048: one
    workgroup per row): the result is the oracle's two-level Lanczos with the
    build's L, the oracle's matrix-free FP64 product and a sparse Q."""
    import scipy.sparse as sp
    from libmobility_amd.workloads import rsa_cube
    N = 3000
    pos, _ = rsa_cube(N, 0.2, seed=17)
    s = _solver(pos, tol=1e-8, ppc=4)
    W = np.random.default_rng(8).normal(size=3 * N)
    u, _ = s.sqrtMdotW_with_noise(W)
    cell_of, _, C = s.coarseCells()
    assert 3 * C > 2048
    L, X = s.coarseFactors()
    counts = np.bincount(cell_of, minlength=C)
    rows = np.arange(3 * N)
    cols = 3 * np.repeat(cell_of, 3) + np.tile(np.arange(3), N)
    Q = sp.csr_matrix((1.0 / np.sqrt(np.repeat(counts[cell_of], 3)), (rows, cols)),
                      shape=(3 * N, 3 * C))

    class Op:
        def __matmul__(self, v):
            return oracle.mdot(pos, ETA, 1.0, F=v)[0]

    ref, it = coarse.TwoLevel(Q, L, X).sqrt(Op(), W, 1e-8)
    rel = np.linalg.norm(u.ravel() - ref) / np.linalg.norm(ref)
    print(f"[coarse] N={N}, 3C={3 * C}: GPU {s.lastLanczosIterations()} / oracle {it} iterations, "
          f"rel {rel:.2e}")
    assert rel <= 1e-6


def test_factors_on_the_two_by_two_tile_path():
    """A coarse block of about 94 tiles: the trailing updates of L and of
    L^{-1} reach the two-by-two-tile kernels (used from 1024 trailing tiles
    up), and L L^T is the oracle's Q^T M Q, X L = I."""
    import scipy.sparse as sp
    from libmobility_amd.workloads import rsa_cube
    N = 4000
    pos, _ = rsa_cube(N, 0.2, seed=21)
    s = _solver(pos, ppc=2)
    s.sqrtMdotW_with_noise(np.random.default_rng(1).normal(size=3 * N))
    cell_of, _, C = s.coarseCells()
    nb = -(-3 * C // 64)
    ends = list(range(4, nb, 4))  # ends of the blocks of 4 tile columns
    assert any((nb - e) * (nb - e + 1) // 2 >= 1024 for e in ends)  # L
    assert any((nb - e) * e >= 1024 for e in ends)  # L^{-1}
    L, X = s.coarseFactors()
    M = oracle.dense_matrix(pos, ETA, 1.0) / (1.0 / (6 * np.pi * ETA))
    counts = np.bincount(cell_of, minlength=C)
    rows = np.arange(3 * N)
    cols = 3 * np.repeat(cell_of, 3) + np.tile(np.arange(3), N)
    Q = sp.csr_matrix((1.0 / np.sqrt(np.repeat(counts[cell_of], 3)), (rows, cols)),
                      shape=(3 * N, 3 * C))
    Mc = np.asarray((Q.T @ M) @ Q)
    del M
    assert np.allclose(L, np.tril(L))
    rel = np.linalg.norm(L @ L.T - Mc) / np.linalg.norm(Mc)
    err = np.abs(X @ L - np.eye(3 * C)).max()
    print(f"[coarse] C={C}, {nb} tiles: |LL^T - Mc|/|Mc| = {rel:.2e}, max|XL - I| = {err:.2e}")
    assert rel <= 1e-5
    assert err <= 1e-11


@pytest.mark.parametrize("wall", [False, True])
def test_fp64_coarse_matrix_matches_the_oracle(wall):
    """The coarse matrix from FP64 pair terms (the fallback for an FP32 matrix
    that rounding leaves indefinite, or asked for): L L^T equals the oracle's
    Q^T M Q to FP64 rounding, where the FP32 build agrees to ~1e-7."""
    from libmobility_amd.workloads import rsa_cube
    N = 1500
    pos = _wall_positions(N, 3) if wall else rsa_cube(N, 0.2, seed=3)[0]
    s = _solver(pos, wall=wall, ppc=40)
    s.setCoarsePairPrecision("fp64")
    s.sqrtMdotW_with_noise(np.random.default_rng(0).normal(size=3 * N))
    assert s.coarseUsedFP64()
    cell_of, grid, C = s.coarseCells()
    L, X = s.coarseFactors()
    M = oracle.dense_matrix(pos, ETA, 1.0, wall=wall) / (1.0 / (6 * np.pi * ETA))
    Mc = coarse.coarse_matrix(M, cell_of, C)
    rel = np.linalg.norm(L @ L.T - Mc) / np.linalg.norm(Mc)
    assert rel <= 1e-12, rel
    assert np.abs(X @ L - np.eye(3 * C)).max() <= 1e-11
    s32 = _solver(pos, wall=wall, ppc=40)
    s32.sqrtMdotW_with_noise(np.random.default_rng(0).normal(size=3 * N))
    assert not s32.coarseUsedFP64()


def test_more_than_8192_cells():
    """A coarse level of 15625 cells (open RSA packing, N = 131072, 8 per
    cell, dimension cap raised to 49152): its coarse matrix has 1.2e8 cell
    pairs, more than one dispatch of one wave per pair can hold (2^32
    work-items), so the kernel strides over them.  Free-space RPY is SPD for
    any configuration, so the build must succeed, and the finer cells must
    cut the iterations below those of the default 40 per cell (38 at this
    size)."""
    from libmobility_amd.workloads import config_c3
    pos = config_c3(131072, 0.2)[0]
    W = np.random.default_rng(1).normal(size=3 * pos.shape[0])
    s = _solver(pos, tol=1e-6, ppc=8)
    s.setPairPrecision("fp32")
    s.setCoarseMaxDim(49152)
    s.sqrtMdotW_with_noise(W)
    C = s.coarseCells()[2]
    assert 8192 < C <= 16384, C
    fine = s.lastLanczosIterations()
    s40 = _solver(pos, tol=1e-6, ppc=40)
    s40.setPairPrecision("fp32")
    s40.sqrtMdotW_with_noise(W)
    assert fine < s40.lastLanczosIterations(), (fine, s40.lastLanczosIterations())
