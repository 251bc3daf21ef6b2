"""Symmetric evaluation (extension): each unordered pair is evaluated once and
applied to both particles, M_ji = M_ij^T blockwise.  The product must equal
the reference's ordered evaluation (NbodyRPY.cu:27-97 with the blocks of
hydrodynamicKernels.cuh: open RPY :41-173, wall terms :187-431) as restated
by the oracle.  Covered: every tile count (one tile, even and odd counts, a
partial last tile), every tiling, open and single-wall systems, every block
set (F, F with angular output, T, F and T), heights below the radius,
overlapping and coincident particles in different tiles, and the ring split
into passes; FP64 1e-12, FP32 pair terms 1e-6 (translational) / 2e-6
(angular blocks) relative in norm."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ETA, A = 0.7, 1.3


def _solver(precision="fp64", sym=1, tpl=0, nch=0, tol=1e-6):
    from libmobility_amd import NBody
    s = NBody("open", "open", "open")
    s.setParameters()
    s.setPairPrecision(precision)
    s.setPairSymmetric(sym)
    s.setPairTiling(tpl, nch)
    s.initialize(viscosity=ETA, hydrodynamicRadius=A, tolerance=tol)
    return s


def _cloud(N, seed, phi=0.15):
    rng = np.random.default_rng(seed)
    L = (N * 4 * np.pi * A ** 3 / 3 / phi) ** (1 / 3)
    return rng.uniform(0, L, (N, 3))


def _rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


# (N, targets per lane, offset chunks): tile counts 1 .. 9, odd and even, and
# partial last tiles
CASES = [
    (100, 1, 0), (256, 1, 0), (300, 1, 0), (700, 1, 0), (1000, 1, 0), (1025, 1, 1),
    (1200, 1, 2), (2300, 1, 4), (2300, 2, 0), (2300, 4, 0), (5000, 2, 3), (5000, 4, 1),
]


@pytest.mark.parametrize("N,tpl,nch", CASES)
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_symmetric_product_vs_oracle(N, tpl, nch, precision, tol):
    pos = _cloud(N, N)
    F = np.random.default_rng(N + 1).normal(size=(N, 3))
    s = _solver(precision, 1, tpl, nch)
    s.setPositions(pos)
    U, _ = s.Mdot(F)
    assert s.lastPairSymmetric()
    assert s.lastPairTiling()[0] == (tpl or 1)
    Uo, _ = oracle.mdot(pos, ETA, A, F=F)
    assert _rel(U, Uo) <= tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-14), ("fp32", 1e-6)])
def test_symmetric_equals_ordered_evaluation(precision, tol):
    N = 9000
    pos = _cloud(N, 3)
    F = np.random.default_rng(4).normal(size=(N, 3))
    ordered = _solver(precision, 0)
    ordered.setPositions(pos)
    U0, _ = ordered.Mdot(F)
    assert not ordered.lastPairSymmetric()
    s = _solver(precision, 1)
    s.setPositions(pos)
    U1, _ = s.Mdot(F)
    assert _rel(U1, U0) <= tol
    U2, _ = s.Mdot(F)
    assert np.array_equal(U1, U2)  # fixed summation order: bitwise reproducible


def test_overlapping_and_coincident_particles_in_different_tiles():
    """Copies of particles of tile 0 placed in the last tile (r = 0 between
    distinct particles: the overlap branch's value at r = 0) and a dense
    cluster of overlapping spheres spread over all tiles."""
    N = 1500  # 6 tiles of 256
    pos = _cloud(N, 5)
    pos[1400:1420] = pos[0:20]
    pos[300:340] = pos[300] + np.random.default_rng(6).normal(scale=0.8, size=(40, 3))
    pos[900:940] = pos[300] + np.random.default_rng(7).normal(scale=0.8, size=(40, 3))
    F = np.random.default_rng(8).normal(size=(N, 3))
    for tpl in (1, 2):
        s = _solver("fp64", 1, tpl)
        s.setPositions(pos)
        U, _ = s.Mdot(F)
        Uo, _ = oracle.mdot(pos, ETA, A, F=F)
        assert _rel(U, Uo) <= 1e-12, tpl


def test_symmetric_matrix_is_symmetric_and_matches_dense():
    N = 600  # 3 tiles of 256 (odd count)
    pos = _cloud(N, 9, phi=0.3)
    s = _solver("fp64", 1, 1)
    s.setPositions(pos)
    n = 3 * N
    cols = np.random.default_rng(10).choice(n, 40, replace=False)
    M = oracle.dense_matrix(pos, ETA, A)
    for c in cols:
        e = np.zeros(n)
        e[c] = 1.0
        u, _ = s.Mdot(e)
        assert np.abs(u.ravel() - M[:, c]).max() <= 1e-12 * np.abs(M).max(), c


def test_auto_mode_picks_the_symmetric_path_for_large_open_systems():
    N = 40000
    pos = _cloud(N, 11, phi=0.2)
    F = np.random.default_rng(12).normal(size=(N, 3))
    s = _solver("fp32", -1)
    s.setPositions(pos)
    U, _ = s.Mdot(F)
    assert s.lastPairSymmetric()
    rows = np.random.default_rng(13).choice(N, 2000, replace=False)
    Uo, _ = oracle.mdot_rows(pos, ETA, A, rows, F=F)
    assert _rel(U[rows], Uo) <= 1e-6
    small = _solver("fp32", -1)
    small.setPositions(pos[:1000])
    small.Mdot(F[:1000])
    assert not small.lastPairSymmetric()
    # the automatic rule starts at 16384 particles
    for n, expect in ((16383, False), (16384, True)):
        s = _solver("fp32", -1)
        s.setPositions(pos[:n])
        s.Mdot(F[:n])
        assert s.lastPairSymmetric() == expect, n


def _wall_solver(precision="fp64", sym=1, tpl=0, nch=0, angular=False, wall=True, memory=0):
    from libmobility_amd import NBody
    s = NBody("open", "open", "single_wall" if wall else "open")
    s.setParameters(wallHeight=0.0) if wall else s.setParameters()
    s.setPairPrecision(precision)
    s.setPairSymmetric(sym, memory)
    s.setPairTiling(tpl, nch)
    s.initialize(viscosity=ETA, hydrodynamicRadius=A, includeAngular=angular)
    return s


def _wall_cloud(N, seed, wall=True):
    pos = _cloud(N, seed)
    if wall:
        pos[:, 2] = np.random.default_rng(seed + 99).uniform(0.4, 9.0, N)  # some below the radius
    return pos


# (name, wall, forces, torques, angular)
BLOCK_SETS = [
    ("wall_F", True, True, False, False),
    ("wall_F_angular", True, True, False, True),
    ("wall_T", True, False, True, True),
    ("wall_FT", True, True, True, True),
    ("open_F_angular", False, True, False, True),
    ("open_T", False, False, True, True),
    ("open_FT", False, True, True, True),
]


@pytest.mark.parametrize("name,wall,hasF,hasT,angular", BLOCK_SETS)
@pytest.mark.parametrize("N,tpl,nch", [(300, 1, 0), (1100, 1, 2), (2300, 2, 3)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_symmetric_blocks_vs_oracle(name, wall, hasF, hasT, angular, N, tpl, nch, precision):
    pos = _wall_cloud(N, N + 7, wall)
    pos[N - 9:N] = pos[0:9]  # coincident with particles of the first tile
    rng = np.random.default_rng(N)
    F = rng.normal(size=(N, 3)) if hasF else None
    T = rng.normal(size=(N, 3)) if hasT else None
    s = _wall_solver(precision, 1, tpl, nch, angular, wall)
    s.setPositions(pos)
    U, W = s.Mdot(F, T)
    assert s.lastPairSymmetric()
    Uo, Wo = oracle.mdot(pos, ETA, A, F=F, T=T, angular=angular, wall=wall)
    tol_u = 1e-12 if precision == "fp64" else 1e-6
    tol_w = 1e-12 if precision == "fp64" else 2e-6
    assert _rel(U, Uo) <= tol_u
    if angular:
        assert _rel(W, Wo) <= tol_w


def test_symmetric_wall_matrix_matches_dense():
    N = 200  # one tile of 256: diagonal item only; then 600 (3 tiles)
    for N in (200, 600):
        pos = _wall_cloud(N, 21 + N)
        s = _wall_solver("fp64", 1, 1, 0, True)
        s.setPositions(pos)
        M = oracle.dense_matrix(pos, ETA, A, angular=True, wall=True)
        n = 3 * N
        for c in np.random.default_rng(N).choice(6 * N, 30, replace=False):
            e = np.zeros(6 * N)
            e[c] = 1.0
            u, w = s.Mdot(e[:n], e[n:])
            col = np.concatenate([u.ravel(), w.ravel()])
            assert np.abs(col - M[:, c]).max() <= 1e-12 * np.abs(M).max(), (N, c)


@pytest.mark.parametrize("wall,angular", [(False, False), (True, True)])
def test_symmetric_passes_under_a_small_memory_budget(wall, angular):
    """A budget of a few slabs splits the ring of offsets into many passes;
    the sum is the same product (to rounding) and still deterministic."""
    N = 3000  # 12 tiles of 256: offsets 1..6
    pos = _wall_cloud(N, 31, wall)
    rng = np.random.default_rng(32)
    F = rng.normal(size=(N, 3))
    T = rng.normal(size=(N, 3)) if angular else None
    one = _wall_solver("fp64", 1, 1, 0, angular, wall)
    one.setPositions(pos)
    U1, W1 = one.Mdot(F, T)
    slab = 8 * 3 * N * (2 if angular else 1)
    many = _wall_solver("fp64", 1, 1, 0, angular, wall, memory=4 * slab)
    many.setPositions(pos)
    U2, W2 = many.Mdot(F, T)
    assert _rel(U2, U1) <= 1e-14
    if angular:
        assert _rel(W2, W1) <= 1e-14
    U3, _ = many.Mdot(F, T)
    assert np.array_equal(U2, U3)
    Uo, _ = oracle.mdot(pos, ETA, A, F=F, T=T, angular=angular, wall=wall)
    assert _rel(U2, Uo) <= 1e-12


def test_symmetric_path_is_not_used_when_off():
    pos = _wall_cloud(600, 14)
    off = _wall_solver("fp64", 0, 0, 0, True, True)
    off.setPositions(pos)
    off.Mdot(np.ones((600, 3)), np.ones((600, 3)))
    assert not off.lastPairSymmetric()


# ragged ensembles: tile counts 1 (tiny and one-tile systems), 2, 3 and more,
# partial last tiles, open and wall systems mixed
ENSEMBLE_SIZES = [5, 256, 300, 700, 1, 1500, 60, 2300]


def _ensemble_solver(precision, tpl=0, nch=0, angular=False, memory=0, kinds=None):
    from libmobility_amd import NBody
    s = NBody("open", "open", "single_wall")
    s.setParameters(wallHeight=0.0)
    s.setPairPrecision(precision)
    s.setPairSymmetric(1, memory)
    s.setPairTiling(tpl, nch)
    s.initialize(viscosity=ETA, hydrodynamicRadius=A, includeAngular=angular, tolerance=1e-6)
    k = kinds if kinds is not None else [i % 2 for i in range(len(ENSEMBLE_SIZES))]
    s.setSystems(ENSEMBLE_SIZES, ["single_wall" if x else "open" for x in k])
    return s, k


def _ensemble_positions(seed, kinds):
    parts = []
    for i, (n, k) in enumerate(zip(ENSEMBLE_SIZES, kinds)):
        parts.append(_wall_cloud(n, seed + i, wall=bool(k)))
    return np.concatenate(parts)


@pytest.mark.parametrize("hasF,hasT,angular", [(True, False, False), (True, False, True),
                                                (False, True, True), (True, True, True)])
@pytest.mark.parametrize("tpl,nch", [(1, 0), (2, 2)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_symmetric_ensemble_vs_oracle(hasF, hasT, angular, tpl, nch, precision):
    s, kinds = _ensemble_solver(precision, tpl, nch, angular)
    pos = _ensemble_positions(40, kinds)
    N = pos.shape[0]
    pos[N - 4:N] = pos[N - 2300:N - 2296]  # coincident particles inside the last system
    rng = np.random.default_rng(41)
    F = rng.normal(size=(N, 3)) if hasF else None
    T = rng.normal(size=(N, 3)) if hasT else None
    s.setPositions(pos)
    U, W = s.Mdot(F, T)
    assert s.lastPairSymmetric()
    Uo, Wo = oracle.mdot(pos, ETA, A, F=F, T=T, angular=angular, wall=kinds,
                         sys_sizes=ENSEMBLE_SIZES)
    tol_u = 1e-12 if precision == "fp64" else 1e-6
    tol_w = 1e-12 if precision == "fp64" else 2e-6
    assert _rel(U, Uo) <= tol_u
    if angular:
        assert _rel(W, Wo) <= tol_w


def test_symmetric_ensemble_passes_and_determinism():
    s, kinds = _ensemble_solver("fp64", 1, 0, True)
    pos = _ensemble_positions(50, kinds)
    N = pos.shape[0]
    s.setPositions(pos)
    rng = np.random.default_rng(51)
    F, T = rng.normal(size=(N, 3)), rng.normal(size=(N, 3))
    U1, W1 = s.Mdot(F, T)
    small, _ = _ensemble_solver("fp64", 1, 0, True, memory=4 * 8 * 3 * N * 2)
    small.setPositions(pos)
    U2, W2 = small.Mdot(F, T)
    assert _rel(U2, U1) <= 1e-14 and _rel(W2, W1) <= 1e-14
    U3, W3 = small.Mdot(F, T)
    assert np.array_equal(U2, U3) and np.array_equal(W2, W3)


def test_symmetric_per_system_lanczos_matches_separate_solvers():
    """Per-system recurrences (LanczosAlgorithm.cu:171-232 per system) over the
    symmetric product: as systems converge their items leave the launch and
    their rows the reduction; every system still equals its own solver."""
    from tests.test_gpu_ensemble import _rsa_ensemble, _single
    sizes, kinds, pos = _rsa_ensemble(seed=61, sizes=(2100, 40, 900, 1300, 5))
    N = pos.shape[0]
    off = np.concatenate([[0], np.cumsum(sizes)])
    tol = 1e-6
    from libmobility_amd import NBody
    ens = NBody("open", "open", "single_wall")
    ens.setParameters(wallHeight=0.0)
    ens.setPairSymmetric(1)
    ens.initialize(viscosity=0.9, hydrodynamicRadius=1.1, tolerance=tol, seed=3)
    ens.setSystems(sizes, ["single_wall" if k else "open" for k in kinds])
    ens.setLanczosPerSystem(True)
    ens.setPositions(pos)
    singles = [_single(pos[a:b], k, False, tol) for a, b, k in zip(off[:-1], off[1:], kinds)]
    rng = np.random.default_rng(62)
    for call in range(2):
        W = rng.normal(size=3 * N)
        u, _ = ens.sqrtMdotW_with_noise(W)
        assert ens.lastPairSymmetric()
        its = ens.lastLanczosIterationsPerSystem()
        assert len(set(its.tolist())) > 1
        for s, (a, b) in enumerate(zip(off[:-1], off[1:])):
            us, _ = singles[s].sqrtMdotW_with_noise(W[3 * a:3 * b])
            assert abs(its[s] - singles[s].lastLanczosIterations()) <= 6, (call, s)
            assert np.linalg.norm(u.ravel()[3 * a:3 * b] - us.ravel()) <= 10 * tol * np.linalg.norm(us), (call, s)


def test_lanczos_on_the_symmetric_product():
    """sqrtMdotW through the symmetric product: a square root of M as
    accurate as the oracle's own Lanczos run (LanczosAlgorithm.cu:171-232)."""
    from libmobility_amd.workloads import rsa_box, rsa_box_side
    N = 700
    L = rsa_box_side(N, 0.2, A)
    # non-overlapping spheres at tol 1e-6: the oracle's stopping step (38)
    # does not move under 1e-15 perturbations of M v
    pos = rsa_box(N, (L, L, L), 15, a=A)
    tol = 1e-6
    s = _solver("fp64", 1, 1, tol=tol)
    s.setPositions(pos)
    W = np.random.default_rng(16).normal(size=3 * N)
    y, _ = s.sqrtMdotW_with_noise(W)
    M = oracle.dense_matrix(pos, ETA, A)
    ref, it_ref = oracle.LanczosOracle().run(lambda v: M @ v, W, tol)
    exact = oracle.dense_sqrt(M) @ W
    assert abs(s.lastLanczosIterations() - it_ref) <= 6
    assert np.linalg.norm(y.ravel() - exact) <= 3 * np.linalg.norm(ref - exact) + 5 * tol * np.linalg.norm(exact)


def test_mode_is_validated():
    s = _solver("fp64", -1)
    with pytest.raises(RuntimeError):
        s.setPairSymmetric(2)


@pytest.mark.parametrize("wall", [False, True])
def test_symmetric_smallest_systems(wall):
    """One particle (a diagonal item and no ring), two coincident particles,
    and an ensemble with an empty system, all against the oracle."""
    rng = np.random.default_rng(71)
    for pos in (np.array([[1.0, 2.0, 3.0]]), np.array([[1.0, 2.0, 3.0], [1.0, 2.0, 3.0]])):
        s = _wall_solver("fp64", 1, 0, 0, True, wall)
        s.setPositions(pos)
        F, T = rng.normal(size=pos.shape), rng.normal(size=pos.shape)
        U, W = s.Mdot(F, T)
        assert s.lastPairSymmetric()
        Uo, Wo = oracle.mdot(pos, ETA, A, F=F, T=T, angular=True, wall=wall)
        assert _rel(U, Uo) <= 1e-13 and _rel(W, Wo) <= 1e-13
    from libmobility_amd import NBody
    e = NBody("open", "open", "single_wall" if wall else "open")
    e.setParameters(wallHeight=0.0) if wall else e.setParameters()
    e.setPairSymmetric(1)
    e.initialize(viscosity=ETA, hydrodynamicRadius=A)
    sizes = [300, 0, 1, 520]
    e.setSystems(sizes)
    pos = _wall_cloud(sum(sizes), 72, wall)
    e.setPositions(pos)
    F = rng.normal(size=(sum(sizes), 3))
    U, _ = e.Mdot(F)
    assert e.lastPairSymmetric()
    Uo, _ = oracle.mdot(pos, ETA, A, F=F, wall=wall, sys_sizes=sizes)
    assert _rel(U, Uo) <= 1e-12


def test_symmetric_open_ensemble_with_four_targets_per_lane_request():
    """A tiling request of 4 targets per lane is served with 2 for ensembles
    (4 exists for one open system only); the product is unchanged."""
    from libmobility_amd import NBody
    sizes = [1500, 700, 2100]
    pos = _cloud(sum(sizes), 81)
    F = np.random.default_rng(82).normal(size=(sum(sizes), 3))
    s = NBody("open", "open", "open")
    s.setParameters()
    s.setPairSymmetric(1)
    s.setPairTiling(4, 0)
    s.initialize(viscosity=ETA, hydrodynamicRadius=A)
    s.setSystems(sizes)
    s.setPositions(pos)
    U, _ = s.Mdot(F)
    assert s.lastPairSymmetric() and s.lastPairTiling()[0] == 2
    Uo, _ = oracle.mdot(pos, ETA, A, F=F, sys_sizes=sizes)
    assert _rel(U, Uo) <= 1e-12


@pytest.mark.parametrize("sym", [0, 1])
@pytest.mark.parametrize("sizes", [None, [700, 0, 1300]])
def test_new_positions_of_the_same_systems_keep_work_lists_but_move_the_result(sym, sizes):
    """setPositions with the same N and systems (a time loop): the work lists
    are kept, the product follows the new positions."""
    from libmobility_amd import NBody
    s = NBody("open", "open", "single_wall")
    s.setParameters(wallHeight=0.0)
    s.setPairSymmetric(sym)
    s.initialize(viscosity=ETA, hydrodynamicRadius=A)
    N = 2000
    if sizes is not None:
        s.setSystems(sizes)
    rng = np.random.default_rng(91)
    F = rng.normal(size=(N, 3))
    for step in range(3):
        pos = _wall_cloud(N, 92 + step)
        s.setPositions(pos)
        U, _ = s.Mdot(F)
        assert s.lastPairSymmetric() == bool(sym)
        Uo, _ = oracle.mdot(pos, ETA, A, F=F, wall=True, sys_sizes=sizes)
        assert _rel(U, Uo) <= 1e-12, step
