"""GPU checks of the learned block preconditioner (no reference counterpart).

  * kNN features (coordinates and wall features) and the FP64 local targets
    of both block sizes, open and bottom wall, against the oracle's dense
    matrices and their eigen inverse square roots;
  * MFMA forward pass and gradients vs the numpy oracle (oracle/mlp.py),
    tolerance for BF16 operands with FP32 accumulation, 3x3 and 6x6 heads;
  * training reduces the loss;
  * the preconditioned sqrtMdotW is a square root of M: ||B B^T - M|| / ||M||
    <= 1e-5 at small N (3x3 and 6x6 blocks), and its fluctuations have
    covariance M;
  * ablations: on bottom-wall slabs the trained blocks cut the iterations of
    the coarse-level square root by at least 15 % (forces, 3x3 blocks; forces
    and torques, 6x6 blocks); on open packings they neither gain nor lose
    (within 10 %): there the collective modes the coarse level whitens set
    the iteration count.
"""
import numpy as np
import pytest

import oracle
from oracle import mlp

pytestmark = pytest.mark.gpu

ETA = 1.0 / (6 * np.pi)


def _open_solver(pos, tol=1e-6, eta=ETA, angular=False):
    from libmobility_amd import NBody
    s = NBody("open", "open", "open")
    s.setParameters()
    s.initialize(viscosity=eta, hydrodynamicRadius=1.0, tolerance=tol, includeAngular=angular)
    s.setPositions(pos)
    return s


def _wall_solver(pos, tol=1e-6, angular=False):
    from libmobility_amd import NBody
    s = NBody("open", "open", "single_wall")
    s.setParameters(wallHeight=0.0)
    s.initialize(viscosity=ETA, hydrodynamicRadius=1.0, tolerance=tol, includeAngular=angular)
    s.setPositions(pos)
    return s


def _slab(N, Lxy, zlo, zhi, seed):
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(0, Lxy, (N, 2)), rng.uniform(zlo, zhi, N)])


def _local_target(cluster, wall, d):
    """Center d x d block of (M_local/t0)^{-1/2} from the oracle's dense
    matrix (6x6: linear entries over sqrt(t0), angular over sqrt(r0))."""
    K = cluster.shape[0]
    M = oracle.dense_matrix(cluster, ETA, 1.0, angular=d == 6, wall=wall)  # t0 = 1
    if d == 6:
        r0 = 1.0 / (8 * np.pi * ETA)
        sc = np.concatenate([np.ones(3 * K), np.full(3 * K, 1.0 / np.sqrt(r0))])
        M = M * sc[:, None] * sc[None, :]
    lam, V = np.linalg.eigh(M)
    Mi = (V / np.sqrt(lam)) @ V.T
    idx = np.r_[0:3, 3 * K:3 * K + 3] if d == 6 else np.r_[0:3]
    return Mi[np.ix_(idx, idx)]


@pytest.mark.parametrize("wall,d", [(False, 3), (True, 3), (True, 6), (False, 6)])
def test_features_and_local_targets(wall, d):
    from libmobility_amd import _ffi
    from libmobility_amd.precond import dataset
    from libmobility_amd.workloads import rsa_cube
    if wall:
        pos = _slab(3000, 60.0, 0.4, 9.0, seed=11)
        s = _wall_solver(pos)
    else:
        pos, _ = rsa_cube(3000, 0.2, seed=11)
        s = _open_solver(pos)
    N = pos.shape[0]
    nbr = np.zeros((N, 16), dtype=np.int32)
    _ffi.check(_ffi.load().lm_nbody_precond_neighbours(s._h, nbr.ctypes.data))
    # the kNN geometry uses the records: heights clamped to one radius
    geo = pos.copy()
    if wall:
        geo[:, 2] = np.maximum(geo[:, 2], 1.0)
    rel_all = geo[nbr] - geo[:, None, :]
    X, T = dataset(s, d, 48, seed=5)
    assert X.shape == (48, 64) and T.shape == (48, d * d)
    for c in range(X.shape[0]):
        rel = X[c, :48].reshape(16, 3).astype(np.float64)
        dist = np.linalg.norm(rel, axis=1)
        assert np.all(np.diff(dist) >= -1e-5)  # sorted by distance
        i = int(np.argmin(np.abs(rel_all - rel[None]).sum(axis=(1, 2))))
        assert np.allclose(rel_all[i], rel, atol=1e-5)
        if wall:
            z = pos[i, 2]
            u = 1.0 / max(z, 1.0)
            assert np.allclose(X[c, 48:51], [u, min(max(z, 0.0), 1.0), u ** 3], atol=1e-6)
        else:
            assert np.array_equal(X[c, 48:51], [0.0, 1.0, 0.0])
        assert np.all(X[c, 51:] == 0)
        k = 16 if d == 3 else 8
        cluster = pos[np.r_[i, nbr[i, :k]]]
        ref = _local_target(cluster, wall, d)
        assert np.allclose(T[c].reshape(d, d), ref, rtol=0, atol=2e-6 * np.abs(ref).max())


def _net(seed=3, d=3):
    from libmobility_amd.precond import LearnedPreconditioner
    return LearnedPreconditioner(seed=seed, block_dim=d)


@pytest.mark.parametrize("d", [3, 6])
def test_forward_and_gradients_vs_numpy_oracle(d):
    rng = np.random.default_rng(0)
    B = 300
    X = rng.normal(scale=2.0, size=(B, mlp.NIN)).astype(np.float32)
    A = rng.normal(size=(B, d, d))
    T = (A @ np.transpose(A, (0, 2, 1)) * 0.1 + 1.3 * np.eye(d)).reshape(B, d * d)
    T = T.astype(np.float32)
    net = _net(d=d)
    assert net.num_weights == mlp.num_weights(d)
    w = net.get_weights().astype(np.float64)
    out = net.forward(X)
    ref, _ = mlp.forward(w, X, d)
    # BF16 operands (8-bit mantissa), FP32 sums
    assert np.linalg.norm(out - ref) <= 2e-2 * np.linalg.norm(ref)
    loss, g = net.loss_grad(X, T)
    lref, gref = mlp.loss_grad(w, X, T, d)
    assert abs(loss - lref) <= 3e-2 * abs(lref)
    n1 = mlp.NIN * mlp.NH
    for name, sl in (("W1", slice(0, n1)), ("W2", slice(n1 + 128, n1 + 128 + 128 * 128)),
                     ("all", slice(None))):
        assert np.linalg.norm(g[sl] - gref[sl]) <= 5e-2 * np.linalg.norm(gref[sl]), name
    cos = g @ gref / (np.linalg.norm(g) * np.linalg.norm(gref))
    assert cos > 0.999


@pytest.mark.parametrize("d", [3, 6])
def test_training_reduces_loss(d):
    from libmobility_amd.precond import dataset
    s = _wall_solver(_slab(20000, 200.0, 0.8, 15.0, seed=2))
    X, T = dataset(s, d, 8192, seed=2)
    net = _net(seed=4, d=d)
    l0, _ = net.loss_grad(X, T)
    hist = net.train(X, T, epochs=15, batch=1024, lr=2e-3, seed=1)
    l1, _ = net.loss_grad(X, T)
    assert l1 < 0.5 * l0 and hist[-1] < hist[0]


def _columns(s, n, angular):
    Bm = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        u, w = s.sqrtMdotW_with_noise(e)
        Bm[:, k] = np.concatenate([u.ravel(), w.ravel()]) if angular else u.ravel()
    return Bm


@pytest.mark.parametrize("case", ["open-untrained", "open-trained", "wall-trained",
                                  "wall-angular-untrained", "wall-angular-trained"])
def test_preconditioned_sqrt_is_a_square_root(case):
    import os
    from libmobility_amd.precond import DEFAULT_WEIGHTS, DEFAULT_WEIGHTS6, default_network
    from libmobility_amd.workloads import rsa_cube
    angular = "angular" in case
    wall = case.startswith("wall")
    trained = case.endswith("-trained")
    if trained and not os.path.exists(DEFAULT_WEIGHTS6 if angular else DEFAULT_WEIGHTS):
        pytest.skip("trained weights not packaged")
    N = 40
    if wall:
        pos = _slab(N, 14.0, 0.6, 6.0, seed=4)
        s = _wall_solver(pos, tol=1e-8, angular=angular)
    else:
        pos, _ = rsa_cube(N, 0.2, seed=4)
        s = _open_solver(pos, tol=1e-8)
    net = default_network(angular) if trained else _net(seed=9, d=6 if angular else 3)
    s.setPreconditioner(net)
    n = (6 if angular else 3) * N
    Bm = _columns(s, n, angular)
    M = oracle.dense_matrix(pos, ETA, 1.0, angular=angular, wall=wall)
    rel = np.linalg.norm(Bm @ Bm.T - M) / np.linalg.norm(M)
    assert rel <= 1e-5, rel
    G = s.preconditionerBlocks()
    d = 6 if angular else 3
    assert G.shape == (N, d, d)
    if d == 3:
        assert np.allclose(G, np.transpose(G, (0, 2, 1)))
        assert np.all(np.linalg.eigvalsh(G) > 0)
    else:
        # G = D S with D SPD and S = diag(1, 1, 1, c, c, c), c = 2a/sqrt(3)
        S = np.diag([1, 1, 1] + [2 / np.sqrt(3)] * 3)
        D = G @ np.linalg.inv(S)
        assert np.allclose(D, np.transpose(D, (0, 2, 1)), atol=1e-12 * np.abs(D).max())
        assert np.all(np.linalg.eigvalsh(D) > 0)


def test_six_by_six_blocks_need_angular_solvers():
    from libmobility_amd.workloads import rsa_cube
    pos, _ = rsa_cube(64, 0.2, seed=3)
    s = _open_solver(pos)
    s.setPreconditioner(_net(d=6))
    with pytest.raises(RuntimeError, match="includeAngular"):
        s.sqrtMdotW_with_noise(np.ones(3 * 64))


def test_preconditioned_fluctuation_dissipation():
    from libmobility_amd.workloads import rsa_cube
    pos, _ = rsa_cube(20, 0.2, seed=6)
    s = _open_solver(pos, tol=1e-8)
    s.initialize(viscosity=ETA, hydrodynamicRadius=1.0, tolerance=1e-8, seed=77)
    s.setPositions(pos)
    s.setPreconditioner(_net(seed=5))
    M = oracle.dense_matrix(pos, ETA, 1.0)
    ns = 3000
    X = np.array([s.sqrtMdotW()[0].ravel() for _ in range(ns)])
    lam, V = np.linalg.eigh(M)
    Z = (X @ V) / np.sqrt(lam)
    cov = np.cov(Z.T)
    assert np.abs(cov - np.eye(60)).max() < 6.0 * np.sqrt(2.0 / ns)


def _iterations(make, net, ppc, Ws):
    s = make()
    if net is not None:
        s.setPreconditioner(net)
    s.setCoarseLevel(ppc)
    its = []
    for w in Ws:
        s.sqrtMdotW_with_noise(w)
        its.append(s.lastLanczosIterations())
    return its


@pytest.mark.parametrize("angular", [False, True])
def test_ablation_wall_slab_trained_blocks_gain_over_the_coarse_level(angular):
    """Bottom-wall slab of the benchmark's density (N = 8192, x, y in
    [0, 90.5a), z in [1.5a, 20a]; tol 1e-6; 40 particles per cell): the coarse
    level alone, and with the trained blocks on top (3x3 for forces, 6x6 for
    forces and torques).  The trained blocks must cut the iterations by at
    least 15 %.  (Ideal blocks -- the exact local inverse square roots the
    network is trained on -- reach 25 % and 21 % on such slabs, a CPU
    measurement at N = 2048 and 4096.)"""
    import os
    from libmobility_amd.precond import DEFAULT_WEIGHTS, DEFAULT_WEIGHTS6, default_network
    if not os.path.exists(DEFAULT_WEIGHTS6 if angular else DEFAULT_WEIGHTS):
        pytest.skip("trained weights not packaged")
    N = 8192
    pos = _slab(N, 256.0 * np.sqrt(N / 65536), 1.5, 20.0, seed=31)
    rng = np.random.default_rng(2)
    n = (6 if angular else 3) * N
    Ws = [rng.normal(size=n) for _ in range(3)]
    make = lambda: _wall_solver(pos, tol=1e-6, angular=angular)
    plain = _iterations(make, None, 0, Ws)
    coarse_only = _iterations(make, None, 40, Ws)
    learned = _iterations(make, default_network(angular), 40, Ws)
    print(f"[ablation] wall angular={angular}: iterations plain {plain} coarse level "
          f"{coarse_only} trained blocks + coarse level {learned}")
    assert np.mean(learned) <= 0.85 * np.mean(coarse_only)
    assert np.mean(coarse_only) < np.mean(plain)


def test_ablation_open_packing_trained_blocks_have_no_effect():
    """Open RSA packing (N = 4096, phi = 0.2, tol 1e-6, 40 per cell): the
    trained 3x3 blocks leave the coarse-level iteration count where it is
    (within 10 %); the coarse level cuts plain Lanczos by at least 25 %."""
    import os
    from libmobility_amd.precond import DEFAULT_WEIGHTS, default_network
    from libmobility_amd.workloads import rsa_cube
    if not os.path.exists(DEFAULT_WEIGHTS):
        pytest.skip("trained weights not packaged")
    N = 4096
    pos, _ = rsa_cube(N, 0.2, seed=22)
    rng = np.random.default_rng(1)
    Ws = [rng.normal(size=3 * N) for _ in range(2)]
    make = lambda: _open_solver(pos, tol=1e-6)
    plain = _iterations(make, None, 0, Ws)
    coarse_only = _iterations(make, None, 40, Ws)
    learned = _iterations(make, default_network(False), 40, Ws)
    print(f"[ablation] open: iterations plain {plain} coarse level {coarse_only} "
          f"trained blocks + coarse level {learned}")
    assert np.mean(coarse_only) <= 0.75 * np.mean(plain)
    assert abs(np.mean(learned) / np.mean(coarse_only) - 1) <= 0.10


@pytest.mark.parametrize("geom", ["cube", "slab", "clusters", "ensemble"])
def test_neighbours_are_the_nearest(geom):
    """The cell-grid neighbour search returns, for every particle, the 16
    nearest particles of its own system, nearest first (distances compared in
    FP64 against a full sort, to FP32 rounding), and -1 past the system's
    other particles."""
    from libmobility_amd import NBody, _ffi
    from libmobility_amd.workloads import rsa_cube
    rng = np.random.default_rng(4)
    sizes, kinds = None, None
    if geom == "cube":
        pos = rsa_cube(2500, 0.2, seed=2)[0]
    elif geom == "slab":
        pos = np.column_stack([rng.uniform(0, 120, (3000, 2)), rng.uniform(1.5, 9.0, 3000)])
    elif geom == "clusters":
        pos = np.vstack([rng.normal(0, 2, (1500, 3)), rng.normal(80, 6, (400, 3)),
                         rng.uniform(-50, 150, (20, 3))])
    else:
        sizes, kinds = [5, 900, 0, 1700, 17, 1], [0, 1, 0, 0, 1, 0]
        parts = []
        for n, k in zip(sizes, kinds):
            p = rng.uniform(0, (n / 0.02) ** (1 / 3) + 1, (n, 3))
            if k:
                p[:, 2] = rng.uniform(1.5, 6.0, n)
            parts.append(p)
        pos = np.concatenate(parts)
    wall = geom == "slab"
    s = NBody("open", "open", "single_wall" if (wall or sizes) else "open")
    s.setParameters(wallHeight=0.0) if (wall or sizes) else s.setParameters()
    s.initialize(viscosity=1.0 / (6 * np.pi), hydrodynamicRadius=1.0)
    if sizes:
        s.setSystems(sizes, ["single_wall" if k else "open" for k in kinds])
    s.setPositions(pos)
    N = pos.shape[0]
    nbr = np.zeros((N, 16), dtype=np.int32)
    _ffi.check(_ffi.load().lm_nbody_precond_neighbours(s._h, nbr.ctypes.data))
    off = np.concatenate([[0], np.cumsum(sizes if sizes else [N])])
    for a, b in zip(off[:-1], off[1:]):
        x = pos[a:b]
        n = b - a
        for i in range(n):
            d = np.sqrt(((x - x[i]) ** 2).sum(1))
            d[i] = np.inf
            ref = np.sort(d)[:min(16, n - 1)]
            got = nbr[a + i]
            m = min(16, n - 1)
            assert np.all(got[m:] == -1), (a + i, got)
            assert np.all((got[:m] >= a) & (got[:m] < b)), (a + i, got)
            dg = np.sqrt(((pos[got[:m]] - x[i]) ** 2).sum(1))
            assert np.allclose(dg, ref, rtol=1e-5, atol=1e-6), (a + i, dg, ref)
