"""GPU checks of the learned block preconditioner (no reference counterpart).

  * kNN features vs scipy's KD-tree;
  * FP64 local targets vs the oracle's dense matrix and eigh inverse sqrt;
  * MFMA forward pass and gradients vs the numpy oracle (oracle/mlp.py),
    tolerance for BF16 operands with FP32 accumulation;
  * training reduces the loss;
  * preconditioned sqrtMdotW is a square root of M: ||B B^T - M|| / ||M||
    <= 1e-5 at small N, and its fluctuations have covariance M.
"""
import numpy as np
import pytest

import oracle
from oracle import mlp

pytestmark = pytest.mark.gpu


def _open_solver(pos, tol=1e-6, eta=1.0 / (6 * np.pi)):
    from libmobility_amd import NBody
    s = NBody("open", "open", "open")
    s.setParameters()
    s.initialize(viscosity=eta, hydrodynamicRadius=1.0, tolerance=tol)
    s.setPositions(pos)
    return s


def test_knn_and_local_targets():
    from libmobility_amd import _ffi
    from libmobility_amd.workloads import rsa_cube
    from scipy.spatial import cKDTree
    pos, _ = rsa_cube(3000, 0.2, seed=11)
    s = _open_solver(pos)
    nc = 64
    X = np.zeros((nc, 48), dtype=np.float32)
    T = np.zeros((nc, 9), dtype=np.float32)
    _ffi.check(_ffi.load().lm_nbody_precond_dataset(s._h, nc, 5, X.ctypes.data, T.ctypes.data))
    tree = cKDTree(pos)
    for c in range(nc):
        d = np.linalg.norm(X[c].reshape(16, 3).astype(np.float64), axis=1)
        assert np.all(np.diff(d) >= -1e-5)  # sorted by distance
    # match centers through their first neighbour offset
    dd, nn = tree.query(pos, k=17)
    first = pos[nn[:, 1]] - pos
    for c in range(nc):
        rel = X[c].reshape(16, 3).astype(np.float64)
        i = int(np.argmin(np.linalg.norm(first - rel[0], axis=1) + np.abs(dd[:, 16] - np.linalg.norm(rel[15]))))
        assert np.allclose(pos[nn[i, 1:]] - pos[i], rel, atol=1e-5)
        cl = pos[nn[i]]
        M = oracle.dense_matrix(cl, 1.0 / (6 * np.pi), 1.0)  # t0 = 1
        lam, V = np.linalg.eigh(M)
        Mi = (V / np.sqrt(lam)) @ V.T
        assert np.allclose(T[c].reshape(3, 3), Mi[:3, :3], rtol=0, atol=2e-6)


def _net(seed=3):
    from libmobility_amd.precond import LearnedPreconditioner
    return LearnedPreconditioner(seed=seed)


def test_forward_and_gradients_vs_numpy_oracle():
    rng = np.random.default_rng(0)
    B = 300
    X = rng.normal(scale=2.0, size=(B, 48)).astype(np.float32)
    A = rng.normal(size=(B, 3, 3))
    T = (A @ np.transpose(A, (0, 2, 1)) * 0.1 + 1.3 * np.eye(3)).reshape(B, 9).astype(np.float32)
    net = _net()
    w = net.get_weights().astype(np.float64)
    out = net.forward(X)
    ref, _ = mlp.forward(w, X)
    # BF16 operands (8-bit mantissa), FP32 sums
    assert np.linalg.norm(out - ref) <= 2e-2 * np.linalg.norm(ref)
    loss, g = net.loss_grad(X, T)
    lref, gref = mlp.loss_grad(w, X, T)
    assert abs(loss - lref) <= 3e-2 * abs(lref)
    for name, sl in (("W1", slice(0, 48 * 128)), ("W2", slice(48 * 128 + 128, 48 * 128 + 128 + 128 * 128)),
                     ("all", slice(None))):
        assert np.linalg.norm(g[sl] - gref[sl]) <= 5e-2 * np.linalg.norm(gref[sl]), name
    cos = g @ gref / (np.linalg.norm(g) * np.linalg.norm(gref))
    assert cos > 0.999


def test_training_reduces_loss():
    from libmobility_amd.precond import make_training_set
    X, T = make_training_set(8192, N=20000, seed=2, n_configs=1)
    net = _net(seed=4)
    l0, _ = net.loss_grad(X, T)
    hist = net.train(X, T, epochs=15, batch=1024, lr=2e-3, seed=1)
    l1, _ = net.loss_grad(X, T)
    assert l1 < 0.5 * l0 and hist[-1] < hist[0]


@pytest.mark.parametrize("trained", [False, True])
def test_preconditioned_sqrt_is_a_square_root(trained):
    from libmobility_amd.precond import DEFAULT_WEIGHTS, LearnedPreconditioner
    import os
    from libmobility_amd.workloads import rsa_cube
    if trained and not os.path.exists(DEFAULT_WEIGHTS):
        pytest.skip("trained weights not packaged")
    pos, _ = rsa_cube(48, 0.2, seed=4)
    s = _open_solver(pos, tol=1e-8)
    net = LearnedPreconditioner(weights=DEFAULT_WEIGHTS) if trained else _net(seed=9)
    s.setPreconditioner(net)
    n = 3 * 48
    Bm = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        u, _ = s.sqrtMdotW_with_noise(e)
        Bm[:, k] = u.ravel()
    M = oracle.dense_matrix(pos, 1.0 / (6 * np.pi), 1.0)
    rel = np.linalg.norm(Bm @ Bm.T - M) / np.linalg.norm(M)
    assert rel <= 1e-5, rel
    G = s.preconditionerBlocks()
    assert np.allclose(G, np.transpose(G, (0, 2, 1)))
    assert np.all(np.linalg.eigvalsh(G) > 0)


def test_preconditioned_fluctuation_dissipation():
    from libmobility_amd.workloads import rsa_cube
    pos, _ = rsa_cube(20, 0.2, seed=6)
    s = _open_solver(pos, tol=1e-8)
    s.initialize(viscosity=1.0 / (6 * np.pi), hydrodynamicRadius=1.0, tolerance=1e-8, seed=77)
    s.setPositions(pos)
    s.setPreconditioner(_net(seed=5))
    M = oracle.dense_matrix(pos, 1.0 / (6 * np.pi), 1.0)
    ns = 3000
    X = np.array([s.sqrtMdotW()[0].ravel() for _ in range(ns)])
    lam, V = np.linalg.eigh(M)
    Z = (X @ V) / np.sqrt(lam)
    cov = np.cov(Z.T)
    assert np.abs(cov - np.eye(60)).max() < 6.0 * np.sqrt(2.0 / ns)


def test_ablation_trained_network_vs_plain():
    """What the trained network gains over plain Lanczos (and over the same
    network untrained) on an RSA packing (N = 2048, phi = 0.2, tol 1e-6).  The blocks act on the
    local (near-field) part of the spectrum; the iteration count of this
    operator is set by its collective long-wavelength modes, so the expected
    gain is ~1x.  Asserted: the trained blocks are closer to their targets
    than the untrained ones, and the trained preconditioner costs at most 20 %
    more iterations than plain Lanczos (that the preconditioned result is a
    square root of M is tested separately)."""
    import os
    from libmobility_amd import _ffi
    from libmobility_amd.precond import DEFAULT_WEIGHTS, LearnedPreconditioner
    from libmobility_amd.workloads import rsa_cube
    if not os.path.exists(DEFAULT_WEIGHTS):
        pytest.skip("trained weights not packaged")
    pos, _ = rsa_cube(2048, 0.2, seed=21)
    rng = np.random.default_rng(0)
    Ws = [rng.normal(size=3 * 2048) for _ in range(3)]

    def run(net):
        s = _open_solver(pos, tol=1e-6)
        if net is not None:
            s.setPreconditioner(net)
        its = []
        for w in Ws:
            s.sqrtMdotW_with_noise(w)
            its.append(s.lastLanczosIterations())
        return its, s

    plain, _ = run(None)
    trained_net = LearnedPreconditioner(weights=DEFAULT_WEIGHTS)
    trained, st = run(trained_net)
    untrained, _ = run(LearnedPreconditioner(seed=123))
    # block error against the FP64 targets on this configuration
    X = np.zeros((512, 48), dtype=np.float32)
    T = np.zeros((512, 9), dtype=np.float32)
    _ffi.check(_ffi.load().lm_nbody_precond_dataset(st._h, 512, 3, X.ctypes.data, T.ctypes.data))
    err_t, _ = trained_net.loss_grad(X, T)
    err_u, _ = LearnedPreconditioner(seed=123).loss_grad(X, T)
    print(f"[ablation] iterations plain {plain} trained {trained} untrained {untrained}; "
          f"block loss trained {err_t:.3e} untrained {err_u:.3e}")
    assert err_t < 0.5 * err_u
    assert np.mean(trained) <= 1.2 * np.mean(plain)


def test_ablation_with_the_coarse_level():
    """What the trained blocks add to the coarse level (RSA packing, N = 4096,
    phi = 0.2, tol 1e-6, 48 particles per cell): the coarse level sets the
    iteration count (well under plain Lanczos); the trained blocks on top
    change it by at most 20 %, and both stay exact square roots (tested
    separately)."""
    import os
    from libmobility_amd.precond import DEFAULT_WEIGHTS, LearnedPreconditioner
    from libmobility_amd.workloads import rsa_cube
    if not os.path.exists(DEFAULT_WEIGHTS):
        pytest.skip("trained weights not packaged")
    N = 4096
    pos, _ = rsa_cube(N, 0.2, seed=22)
    rng = np.random.default_rng(1)
    Ws = [rng.normal(size=3 * N) for _ in range(2)]

    def run(net, ppc):
        s = _open_solver(pos, tol=1e-6)
        if net is not None:
            s.setPreconditioner(net)
        s.setCoarseLevel(ppc)
        its = []
        for w in Ws:
            s.sqrtMdotW_with_noise(w)
            its.append(s.lastLanczosIterations())
        return its

    plain = run(None, 0)
    coarse_only = run(None, 48)
    two_level = run(LearnedPreconditioner(weights=DEFAULT_WEIGHTS), 48)
    print(f"[ablation] iterations plain {plain} coarse level {coarse_only} "
          f"trained blocks + coarse level {two_level}")
    assert np.mean(coarse_only) <= 0.75 * np.mean(plain)
    assert np.mean(two_level) <= 1.2 * np.mean(coarse_only)


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
