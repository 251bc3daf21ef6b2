"""CPU checks that pin the coarse-level oracle (oracle/coarse.py).

  * the cell rule against a brute-force box test;
  * Q^T M Q by block sums against explicit dense Q;
  * Gc Gc^{-1} = I and Q^T (Gc M Gc^T) Q = I (the coarse space is whitened);
  * G^{-1} (G M G^T)^{1/2} is a square root of M (dense columns), with and
    without blocks D;
  * the coarse level lowers the Lanczos iteration count on an RSA packing.
"""
import numpy as np

import oracle
from oracle import coarse
from libmobility_amd.workloads import rsa_cube

ETA = 1.0 / (6 * np.pi)


def test_cell_rule_brute_force():
    rng = np.random.default_rng(1)
    xyz = rng.uniform(size=(2000, 3)) * np.array([40.0, 25.0, 7.0]) + np.array([-3.0, 5.0, 1.0])
    cell_of, g, C = coarse.coarse_cells(xyz, 16)
    assert cell_of.shape == (2000,) and C == cell_of.max() + 1
    assert np.all(np.bincount(cell_of) > 0)
    # about N / ppc cells, shaped after the box
    assert 0.5 * 2000 / 16 <= np.prod(g) <= 2.0 * 2000 / 16
    assert g[0] > g[1] > g[2]
    lo, hi = xyz.min(0), xyz.max(0)
    edges = [np.linspace(lo[k], hi[k], g[k] + 1) for k in range(3)]
    raw = np.zeros(2000, dtype=np.int64)
    for k in range(3):
        q = np.clip(np.searchsorted(edges[k], xyz[:, k], side="right") - 1, 0, g[k] - 1)
        raw = raw * g[k] + q
    # same partition (compact numbering in raw order)
    assert np.array_equal(np.searchsorted(np.unique(raw), raw), cell_of)
    # the coarse dimension is capped
    _, _, C2 = coarse.coarse_cells(xyz, 1, max_dim=300)
    assert 3 * C2 <= 300


def test_coarse_matrix_equals_dense_product():
    rng = np.random.default_rng(2)
    N = 60
    A = rng.normal(size=(3 * N, 3 * N))
    M = A @ A.T
    cell_of = rng.integers(0, 7, size=N)
    cell_of[:7] = np.arange(7)
    Q = coarse.coarse_Q(cell_of, 7)
    assert np.allclose(Q.T @ Q, np.eye(21))
    assert np.allclose(coarse.coarse_matrix(M, cell_of, 7), Q.T @ M @ Q, rtol=1e-12, atol=1e-10)


def _setup(N, ppc, seed):
    pos, _ = rsa_cube(N, 0.2, seed=seed)
    M = oracle.dense_matrix(pos, ETA, 1.0)
    cell_of, _, C = coarse.coarse_cells(pos, ppc)
    Q = coarse.coarse_Q(cell_of, C)
    L, X = coarse.coarse_factors(coarse.coarse_matrix(M, cell_of, C))
    return pos, M, Q, L, X


def test_whitening_and_exact_inverse():
    _, M, Q, L, X = _setup(64, 8, 5)
    T = coarse.TwoLevel(Q, L, X)
    n = M.shape[0]
    G = np.array([T.back(e) for e in np.eye(n)]).T
    Gi = np.array([T.inverse(e) for e in np.eye(n)]).T
    assert np.allclose(G @ Gi, np.eye(n), atol=1e-12)
    assert np.allclose(Q.T @ (G @ M @ G.T) @ Q, np.eye(Q.shape[1]), atol=1e-10)
    Gt = np.array([T.forward(e) for e in np.eye(n)]).T
    assert np.allclose(Gt, G.T, atol=1e-14)


def test_two_level_sqrt_is_a_square_root():
    pos, M, Q, L, X = _setup(30, 4, 6)
    rng = np.random.default_rng(3)
    A = rng.normal(size=(30, 3, 3)) * 0.2
    D = np.eye(3)[None] + np.einsum("nij,nkj->nik", A, A)  # SPD blocks
    n = 90
    for blocks in (None, D):
        T = coarse.TwoLevel(Q, L, X, D=blocks)
        B = np.array([T.sqrt(M, e, 1e-10)[0] for e in np.eye(n)]).T
        assert np.linalg.norm(B @ B.T - M) / np.linalg.norm(M) < 1e-8


def test_coarse_level_lowers_iterations():
    pos, M, Q, L, X = _setup(1024, 16, 7)
    W = np.random.default_rng(4).normal(size=M.shape[0])
    _, plain = oracle.LanczosOracle(1000).run(lambda v: M @ v, W, 1e-6)
    _, two = coarse.TwoLevel(Q, L, X).sqrt(M, W, 1e-6)
    assert two < 0.75 * plain, (two, plain)
