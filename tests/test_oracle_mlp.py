"""Pins the numpy network oracle (oracle/mlp.py): its hand-written backward
pass against central finite differences, and G = L L^T is SPD, for 3x3 and
6x6 blocks."""
import numpy as np
import pytest

from oracle import mlp


def _setup(d, seed=0, B=5):
    rng = np.random.default_rng(seed)
    w = rng.normal(scale=0.15, size=mlp.num_weights(d))
    X = rng.normal(scale=2.0, size=(B, mlp.NIN))
    A = rng.normal(size=(B, d, d))
    T = A @ np.transpose(A, (0, 2, 1)) / d + np.eye(d)
    return w, X, T.reshape(B, d * d)


@pytest.mark.parametrize("d", [3, 6])
def test_backward_matches_finite_differences(d):
    w, X, T = _setup(d)
    l, g = mlp.loss_grad(w, X, T, d)
    assert np.isclose(l, mlp.loss(w, X, T, d), rtol=1e-12)
    n = mlp.num_weights(d)
    rng = np.random.default_rng(1)
    idx = rng.choice(n, size=60, replace=False)
    # include every output weight's bias and a few first-layer biases
    idx = np.unique(np.concatenate([idx, np.arange(n - mlp.nout(d), n),
                                    np.arange(mlp.NIN * mlp.NH, mlp.NIN * mlp.NH + 5)]))
    h = 1e-5  # the 6x6 loss is ~2e4 here: a smaller step meets rounding
    for k in idx:
        wp = w.copy()
        wp[k] += h
        wm = w.copy()
        wm[k] -= h
        fd = (mlp.loss(wp, X, T, d) - mlp.loss(wm, X, T, d)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


@pytest.mark.parametrize("d", [3, 6])
def test_blocks_are_spd(d):
    w, X, _ = _setup(d, seed=3, B=50)
    G = mlp.blocks(w, X, d)
    assert np.allclose(G, np.transpose(G, (0, 2, 1)))
    assert np.all(np.linalg.eigvalsh(G) > 0)


def test_factor_layout():
    """Outputs 0..d-1 are the diagonal (softplus), the rest the strictly
    lower entries in row order."""
    O = np.arange(21, dtype=np.float64)[None, :] + 30.0
    L = mlp.factor(O, 6)[0]
    assert np.allclose(np.diag(L), O[0, :6] + mlp.EPS)
    assert L[1, 0] == O[0, 6] and L[2, 0] == O[0, 7] and L[2, 1] == O[0, 8]
    assert L[5, 4] == O[0, 20] and np.all(np.triu(L, 1) == 0)
