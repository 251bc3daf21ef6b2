"""Pins the numpy network oracle (oracle/mlp.py): its hand-written backward
pass against central finite differences, and G = L L^T is SPD."""
import numpy as np

from oracle import mlp


def _setup(seed=0, B=5):
    rng = np.random.default_rng(seed)
    w = rng.normal(scale=0.15, size=mlp.NUM_WEIGHTS)
    X = rng.normal(scale=2.0, size=(B, mlp.NIN))
    A = rng.normal(size=(B, 3, 3))
    T = A @ np.transpose(A, (0, 2, 1)) + np.eye(3)
    return w, X, T.reshape(B, 9)


def test_backward_matches_finite_differences():
    w, X, T = _setup()
    l, g = mlp.loss_grad(w, X, T)
    assert np.isclose(l, mlp.loss(w, X, T), rtol=1e-12)
    rng = np.random.default_rng(1)
    idx = rng.choice(mlp.NUM_WEIGHTS, size=60, replace=False)
    # include every bias and the output layer
    idx = np.unique(np.concatenate([idx, np.arange(mlp.NUM_WEIGHTS - 6, mlp.NUM_WEIGHTS),
                                    np.arange(48 * 128, 48 * 128 + 5)]))
    h = 1e-6
    for k in idx:
        wp = w.copy(); wp[k] += h
        wm = w.copy(); wm[k] -= h
        fd = (mlp.loss(wp, X, T) - mlp.loss(wm, X, T)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_blocks_are_spd():
    w, X, _ = _setup(seed=3, B=50)
    G = mlp.blocks(w, X)
    assert np.allclose(G, np.transpose(G, (0, 2, 1)))
    assert np.all(np.linalg.eigvalsh(G) > 0)
