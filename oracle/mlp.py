"""TEST INFRASTRUCTURE ONLY -- numpy restatement of the learned block
preconditioner's network, loss and gradient.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import this module.  The network has no reference counterpart; this file is
the CPU check of the GPU forward/backward kernels.  It is pinned by central
finite differences (tests/test_oracle_mlp.py).

Block size d = 3 (6 outputs) or 6 (21 outputs).  Layout of the flat weight
vector: W1[64][128] b1[128] W2[128][128] b2[128] W3[128][out] b3[out];
G = L L^T with L lower triangular, L_ii = softplus(o_i) + 1e-3 and the
strictly lower entries o_d, o_{d+1}, ... in row order (L_10, L_20, L_21, ...);
loss = mean_b sum_ij (G_b - T_b)_ij^2.
"""
import numpy as np
from scipy.special import erf

NIN, NH = 64, 128
EPS = 1e-3


def nout(d):
    return d * (d + 1) // 2


def shapes(d=3):
    return [("W1", (NIN, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
            ("W3", (NH, nout(d))), ("b3", (nout(d),))]


def num_weights(d=3):
    return sum(int(np.prod(s)) for _, s in shapes(d))


NUM_WEIGHTS = num_weights(3)


def unpack(w, d=3):
    out, o = {}, 0
    for name, shp in shapes(d):
        n = int(np.prod(shp))
        out[name] = np.asarray(w[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_prime(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)


def softplus(x):
    return np.where(x > 20, x, np.log1p(np.exp(np.minimum(x, 20))))


def forward(w, X, d=3):
    p = unpack(w, d)
    X = np.asarray(X, dtype=np.float64)
    Z1 = X @ p["W1"] + p["b1"]
    H1 = gelu(Z1)
    Z2 = H1 @ p["W2"] + p["b2"]
    H2 = gelu(Z2)
    O = H2 @ p["W3"] + p["b3"]
    return O, (X, Z1, H1, Z2, H2)


def _lower(d):
    return [(i, j) for i in range(d) for j in range(i)]


def factor(O, d=3):
    B = O.shape[0]
    L = np.zeros((B, d, d))
    for i in range(d):
        L[:, i, i] = softplus(O[:, i]) + EPS
    for q, (i, j) in enumerate(_lower(d)):
        L[:, i, j] = O[:, d + q]
    return L


def blocks(w, X, d=3):
    L = factor(forward(w, X, d)[0], d)
    return L @ np.transpose(L, (0, 2, 1))


def loss(w, X, T, d=3):
    G = blocks(w, X, d)
    T = np.asarray(T, dtype=np.float64).reshape(-1, d, d)
    return float(np.mean(np.sum((G - T) ** 2, axis=(1, 2))))


def loss_grad(w, X, T, d=3):
    O, (X, Z1, H1, Z2, H2) = forward(w, X, d)
    p = unpack(w, d)
    B = O.shape[0]
    L = factor(O, d)
    G = L @ np.transpose(L, (0, 2, 1))
    T = np.asarray(T, dtype=np.float64).reshape(-1, d, d)
    R = G - T
    l = float(np.mean(np.sum(R * R, axis=(1, 2))))
    dG = 2.0 * R / B                        # d loss / d G
    dL = (dG + np.transpose(dG, (0, 2, 1))) @ L
    sig = lambda x: 1.0 / (1.0 + np.exp(-x))
    dO = np.zeros_like(O)
    for i in range(d):
        dO[:, i] = dL[:, i, i] * sig(O[:, i])
    for q, (i, j) in enumerate(_lower(d)):
        dO[:, d + q] = dL[:, i, j]
    g = {}
    g["W3"] = H2.T @ dO
    g["b3"] = dO.sum(0)
    dZ2 = (dO @ p["W3"].T) * gelu_prime(Z2)
    g["W2"] = H1.T @ dZ2
    g["b2"] = dZ2.sum(0)
    dZ1 = (dZ2 @ p["W2"].T) * gelu_prime(Z1)
    g["W1"] = X.T @ dZ1
    g["b1"] = dZ1.sum(0)
    return l, np.concatenate([g[name].ravel() for name, _ in shapes(d)])
