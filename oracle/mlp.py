"""TEST INFRASTRUCTURE ONLY -- numpy restatement of the learned block
preconditioner's network, loss and gradient.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import this module.  The network has no reference counterpart; this file is
the CPU check of the GPU forward/backward kernels.  It is pinned by central
finite differences (tests/test_oracle_mlp.py).

Layout of the flat weight vector: W1[48][128] b1[128] W2[128][128] b2[128]
W3[128][6] b3[6]; G = L L^T with L = [[s0,0,0],[o3,s1,0],[o4,o5,s2]],
s_i = softplus(o_i) + 1e-3; loss = mean_b sum_ij (G_b - T_b)_ij^2.
"""
import numpy as np
from scipy.special import erf

NIN, NH, NOUT = 48, 128, 6
EPS = 1e-3
SHAPES = [("W1", (NIN, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
          ("W3", (NH, NOUT)), ("b3", (NOUT,))]
NUM_WEIGHTS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(w):
    out, o = {}, 0
    for name, shp in SHAPES:
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


def forward(w, X):
    p = unpack(w)
    X = np.asarray(X, dtype=np.float64)
    Z1 = X @ p["W1"] + p["b1"]
    H1 = gelu(Z1)
    Z2 = H1 @ p["W2"] + p["b2"]
    H2 = gelu(Z2)
    O = H2 @ p["W3"] + p["b3"]
    return O, (X, Z1, H1, Z2, H2)


def factor(O):
    B = O.shape[0]
    L = np.zeros((B, 3, 3))
    L[:, 0, 0] = softplus(O[:, 0]) + EPS
    L[:, 1, 1] = softplus(O[:, 1]) + EPS
    L[:, 2, 2] = softplus(O[:, 2]) + EPS
    L[:, 1, 0] = O[:, 3]
    L[:, 2, 0] = O[:, 4]
    L[:, 2, 1] = O[:, 5]
    return L


def blocks(w, X):
    L = factor(forward(w, X)[0])
    return L @ np.transpose(L, (0, 2, 1))


def loss(w, X, T9):
    G = blocks(w, X)
    T = np.asarray(T9, dtype=np.float64).reshape(-1, 3, 3)
    return float(np.mean(np.sum((G - T) ** 2, axis=(1, 2))))


def loss_grad(w, X, T9):
    O, (X, Z1, H1, Z2, H2) = forward(w, X)
    p = unpack(w)
    B = O.shape[0]
    L = factor(O)
    G = L @ np.transpose(L, (0, 2, 1))
    T = np.asarray(T9, dtype=np.float64).reshape(-1, 3, 3)
    R = G - T
    l = float(np.mean(np.sum(R * R, axis=(1, 2))))
    dG = 2.0 * R / B                        # d loss / d G
    dL = (dG + np.transpose(dG, (0, 2, 1))) @ L
    sig = lambda x: 1.0 / (1.0 + np.exp(-x))
    dO = np.zeros_like(O)
    dO[:, 0] = dL[:, 0, 0] * sig(O[:, 0])
    dO[:, 1] = dL[:, 1, 1] * sig(O[:, 1])
    dO[:, 2] = dL[:, 2, 2] * sig(O[:, 2])
    dO[:, 3] = dL[:, 1, 0]
    dO[:, 4] = dL[:, 2, 0]
    dO[:, 5] = dL[:, 2, 1]
    g = {}
    g["W3"] = H2.T @ dO
    g["b3"] = dO.sum(0)
    dZ2 = (dO @ p["W3"].T) * gelu_prime(Z2)
    g["W2"] = H1.T @ dZ2
    g["b2"] = dZ2.sum(0)
    dZ1 = (dZ2 @ p["W2"].T) * gelu_prime(Z1)
    g["W1"] = X.T @ dZ1
    g["b1"] = dZ1.sum(0)
    return l, np.concatenate([g[name].ravel() for name, _ in SHAPES])
