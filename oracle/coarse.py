"""TEST INFRASTRUCTURE ONLY -- numpy restatement of the coarse level of the
two-level sqrtMdotW preconditioner.  The reference has no counterpart: its
sqrtMdotW is plain Lanczos (include/lanczos/LanczosAlgorithm.cu:26-232),
which ``oracle.LanczosOracle`` restates and which this module reuses on the
preconditioned operator.

Only tests/ may import this module.

Definitions (one system, positions in units of a):
  * cells: g_k = max(1, round(max(ext_k, 1) / s)) cells along axis k, with
    s = cbrt(prod_k max(ext_k, 1) / max(1, N / ppc)) grown by 3 % until the
    non-empty cells C satisfy 3C <= max_dim; particle i is in cell
    q_k = min(g_k - 1, max(0, floor((x_k - lo_k) / ext_k * g_k))); empty cells
    are dropped and the rest numbered in raw order (x slowest).
  * Q (3N x 3C): Q[3i+a, 3c+a] = 1 / sqrt(n_c).
  * Mc = Q^T M Q = L L^T (M: translational block), X = L^{-1}.
  * Gc = I + Q (X - I) Q^T,  Gc^{-1} = I + Q (L - I) Q^T,  G = D Gc.
  * B W = G^{-1} (G M G^T)^{1/2} W, so B B^T = M.

Pinning (tests/test_oracle_coarse.py): the cell rule against a brute-force
box test; Mc against explicit dense Q; Gc Gc^{-1} = I; Q^T (Gc M Gc^T) Q = I;
B B^T = M from dense columns; the iteration count drops against plain
Lanczos on an RSA packing.
"""
import numpy as np

from . import LanczosOracle


def coarse_cells(xyz, ppc, max_dim=24576):
    """(cell_of (N,), grid (3,), C) by the rule in the module docstring."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    N = xyz.shape[0]
    lo = xyz.min(axis=0)
    hi = xyz.max(axis=0)
    ext = hi - lo
    vol = 1.0
    for k in range(3):
        vol *= max(ext[k], 1.0)
    want = max(1.0, N / max(ppc, 1))
    side = np.cbrt(vol / want)
    while True:
        g = np.array([max(1, int(np.floor(max(ext[k], 1.0) / side + 0.5))) for k in range(3)])
        q = np.zeros((N, 3), dtype=np.int64)
        for k in range(3):
            if ext[k] > 0.0:
                f = np.floor((xyz[:, k] - lo[k]) / ext[k] * g[k])
                q[:, k] = np.minimum(g[k] - 1, np.maximum(0.0, f)).astype(np.int64)
        raw = (q[:, 0] * g[1] + q[:, 1]) * g[2] + q[:, 2]
        used = np.unique(raw)
        if 3 * used.size <= max_dim:
            break
        side *= 1.03
    cell_of = np.searchsorted(used, raw)
    return cell_of.astype(np.int32), g, used.size


def coarse_Q(cell_of, C):
    """Dense Q (3N x 3C)."""
    cell_of = np.asarray(cell_of)
    N = cell_of.size
    counts = np.bincount(cell_of, minlength=C)
    Q = np.zeros((3 * N, 3 * C))
    for a in range(3):
        Q[3 * np.arange(N) + a, 3 * cell_of + a] = 1.0 / np.sqrt(counts[cell_of])
    return Q


def coarse_matrix(M, cell_of, C):
    """Q^T M Q for the translational block M (3N x 3N), by cell-block sums."""
    cell_of = np.asarray(cell_of)
    N = cell_of.size
    counts = np.bincount(cell_of, minlength=C).astype(np.float64)
    S = np.zeros((3 * N, 3 * C))
    # columns: sum of M over the particles of each cell
    for a in range(3):
        np.add.at(S.T, 3 * cell_of + a, M[:, 3 * np.arange(N) + a].T)
    Mc = np.zeros((3 * C, 3 * C))
    for a in range(3):
        np.add.at(Mc, 3 * cell_of + a, S[3 * np.arange(N) + a, :])
    sc = np.repeat(1.0 / np.sqrt(counts), 3)
    return Mc * sc[:, None] * sc[None, :]


def coarse_factors(Mc):
    """(L, X = L^{-1})."""
    L = np.linalg.cholesky(Mc)
    X = np.linalg.solve(L, np.eye(L.shape[0]))
    return L, np.tril(X)


class TwoLevel:
    """G = D Gc on vectors of length n >= 3N (entries past 3N pass through)."""

    def __init__(self, Q, L, X=None, D=None):
        self.Q = Q
        self.L = L
        self.X = np.linalg.solve(L, np.eye(L.shape[0])) if X is None else X
        self.D = D  # (N, 3, 3) or None
        self.n3 = Q.shape[0]

    def _coarse(self, T, v):
        v = v.copy()
        c = self.Q.T @ v[: self.n3]
        v[: self.n3] += self.Q @ (T @ c - c)
        return v

    def _blocks(self, B, v):
        v = v.copy()
        x = v[: self.n3].reshape(-1, 3)
        v[: self.n3] = np.einsum("nij,nj->ni", B, x).ravel()
        return v

    def forward(self, v):  # G^T v = Gc^T D v
        if self.D is not None:
            v = self._blocks(self.D, v)
        return self._coarse(self.X.T, v)

    def back(self, u):  # G u = D Gc u
        u = self._coarse(self.X, u)
        return self._blocks(self.D, u) if self.D is not None else u

    def inverse(self, y):  # G^{-1} y = Gc^{-1} D^{-1} y
        if self.D is not None:
            y = self._blocks(np.linalg.inv(self.D), y)
        return self._coarse(self.L, y)

    def sqrt(self, M, W, tol, limit=1000):
        """(G^{-1} (G M G^T)^{1/2} W, iterations) with the Lanczos oracle."""
        lz = LanczosOracle(limit)
        y, it = lz.run(lambda v: self.back(M @ self.forward(v)), W, tol)
        return self.inverse(y), it
