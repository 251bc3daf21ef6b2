"""Synthetic particle configurations of the benchmark workloads.

All generators are seeded numpy code (host side), used by bench.py and the
tests to build inputs of the documented shapes:
  * ``uniform_cube``      i.i.d. uniform positions in a cube (overlaps allowed)
  * ``wall_slab``         x, y uniform in [0, Lxy), z uniform in [zmin, zmax]
  * ``rsa_cube``          random sequential addition of non-overlapping spheres
                          of radius a in a cube sized for volume fraction phi
"""
import math

import numpy as np


def uniform_cube(N, L, seed):
    return np.random.default_rng(seed).uniform(0.0, L, (N, 3))


def wall_slab(N, Lxy, zmin, zmax, seed):
    rng = np.random.default_rng(seed)
    pos = np.empty((N, 3))
    pos[:, 0] = rng.uniform(0.0, Lxy, N)
    pos[:, 1] = rng.uniform(0.0, Lxy, N)
    pos[:, 2] = rng.uniform(zmin, zmax, N)
    return pos


def rsa_box_side(N, phi, a=1.0):
    return (N * 4.0 * math.pi * a ** 3 / 3.0 / phi) ** (1.0 / 3.0)


def rsa_cube(N, phi, seed, a=1.0, batch=None):
    """Non-overlapping spheres (centre distance >= 2a) in [0, L)^3.

    Candidates are drawn in batches; a candidate is kept when it is at least
    2a from every accepted sphere and from every lower-indexed candidate of its
    batch.  Uses a cell grid of side 2a/sqrt(3) (at most one centre per cell).
    """
    rng = np.random.default_rng(seed)
    L = rsa_box_side(N, phi, a)
    cell = 2.0 * a / math.sqrt(3.0)
    nc = max(1, int(math.ceil(L / cell)))
    grid = -np.ones((nc, nc, nc), dtype=np.int64)
    pos = np.empty((N, 3))
    count = 0
    offs = np.array([(i, j, k) for i in range(-2, 3) for j in range(-2, 3) for k in range(-2, 3)])
    d2min = (2.0 * a) ** 2
    batch = batch or max(1024, N // 8)
    guard = 0
    while count < N:
        guard += 1
        if guard > 10000:
            raise RuntimeError("rsa_cube: volume fraction too high for sequential addition")
        B = min(batch, 2 * (N - count) + 64)
        cand = rng.uniform(0.0, L, (B, 3))
        ci = np.minimum((cand / cell).astype(np.int64), nc - 1)
        # against accepted spheres
        nb = ci[:, None, :] + offs[None, :, :]
        valid = np.all((nb >= 0) & (nb < nc), axis=2)
        nbc = np.clip(nb, 0, nc - 1)
        ids = grid[nbc[..., 0], nbc[..., 1], nbc[..., 2]]
        ids = np.where(valid, ids, -1)
        has = ids >= 0
        d = cand[:, None, :] - pos[np.maximum(ids, 0)]
        close = has & (np.einsum("bkc,bkc->bk", d, d) < d2min)
        ok = ~close.any(axis=1)
        cand = cand[ok]
        ci = ci[ok]
        if cand.shape[0] == 0:
            continue
        # Within the batch: drop every candidate that conflicts with a
        # lower-indexed one (two centres in one cell always conflict).
        nbat = cand.shape[0]
        key = (ci[:, 0] * nc + ci[:, 1]) * nc + ci[:, 2]
        first = np.full(nc * nc * nc, nbat, dtype=np.int64)
        np.minimum.at(first, key, np.arange(nbat))
        keep = first[key] == np.arange(nbat)
        nb = ci[:, None, :] + offs[None, :, :]
        valid = np.all((nb >= 0) & (nb < nc), axis=2)
        nbc = np.clip(nb, 0, nc - 1)
        nkey = (nbc[..., 0] * nc + nbc[..., 1]) * nc + nbc[..., 2]
        other = np.where(valid, first[nkey], nbat)
        lower = other < np.arange(nbat)[:, None]
        d = cand[:, None, :] - cand[np.minimum(other, nbat - 1)]
        clash = lower & (np.einsum("bkc,bkc->bk", d, d) < d2min)
        keep &= ~clash.any(axis=1)
        cand = cand[keep]
        ci = ci[keep]
        take = min(cand.shape[0], N - count)
        cand = cand[:take]
        ci = ci[:take]
        pos[count:count + take] = cand
        grid[ci[:, 0], ci[:, 1], ci[:, 2]] = np.arange(count, count + take)
        count += take
    return pos, L
