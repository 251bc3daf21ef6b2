"""Synthetic particle configurations of the benchmark workloads.

All generators are seeded numpy code (host side), used by bench.py and the
tests to build inputs of the documented shapes:
  * ``uniform_cube``      i.i.d. uniform positions in a cube (overlaps allowed)
  * ``wall_slab``         x, y uniform in [0, Lxy), z uniform in [zmin, zmax]
  * ``rsa_cube``          random sequential addition of non-overlapping spheres
                          of radius a in a cube sized for volume fraction phi
  * ``rsa_box``           the same in a rectangular box
  * ``ensemble_c4``       512 ragged systems, half open, half above a wall

and the five benchmark configurations with their seeds (``config_c1`` ...
``config_c5``), shared by bench.py and the full-size tests.
"""
import math

import numpy as np


def uniform_cube(N, L, seed):
    return np.random.default_rng(seed).uniform(0.0, L, (N, 3))


def wall_slab(N, Lxy, zmin, zmax, seed):
    rng = np.random.default_rng(seed)
    pos = np.empty((N, 3))
    pos[:, :2] = rng.uniform(0.0, Lxy, (N, 2))
    pos[:, 2] = rng.uniform(zmin, zmax, N)
    return pos


def rsa_box_side(N, phi, a=1.0):
    return (N * 4.0 * math.pi * a ** 3 / 3.0 / phi) ** (1.0 / 3.0)


def rsa_cube(N, phi, seed, a=1.0, batch=None):
    """Non-overlapping spheres in a cube [0, L)^3 sized for volume fraction phi;
    returns (positions, L)."""
    L = rsa_box_side(N, phi, a)
    return rsa_box(N, (L, L, L), seed, a=a, batch=batch), L


def rsa_box(N, dims, seed, a=1.0, batch=None, origin=(0.0, 0.0, 0.0)):
    """Non-overlapping spheres (centre distance >= 2a) in origin + [0, dims).

    Candidates are drawn in batches; a candidate is kept when it is at least
    2a from every accepted sphere and from every lower-indexed candidate of its
    batch.  Uses a cell grid of side 2a/sqrt(3) (at most one centre per cell).
    """
    rng = np.random.default_rng(seed)
    dims = np.asarray(dims, dtype=np.float64)
    cell = 2.0 * a / math.sqrt(3.0)
    nc = np.maximum(1, np.ceil(dims / cell).astype(np.int64))
    grid = -np.ones(tuple(nc), dtype=np.int64)
    pos = np.empty((N, 3))
    count = 0
    offs = np.array([(i, j, k) for i in range(-2, 3) for j in range(-2, 3) for k in range(-2, 3)])
    d2min = (2.0 * a) ** 2
    batch = batch or max(1024, N // 8)
    guard = 0
    while count < N:
        guard += 1
        if guard > 10000:
            raise RuntimeError("rsa_box: volume fraction too high for sequential addition")
        B = min(batch, 2 * (N - count) + 64)
        cand = rng.uniform(0.0, 1.0, (B, 3)) * dims
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
        key = (ci[:, 0] * nc[1] + ci[:, 1]) * nc[2] + ci[:, 2]
        first = np.full(int(nc.prod()), nbat, dtype=np.int64)
        np.minimum.at(first, key, np.arange(nbat))
        keep = first[key] == np.arange(nbat)
        nb = ci[:, None, :] + offs[None, :, :]
        valid = np.all((nb >= 0) & (nb < nc), axis=2)
        nbc = np.clip(nb, 0, nc - 1)
        nkey = (nbc[..., 0] * nc[1] + nbc[..., 1]) * nc[2] + nbc[..., 2]
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
    return pos + np.asarray(origin, dtype=np.float64)


def ensemble_c4(nsys=512, nmin=64, nmax=8192, seed=2024, a=1.0):
    """Ragged ensemble: N_s log-uniform in [nmin, nmax], even systems open
    (RSA cube), odd systems above a wall at z = 0 (RSA slab, z in [1.5a, 10a]),
    volume fraction uniform in [0.05, 0.25].  Returns (sizes, kinds, positions)."""
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(math.log(nmin), math.log(nmax), nsys))).astype(np.int64)
    phis = rng.uniform(0.05, 0.25, nsys)
    kinds = np.arange(nsys) % 2
    pos = []
    for s, (n, phi, k) in enumerate(zip(sizes, phis, kinds)):
        vol = n * 4.0 * math.pi * a ** 3 / 3.0 / phi
        if k:
            h = 8.5 * a
            lxy = math.sqrt(vol / h)
            pos.append(rsa_box(int(n), (lxy, lxy, h), seed * 7 + s, a=a, origin=(0.0, 0.0, 1.5 * a)))
        else:
            side = vol ** (1.0 / 3.0)
            pos.append(rsa_box(int(n), (side, side, side), seed * 7 + s, a=a))
    return sizes, kinds, np.concatenate(pos)


# ---- the five benchmark configurations (a = 1, seeds fixed) --------------------


def config_c1(N=1024):
    """Open RPY, N i.i.d. uniform in a cube of side 24a (overlaps kept); forces
    and noise N(0,1).  Returns (positions, forces, noise)."""
    rng = np.random.default_rng(0)
    pos = rng.uniform(0, 24.0, (N, 3))
    return pos, rng.normal(size=(N, 3)), rng.normal(size=3 * N)


def config_c2(N=65536):
    """Single wall, x, y uniform in [0, 256a), z uniform in [1.5a, 20a]; forces
    and torques N(0,1).  Returns (positions, forces, torques, noise6N)."""
    pos = wall_slab(N, 256.0, 1.5, 20.0, seed=65536)
    rng = np.random.default_rng(8)
    F = rng.normal(size=(N, 3))
    T = rng.normal(size=(N, 3))
    return pos, F, T, rng.normal(size=6 * N)


def config_c3(N=131072, phi=0.20):
    """Open RPY, RSA packing at volume fraction phi in a cube; forces and noise
    N(0,1).  Returns (positions, side, forces, noise)."""
    pos, L = rsa_cube(N, phi, seed=2026)
    rng = np.random.default_rng(7)
    return pos, L, rng.normal(size=(N, 3)), rng.normal(size=3 * N)


def config_c4():
    """512 ragged systems (``ensemble_c4``); forces and noise N(0,1).  Returns
    (sizes, kinds, positions, forces, noise)."""
    sizes, kinds, pos = ensemble_c4()
    N = int(sizes.sum())
    rng = np.random.default_rng(9)
    return sizes, kinds, pos, rng.normal(size=(N, 3)), rng.normal(size=3 * N)


def config_c5(N=1048576):
    """Single wall, x, y uniform in [0, 1024a sqrt(N/2^20)), z uniform in
    [1.5a, 30a]; forces and torques N(0,1).  Returns (positions, forces,
    torques, Lxy)."""
    Lxy = 1024.0 * math.sqrt(N / 1048576.0)
    pos = wall_slab(N, Lxy, 1.5, 30.0, seed=1048576)
    rng = np.random.default_rng(10)
    return pos, rng.normal(size=(N, 3)), rng.normal(size=(N, 3)), Lxy
