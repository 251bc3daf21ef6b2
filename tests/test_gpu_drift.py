"""GPU parity of the thermal drift divM (random finite differences) per sample.

``divM_with_noise`` runs NBody::divM (solvers/NBody/mobility.h:184-236,
include/MobilityInterface/random_finite_differences.h:30-63) with the random
displacement W given, so each sample is compared with the oracle's
``rfd_drift`` on the same W: (prefactor/delta) [M(q + delta/2 W) W -
M(q - delta/2 W) W], delta in units of the radius, accumulated into the
outputs.  The drift's two products run in FP64 pairs by default, whatever the
pair precision of the solver (FP32 pair terms lose ~1e-7 relative per product
and the difference is divided by delta).

Tolerance: 1e-9 relative in FP64 (rounding of two products ~1e-15, amplified
by 1/delta = 1e3 or 1e4).
"""
import ctypes

import numpy as np
import pytest

import oracle
from tests.test_gpu_mdot import _solver

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


def _slab(N, Lxy, zlo, zhi, seed):
    rng = np.random.default_rng(seed)
    pos = np.empty((N, 3))
    pos[:, :2] = rng.uniform(0, Lxy, (N, 2))
    pos[:, 2] = rng.uniform(zlo, zhi, N)
    return pos


@pytest.mark.parametrize("angular", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("wall_height", [0.0, 5.4])
def test_divm_with_noise_matches_oracle_rfd(angular, precision, wall_height):
    """One RFD sample against the oracle on the same W (near-wall heights from
    0.3a, so the damped region is crossed); the pair precision of the solver
    does not change the drift, which runs in FP64 pairs."""
    N, a, eta, delta = 300, 1.3, 0.8, 1e-3
    pos = _slab(N, 30.0, 0.3 * a, 6.0, seed=21)
    s = _solver(True, a, eta, angular=angular, precision=precision, wallHeight=wall_height,
                delta=delta)
    s.setPositions(pos + np.array([0, 0, wall_height]))
    W = np.random.default_rng(2).normal(size=3 * N)
    u, w = s.divM_with_noise(W, prefactor=0.7)
    lin, ang = oracle.rfd_drift(pos, W, eta, a, delta * a, prefactor=0.7, angular=angular)
    assert _rel(u, lin) <= 1e-9
    if angular:
        assert _rel(w, ang) <= 1e-9
    else:
        assert w is None


def test_divm_with_noise_is_independent_of_the_pair_precision():
    N, a = 500, 1.0
    pos = _slab(N, 60.0, 0.5, 8.0, seed=3)
    W = np.random.default_rng(9).normal(size=3 * N)
    out = []
    for prec in ("fp64", "fp32"):
        s = _solver(True, a, 1.0, angular=True, precision=prec, delta=1e-4)
        s.setPositions(pos)
        out.append(s.divM_with_noise(W))
    assert _rel(out[1][0], out[0][0]) <= 1e-13
    assert _rel(out[1][1], out[0][1]) <= 1e-13


def test_divm_fp32_drift_is_the_opt_in_and_stays_close_at_small_extent():
    """setDriftPrecision('fp32') runs the two products in FP32 pairs: at a
    system extent of 20a and delta = 1e-3 a the drift loses about 1e-4..1e-3
    relative (each product ~1e-7, divided by delta)."""
    N, a = 400, 1.0
    pos = _slab(N, 20.0, 1.2, 6.0, seed=5)
    W = np.random.default_rng(1).normal(size=3 * N)
    s = _solver(True, a, 1.0, precision="fp32", delta=1e-3)
    s.setDriftPrecision("fp32")
    s.setPositions(pos)
    u, _ = s.divM_with_noise(W)
    lin, _ = oracle.rfd_drift(pos, W, 1.0, a, 1e-3)
    assert _rel(u, lin) <= 1e-2


def test_divm_with_noise_at_wide_slab_small_step():
    """C5's slab width (1024a) and step (1e-4 a) at N = 8192: a displacement of
    5e-5 a is below the FP32 spacing at 512a from the centre, so only FP64
    records resolve it.  The default drift matches the oracle to 1e-8."""
    N, a = 8192, 1.0
    pos = _slab(N, 1024.0, 1.5, 30.0, seed=1048576)
    W = np.random.default_rng(4).normal(size=3 * N)
    s = _solver(True, a, 1.0 / (6 * np.pi), angular=True, precision="fp32", delta=1e-4)
    s.setPositions(pos)
    u, w = s.divM_with_noise(W)
    lin, ang = oracle.rfd_drift(pos, W, 1.0 / (6 * np.pi), a, 1e-4, angular=True)
    assert _rel(u, lin) <= 1e-8
    assert _rel(w, ang) <= 1e-8


def test_divm_with_noise_mixed_ensemble_ignores_open_systems():
    """Open systems have no drift (mobility.h:189-190): their W entries move
    nothing and their outputs stay zero; wall systems match the oracle with
    the open entries of W set to zero."""
    sizes = np.array([40, 25, 60, 33])
    kinds = np.array([0, 1, 0, 1])
    rng = np.random.default_rng(6)
    pos = []
    for n, k in zip(sizes, kinds):
        p = rng.uniform(0, 12.0, (n, 3))
        if k:
            p[:, 2] = rng.uniform(0.6, 5.0, n)
        pos.append(p)
    pos = np.concatenate(pos)
    from libmobility_amd import NBody
    s = NBody("open", "open", "single_wall")
    s.setParameters(wallHeight=0.0, delta=1e-3)
    s.setSystems(sizes, kinds)
    s.initialize(viscosity=0.9, hydrodynamicRadius=1.1)
    s.setPositions(pos)
    N = pos.shape[0]
    W = rng.normal(size=3 * N)
    u, _ = s.divM_with_noise(W)
    off = np.concatenate([[0], np.cumsum(sizes)])
    Wz = W.copy()
    for k, a0, b0 in zip(kinds, off[:-1], off[1:]):
        if not k:
            Wz[3 * a0:3 * b0] = 0.0
    lin, _ = oracle.rfd_drift(pos, Wz, 0.9, 1.1, 1e-3 * 1.1, wall=list(kinds), sys_sizes=sizes)
    u = u.ravel()
    for k, a0, b0 in zip(kinds, off[:-1], off[1:]):
        if k:
            assert _rel(u[3 * a0:3 * b0], lin[3 * a0:3 * b0]) <= 1e-9
        else:
            assert np.all(u[3 * a0:3 * b0] == 0)


def test_divm_with_noise_accumulates_through_the_c_abi():
    """The outputs are added to (mobility.h:227-235), host and device memory."""
    import torch
    from libmobility_amd import _ffi
    N, a = 64, 1.0
    pos = _slab(N, 10.0, 0.8, 4.0, seed=8)
    W = np.random.default_rng(3).normal(size=3 * N)
    s = _solver(True, a, 1.0, angular=True, delta=1e-3)
    s.setPositions(pos)
    lib = _ffi.load()
    base_u = np.random.default_rng(7).normal(size=3 * N)
    base_w = np.random.default_rng(8).normal(size=3 * N)
    lin, ang = oracle.rfd_drift(pos, W, 1.0, a, 1e-3, prefactor=2.0, angular=True)
    u, w = base_u.copy(), base_w.copy()
    _ffi.check(lib.lm_nbody_div_m_noise(s._h, W.ctypes.data, u.ctypes.data, w.ctypes.data,
                                        ctypes.c_double(2.0), _ffi.LM_MEM_HOST))
    assert _rel(u - base_u, lin) <= 1e-9 and _rel(w - base_w, ang) <= 1e-9
    dW = torch.tensor(W, device="cuda:0")
    du = torch.tensor(base_u, device="cuda:0")
    dw = torch.tensor(base_w, device="cuda:0")
    torch.cuda.synchronize()
    _ffi.check(lib.lm_nbody_div_m_noise(s._h, dW.data_ptr(), du.data_ptr(), dw.data_ptr(),
                                        ctypes.c_double(2.0), _ffi.LM_MEM_DEVICE))
    assert _rel(du.cpu().numpy() - base_u, lin) <= 1e-9
    assert _rel(dw.cpu().numpy() - base_w, ang) <= 1e-9
    with pytest.raises(RuntimeError, match="null noise"):
        _ffi.check(lib.lm_nbody_div_m_noise(s._h, None, u.ctypes.data, w.ctypes.data,
                                            ctypes.c_double(1.0), _ffi.LM_MEM_HOST))
