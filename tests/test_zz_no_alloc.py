"""The allocation-free Mdot (lm_nbody_mdot_noalloc) against lm_nbody_mdot.

The entry point has no counterpart in the reference; it must give
NBody::Mdot's result (solvers/NBody/mobility.h:166-182) with every buffer
beyond inputs and outputs in its static device array.  Checked through the C
ABI on two calls in a row, the second with fewer particles (the array keeps
the first call's contents), against the allocating entry point on the same
inputs: FP64 pairs 1e-12, FP32 pairs 1e-6 relative (the two entry points cut
the source range into different chunks, so FP32 tile sums round
differently).  Outputs start as NaN so that every entry must be written.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = a.detach().cpu().numpy().ravel()
    b = b.detach().cpu().numpy().ravel()
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _solver(wall, angular, precision, nbatch=-1, nperbatch=-1):
    from libmobility_amd import NBody
    s = NBody("open", "open", "single_wall" if wall else "open")
    if wall:
        s.setParameters(wallHeight=0.0, Nbatch=nbatch, NperBatch=nperbatch)
    else:
        s.setParameters(Nbatch=nbatch, NperBatch=nperbatch)
    s.setPairPrecision(precision)
    s.initialize(viscosity=1.0 / (6 * np.pi), hydrodynamicRadius=1.0, includeAngular=angular)
    return s


def _inputs(N, wall, seed):
    import torch
    rng = np.random.default_rng(seed)
    L = (N * 4.19 / 0.1) ** (1 / 3)
    pos = rng.uniform(0, L, (N, 3))
    if wall:
        pos[:, 2] = rng.uniform(0.5, 12.0, N)
    dev = "cuda:0"
    return (torch.tensor(pos, device=dev), torch.tensor(rng.normal(size=(N, 3)), device=dev),
            torch.tensor(rng.normal(size=(N, 3)), device=dev))


def _noalloc(s, F, T, angular):
    import torch
    from libmobility_amd import _ffi
    lib = _ffi.load()
    N = (F if F is not None else T).shape[0]
    U = torch.full((N, 3), float("nan"), dtype=torch.float64, device="cuda:0")
    W = torch.full((N, 3), float("nan"), dtype=torch.float64, device="cuda:0") if angular else None
    torch.cuda.synchronize()
    _ffi.check(lib.lm_nbody_mdot_noalloc(s._h, F.data_ptr() if F is not None else None,
                                         T.data_ptr() if T is not None else None, U.data_ptr(),
                                         W.data_ptr() if W is not None else None))
    torch.cuda.synchronize()
    return U, W


@pytest.mark.parametrize("precision,bound", [("fp64", 1e-12), ("fp32", 1e-6)])
@pytest.mark.parametrize("wall,angular", [(False, False), (False, True), (True, False), (True, True)])
def test_noalloc_mdot_equals_mdot_on_two_calls(precision, bound, wall, angular):
    s = _solver(wall, angular, precision)
    for N, seed in ((6000, 1), (1500, 2)):
        pos, F, T = _inputs(N, wall, seed)
        s.setPositions(pos)
        Tin = T if angular else None
        ref_u, ref_w = s.Mdot(F, Tin)
        u, w = _noalloc(s, F, Tin, angular)
        assert _rel(u, ref_u) <= bound
        if angular:
            assert _rel(w, ref_w) <= bound
            # torques only, then forces only
            ref_u, ref_w = s.Mdot(None, T)
            u, w = _noalloc(s, None, T, angular)
            assert _rel(u, ref_u) <= bound and _rel(w, ref_w) <= bound
            ref_u, ref_w = s.Mdot(F)
            u, w = _noalloc(s, F, None, angular)
            assert _rel(u, ref_u) <= bound and _rel(w, ref_w) <= bound


def test_noalloc_mdot_cuts_the_slabs_to_the_scratch_and_handles_batches():
    """N = 131072 in FP32: the automatic 32 source chunks would need 100 MB of
    slabs, so the call cuts them to what its 64 MiB hold; then a smaller call
    over three equal batches."""
    s = _solver(False, False, "fp32")
    pos, F, _ = _inputs(131072, False, 3)
    s.setPositions(pos)
    ref, _ = s.Mdot(F)
    u, _ = _noalloc(s, F, None, False)
    assert _rel(u, ref) <= 1e-6
    b = _solver(True, True, "fp64", nbatch=3, nperbatch=700)
    pos, F, T = _inputs(2100, True, 4)
    b.setPositions(pos)
    ref_u, ref_w = b.Mdot(F, T)
    u, w = _noalloc(b, F, T, True)
    assert _rel(u, ref_u) <= 1e-12 and _rel(w, ref_w) <= 1e-12


def test_noalloc_mdot_refusals():
    import torch
    from libmobility_amd import _ffi
    lib = _ffi.load()
    s = _solver(False, False, "fp64")
    s.setSystems([100, 50])
    pos, F, T = _inputs(150, False, 5)
    s.setPositions(pos)
    U = torch.zeros((150, 3), dtype=torch.float64, device="cuda:0")
    assert lib.lm_nbody_mdot_noalloc(s._h, F.data_ptr(), None, U.data_ptr(), None) != 0
    assert "not ragged systems" in lib.lm_last_error().decode()
    t = _solver(False, False, "fp64")
    t.setPositions(pos)
    assert lib.lm_nbody_mdot_noalloc(t._h, F.data_ptr(), T.data_ptr(), U.data_ptr(), None) != 0
    assert "includeAngular" in lib.lm_last_error().decode()
    assert lib.lm_nbody_mdot_noalloc(t._h, F.data_ptr(), None, None, None) != 0
