"""Drop-in behaviour of the Python NBody class (and, for the generic tests,
SelfMobility) on the GPU, following the reference's own interface tests
(tests/test_interface.py,
test_initialization.py, test_precision.py, test_array_origins.py,
test_thermal_drift.py:77-96): shapes and types of outputs, error cases,
parameter handling; plus edge cases of the pair kernel (coincident distinct
particles, which the reference treats as a self pair because it tests r == 0,
hydrodynamicKernels.cuh:143-150,386-388; empty systems of an ensemble)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

GEOMS = ["open", "single_wall"]
# the generic interface tests also run on SelfMobility (solver_configs_all,
# tests/utils.py:32-41, with its sane parameter of utils.py:17)
ALL = GEOMS + ["self"]


def _solver(geom, includeAngular=False, seed=None, **params):
    from libmobility_amd import NBody, SelfMobility
    if geom == "self":
        s = SelfMobility("open", "open", "open")
        s.setParameters(parameter=5.0)
        s.initialize(viscosity=1.0, hydrodynamicRadius=1.0, includeAngular=includeAngular,
                     seed=seed)
        return s
    s = NBody("open", "open", geom)
    if geom == "single_wall":
        params.setdefault("wallHeight", 0.0)
    s.setParameters(algorithm="advise", delta=1e-3, **params)
    s.initialize(viscosity=1.0, hydrodynamicRadius=1.0, includeAngular=includeAngular, seed=seed)
    return s


def _pos(n, rng):
    # inside [0, 10)^3 with z >= 0.5 (above a wall at z = 0)
    p = rng.random((n, 3)) * 10.0
    p[:, 2] = 0.5 + rng.random(n) * 9.5
    return p


@pytest.mark.parametrize("geom", ALL)
def test_non_contiguous_positions(geom):
    """test_interface.py:15-30"""
    rng = np.random.default_rng(0)
    s = _solver(geom)
    m = rng.random((3, 3)) + 1.0
    col = m[:, 0]  # strided view of 3 values = one particle
    assert not col.flags["C_CONTIGUOUS"]
    s.setPositions(col)
    u, _ = s.Mdot(np.ones(3))
    assert u.shape == (3,) and np.all(np.isfinite(u))


@pytest.mark.parametrize("geom", ALL)
def test_linear_only_shapes(geom):
    """test_interface.py:33-57: without includeAngular only U, shaped like F."""
    rng = np.random.default_rng(1)
    s = _solver(geom)
    pos = _pos(2, rng)
    f = rng.random((2, 3))
    s.setPositions(pos)
    u, w = s.Mdot(f)
    assert u.shape == (2, 3) and w is None
    u, w = s.Mdot(f.reshape(6))
    assert u.shape == (6,) and w is None
    s.setPositions(pos.reshape(6))
    u, w = s.Mdot(f.reshape(6))
    assert u.shape == (6,) and w is None


@pytest.mark.parametrize("geom", ALL)
@pytest.mark.parametrize("which", ["torques", "forces"])
def test_angular_returns_both(geom, which):
    """test_interface.py:60-123: a torque (or force) alone gives both U and W,
    both nonzero (hydrodynamic coupling), shaped like the input."""
    rng = np.random.default_rng(2)
    s = _solver(geom, includeAngular=True)
    pos = _pos(5, rng)
    v = rng.random((5, 3))
    for p, x in ((pos, v), (pos, v.reshape(15)), (pos.reshape(15), v.reshape(15))):
        s.setPositions(p)
        u, w = s.Mdot(**{which: x})
        assert u.shape == x.shape and w.shape == x.shape
        # the driven block is nonzero; the coupled one too except for
        # SelfMobility, which has no hydrodynamic coupling
        # (test_interface.py:66-67, :100-101)
        driven, coupled = (w, u) if which == "torques" else (u, w)
        assert np.linalg.norm(driven) > 0
        if geom != "self":
            assert np.linalg.norm(coupled) > 0


@pytest.mark.parametrize("geom", ALL)
def test_mixed_input_shapes(geom):
    """test_interface.py:126-154: U follows the forces' shape, W the torques'."""
    rng = np.random.default_rng(3)
    s = _solver(geom, includeAngular=True)
    for fs, ts in (((3,), (1, 3)), ((1, 3), (3,)), ((1, 3), (1, 3)), ((3,), (3,))):
        pos = _pos(1, rng)
        f, t = rng.random(fs), rng.random(ts)
        for p in (pos, pos.reshape(3)):
            s.setPositions(p)
            u, w = s.Mdot(f, t)
            assert u.shape == f.shape and w.shape == t.shape
            assert np.linalg.norm(u) > 0 and np.linalg.norm(w) > 0


@pytest.mark.parametrize("geom", ALL)
def test_noise_outputs_follow_positions_shape(geom):
    """test_interface.py:159-251: sqrtMdotW, divM and LangevinVelocities
    without forces take the shape of the positions; with forces, the forces'."""
    rng = np.random.default_rng(4)
    s = _solver(geom)
    pos = _pos(1, rng)
    for p in (pos, pos.reshape(3)):
        s.setPositions(p)
        assert s.sqrtMdotW()[0].shape == p.shape
        assert s.divM()[0].shape == p.shape
        assert s.LangevinVelocities(1.0, 1.0)[0].shape == p.shape
        for fs in ((3,), (1, 3)):
            f = rng.random(fs)
            assert s.LangevinVelocities(1.0, 1.0, f)[0].shape == fs
    sa = _solver(geom, includeAngular=True)
    for fs, ts in (((3,), (1, 3)), ((1, 3), (3,))):
        sa.setPositions(pos)
        v, w = sa.LangevinVelocities(dt=1.0, kbt=1.0, forces=rng.random(fs), torques=rng.random(ts))
        assert v.shape == fs and w.shape == ts


@pytest.mark.parametrize("geom", ALL)
def test_error_cases(geom):
    """test_interface.py:254-333: torques without includeAngular, calls before
    setPositions, malformed positions, forces/torques of the wrong size."""
    rng = np.random.default_rng(5)
    s = _solver(geom)
    with pytest.raises(RuntimeError):
        s.Mdot(rng.random((1, 3)))
    with pytest.raises(RuntimeError):
        s.sqrtMdotW()
    with pytest.raises(RuntimeError):
        s.LangevinVelocities(dt=1.0, kbt=1.0)
    for shape in ((4, 6), (3, 1), (1, 4)):
        with pytest.raises(RuntimeError):
            s.setPositions(rng.random(shape) + 1.0)
    for n_wrong in (4, 6):
        with pytest.raises(RuntimeError):
            s.Mdot(forces=rng.random(3 * n_wrong))
    s.setPositions(_pos(1, rng))
    with pytest.raises(RuntimeError):
        s.Mdot(rng.random((1, 3)), rng.random((1, 3)))
    sa = _solver(geom, includeAngular=True)
    sa.setPositions(_pos(5, rng))
    for n_wrong in (4, 6):
        with pytest.raises(RuntimeError):
            sa.Mdot(torques=rng.random(3 * n_wrong))


@pytest.mark.parametrize("geom", ALL)
@pytest.mark.parametrize("includeAngular", [False, True])
def test_changing_number_of_particles(geom, includeAngular):
    """test_interface.py:360-383"""
    rng = np.random.default_rng(6)
    s = _solver(geom, includeAngular=includeAngular)
    for n in (1, 2, 3):
        s.setPositions(_pos(n, rng))
        args = (rng.random((n, 3)), rng.random((n, 3))) if includeAngular else (rng.random((n, 3)),)
        mf, mt = s.Mdot(*args)
        dw, dt_ = s.sqrtMdotW()
        vf, vt = s.LangevinVelocities(1.0, 1.0, *args)
        for x in (mf, dw, vf) + ((mt, dt_, vt) if includeAngular else ()):
            assert x.shape == (n, 3)


@pytest.mark.parametrize("geom", ALL)
@pytest.mark.parametrize("includeAngular", [False, True])
def test_zero_temperature_langevin_is_mdot(geom, includeAngular):
    """test_interface.py:336-357"""
    rng = np.random.default_rng(7)
    s = _solver(geom, includeAngular=includeAngular)
    s.setPositions(_pos(1, rng))
    args = (rng.random((1, 3)), rng.random((1, 3))) if includeAngular else (rng.random((1, 3)),)
    mf, mt = s.Mdot(*args)
    lf, lt = s.LangevinVelocities(1.0, 0.0, *args)
    assert np.allclose(mf, lf, atol=1e-12)
    if includeAngular:
        assert np.allclose(mt, lt, atol=1e-12)


@pytest.mark.parametrize("nb,npb", [(1, 12), (12, 1)])
def test_batch_parameters(nb, npb):
    """test_initialization.py:35-52: equal batches; each batch is an
    independent system."""
    rng = np.random.default_rng(8)
    from libmobility_amd import NBody
    s = NBody("open", "open", "open")
    s.setParameters("advise", nb, npb)
    s.initialize(viscosity=1.0, hydrodynamicRadius=1.5)
    pos = rng.random((nb * npb, 3)) * 6
    s.setPositions(pos)
    f = rng.normal(size=(nb * npb, 3))
    u, _ = s.Mdot(f)
    ref, _ = oracle.mdot(pos, 1.0, 1.5, F=f, sys_sizes=[npb] * nb)
    assert np.allclose(u.ravel(), ref, rtol=1e-12, atol=1e-14)


def test_default_and_missing_parameters():
    """test_initialization.py:55-62,79-88: setParameters may be omitted."""
    rng = np.random.default_rng(9)
    from libmobility_amd import NBody
    for configure in (True, False):
        s = NBody("open", "open", "open")
        if configure:
            s.setParameters("advise")
        s.initialize(viscosity=1.0, hydrodynamicRadius=1.5)
        pos = rng.random((10, 3))
        s.setPositions(pos)
        u, _ = s.Mdot(pos)
        assert np.all(np.isfinite(u)) and np.linalg.norm(u) > 0


@pytest.mark.parametrize("nb,npb,n", [(5, 1, 10), (2, -1, 8)])
def test_bad_batch_parameters(nb, npb, n):
    """test_initialization.py:91-106: a batch layout that does not match the
    number of particles raises by setPositions at the latest."""
    from libmobility_amd import NBody
    s = NBody("open", "open", "open")
    s.setParameters("advise", nb, npb)
    with pytest.raises(RuntimeError):
        s.initialize(viscosity=1.0, hydrodynamicRadius=1.5)
        s.setPositions(np.random.default_rng(10).random((n, 3)))


@pytest.mark.parametrize("geom", ALL)
def test_single_precision_inputs(geom):
    """test_precision.py:9-62: float32 inputs are converted; the product of
    ones is nonzero in every entry."""
    rng = np.random.default_rng(11)
    s = _solver(geom)
    pos = _pos(1, rng).astype(np.float32)
    s.setPositions(pos)
    s.setPositions(pos.astype(np.float64))
    u, _ = s.Mdot(np.ones(3, dtype=np.float32))
    assert np.all(u != 0)


@pytest.mark.parametrize("geom", ALL)
@pytest.mark.parametrize("use_torques", [False, True])
@pytest.mark.parametrize("method", ["Mdot", "divM", "sqrtMdotW", "LangevinVelocities"])
@pytest.mark.parametrize("framework", ["numpy", "torch_cpu", "torch_gpu"])
def test_array_origins(geom, use_torques, method, framework):
    """test_array_origins.py:64-96: outputs come back in the input's framework
    and on its device."""
    import torch
    rng = np.random.default_rng(12)
    s = _solver(geom, includeAngular=use_torques)
    pos, f = _pos(10, rng), rng.random((10, 3))
    t = rng.random((10, 3)) if use_torques else None
    if framework != "numpy":
        dev = "cuda:0" if framework == "torch_gpu" else "cpu"
        pos = torch.from_numpy(pos).to(dev)
        f = torch.from_numpy(f).to(dev)
        t = torch.from_numpy(t).to(dev) if use_torques else None
    s.setPositions(pos)
    if method == "Mdot":
        mf, mt = s.Mdot(f, t)
    elif method == "divM":
        mf, mt = s.divM()
    elif method == "sqrtMdotW":
        mf, mt = s.sqrtMdotW()
    else:
        mf, mt = s.LangevinVelocities(1.0, 1.0, f, t)
    assert type(mf) is type(pos)
    if use_torques:
        assert type(mt) is type(pos)
    if framework != "numpy":
        assert mf.device == pos.device
        if use_torques:
            assert mt.device == pos.device
        mf_h = mf.cpu().numpy()
    else:
        mf_h = mf
    assert np.all(np.isfinite(mf_h))


def test_divm_draws_new_numbers_each_call():
    """test_thermal_drift.py:77-96"""
    rng = np.random.default_rng(13)
    s = _solver("single_wall", seed=1234)
    s.setPositions(_pos(4, rng))
    d1, _ = s.divM()
    d2, _ = s.divM()
    assert not np.allclose(d1, d2)


@pytest.mark.parametrize("geom", GEOMS)
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_coincident_particles_are_a_self_pair(geom, precision):
    """Two distinct particles at the same position: the reference detects the
    self pair by r == 0, so each gets the self block of the other's force."""
    rng = np.random.default_rng(14)
    pos = np.array([[3.0, 4.0, 2.5], [3.0, 4.0, 2.5], [5.0, 1.0, 3.5]])
    f = rng.normal(size=(3, 3))
    t = rng.normal(size=(3, 3))
    s = _solver(geom, includeAngular=True)
    s.setPairPrecision(precision)
    s.setPositions(pos)
    u, w = s.Mdot(f, t)
    uo, wo = oracle.mdot(pos, 1.0, 1.0, F=f, T=t, angular=True, wall=(geom == "single_wall"))
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert np.linalg.norm(u.ravel() - uo) <= tol * np.linalg.norm(uo)
    assert np.linalg.norm(w.ravel() - wo) <= tol * np.linalg.norm(wo)
    # both coincident particles feel S (F0 + F1) + M(x0, x2) F2: equal velocities
    u = u.reshape(3, 3)
    assert np.abs(u[0] - u[1]).max() <= tol * np.abs(u).max()


def test_empty_system_in_ensemble():
    """A zero-size system in a ragged ensemble contributes nothing and leaves
    the others unchanged."""
    rng = np.random.default_rng(15)
    from libmobility_amd import NBody
    sizes = [5, 0, 7]
    pos = _pos(12, rng)
    f = rng.normal(size=(12, 3))
    s = NBody("open", "open", "open")
    s.setParameters()
    s.initialize(viscosity=1.0, hydrodynamicRadius=1.0, tolerance=1e-8)
    s.setSystems(sizes)
    s.setPositions(pos)
    u, _ = s.Mdot(f)
    ref, _ = oracle.mdot(pos, 1.0, 1.0, F=f, sys_sizes=sizes)
    assert np.allclose(u.ravel(), ref, rtol=1e-12, atol=1e-14)
    s.setLanczosPerSystem(True)
    W = rng.normal(size=36)
    y, _ = s.sqrtMdotW_with_noise(W)
    M = oracle.dense_matrix(pos, 1.0, 1.0, sys_sizes=sizes)
    exact = oracle.dense_sqrt(M) @ W
    assert np.linalg.norm(y.ravel() - exact) <= 1e-5 * np.linalg.norm(exact)
    assert list(s.lastLanczosIterationsPerSystem()[[1]]) == [0]


class _DLWrap:
    """A device array of a framework the library does not know, speaking only
    DLPack (python_tensor.h:30-93 takes any nanobind framework)."""

    def __init__(self, a):
        self._a = a

    def __dlpack__(self, *args, **kwargs):
        return self._a.__dlpack__(*args, **kwargs)

    def __dlpack_device__(self):
        return self._a.__dlpack_device__()


@pytest.mark.parametrize("geom", GEOMS)
def test_dlpack_device_arrays_without_a_copy(geom):
    """Device arrays that only speak DLPack reach the library as the
    producer's own device pointer (no copy, no host round trip), and the
    products equal those of the same torch tensors passed directly."""
    import torch
    from libmobility_amd import nbody
    rng = np.random.default_rng(21)
    pos = torch.from_numpy(_pos(64, rng)).to("cuda:0")
    f = torch.from_numpy(rng.normal(size=(64, 3))).to("cuda:0")
    t = torch.from_numpy(rng.normal(size=(64, 3))).to("cuda:0")
    for x in (pos, f, t):
        arr = nbody._Arr(_DLWrap(x))
        assert arr.ptr == x.data_ptr()
        assert arr.mem == nbody._ffi.LM_MEM_DEVICE and arr.device == str(x.device)
    s_dl = _solver(geom, includeAngular=True)
    s_dl.setPositions(_DLWrap(pos))
    u_dl, w_dl = s_dl.Mdot(_DLWrap(f), _DLWrap(t))
    s_t = _solver(geom, includeAngular=True)
    s_t.setPositions(pos)
    u_t, w_t = s_t.Mdot(f, t)
    assert torch.is_tensor(u_dl) and u_dl.device == pos.device
    assert torch.equal(u_dl, u_t) and torch.equal(w_dl, w_t)
