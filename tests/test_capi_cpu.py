"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/libmobility_amd.h declares, and its pair/generator arithmetic (run on
the host through the self-check entry points) agrees with the oracle and
with the published Philox4x32-10 known-answer vectors.  No GPU work here."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "libmobility_amd.h")


def _lib():
    from libmobility_amd import _ffi
    return _ffi.load()


def _declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lm_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    from libmobility_amd import _ffi
    declared = set(_declared_symbols())
    assert declared == set(_ffi.SIGNATURES), declared ^ set(_ffi.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = _lib()
    for name in _declared_symbols():
        assert hasattr(lib, name), name


def test_version_precision():
    lib = _lib()
    assert lib.lm_version() == b"1.1"
    assert lib.lm_precision() == b"double"


def _blk(wall, fp32, pi, pj, a):
    lib = _lib()
    dp = ctypes.POINTER(ctypes.c_double)
    pi = np.ascontiguousarray(pi, float)
    pj = np.ascontiguousarray(pj, float)
    out = np.zeros(36)
    assert lib.lm_selftest_pair_block_host(wall, fp32, pi.ctypes.data_as(dp), pj.ctypes.data_as(dp),
                                           a, out.ctypes.data_as(dp)) == 0
    return out.reshape(6, 6)


@pytest.mark.parametrize("wall", [0, 1])
def test_kernel_pair_arithmetic_fp64(wall):
    """The regrouped wall/open arithmetic of pair_math.hpp vs the oracle's
    reference-form matrices, incl. self pairs, overlaps and z < a."""
    rng = np.random.default_rng(10 + wall)
    for t in range(3000):
        a = rng.uniform(0.5, 2.0)
        pi = rng.uniform(-4, 4, 3) * a
        pj = rng.uniform(-4, 4, 3) * a
        if wall:
            pi[2] = rng.uniform(-0.2, 4) * a
            pj[2] = rng.uniform(-0.2, 4) * a
        if t % 40 == 0:
            pj = pi.copy()
        B = _blk(wall, 0, pi, pj, a)
        O = oracle.pair_block(wall, pi, pj, a)
        assert np.abs(B - O).max() <= 1e-13 * max(np.abs(O).max(), 1e-300)


@pytest.mark.parametrize("wall", [0, 1])
def test_symmetric_pair_arithmetic_both_directions(wall):
    """One evaluation of a pair (pair_both in pair_math.hpp) gives target i the
    reference block M_ij and source j the block M_ji, i.e. the transposed
    forms of every block are right (incl. self pairs, overlaps and z < a)."""
    lib = _lib()
    dp = ctypes.POINTER(ctypes.c_double)
    rng = np.random.default_rng(30 + wall)
    for t in range(2000):
        a = rng.uniform(0.5, 2.0)
        pi = rng.uniform(-4, 4, 3) * a
        pj = rng.uniform(-4, 4, 3) * a
        if wall:
            pi[2] = rng.uniform(-0.2, 4) * a
            pj[2] = rng.uniform(-0.2, 4) * a
        if t % 40 == 0:
            pj = pi.copy()
        if t % 40 == 1:
            pj = pi + rng.uniform(-0.5, 0.5, 3) * a  # overlapping
        fwd, rev = np.zeros(36), np.zeros(36)
        assert lib.lm_selftest_pair_both_host(
            wall, pi.ctypes.data_as(dp), pj.ctypes.data_as(dp), a, fwd.ctypes.data_as(dp),
            rev.ctypes.data_as(dp)) == 0
        Oij = oracle.pair_block(wall, pi, pj, a)
        Oji = oracle.pair_block(wall, pj, pi, a)
        scale = max(np.abs(Oij).max(), np.abs(Oji).max(), 1e-300)
        assert np.abs(fwd.reshape(6, 6) - Oij).max() <= 1e-13 * scale, t
        assert np.abs(rev.reshape(6, 6) - Oji).max() <= 1e-13 * scale, t


@pytest.mark.parametrize("wall", [0, 1])
def test_kernel_pair_arithmetic_fp32(wall):
    rng = np.random.default_rng(20 + wall)
    worst = 0.0
    for t in range(1000):
        pi = rng.uniform(-6, 6, 3)
        pj = rng.uniform(-6, 6, 3)
        if wall:
            pi[2] = rng.uniform(0.5, 6)
            pj[2] = rng.uniform(0.5, 6)
        B = _blk(wall, 1, pi, pj, 1.0)
        O = oracle.pair_block(wall, pi, pj, 1.0)
        worst = max(worst, np.abs(B - O).max() / np.abs(O).max())
    assert worst < 1e-5


def test_philox_known_answers():
    """Published Philox4x32-10 test vectors (Random123 kat_vectors)."""
    lib = _lib()
    U4 = ctypes.c_uint32 * 4
    K2 = ctypes.c_uint32 * 2
    cases = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in cases:
        out = U4()
        assert lib.lm_selftest_philox_host(U4(*ctr), K2(*key), out) == 0
        assert list(out) == want


def test_python_errors_before_device_work():
    """Configuration errors of pythonify.h / mobility.h raise RuntimeError."""
    from libmobility_amd import NBody
    with pytest.raises(RuntimeError, match="Invalid periodicity"):
        NBody("periodicasdas", "periodic", "open")
    with pytest.raises(RuntimeError):
        NBody("periodic", "periodic", "open")
    s = NBody("open", "open", "open")
    with pytest.raises(RuntimeError, match="only valid for bottom wall"):
        s.setParameters(wallHeight=0.5)
    s.setParameters("advise")
    w = NBody("open", "open", "single_wall")
    with pytest.raises(RuntimeError, match="required for a bottom wall"):
        w.setParameters("advise")
    w.setParameters(wallHeight=0.5)
    with pytest.raises(RuntimeError, match="Invalid algorithm"):
        s.setParameters("fmm")
    assert NBody.precision == "double"


def test_outputs_are_sized_by_the_solvers_own_particle_count(monkeypatch):
    """sqrtMdotW/divM outputs follow the module-global last positions shape in
    form ((N,3) or (3N,), pythonify.h:34-36) but always hold this solver's 3N
    values: positions set on another solver since must not size them (the
    library writes 3N entries)."""
    from libmobility_amd import nbody
    s = nbody.NBody("open", "open", "open")
    monkeypatch.setattr(s, "getNumberParticles", lambda: 5)
    monkeypatch.setattr(nbody, "_last_shape", (7, 3))
    lin, ang, _ = s._outputs_like_positions()
    assert lin.shape == (5, 3) and ang is None
    monkeypatch.setattr(nbody, "_last_shape", (21,))
    assert s._outputs_like_positions()[0].shape == (15,)


def test_output_buffers_smaller_than_3N_are_refused(monkeypatch):
    """Regression for a host heap overrun: an output sized from another
    solver's positions (7 particles) handed to a solver of 165 particles would
    receive 495 values in a 21-value buffer.  Every output pointer passed to
    the library is checked against this solver's 3N first."""
    from libmobility_amd import nbody
    s = nbody.NBody("open", "open", "open")
    monkeypatch.setattr(s, "getNumberParticles", lambda: 165)
    with pytest.raises(RuntimeError, match="holds 21 values, the solver writes 495"):
        s._out(np.zeros((7, 3)))
    assert s._out(np.zeros((165, 3))) is not None
    assert s._out(None) is None
    with pytest.raises(RuntimeError, match="holds 495 values, the solver writes 990"):
        s._out(np.zeros((165, 3)), nvec=2)


def test_libmobility_surface_and_selfmobility_configuration():
    """solvers/__init__.py:1-6 and SelfMobility's constructor
    (solvers/SelfMobility/mobility.h:31-36): open boundaries only; PSE and
    DPStokes are not part of this library and raise."""
    import libMobility
    assert set(libMobility.__all__) == {"PSE", "SelfMobility", "NBody", "DPStokes"}
    for cls, args in ((libMobility.PSE, ("periodic",) * 3),
                      (libMobility.DPStokes, ("periodic", "periodic", "open"))):
        with pytest.raises(RuntimeError, match="not part of this library"):
            cls(*args)
    with pytest.raises(RuntimeError, match="open boundary solver"):
        libMobility.SelfMobility("open", "open", "single_wall")
    with pytest.raises(RuntimeError, match="Invalid periodicity"):
        libMobility.SelfMobility("open", "open", "nope")
    s = libMobility.SelfMobility("open", "open", "open")
    s.setParameters(parameter=5.0)


@pytest.mark.parametrize("shape", ["cube", "slab", "plane", "one", "capped"])
def test_coarse_cells_match_the_oracle_rule(shape):
    """The library's cell assignment of the coarse level is the oracle's rule,
    integer for integer (oracle/coarse.py)."""
    from oracle import coarse
    from libmobility_amd.workloads import rsa_cube
    lib = _lib()
    rng = np.random.default_rng(11)
    ppc = 32
    if shape == "cube":
        xyz, _ = rsa_cube(3000, 0.2, seed=3)
    elif shape == "slab":
        xyz = rng.uniform(size=(5000, 3)) * np.array([256.0, 256.0, 18.5]) + np.array([0, 0, 1.5])
    elif shape == "plane":
        xyz = rng.uniform(size=(700, 3)) * np.array([30.0, 50.0, 0.0])
    elif shape == "one":
        xyz = np.array([[1.0, 2.0, 3.0]])
    else:
        xyz = rng.uniform(size=(30000, 3)) * 100.0
        ppc = 1
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    N = xyz.shape[0]
    nc = ctypes.c_int(0)
    grid = np.zeros(3, dtype=np.int32)
    cell_of = np.zeros(N, dtype=np.int32)
    st = lib.lm_selftest_coarse_cells_host(xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                           N, ppc, ctypes.addressof(nc), grid.ctypes.data,
                                           cell_of.ctypes.data)
    assert st == 0
    ref, g, C = coarse.coarse_cells(xyz, ppc)
    assert nc.value == C and list(grid) == list(g)
    assert np.array_equal(cell_of, ref)
    assert 3 * C <= 24576


class _DLWrap:
    """An array of an unknown framework that only speaks DLPack."""

    def __init__(self, a):
        self._a = a

    def __dlpack__(self, *args, **kwargs):
        return self._a.__dlpack__(*args, **kwargs)

    def __dlpack_device__(self):
        return self._a.__dlpack_device__()


def test_dlpack_arrays_are_taken_without_a_copy():
    """Any DLPack producer (python_tensor.h:60-93 accepts every nanobind
    framework) is adopted through torch.from_dlpack: the library receives the
    producer's own pointer (no copy) when it is FP64 and contiguous; other
    dtypes are converted, as nanobind converts them (test_precision.py:38-62)."""
    from libmobility_amd import nbody
    a = np.random.default_rng(0).normal(size=(50, 3))
    arr = nbody._Arr(_DLWrap(a))
    assert arr.ptr == a.ctypes.data
    assert arr.framework == "torch" and arr.mem == nbody._ffi.LM_MEM_HOST and arr.shape == (50, 3)
    b = a.astype(np.float32)
    arr32 = nbody._Arr(_DLWrap(b))
    assert arr32.ptr != b.ctypes.data and np.allclose(arr32.obj.numpy(), b)
