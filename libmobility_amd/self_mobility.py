"""Python ``SelfMobility`` class: libMobility's SelfMobility solver
(solvers/SelfMobility/mobility.h:17-111, python_wrapper.cu:6-29) on the GPU.

No hydrodynamic interactions: Mdot scales forces by 1/(6 pi eta a) and
torques by 1/(8 pi eta a^3); sqrtMdotW adds the square root of that diagonal
times Gaussian noise drawn on the device; the drift (divM) is zero.  The
array rules, output framework/device and error texts are those of the NBody
class (pythonify.h), shared through ``_Solver``.
"""
import ctypes

from . import _ffi
from .nbody import _Solver


class SelfMobility(_Solver):
    """The mobility is (1/(6 pi eta a)) I for linear and (1/(8 pi eta a^3)) I
    for angular velocities; open boundaries in the three directions only."""

    precision = "double"
    version = "1.1"
    _prefix = "self"

    def __init__(self, periodicityX, periodicityY, periodicityZ):
        per = []
        for p in (periodicityX, periodicityY, periodicityZ):
            if p not in _ffi.PERIODICITY:
                raise RuntimeError("[libMobility] Invalid periodicity")
            per.append(_ffi.PERIODICITY[p])
        self._lib = _ffi.load()
        h = ctypes.c_void_p()
        _ffi.check(self._lib.lm_self_create(per[0], per[1], per[2], ctypes.byref(h)))
        self._h = h
        self._include_angular = False
        self._stream_set = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.lm_self_destroy(h)
            except Exception:
                pass
            self._h = None

    def setParameters(self, parameter):
        """python_wrapper.cu:21-27: an example parameter; it is ignored."""
        float(parameter)

    def initialize(self, viscosity, hydrodynamicRadius, includeAngular=False, tolerance=1e-4,
                   seed=None, lanczos_callback=None):
        """pythonify.h:241-260 -> SelfMobility::initialize (mobility.h:38-50).
        The square root of a diagonal needs no Lanczos: tolerance and
        callback are accepted and unused."""
        s = 0 if seed is None else int(seed)
        _ffi.check(self._lib.lm_self_initialize(
            self._h, float(viscosity), float(hydrodynamicRadius), int(bool(includeAngular)),
            float(tolerance), ctypes.c_uint64(s & 0xFFFFFFFFFFFFFFFF)))
        self._include_angular = bool(includeAngular)
