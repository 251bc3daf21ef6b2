"""Python ``NBody`` class: the drop-in surface of libMobility's NBody solver.

Mirrors the class that ``MOBILITY_PYTHONIFY_WITH_EXTRA_CODE`` generates
(include/MobilityInterface/pythonify.h:364-407) plus NBody's
``setParameters`` (solvers/NBody/python_wrapper.cu:50-61): same method names,
argument names and defaults, shape rules, output framework/device, and
``RuntimeError`` texts.  All arithmetic runs in the HIP library through the C
ABI of include/libmobility_amd.h; arrays that already live on the GPU (torch
ROCm tensors, or any DLPack producer's device arrays) are passed by device
pointer without copies.

Extensions (no reference counterpart): ``setPairPrecision``,
``sqrtMdotW_with_noise``, ``hydrodynamicVelocities``, ``setSystems``,
``lastLanczosIterations``.
"""
import ctypes

import numpy as np

from . import _ffi

# Module-level "last positions" state, as pythonify.h:34-36 keeps it: the
# output shape/framework/device of sqrtMdotW and divM come from the last
# setPositions (or Mdot) call of any solver in the process.
_last_shape = ()
_last_framework = "numpy"
_last_device = "cpu"


def _is_torch(x):
    try:
        import torch
    except ImportError:
        return False
    return isinstance(x, torch.Tensor)


def _adopt_dlpack(x):
    """Arrays of other frameworks that speak DLPack (``__dlpack__`` and
    ``__dlpack_device__``: CuPy, JAX, ...), host or ROCm device memory, are
    taken over without a copy as torch tensors sharing their memory
    (torch.from_dlpack); outputs are then torch tensors on the same device.
    numpy arrays and torch tensors pass unchanged."""
    if x is None or isinstance(x, np.ndarray) or _is_torch(x):
        return x
    if hasattr(x, "__dlpack__") and hasattr(x, "__dlpack_device__"):
        import torch
        return torch.from_dlpack(x)
    return x


class _Arr:
    """An FP64, C-contiguous view of a caller array plus where it lives."""

    __slots__ = ("obj", "framework", "device", "shape", "size", "ndim", "ptr", "mem")

    def __init__(self, x):
        if x is None:
            x = np.zeros(0)
        x = _adopt_dlpack(x)
        if _is_torch(x):
            import torch

            t = x.detach()
            if t.dtype != torch.float64:
                t = t.to(torch.float64)
            t = t.contiguous()
            self.obj = t
            self.framework = "torch"
            self.device = str(t.device)
            self.shape = tuple(t.shape)
            self.size = t.numel()
            self.ndim = t.dim()
            self.ptr = t.data_ptr() if self.size else None
            self.mem = _ffi.LM_MEM_DEVICE if t.is_cuda else _ffi.LM_MEM_HOST
        else:
            a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
            self.obj = a
            self.framework = "numpy"
            self.device = "cpu"
            self.shape = tuple(a.shape)
            self.size = a.size
            self.ndim = a.ndim
            self.ptr = a.ctypes.data if a.size else None
            self.mem = _ffi.LM_MEM_HOST

    def moved_to(self, mem, device):
        """Same data in the given memory space (torch tensors only move)."""
        if self.size == 0 or self.mem == mem:
            return self
        import torch

        src = self.obj if self.framework == "torch" else torch.from_numpy(self.obj)
        return _Arr(src.to(device if mem == _ffi.LM_MEM_DEVICE else "cpu"))


def _check_and_get_shape(arr):
    """pythonify.h:83-107."""
    if arr.size == 0:
        return ()
    if arr.size % 3 != 0:
        raise RuntimeError(
            "[libMobility] The input arrays for positions, forces, and torques must each have "
            "total size 3*N.")
    n = arr.size // 3
    err = False
    if arr.ndim == 1 and arr.shape[0] != 3 * n:
        err = True
    elif arr.ndim == 2 and arr.shape[1] != 3:
        err = True
    elif arr.ndim not in (1, 2):
        err = True
    if err:
        raise RuntimeError(
            "[libMobility] The input arrays for positions, forces, and torques must each have "
            "shape (N, 3) or (3*N).")
    return arr.shape


def _zeros(shape, framework, device):
    """python_tensor.h:30-56: output in the caller's framework and device."""
    if framework == "torch":
        import torch

        return torch.zeros(shape, dtype=torch.float64, device=device)
    return np.zeros(shape, dtype=np.float64)


def _ptr(x):
    if x is None:
        return None
    if _is_torch(x):
        return x.data_ptr() if x.numel() else None
    return x.ctypes.data if x.size else None


def _mem_of(framework, device):
    return _ffi.LM_MEM_DEVICE if (framework == "torch" and device.startswith("cuda")) else _ffi.LM_MEM_HOST


class _Solver:
    """The part of the pythonify.h class every solver shares (positions,
    array rules, Mdot, sqrtMdotW, divM, LangevinVelocities, clean), over the
    C-ABI functions lm_<prefix>_*."""

    _prefix = ""

    def _f(self, name):
        return getattr(self._lib, "lm_" + self._prefix + "_" + name)

    # -- helpers ------------------------------------------------------------------
    def _out(self, x, nvec=1):
        """Pointer of an output array after checking that it holds exactly the
        nvec * 3N values the library writes into it (a smaller buffer would be
        overrun).  Mdot and LangevinVelocities need no such check: the library
        refuses forces and torques of another size than 3N before writing."""
        if x is None:
            return None
        n = nvec * 3 * self.getNumberParticles()
        size = x.numel() if _is_torch(x) else x.size
        if size != n:
            raise RuntimeError(f"[libmobility_amd] output array holds {size} values, the solver "
                               f"writes {n}")
        return _ptr(x)

    def _bind_stream(self, mem, device):
        if mem != _ffi.LM_MEM_DEVICE:
            return
        import torch

        st = torch.cuda.current_stream(torch.device(device)).cuda_stream
        if st != self._stream_set:
            _ffi.check(self._f("set_stream")(self._h, ctypes.c_void_p(st)))
            self._stream_set = st

    def getNumberParticles(self):
        return int(self._f("num_particles")(self._h))

    # -- interface ----------------------------------------------------------------
    def setPositions(self, positions):
        """pythonify.h:262-271 -> NBody::setPositions (mobility.h:116-162)."""
        global _last_shape, _last_framework, _last_device
        arr = _Arr(positions)
        shape = _check_and_get_shape(arr)
        _last_shape = shape
        _last_framework = arr.framework
        _last_device = arr.device
        self._bind_stream(arr.mem, arr.device)
        _ffi.check(self._f("set_positions")(self._h, arr.ptr, arr.size, arr.mem))
        _last_shape = arr.shape

    def _setup_arrays(self, forces, torques):
        """pythonify.h:109-149."""
        global _last_framework, _last_device
        N = self.getNumberParticles()
        f = _Arr(forces)
        t = _Arr(torques)
        if f.size > 0 and f.size != 3 * N:
            raise RuntimeError("The forces array must have size 3*N.")
        if t.size > 0 and t.size != 3 * N:
            raise RuntimeError("The torques array must have size 3*N.")
        lead = f if f.size > 0 else (t if t.size > 0 else f)
        framework, device, mem = lead.framework, lead.device, lead.mem
        _last_framework = framework
        _last_device = device
        f = f.moved_to(mem, device)
        t = t.moved_to(mem, device)
        f_shape = _check_and_get_shape(f) if f.size > 0 else _check_and_get_shape(t)
        mf = _zeros(f_shape, framework, device)
        mt = None
        if not self._include_angular:
            if t.size > 0:
                raise RuntimeError(
                    "The solver was configured without including angular velocities. Set "
                    "includeAngular to true when initializing if you want to use torques")
        else:
            t_shape = f_shape if t.size == 0 else _check_and_get_shape(t)
            mt = _zeros(t_shape, framework, device)
        return f, t, mf, mt, mem, device

    def Mdot(self, forces=None, torques=None):
        """pythonify.h:208-216 -> NBody::Mdot (mobility.h:166-182)."""
        f, t, mf, mt, mem, device = self._setup_arrays(forces, torques)
        self._bind_stream(mem, device)
        _ffi.check(self._f("mdot")(self._h, f.ptr, f.size, t.ptr, t.size, _ptr(mf),
                                           _ptr(mt), mem))
        return mf, mt

    def _own_shape(self):
        """The module-global last shape (pythonify.h:34-36) in form -- (3N,) or
        (N, 3) -- but sized for this solver's N: another solver's positions
        set since must not size this solver's outputs (they receive 3N
        values)."""
        N = self.getNumberParticles()
        return (3 * N,) if len(_last_shape) == 1 else (N, 3)

    def _outputs_like_positions(self):
        shape = self._own_shape()
        linear = _zeros(shape, _last_framework, _last_device)
        angular = _zeros(shape, _last_framework, _last_device) if self._include_angular else None
        return linear, angular, _mem_of(_last_framework, _last_device)

    def sqrtMdotW(self, prefactor=1.0):
        """pythonify.h:170-185 -> Mobility::sqrtMdotW (MobilityInterface.h:132-188)."""
        linear, angular, mem = self._outputs_like_positions()
        self._bind_stream(mem, _last_device)
        _ffi.check(self._f("sqrt_mdot_w")(self._h, self._out(linear), self._out(angular),
                                                  float(prefactor), mem))
        return linear, angular

    def divM(self, prefactor=1):
        """pythonify.h:329-345 -> NBody::divM (mobility.h:184-236)."""
        if self.getNumberParticles() <= 0:
            raise RuntimeError(
                "[libMobility] The number of particles is not set. Did you forget to call "
                "setPositions?")
        linear, angular, mem = self._outputs_like_positions()
        self._bind_stream(mem, _last_device)
        _ffi.check(self._f("div_m")(self._h, self._out(linear), self._out(angular), float(prefactor),
                                            mem))
        return linear, angular

    def LangevinVelocities(self, dt, kbt, forces=None, torques=None):
        """pythonify.h:273-294 -> Mobility::LangevinVelocities (MobilityInterface.h:191-206)."""
        f, t, mf, mt, mem, device = self._setup_arrays(forces, torques)
        if f.size == 0:
            mf = _zeros(self._own_shape(), _last_framework, _last_device)
            mem = _mem_of(_last_framework, _last_device)
            device = _last_device
        if self._include_angular and t.size == 0:
            mt = _zeros(self._own_shape(), _last_framework, _last_device)
        f = f.moved_to(mem, device)
        t = t.moved_to(mem, device)
        self._bind_stream(mem, device)
        _ffi.check(self._f("langevin_velocities")(
            self._h, float(dt), float(kbt), f.ptr, f.size, t.ptr, t.size, _ptr(mf), _ptr(mt), mem))
        return mf, mt

    def hydrodynamicVelocities(self, forces=None, torques=None):
        """Mdot + M^{1/2} W + div M with unit prefactors (LangevinVelocities, dt=2, kbt=1)."""
        return self.LangevinVelocities(2.0, 1.0, forces, torques)

    def clean(self):
        """Mobility::clean (MobilityInterface.h:220-223)."""
        _ffi.check(self._f("clean")(self._h))


class NBody(_Solver):
    """O(N^2) hydrodynamic interactions on one MI355X.

    Open boundaries in x and y; z is ``open`` (Rotne-Prager-Yamakawa) or
    ``single_wall`` (Rotne-Prager-Blake with Swan-Brady corrections, wall at
    ``wallHeight``).
    """

    precision = "double"
    version = "1.1"
    _prefix = "nbody"

    def __init__(self, periodicityX, periodicityY, periodicityZ):
        # string2Periodicity, pythonify.h:38-52
        per = []
        for p in (periodicityX, periodicityY, periodicityZ):
            if p not in _ffi.PERIODICITY:
                raise RuntimeError("[libMobility] Invalid periodicity")
            per.append(_ffi.PERIODICITY[p])
        self._lib = _ffi.load()
        h = ctypes.c_void_p()
        _ffi.check(self._lib.lm_nbody_create(per[0], per[1], per[2], ctypes.byref(h)))
        self._h = h
        self._cb = None
        self._precond = None
        self._include_angular = False
        self._stream_set = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.lm_nbody_destroy(h)
            except Exception:
                pass
            self._h = None

    # -- configuration ---------------------------------------------------------
    def setParameters(self, algorithm="advise", Nbatch=-1, NperBatch=-1, wallHeight=None,
                      delta=1e-3):
        """NBody.setParameters, solvers/NBody/python_wrapper.cu:50-61."""
        if algorithm not in _ffi.ALGORITHM:
            raise RuntimeError("Invalid algorithm selected")
        has_wall = wallHeight is not None
        _ffi.check(self._lib.lm_nbody_set_parameters(
            self._h, _ffi.ALGORITHM[algorithm], int(Nbatch), int(NperBatch), int(has_wall),
            float(wallHeight) if has_wall else 0.0, float(delta)))

    def setPairPrecision(self, precision="fp64"):
        """Extension: 'fp64' (default) or 'fp32' (FP32 pair terms, FP64 accumulation)."""
        if precision not in _ffi.PAIR_PRECISION:
            raise RuntimeError("[libmobility_amd] pair precision must be 'fp64' or 'fp32'")
        _ffi.check(self._lib.lm_nbody_set_pair_precision(self._h, _ffi.PAIR_PRECISION[precision]))

    def setDriftPrecision(self, precision="fp64"):
        """Extension: pair arithmetic of divM's two products, 'fp64' (default,
        whatever the pair precision) or 'fp32'."""
        if precision not in _ffi.PAIR_PRECISION:
            raise RuntimeError("[libmobility_amd] pair precision must be 'fp64' or 'fp32'")
        _ffi.check(self._lib.lm_nbody_set_drift_precision(self._h, _ffi.PAIR_PRECISION[precision]))

    def setSystems(self, sizes, kinds=None):
        """Extension: independent systems of the given (ragged) sizes.

        ``kinds``: per system "open"/"single_wall" (or 0/1); default: the
        solver's periodicity.  Mixed ensembles run in one persistent launch.
        """
        sizes = np.asarray(sizes, dtype=np.int64)
        off = np.zeros(sizes.size + 1, dtype=np.int64)
        off[1:] = np.cumsum(sizes)
        self._sys_off = off
        kp = None
        if kinds is not None:
            k = np.array([{"open": 0, "single_wall": 1}.get(x, x) for x in kinds], dtype=np.int32)
            if k.size != sizes.size:
                raise RuntimeError("[libmobility_amd] one kind per system is required")
            self._sys_kinds = k
            kp = k.ctypes.data
        _ffi.check(self._lib.lm_nbody_set_systems(self._h, off.ctypes.data, int(sizes.size), kp))

    def setPairTiling(self, targetsPerLane=0, sourceChunks=0):
        """Extension: fix the pair-kernel tiling (0 = automatic)."""
        _ffi.check(self._lib.lm_nbody_set_pair_tiling(self._h, int(targetsPerLane), int(sourceChunks)))

    def lastPairTiling(self):
        """Extension: (targets per lane, source chunks) of the last pair launch."""
        tpl, nch = ctypes.c_int(0), ctypes.c_int(0)
        _ffi.check(self._lib.lm_nbody_get_pair_tiling(self._h, ctypes.byref(tpl), ctypes.byref(nch)))
        return tpl.value, nch.value

    def setPairSymmetric(self, mode=-1, memoryBytes=0):
        """Extension: evaluate each unordered pair once for a single open or
        single-wall system (-1 auto, 0 off, 1 on); memoryBytes bounds its
        partial sums (0 = 4 GiB; the work is split into passes that fit)."""
        _ffi.check(self._lib.lm_nbody_set_pair_symmetric(self._h, int(mode)))
        _ffi.check(self._lib.lm_nbody_set_pair_symmetric_memory(self._h, int(memoryBytes)))

    def lastPairSymmetric(self):
        """Extension: whether the last pair product used the symmetric path."""
        used = ctypes.c_int(0)
        _ffi.check(self._lib.lm_nbody_last_pair_symmetric(self._h, ctypes.byref(used)))
        return bool(used.value)

    def setPersistent(self, mode=1):
        """Extension: persistent pair launches (-1 auto, 0 off, 1 on)."""
        _ffi.check(self._lib.lm_nbody_set_persistent(self._h, int(mode)))

    def setLanczosPerSystem(self, on=True):
        """Ragged ensembles: one Lanczos recurrence per system (each converging
        on its own, as separate per-system solvers would) instead of one over
        the whole block-diagonal operator (the reference's batched mode)."""
        _ffi.check(self._lib.lm_nbody_set_lanczos_per_system(self._h, int(bool(on))))

    def lastLanczosIterationsPerSystem(self):
        nsys = int(self._lib.lm_nbody_num_systems(self._h))
        out = np.zeros(max(nsys, 0), dtype=np.int32)
        _ffi.check(self._lib.lm_nbody_last_lanczos_iterations_per_system(
            self._h, out.ctypes.data, nsys))
        return out

    def setLanczosIterationLimit(self, limit=200):
        """Iterations before "[Lanczos] Could not converge" (the public
        ``iterationHardLimit`` of the reference's lanczos::Solver,
        LanczosAlgorithm.h:49; default 200, at most 1024)."""
        _ffi.check(self._lib.lm_nbody_set_lanczos_iteration_limit(self._h, int(limit)))

    def setLanczosReorthogonalization(self, on=True):
        """Extension: orthogonalise every new Lanczos basis vector against all
        earlier ones in FP64 (the reference's recurrence never does).  Off by
        default; applies to single-recurrence solves."""
        _ffi.check(self._lib.lm_nbody_set_lanczos_reorthogonalization(self._h, 1 if on else 0))

    def setPreconditioner(self, preconditioner="learned"):
        """Extension: attach a learned block preconditioner to sqrtMdotW.

        ``"learned"`` loads the packaged trained weights (3x3 blocks, or 6x6
        blocks when the solver was initialised with includeAngular); a
        :class:`libmobility_amd.precond.LearnedPreconditioner` is used as is;
        ``None`` returns to plain Lanczos.
        """
        from .precond import default_network
        if preconditioner is None:
            self._precond = None
            _ffi.check(self._lib.lm_nbody_set_preconditioner(self._h, None))
            return
        if isinstance(preconditioner, str):
            if preconditioner != "learned":
                raise RuntimeError("[libmobility_amd] preconditioner must be 'learned', an object or None")
            preconditioner = default_network(self._include_angular)
        self._precond = preconditioner  # keep alive while attached
        _ffi.check(self._lib.lm_nbody_set_preconditioner(self._h, preconditioner._h))

    def preconditionerBlocks(self):
        """Extension: the blocks G_i (N, d, d) of the attached preconditioner
        (d = 3, or 6 acting on a particle's linear and angular entries)."""
        N = self.getNumberParticles()
        d = getattr(self._precond, "block_dim", 3)
        G = np.zeros((N, d, d))
        _ffi.check(self._lib.lm_nbody_precond_blocks(self._h, G.ctypes.data))
        return G

    def setCoarseMaxDim(self, maxDim=24576):
        """Extension: the largest coarse dimension 3C of one system (default
        24576); denser requests get bigger cells.  The build holds three dense
        (3C)^2 FP64 matrices per system."""
        _ffi.check(self._lib.lm_nbody_set_coarse_max_dim(self._h, int(maxDim)))

    def setCoarsePairPrecision(self, precision="fp32"):
        """Extension: pair arithmetic of the coarse matrix Q^T M Q; 'fp32'
        (default) falls back to FP64 by itself when the FP32 matrix turns out
        indefinite, 'fp64' always uses FP64."""
        if precision not in _ffi.PAIR_PRECISION:
            raise RuntimeError("[libmobility_amd] pair precision must be 'fp64' or 'fp32'")
        _ffi.check(self._lib.lm_nbody_set_coarse_pair_precision(self._h,
                                                                 _ffi.PAIR_PRECISION[precision]))

    def coarseUsedFP64(self):
        """Extension: whether the last coarse build's matrix came from FP64 pairs."""
        used = ctypes.c_int(0)
        _ffi.check(self._lib.lm_nbody_coarse_used_fp64(self._h, ctypes.byref(used)))
        return bool(used.value)

    def setCoarseLevel(self, particlesPerCell=40):
        """Extension: coarse level of the sqrtMdotW preconditioner (every system
        of an ensemble gets its own cells and coarse block).

        Particles are grouped into spatial cells of about ``particlesPerCell``
        particles.  Fewer particles per cell cut the iterations and cost a
        larger build (cubic in the cell count); at N = 131072, 40 per cell gave
        the lowest time per call both with a rebuild per call and without.  With Q the orthonormal cell columns and ``Q^T M Q = L L^T``,
        Lanczos runs on ``G M G^T`` with ``G = D (I + Q (L^{-1} - I) Q^T)`` (D the
        learned blocks when a network is attached) and the result is
        ``G^{-1} (G M G^T)^{1/2} W``.  ``0`` or ``None`` switches it off.
        """
        ppc = 0 if particlesPerCell is None else int(particlesPerCell)
        _ffi.check(self._lib.lm_nbody_set_coarse_level(self._h, ppc))

    def coarseCells(self):
        """Extension: (cell of each particle (N,), cells per axis (3,), number of
        non-empty cells) of the last coarse build."""
        N = self.getNumberParticles()
        cell_of = np.zeros(N, dtype=np.int32)
        grid = np.zeros(3, dtype=np.int32)
        nc = ctypes.c_int(0)
        _ffi.check(self._lib.lm_nbody_coarse_cells(self._h, ctypes.addressof(nc),
                                                   grid.ctypes.data, cell_of.ctypes.data))
        return cell_of, grid, nc.value

    def coarseFactors(self, system=0):
        """Extension: (L, L^{-1}) of one system's coarse block, (3C_s, 3C_s) each."""
        nc = ctypes.c_int(0)
        _ffi.check(self._lib.lm_nbody_coarse_system_cells(self._h, int(system),
                                                          ctypes.addressof(nc)))
        L = np.zeros((3 * nc.value, 3 * nc.value))
        X = np.zeros((3 * nc.value, 3 * nc.value))
        _ffi.check(self._lib.lm_nbody_coarse_factors(self._h, int(system), L.ctypes.data,
                                                     X.ctypes.data))
        return L, X

    def coarseBuildMs(self):
        """Extension: wall-clock milliseconds of the last coarse build."""
        ms = ctypes.c_double(0.0)
        _ffi.check(self._lib.lm_nbody_coarse_build_ms(self._h, ctypes.addressof(ms)))
        return ms.value

    def coarseBuilds(self):
        """Extension: number of coarse builds so far."""
        n = ctypes.c_int64(0)
        _ffi.check(self._lib.lm_nbody_coarse_builds(self._h, ctypes.addressof(n)))
        return n.value

    def setPreconditionerRefresh(self, interval=1):
        """Extension: keep the preconditioner (learned blocks and coarse level)
        for ``interval`` setPositions calls of the same systems before
        rebuilding it.  The result stays an exact square root of the current M
        (G keeps its exact inverse); only the iteration count can drift."""
        _ffi.check(self._lib.lm_nbody_set_preconditioner_refresh(self._h, int(interval)))

    def initialize(self, viscosity, hydrodynamicRadius, includeAngular=False, tolerance=1e-4,
                   seed=None, lanczos_callback=None):
        """pythonify.h:241-260 -> NBody::initialize (mobility.h:99-114)."""
        cb = _ffi.LANCZOS_CALLBACK()
        if lanczos_callback is not None:
            user_cb = lanczos_callback

            def _tramp(i, err, _user):
                user_cb(int(i), float(err))

            cb = _ffi.LANCZOS_CALLBACK(_tramp)
        self._cb = cb  # keep alive
        s = 0 if seed is None else int(seed)
        _ffi.check(self._lib.lm_nbody_initialize(
            self._h, float(viscosity), float(hydrodynamicRadius), int(bool(includeAngular)),
            float(tolerance), ctypes.c_uint64(s & 0xFFFFFFFFFFFFFFFF), cb, None))
        self._include_angular = bool(includeAngular)

    def sqrtMdotW_with_noise(self, noise, prefactor=1.0):
        """Extension: prefactor * M^{1/2} W for a given W (3N, or 6N with angular)."""
        linear, angular, mem = self._outputs_like_positions()
        w = _Arr(noise).moved_to(mem, _last_device)
        n = self.getNumberParticles() * (6 if self._include_angular else 3)
        if w.size != n:
            raise RuntimeError(f"[libmobility_amd] noise must have {n} entries")
        self._bind_stream(mem, _last_device)
        _ffi.check(self._lib.lm_nbody_sqrt_mdot_w_noise(self._h, w.ptr, self._out(linear), self._out(angular),
                                                        float(prefactor), mem))
        return linear, angular

    def divM_with_noise(self, noise, prefactor=1.0):
        """Extension: the random finite difference of divM for a given
        displacement W (3N values, random_finite_differences.h:46-62)."""
        if self.getNumberParticles() <= 0:
            raise RuntimeError(
                "[libMobility] The number of particles is not set. Did you forget to call "
                "setPositions?")
        linear, angular, mem = self._outputs_like_positions()
        w = _Arr(noise).moved_to(mem, _last_device)
        n = 3 * self.getNumberParticles()
        if w.size != n:
            raise RuntimeError(f"[libmobility_amd] noise must have {n} entries")
        self._bind_stream(mem, _last_device)
        _ffi.check(self._lib.lm_nbody_div_m_noise(self._h, w.ptr, self._out(linear), self._out(angular),
                                                  float(prefactor), mem))
        return linear, angular

    # -- several vectors per pass (extension) ----------------------------------------
    def _stacked(self, x, what):
        a = _Arr(x)
        n3 = 3 * self.getNumberParticles()
        if n3 == 0 or a.size == 0 or a.size % n3 != 0 or a.ndim < 2 or a.shape[0] * n3 != a.size:
            raise RuntimeError(f"[libmobility_amd] {what} must have shape (nrhs, N, 3) or (nrhs, 3N)")
        return a, a.size // n3

    def MdotMulti(self, forces):
        """Extension: the translational product for several force vectors in
        one pass over the pairs.  ``forces`` has shape (nrhs, N, 3) or
        (nrhs, 3N); the result has the same shape and framework."""
        f, nrhs = self._stacked(forces, "forces")
        out = _zeros(f.shape, f.framework, f.device)
        self._bind_stream(f.mem, f.device)
        _ffi.check(self._lib.lm_nbody_mdot_multi(self._h, f.ptr, nrhs, self._out(out, nrhs), f.mem))
        return out

    def sqrtMdotWMulti(self, nrhs, prefactor=1.0):
        """Extension: ``nrhs`` independent square roots (the noise that ``nrhs``
        successive sqrtMdotW calls would draw) sharing every pass over the
        pairs; shape (nrhs,) + the positions' form."""
        shape = (int(nrhs),) + self._own_shape()
        out = _zeros(shape, _last_framework, _last_device)
        mem = _mem_of(_last_framework, _last_device)
        self._bind_stream(mem, _last_device)
        _ffi.check(self._lib.lm_nbody_sqrt_mdot_w_multi(self._h, None, int(nrhs), self._out(out, int(nrhs)),
                                                        float(prefactor), mem))
        return out

    def sqrtMdotWMulti_with_noise(self, noise, prefactor=1.0):
        """Extension: prefactor * M^{1/2} W_r for given W of shape (nrhs, 3N) or (nrhs, N, 3)."""
        w, nrhs = self._stacked(noise, "noise")
        out = _zeros(w.shape, w.framework, w.device)
        self._bind_stream(w.mem, w.device)
        _ffi.check(self._lib.lm_nbody_sqrt_mdot_w_multi(self._h, w.ptr, nrhs, self._out(out, nrhs),
                                                        float(prefactor), w.mem))
        return out

    def lastLanczosIterationsMulti(self, nrhs):
        out = np.zeros(int(nrhs), dtype=np.int32)
        _ffi.check(self._lib.lm_nbody_last_lanczos_iterations_multi(self._h, out.ctypes.data,
                                                                    int(nrhs)))
        return out

    def lastLanczosIterations(self):
        """Extension: iteration index at which the last sqrtMdotW converged."""
        return int(self._lib.lm_nbody_last_lanczos_iterations(self._h))
