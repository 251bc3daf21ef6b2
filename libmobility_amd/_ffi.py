"""ctypes binding of the C ABI in include/libmobility_amd.h.

The shared library is built in-tree (libmobility_amd/lib/liblmamd.so).  There
is no fallback: if it is missing, importing the solver raises ImportError.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "liblmamd.so")

LM_OK = 0
LM_MEM_HOST = 0
LM_MEM_DEVICE = 1

PERIODICITY = {"single_wall": 0, "two_walls": 1, "open": 2, "periodic": 3, "unspecified": 4}
ALGORITHM = {"fast": 0, "naive": 1, "block": 2, "advise": 3}
PAIR_PRECISION = {"fp64": 0, "fp32": 1}

LANCZOS_CALLBACK = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_double, ctypes.c_void_p)

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64

# name -> (restype, argtypes); the list is also what the CPU tests check
# against the declarations in include/libmobility_amd.h.
SIGNATURES = {
    "lm_last_error": (ctypes.c_char_p, []),
    "lm_version": (ctypes.c_char_p, []),
    "lm_precision": (ctypes.c_char_p, []),
    "lm_nbody_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "lm_nbody_destroy": (ctypes.c_int, [_vp]),
    "lm_nbody_set_parameters": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_double, ctypes.c_double]),
    "lm_nbody_set_pair_precision": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lm_nbody_set_drift_precision": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lm_nbody_set_stream": (ctypes.c_int, [_vp, _vp]),
    "lm_nbody_set_pair_tiling": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "lm_nbody_set_pair_symmetric": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lm_nbody_set_pair_symmetric_memory": (ctypes.c_int, [_vp, ctypes.c_int64]),
    "lm_nbody_last_pair_symmetric": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "lm_nbody_set_autotune": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lm_nbody_get_pair_tiling": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int),
                                                ctypes.POINTER(ctypes.c_int)]),
    "lm_nbody_initialize": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                           ctypes.c_double, ctypes.c_uint64, LANCZOS_CALLBACK, _vp]),
    "lm_nbody_set_positions": (ctypes.c_int, [_vp, _vp, _i64, ctypes.c_int]),
    "lm_nbody_set_systems": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp]),
    "lm_nbody_set_persistent": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lm_nbody_set_lanczos_iteration_limit": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lm_nbody_set_lanczos_reorthogonalization": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lm_nbody_set_lanczos_per_system": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lm_nbody_last_lanczos_iterations_per_system": (ctypes.c_int, [_vp, _vp, ctypes.c_int]),
    "lm_nbody_num_systems": (ctypes.c_int, [_vp]),
    "lm_nbody_mdot_multi": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, ctypes.c_int]),
    "lm_nbody_sqrt_mdot_w_multi": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, ctypes.c_double,
                                                  ctypes.c_int]),
    "lm_nbody_last_lanczos_iterations_multi": (ctypes.c_int, [_vp, _vp, ctypes.c_int]),
    "lm_nbody_num_particles": (_i64, [_vp]),
    "lm_nbody_mdot": (ctypes.c_int, [_vp, _vp, _i64, _vp, _i64, _vp, _vp, ctypes.c_int]),
    "lm_nbody_sqrt_mdot_w": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, ctypes.c_int]),
    "lm_nbody_sqrt_mdot_w_noise": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int]),
    "lm_nbody_mdot_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "lm_nbody_div_m": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, ctypes.c_int]),
    "lm_nbody_div_m_noise": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int]),
    "lm_nbody_langevin_velocities": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, _vp, _i64,
                                                    _vp, _i64, _vp, _vp, ctypes.c_int]),
    "lm_nbody_hydrodynamic_velocities": (ctypes.c_int, [_vp, _vp, _i64, _vp, _i64, _vp, _vp, ctypes.c_int]),
    "lm_nbody_clean": (ctypes.c_int, [_vp]),
    "lm_nbody_last_lanczos_iterations": (ctypes.c_int, [_vp]),
    "lm_nbody_synchronize": (ctypes.c_int, [_vp]),
    "lm_precond_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
    "lm_precond_create_blocks": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "lm_precond_describe": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "lm_precond_destroy": (ctypes.c_int, [_vp]),
    "lm_precond_num_weights": (_i64, []),
    "lm_precond_init_weights": (ctypes.c_int, [_vp, ctypes.c_uint64]),
    "lm_precond_set_weights": (ctypes.c_int, [_vp, _vp, _i64]),
    "lm_precond_get_weights": (ctypes.c_int, [_vp, _vp, _i64]),
    "lm_precond_train": (ctypes.c_int, [_vp, _vp, _vp, _i64, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_float, ctypes.c_uint64, _vp]),
    "lm_precond_forward": (ctypes.c_int, [_vp, _vp, _i64, _vp]),
    "lm_precond_loss_grad": (ctypes.c_int, [_vp, _vp, _vp, _i64, _vp, _vp]),
    "lm_nbody_precond_dataset": (ctypes.c_int, [_vp, ctypes.c_int, _i64, ctypes.c_uint64, _vp,
                                                _vp, _vp]),
    "lm_nbody_set_preconditioner": (ctypes.c_int, [_vp, _vp]),
    "lm_nbody_precond_blocks": (ctypes.c_int, [_vp, _vp]),
    "lm_nbody_precond_neighbours": (ctypes.c_int, [_vp, _vp]),
    "lm_nbody_set_coarse_level": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lm_nbody_set_coarse_max_dim": (ctypes.c_int, [_vp, ctypes.c_int64]),
    "lm_nbody_set_coarse_pair_precision": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lm_nbody_coarse_used_fp64": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "lm_nbody_coarse_cells": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "lm_nbody_coarse_factors": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
    "lm_nbody_coarse_system_cells": (ctypes.c_int, [_vp, ctypes.c_int, _vp]),
    "lm_nbody_coarse_build_ms": (ctypes.c_int, [_vp, _vp]),
    "lm_nbody_coarse_builds": (ctypes.c_int, [_vp, _vp]),
    "lm_nbody_set_preconditioner_refresh": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lm_selftest_pair_block_host": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp,
                                                   ctypes.c_double, _dp]),
    "lm_self_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "lm_self_destroy": (ctypes.c_int, [_vp]),
    "lm_self_set_stream": (ctypes.c_int, [_vp, _vp]),
    "lm_self_initialize": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_uint64]),
    "lm_self_set_positions": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_int]),
    "lm_self_num_particles": (ctypes.c_int64, [_vp]),
    "lm_self_mdot": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, _vp, ctypes.c_int64, _vp, _vp,
                                    ctypes.c_int]),
    "lm_self_sqrt_mdot_w": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, ctypes.c_int]),
    "lm_self_div_m": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, ctypes.c_int]),
    "lm_self_langevin_velocities": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, _vp,
                                                   ctypes.c_int64, _vp, ctypes.c_int64, _vp, _vp,
                                                   ctypes.c_int]),
    "lm_self_clean": (ctypes.c_int, [_vp]),
    "lm_selftest_pair_both_host": (ctypes.c_int, [ctypes.c_int, _dp, _dp, ctypes.c_double, _dp,
                                                  _dp]),
    "lm_selftest_tridiag_sqrt": (ctypes.c_int, [_dp, _dp, ctypes.c_int, _dp]),
    "lm_selftest_philox_host": (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint32),
                                               ctypes.POINTER(ctypes.c_uint32),
                                               ctypes.POINTER(ctypes.c_uint32)]),
    "lm_selftest_coarse_cells_host": (ctypes.c_int, [_dp, ctypes.c_int64, ctypes.c_int, _vp, _vp,
                                                     _vp]),
}

_lib = None


def _share_torch_hip_runtime():
    """PyTorch-ROCm wheels carry their own libamdhip64 (same SONAME as the
    system one).  Loading torch first lets this library bind to that runtime;
    loading the system runtime first would put a second HIP runtime in the
    process, after which torch finds no GPU."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load():
    """Load the HIP library (raises ImportError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    _share_torch_hip_runtime()
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libmobility_amd: {LIB_PATH} is missing; build it with "
            "`python -m libmobility_amd.build` (hipcc, gfx950). There is no CPU fallback.")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status != LM_OK:
        msg = load().lm_last_error().decode("utf-8", "replace")
        raise RuntimeError(msg)
