"""Benchmark: NBody Mdot throughput and Lanczos sqrtMdotW time on one MI355X.

Default workload (one GPU): open-boundary RPY, N = 131072 particles of radius
a = 1, random sequential addition at volume fraction 0.20 in a cube of side
140.0a, viscosity 1/(6 pi), forces and Brownian noise W i.i.d. N(0,1).

A "step" is one Mdot over the whole system (N^2 ordered pairs, self pair
included) in FP32-pair / FP64-accumulate mode, inputs and outputs resident in
HBM.  K steps are timed with HIP events on the solver's stream.  The Lanczos
square root (tol 1e-6) is timed separately (median of a few calls) and
reported with its iteration count, for the plain FP32-pair and FP64-pair
operators and the two-level preconditioned FP32-pair operator.  The CPU baseline
is the oracle's OpenMP FP64 matrix-free product on a bounded sample of target
rows, run on this host's cores.

Other workloads (``--workload``), each printing one JSON line of its own:
  c1  open RPY, N = 1024 uniform in a cube of side 24a (overlaps), FP64 Mdot
      and sqrtMdotW (tol 1e-6) against the dense FP64 matrix and its
      eigen-square-root; CPU baseline = dense build + matvec + eigh square root
  c2  single wall, N = 65536, x,y in [0,256a), z in [1.5a,20a], forces and
      torques: FP64 vs FP32-pair Mdot, error per block (UF, UT, WF, WT)
  c4  512 ragged systems (N log-uniform in [64, 8192]), half open, half
      single wall: one persistent launch vs per-system solves
  c5  single wall, N = 1048576, x,y in [0,1024a), z in [1.5a,30a], forces and
      torques: FP32/FP64 Mdot and hydrodynamicVelocities (Mdot + sqrtMdotW at
      tol 1e-6 + RFD drift with delta 1e-4 a), FP32 pairs, two-level
      preconditioner (or plain Lanczos with --c5-plain)

Prints exactly one JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes
import json
import math
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from libmobility_amd import workloads  # noqa: E402  (numpy generators only)

METRIC = "Mdot Gpair/s + sqrtMdotW ms at N=131072 tol 1e-6: learned-prec vs plain"
# coarse level of the preconditioned sqrtMdotW: target particles per cell
COARSE_PPC = 40
# C5: the coarse dimension cap raised from the library's default 24576 to
# 49152 (16200 cells instead of 7938: 53 Lanczos iterations instead of 58,
# build 3.9 s, three dense 49152^2 FP64 matrices = 58 GB of HBM)
C5_COARSE_MAX_DIM = 49152
PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 peak
PEAK_FP64_TFLOPS = 78.6
ETA = 1.0 / (6.0 * np.pi)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def flops_per_pair():
    with open(os.path.join(ROOT, "libmobility_amd", "flop_model.json")) as f:
        return json.load(f)["flops_per_pair"]


class Gpu:
    """Device buffers, the torch stream and C-ABI solver handles."""

    def __init__(self):
        import torch
        from libmobility_amd import _ffi
        self.torch = torch
        self.ffi = _ffi
        self.lib = _ffi.load()
        self.dev = torch.device("cuda:0")
        self.stream = torch.cuda.current_stream(self.dev)

    def put(self, a):
        a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
        return self.torch.from_numpy(a.copy()).to(self.dev)

    def zeros(self, n):
        return self.torch.zeros(n, dtype=self.torch.float64, device=self.dev)

    def sync(self):
        self.torch.cuda.synchronize()

    def solver(self, pos_d, prec, wall=False, angular=False, tol=1e-6, delta=1e-3, autotune=True,
               tiling=None, sizes=None, kinds=None, persistent=-1, iteration_limit=None,
               seed=12345, callback=None):
        """callback(i, err): the Lanczos callback of lanczos::Solver::run; with it
        set, the estimate and its error are computed at every iteration."""
        ffi, lib = self.ffi, self.lib
        cb = ffi.LANCZOS_CALLBACK()
        if callback is not None:
            cb = ffi.LANCZOS_CALLBACK(lambda i, e, _u: callback(int(i), float(e)))
            self._callbacks = getattr(self, "_callbacks", []) + [cb]  # keep alive
        h = ctypes.c_void_p()
        ffi.check(lib.lm_nbody_create(2, 2, 0 if wall else 2, ctypes.byref(h)))
        ffi.check(lib.lm_nbody_set_parameters(h, 3, -1, -1, int(wall), 0.0, delta))
        ffi.check(lib.lm_nbody_set_pair_precision(h, ffi.PAIR_PRECISION[prec]))
        ffi.check(lib.lm_nbody_set_stream(h, ctypes.c_void_p(self.stream.cuda_stream)))
        ffi.check(lib.lm_nbody_initialize(h, ETA, 1.0, int(angular), tol, seed, cb, None))
        if sizes is not None:
            off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
            k = np.ascontiguousarray(kinds, dtype=np.int32)
            ffi.check(lib.lm_nbody_set_systems(h, off.ctypes.data, len(sizes), k.ctypes.data))
            ffi.check(lib.lm_nbody_set_persistent(h, persistent))
        if iteration_limit is not None:
            ffi.check(lib.lm_nbody_set_lanczos_iteration_limit(h, iteration_limit))
        ffi.check(lib.lm_nbody_set_positions(h, pos_d.data_ptr(), pos_d.numel(), ffi.LM_MEM_DEVICE))
        if tiling is not None:
            ffi.check(lib.lm_nbody_set_pair_tiling(h, tiling[0], tiling[1]))
        else:
            ffi.check(lib.lm_nbody_set_autotune(h, int(autotune)))
        return h

    def evaluation(self, h):
        """'symmetric' if the last product evaluated each unordered pair once."""
        used = ctypes.c_int()
        self.ffi.check(self.lib.lm_nbody_last_pair_symmetric(h, ctypes.byref(used)))
        return "symmetric" if used.value else "ordered"

    def tiling(self, h):
        t, c = ctypes.c_int(), ctypes.c_int()
        self.ffi.check(self.lib.lm_nbody_get_pair_tiling(h, ctypes.byref(t), ctypes.byref(c)))
        return {"targets_per_lane": t.value, "source_chunks": c.value}

    def mdot(self, h, F=None, T=None, U=None, W=None):
        p = lambda x: None if x is None else x.data_ptr()
        self.ffi.check(self.lib.lm_nbody_mdot(h, p(F), 0 if F is None else F.numel(), p(T),
                                              0 if T is None else T.numel(), p(U), p(W),
                                              self.ffi.LM_MEM_DEVICE))

    def time_events(self, fn, steps, warmup):
        """ms per call of fn, HIP events on the solver stream."""
        for _ in range(warmup):
            fn()
        self.sync()
        e0 = self.torch.cuda.Event(enable_timing=True)
        e1 = self.torch.cuda.Event(enable_timing=True)
        e0.record(self.stream)
        for _ in range(steps):
            fn()
        e1.record(self.stream)
        self.sync()
        return e0.elapsed_time(e1) / max(steps, 1)

    def coarse_cells(self, h):
        nc = ctypes.c_int()
        grid = (ctypes.c_int * 3)()
        self.ffi.check(self.lib.lm_nbody_coarse_cells(h, ctypes.byref(nc), grid, None))
        return nc.value

    def coarse_build_ms(self, h):
        ms = ctypes.c_double()
        self.ffi.check(self.lib.lm_nbody_coarse_build_ms(h, ctypes.byref(ms)))
        return ms.value

    def sqrt_noise(self, h, Wn, U, W=None, prefactor=1.0):
        self.ffi.check(self.lib.lm_nbody_sqrt_mdot_w_noise(
            h, Wn.data_ptr(), U.data_ptr(), None if W is None else W.data_ptr(), prefactor,
            self.ffi.LM_MEM_DEVICE))

    def time_lanczos(self, h, Wn, n_out, calls, angular=False, fresh_pos=None):
        """Median ms per sqrtMdotW call after one untimed call.  With fresh_pos,
        each timed call starts with setPositions(fresh_pos), so it includes
        what new positions cost (records, and with a preconditioner the kNN
        graph, the network and the coarse level)."""
        U = self.zeros(n_out)
        W = self.zeros(n_out) if angular else None
        times, iters = [], []
        for c in range(calls + 1):
            self.sync()
            t = time.perf_counter()
            if fresh_pos is not None:
                self.ffi.check(self.lib.lm_nbody_set_positions(
                    h, fresh_pos.data_ptr(), fresh_pos.numel(), self.ffi.LM_MEM_DEVICE))
            self.sqrt_noise(h, Wn, U, W)
            self.sync()
            dt = (time.perf_counter() - t) * 1e3
            it = self.lib.lm_nbody_last_lanczos_iterations(h)
            if c > 0 or calls == 0:
                times.append(dt)
                iters.append(it)
            log(f"  sqrtMdotW call {c}: {dt:.1f} ms, iterations {it}")
        return statistics.median(times), iters

    def destroy(self, *hs):
        for h in hs:
            self.lib.lm_nbody_destroy(h)


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def result_line(metric, value, unit, steps, warmup, ms, higher, dtype, config, roofline, extra,
                cpu):
    return {
        "metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": ms, "higher_is_better": higher, "scaling": "weak",
        "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config,
        "roofline": roofline, "cpu_baseline": cpu, "extra": extra,
    }


# HBM bytes per FP32 pair-kernel launch, from rocprofv3 --pmc FETCH_SIZE (x 2.00)
# and WRITE_SIZE (x 1.00); the scales come from a 2 GiB device copy.  Keyed by
# (workload, N, pair evaluation, targets per lane, source chunks) of the
# profiled configuration.
# C3 tiling of the bench line and its profiles: 4 targets per lane, 32
# offset chunks
C3_TILING = (4, 32)
# C1 tiling (the autotuner's usual pick there), pinned so that profiles of
# the same command run the same kernels: 1 target per lane, 16 source chunks
C1_TILING = (1, 16)

# FP32 VALU flops issued per launch, SQ_INSTS_VALU_FLOPS_FP32 x 64 lanes from
# rocprofv3 --pmc, same keys as MEASURED_TRAFFIC_BYTES.
MEASURED_FP32_FLOPS = {
    ("c3", 131072, "symmetric", 4, 64): 7.134e9 * 64,
    ("c3", 131072, "symmetric", 4, 32): 7.00029e9 * 64,
    ("c2", 65536, "symmetric", 1, 64): 1.25718e10 * 64,
    ("c4", 816182, "symmetric", 2, 8): 3.19152e9 * 64,
}

MEASURED_TRAFFIC_BYTES = {
    ("c3", 131072, "ordered", 2, 32): (10454.1 * 2.0 + 98304.0) * 1024.0,
    ("c3", 131072, "symmetric", 4, 64): (152410.0 * 2.0 + 394750.0) * 1024.0,
    ("c3", 131072, "symmetric", 4, 32): (104635.2 * 2.0 + 296448.0) * 1024.0,
    ("c3", 131072, "symmetric", 2, 64): (177902.1 * 2.0 + 591360.0) * 1024.0,
    ("c2", 65536, "symmetric", 1, 64): (130223.0 * 2.0 + 591360.0) * 1024.0,
    ("c4", 816182, "symmetric", 2, 8): (108844.0 * 2.0 + 165989.0) * 1024.0,
    ("c1", 1024, "ordered", 1, 16): (173.0 * 2.0 + 384.0) * 1024.0,
}


def compute_roofline(flops_per_launch, ms, peak_tflops, traffic=None):
    achieved = flops_per_launch / (ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": peak_tflops, "unit": "TFLOP/s",
            "frac": achieved / peak_tflops, "traffic": traffic}


def counted_roofline(key, flops_per_launch, ms):
    """FP32 roofline of a pair product, with the HBM traffic and the FP32 flops
    the hardware issues (SQ counters) where they were measured for `key`."""
    roof = compute_roofline(flops_per_launch, ms, PEAK_FP32_TFLOPS, MEASURED_TRAFFIC_BYTES.get(key))
    issued = MEASURED_FP32_FLOPS.get(key)
    if issued is not None:
        roof["hw_achieved"] = issued / (ms * 1e-3) / 1e12
        roof["hw_frac"] = roof["hw_achieved"] / PEAK_FP32_TFLOPS
    return roof


# ---------------------------------------------------------------------------
# C3 (default): open RPY, N = 131072, RSA phi = 0.2


def cpu_baseline_rows(pos, F, target_seconds=12.0, first=256):
    """CPU baseline of C3: the oracle's OpenMP FP64 open-RPY UF product over a
    bounded sample of target rows (rows 0..R-1, R sized for ~12 s).  Returns
    the baseline record and the sampled rows' velocities (the checker)."""
    import oracle
    N = pos.shape[0]
    t = time.perf_counter()
    oracle.mdot_open_uf_rows(pos, ETA, 1.0, F, 0, first)
    dt = time.perf_counter() - t
    rows = int(min(N, max(first, first * target_seconds / max(dt, 1e-6))))
    t = time.perf_counter()
    U = oracle.mdot_open_uf_rows(pos, ETA, 1.0, F, 0, rows)
    dt = time.perf_counter() - t
    pairs = rows * N
    cpu = {
        "value": pairs / dt / 1e9, "unit": "Gpair/s", "cores": oracle.num_threads(),
        "kind": "port",
        "sample": f"FP64 open-RPY UF product, target rows 0..{rows - 1} x all {N} sources "
                  f"({pairs:.3e} pairs, {dt:.1f} s), OpenMP C oracle",
    }
    return cpu, U


def bbt_errors(g, net, N=512, phi=0.20):
    """||B B^T - M||_F / ||M||_F of the FP32-pair Lanczos square root, plain and
    with the two-level preconditioner (learned blocks + coarse level), at tol
    1e-4 and 1e-6, on an RSA packing of
    N particles.  Column k of B is sqrtMdotW applied to e_k; M is the solver's
    own FP64-pair matrix (Mdot of e_k), which the tests hold to 1e-12 of the
    oracle's dense matrix."""
    from libmobility_amd.workloads import rsa_cube
    pos, _ = rsa_cube(N, phi, seed=5)
    n = 3 * N
    pos_d = g.put(pos)
    E = g.torch.eye(n, dtype=g.torch.float64, device=g.dev)
    h64 = g.solver(pos_d, "fp64", autotune=False)
    Mcols = g.zeros(n * n).view(n, n)
    for k in range(n):
        g.mdot(h64, E[k], None, Mcols[k])
    g.sync()
    g.destroy(h64)
    M = Mcols.cpu().numpy().T
    out = {"N": N, "phi": phi}
    for label, use_net in (("plain", False), ("learned", True)):
        for tol in (1e-4, 1e-6):
            h = g.solver(pos_d, "fp32", tol=tol, autotune=False)
            if use_net:
                g.ffi.check(g.lib.lm_nbody_set_preconditioner(h, net._h))
                g.ffi.check(g.lib.lm_nbody_set_coarse_level(h, COARSE_PPC))
            Bcols = g.zeros(n * n).view(n, n)
            its = []
            for k in range(n):
                g.sqrt_noise(h, E[k], Bcols[k])
                its.append(g.lib.lm_nbody_last_lanczos_iterations(h))
            g.sync()
            g.destroy(h)
            B = Bcols.cpu().numpy().T
            err = float(np.linalg.norm(B @ B.T - M) / np.linalg.norm(M))
            out[f"{label}_tol{tol:g}"] = {"rel_err": err, "iterations_mean": float(np.mean(its))}
            log(f"  ||BB^T - M||/||M|| {label} tol {tol:g}: {err:.2e} "
                f"({np.mean(its):.1f} iterations)")
    return out


def multi_vector_leg(g, h, F_d, W_d, N, ms_single, extra, args, nrhs=4):
    """Several vectors per pass: the FP32 product of nrhs force vectors and
    nrhs square roots at tol 1e-6 sharing each pass, against the single-vector
    figures of the same run."""
    n3 = 3 * N
    F4 = F_d.repeat(nrhs).view(nrhs, n3).clone()
    F4 += 0.01 * g.torch.arange(nrhs, dtype=g.torch.float64, device=g.dev).view(nrhs, 1)
    U4 = g.zeros(nrhs * n3)
    log(f"{nrhs}-vector FP32 product")

    def mdot4():
        g.ffi.check(g.lib.lm_nbody_mdot_multi(h, F4.data_ptr(), nrhs, U4.data_ptr(),
                                              g.ffi.LM_MEM_DEVICE))

    ms4 = g.time_events(mdot4, max(3, args.steps // 2), 1)
    out = {
        "mdot_multi4_fp32pair_ms": ms4,
        "mdot_multi4_fp32pair_gpair_vectors_s": nrhs * float(N) ** 2 / (ms4 * 1e-3) / 1e9,
        "mdot_multi4_speedup_vs_single": nrhs * ms_single / ms4,
    }
    log(f"  {ms4:.3f} ms for {nrhs} vectors ({nrhs * ms_single / ms4:.2f}x the single-vector rate)")
    if not args.no_lanczos and "sqrtMdotW_plain_fp32pair_ms" in extra:
        W4 = W_d.repeat(nrhs).view(nrhs, n3).clone()
        W4[1:] = g.torch.randn(nrhs - 1, n3, dtype=g.torch.float64, device=g.dev,
                               generator=g.torch.Generator(device=g.dev).manual_seed(3))
        Y4 = g.zeros(nrhs * n3)
        times = []
        for c in range(2):
            g.sync()
            t = time.perf_counter()
            g.ffi.check(g.lib.lm_nbody_sqrt_mdot_w_multi(h, W4.data_ptr(), nrhs, Y4.data_ptr(), 1.0,
                                                         g.ffi.LM_MEM_DEVICE))
            g.sync()
            times.append((time.perf_counter() - t) * 1e3)
            log(f"  {nrhs}-vector sqrtMdotW call {c}: {times[-1]:.1f} ms")
        its = np.zeros(nrhs, dtype=np.int32)
        g.ffi.check(g.lib.lm_nbody_last_lanczos_iterations_multi(h, its.ctypes.data, nrhs))
        out.update({
            "sqrtMdotW_multi4_fp32pair_ms": times[-1],
            "sqrtMdotW_multi4_iterations": its.tolist(),
            "sqrtMdotW_multi4_speedup_vs_4_single": nrhs * extra["sqrtMdotW_plain_fp32pair_ms"] / times[-1],
        })
    return out


def bench_c3(args):
    g = Gpu()
    N = args.N or 131072
    phi = 0.20
    log(f"generating RSA packing N={N} phi={phi}")
    pos, L, F, Wn = workloads.config_c3(N, phi)
    pos_d, F_d, W_d, U_d = g.put(pos), g.put(F), g.put(Wn), g.zeros(3 * N)
    fpp = flops_per_pair()["open_UF"]
    pairs = float(N) * float(N)

    # pinned tiling (4 targets per lane, 32 offset chunks) unless given: the
    # profile summaries are taken at this tiling
    tiling = args.tiling or C3_TILING
    log(f"FP32-pair Mdot, tiling {tiling}")
    h32 = g.solver(pos_d, "fp32", tiling=tiling)
    ms32 = g.time_events(lambda: g.mdot(h32, F_d, None, U_d), args.steps, args.warmup)
    U32 = U_d.cpu().numpy()
    gpair32 = pairs / (ms32 * 1e-3) / 1e9
    log(f"  {ms32:.3f} ms/Mdot, {gpair32:.1f} Gpair/s")
    log("FP64-pair Mdot")
    h64 = g.solver(pos_d, "fp64", tiling=tiling)
    ms64 = g.time_events(lambda: g.mdot(h64, F_d, None, U_d), max(3, args.steps // 4), 1)
    U64 = U_d.cpu().numpy()
    gpair64 = pairs / (ms64 * 1e-3) / 1e9
    log(f"  {ms64:.3f} ms/Mdot, {gpair64:.1f} Gpair/s")

    extra = {
        "pair_evaluation_fp32pair": g.evaluation(h32),
        "pair_evaluation_fp64pair": g.evaluation(h64),
        "tiling_fp32pair": g.tiling(h32),
        "tiling_fp64pair": g.tiling(h64),
        "mdot_fp32pair_ms": ms32,
        "mdot_fp64pair_ms": ms64,
        "mdot_fp64pair_gpair_s": gpair64,
        "mdot_fp64pair_valu_frac_of_fp64_peak": gpair64 * 1e9 * fpp / (PEAK_FP64_TFLOPS * 1e12),
        "flops_per_pair": fpp,
        "mdot_rel_err_fp32pair_vs_fp64pair": rel(U32, U64),
    }
    if extra["pair_evaluation_fp32pair"] == "symmetric":
        log("FP32-pair Mdot, ordered evaluation (every ordered pair, as the reference)")
        ho = g.solver(pos_d, "fp32", tiling=tiling)
        g.ffi.check(g.lib.lm_nbody_set_pair_symmetric(ho, 0))
        ms_o = g.time_events(lambda: g.mdot(ho, F_d, None, U_d), max(3, args.steps // 4), 1)
        g.destroy(ho)
        extra["mdot_fp32pair_ordered_ms"] = ms_o
        extra["mdot_fp32pair_symmetric_speedup"] = ms_o / ms32
        log(f"  {ms_o:.3f} ms/Mdot ordered; symmetric {ms_o / ms32:.2f}x faster")
    if not args.no_lanczos:
        # Every FP32 Lanczos leg uses the plain solver's tiling, so that plain
        # and preconditioned solves apply the same products.  Tilings of equal
        # speed sum in different orders, and near tol 1e-6 the plain FP32
        # iteration count moves with that rounding (133 or 151 at N=131072).
        t = g.tiling(h32)
        lt = (t["targets_per_lane"], t["source_chunks"])
        extra["lanczos_tiling_fp32pair"] = t
        log("plain Lanczos sqrtMdotW, FP32 pairs, tol 1e-6")
        ms_l32, it32 = g.time_lanczos(h32, W_d, 3 * N, args.lanczos_calls)
        extra["sqrtMdotW_plain_fp32pair_ms"] = ms_l32
        extra["sqrtMdotW_plain_fp32pair_iterations"] = it32
        log("plain Lanczos sqrtMdotW, FP64 pairs, tol 1e-6")
        ms_l64, it64 = g.time_lanczos(h64, W_d, 3 * N, max(1, args.lanczos_calls - 2))
        extra["sqrtMdotW_plain_fp64pair_ms"] = ms_l64
        extra["sqrtMdotW_plain_fp64pair_iterations"] = it64
        y32, y64 = g.zeros(3 * N), g.zeros(3 * N)
        g.sqrt_noise(h32, W_d, y32)
        g.sqrt_noise(h64, W_d, y64)
        g.sync()
        extra["sqrtMdotW_rel_diff_fp32pair_vs_fp64pair"] = rel(y32.cpu().numpy(), y64.cpu().numpy())
        log("Lanczos sqrtMdotW with FP64 reorthogonalisation (extension), FP32 pairs, tol 1e-6")
        hr = g.solver(pos_d, "fp32", tiling=lt)
        g.ffi.check(g.lib.lm_nbody_set_lanczos_reorthogonalization(hr, 1))
        ms_lr, itr = g.time_lanczos(hr, W_d, 3 * N, args.lanczos_calls)
        yr = g.zeros(3 * N)
        g.sqrt_noise(hr, W_d, yr)
        g.sync()
        g.destroy(hr)
        extra["sqrtMdotW_reorth_fp32pair_ms"] = ms_lr
        extra["sqrtMdotW_reorth_fp32pair_iterations"] = itr
        extra["sqrtMdotW_reorth_speedup_vs_plain"] = ms_l32 / ms_lr
        extra["sqrtMdotW_rel_diff_reorth_fp32pair_vs_plain_fp64pair"] = rel(yr.cpu().numpy(),
                                                                          y64.cpu().numpy())
        del yr
        log("plain Lanczos sqrtMdotW, FP32 pairs, tol 1e-6, new positions before each call")
        ms_l32p, _ = g.time_lanczos(h32, W_d, 3 * N, args.lanczos_calls, fresh_pos=pos_d)
        extra["sqrtMdotW_plain_fp32pair_with_setPositions_ms"] = ms_l32p
        from libmobility_amd.precond import default_network
        net = default_network(False)
        log(f"two-level preconditioned Lanczos sqrtMdotW (learned 3x3 blocks + coarse level of "
            f"{COARSE_PPC} particles per cell), FP32 pairs, tol 1e-6")
        hl = g.solver(pos_d, "fp32", tiling=lt)
        g.ffi.check(g.lib.lm_nbody_set_preconditioner(hl, net._h))
        g.ffi.check(g.lib.lm_nbody_set_coarse_level(hl, COARSE_PPC))
        ms_ll, itl = g.time_lanczos(hl, W_d, 3 * N, args.lanczos_calls)
        log("  the same with new positions before each call (kNN, network and coarse build "
            "included)")
        ms_llp, _ = g.time_lanczos(hl, W_d, 3 * N, args.lanczos_calls, fresh_pos=pos_d)
        cells = g.coarse_cells(hl)
        build = g.coarse_build_ms(hl)
        # the last timed call rebuilt the preconditioner; kept for `refresh`
        # setPositions calls, the next `refresh` calls hold exactly one build
        refresh = 10
        log(f"  the same, preconditioner kept for {refresh} setPositions calls (one build in "
            f"{refresh} calls)")
        g.ffi.check(g.lib.lm_nbody_set_preconditioner_refresh(hl, refresh))
        nb0, nb = ctypes.c_int64(), ctypes.c_int64()
        g.ffi.check(g.lib.lm_nbody_coarse_builds(hl, ctypes.byref(nb0)))
        Ur = g.zeros(3 * N)
        g.sync()
        t = time.perf_counter()
        for _ in range(refresh):
            g.ffi.check(g.lib.lm_nbody_set_positions(hl, pos_d.data_ptr(), pos_d.numel(),
                                                     g.ffi.LM_MEM_DEVICE))
            g.sqrt_noise(hl, W_d, Ur)
        g.sync()
        ms_refresh = (time.perf_counter() - t) * 1e3 / refresh
        g.ffi.check(g.lib.lm_nbody_coarse_builds(hl, ctypes.byref(nb)))
        g.destroy(hl)
        log(f"  {ms_refresh:.1f} ms per call, {nb.value - nb0.value} coarse build(s)")
        log("  the same (positions fixed) with FP64 reorthogonalisation")
        hlr = g.solver(pos_d, "fp32", tiling=lt)
        g.ffi.check(g.lib.lm_nbody_set_preconditioner(hlr, net._h))
        g.ffi.check(g.lib.lm_nbody_set_coarse_level(hlr, COARSE_PPC))
        g.ffi.check(g.lib.lm_nbody_set_lanczos_reorthogonalization(hlr, 1))
        ms_llr, itlr = g.time_lanczos(hlr, W_d, 3 * N, max(1, args.lanczos_calls - 1))
        g.destroy(hlr)
        extra["sqrtMdotW_two_level_reorth_fp32pair_ms"] = ms_llr
        extra["sqrtMdotW_two_level_reorth_fp32pair_iterations"] = itlr
        extra["sqrtMdotW_two_level_preconditioner"] = (
            f"learned 3x3 blocks (64-feature MLP, BF16 MFMA, trained on the GPU) + coarse level "
            f"({COARSE_PPC} particles per cell target, {cells} cells); on this open packing the "
            f"speedup comes from the coarse level: the learned blocks neither gain nor lose "
            f"(see sqrtMdotW_ablation)")
        extra["sqrtMdotW_two_level_fp32pair_ms"] = ms_ll
        extra["sqrtMdotW_two_level_fp32pair_iterations"] = itl
        extra["sqrtMdotW_two_level_speedup_vs_plain"] = ms_l32 / ms_ll
        extra["sqrtMdotW_two_level_with_setPositions_ms"] = ms_llp
        extra["sqrtMdotW_two_level_speedup_vs_plain_with_setPositions"] = ms_l32p / ms_llp
        extra["sqrtMdotW_two_level_with_setPositions_refresh10_ms"] = ms_refresh
        extra["sqrtMdotW_two_level_refresh10_coarse_builds"] = nb.value - nb0.value
        extra["sqrtMdotW_two_level_speedup_vs_plain_with_setPositions_refresh10"] = (
            ms_l32p / ms_refresh)
        extra["coarse_build_ms"] = build
        log("tol 1e-4: plain and two-level sqrtMdotW, FP32 pairs")
        hp4 = g.solver(pos_d, "fp32", tol=1e-4, tiling=lt)
        ms_p4, it_p4 = g.time_lanczos(hp4, W_d, 3 * N, max(1, args.lanczos_calls - 1))
        g.destroy(hp4)
        ht4 = g.solver(pos_d, "fp32", tol=1e-4, tiling=lt)
        g.ffi.check(g.lib.lm_nbody_set_preconditioner(ht4, net._h))
        g.ffi.check(g.lib.lm_nbody_set_coarse_level(ht4, COARSE_PPC))
        ms_t4, it_t4 = g.time_lanczos(ht4, W_d, 3 * N, max(1, args.lanczos_calls - 1))
        g.destroy(ht4)
        extra["sqrtMdotW_tol1e-4"] = {
            "plain_fp32pair_ms": ms_p4, "plain_fp32pair_iterations": it_p4,
            "two_level_fp32pair_ms": ms_t4, "two_level_fp32pair_iterations": it_t4,
            "two_level_speedup_vs_plain": ms_p4 / ms_t4}
        log("ablation: coarse level only (no network)")
        hc = g.solver(pos_d, "fp32", tiling=lt)
        g.ffi.check(g.lib.lm_nbody_set_coarse_level(hc, COARSE_PPC))
        ms_c, itc = g.time_lanczos(hc, W_d, 3 * N, max(1, args.lanczos_calls - 1))
        g.destroy(hc)
        log("ablation: learned blocks only (no coarse level)")
        hb = g.solver(pos_d, "fp32", tiling=lt)
        g.ffi.check(g.lib.lm_nbody_set_preconditioner(hb, net._h))
        ms_b, itb = g.time_lanczos(hb, W_d, 3 * N, max(1, args.lanczos_calls - 2))
        g.destroy(hb)
        extra["sqrtMdotW_ablation"] = {
            "coarse_only_ms": ms_c, "coarse_only_iterations": itc,
            "learned_blocks_only_ms": ms_b, "learned_blocks_only_iterations": itb,
            "learned_blocks_gain_over_coarse_only": ms_c / ms_ll}
        if not args.no_bbt:
            log("square-root check at small N (columns of B from unit vectors)")
            extra["bbt_rel_err_small_N"] = bbt_errors(g, net)
    if not args.no_multi:
        extra.update(multi_vector_leg(g, h32, F_d, W_d, N, ms32, extra, args))
    g.destroy(h32, h64)

    config = {
        "workload": f"open-boundary RPY Mdot (UF), N={N}, RSA packing phi={phi} in a cube of "
                    f"side {L:.1f}a, a=1, eta=1/(6 pi), F~N(0,1); FP32 pair terms with FP64 "
                    f"per-tile accumulation; sqrtMdotW tol 1e-6",
        "N": N,
        "pairs_per_step": pairs,
    }
    t32 = extra["tiling_fp32pair"]
    key = ("c3", N, extra["pair_evaluation_fp32pair"], t32["targets_per_lane"], t32["source_chunks"])
    roof = counted_roofline(key, pairs * fpp, ms32)
    res = result_line(METRIC, gpair32, "Gpair/s", args.steps, args.warmup, ms32, True, "f32",
                      config, roof, extra, None)
    if not args.no_cpu:
        log("CPU baseline (oracle, OpenMP FP64)")
        cpu, Uo = cpu_baseline_rows(pos, F)
        res["cpu_baseline"] = cpu
        if "sqrtMdotW_plain_fp64pair_iterations" in extra:
            # CPU plain Lanczos projected from the measured CPU product rate:
            # (iterations + 1) products of N^2 pairs (the vector work is
            # O(N) per iteration and not counted)
            its = int(np.median(extra["sqrtMdotW_plain_fp64pair_iterations"]))
            ms_cpu = (its + 1) * pairs / (cpu["value"] * 1e9) * 1e3
            extra["cpu_sqrtMdotW_plain_projected_ms"] = ms_cpu
            extra["cpu_sqrtMdotW_plain_projection"] = (
                f"projection, not a run: ({its} iterations of the FP64 plain solve + 1) x N^2 "
                f"pairs at the measured CPU product rate {cpu['value']:.2f} Gpair/s on "
                f"{cpu['cores']} threads")
        k = Uo.size
        extra["mdot_rel_err_fp32pair_vs_oracle_rows"] = rel(U32[:k], Uo)
        extra["mdot_rel_err_fp64pair_vs_oracle_rows"] = rel(U64[:k], Uo)
        extra["oracle_rows"] = k // 3
    return res


# ---------------------------------------------------------------------------
# C1: open RPY, N = 1024, uniform in a cube of side 24a, FP64


def cpu_leg_c1(pos, F, Wn):
    """CPU baseline of C1, the reference's CPU path restated: dense FP64 matrix
    (OpenMP C oracle), matvec and eigen-square-root applied to W.  Its outputs
    are also the checker for the GPU results."""
    import oracle
    log("CPU baseline: dense FP64 matrix, matvec, eigen-square-root")
    t = time.perf_counter()
    M = oracle.dense_matrix(pos, ETA, 1.0)
    t_build = time.perf_counter() - t
    t = time.perf_counter()
    Uref = M @ F.ravel()
    t_mv = time.perf_counter() - t
    t = time.perf_counter()
    exact = oracle.dense_sqrt(M) @ Wn
    t_sqrt = time.perf_counter() - t
    ref, it_ref = oracle.LanczosOracle().run(lambda v: M @ v, Wn, 1e-6)
    try:
        from threadpoolctl import threadpool_info
        blas = max([d.get("num_threads", 1) for d in threadpool_info()] or [1])
    except Exception:
        blas = None
    cpu = {
        "value": (t_build + t_mv + t_sqrt) * 1e3, "unit": "ms", "cores": oracle.num_threads(),
        "kind": "port",
        "sample": f"whole workload: dense 3N x 3N FP64 matrix by the OpenMP C oracle "
                  f"({t_build * 1e3:.1f} ms), matvec ({t_mv * 1e3:.2f} ms), eigh square root "
                  f"applied to W ({t_sqrt * 1e3:.1f} ms; numpy/LAPACK with {blas} BLAS threads)",
    }
    return cpu, Uref, exact, ref, it_ref


def bench_c1(args):
    g = Gpu()
    N = args.N or 1024
    pos, F, Wn = workloads.config_c1(N)
    pos_d, F_d, W_d, U_d = g.put(pos), g.put(F), g.put(Wn), g.zeros(3 * N)
    h = g.solver(pos_d, "fp64", tiling=args.tiling or C1_TILING)
    ms_mdot = g.time_events(lambda: g.mdot(h, F_d, None, U_d), args.steps, args.warmup)
    U = U_d.cpu().numpy()
    ms_sqrt, its = g.time_lanczos(h, W_d, 3 * N, args.lanczos_calls)
    Uq = g.zeros(3 * N)
    g.sqrt_noise(h, W_d, Uq)
    g.sync()
    sq = Uq.cpu().numpy()
    tiling = g.tiling(h)
    g.destroy(h)
    pairs = float(N) ** 2
    fpp = flops_per_pair()["open_UF"]
    extra = {
        "mdot_ms": ms_mdot,
        "sqrtMdotW_ms": ms_sqrt,
        "sqrtMdotW_iterations": its,
        "tiling": tiling,
    }
    cpu = None
    if not args.no_cpu:
        cpu, Uref, exact, ref, it_ref = cpu_leg_c1(pos, F, Wn)
        extra.update({
            "oracle_lanczos_iterations": it_ref,
            "mdot_rel_err_vs_dense_fp64": rel(U, Uref),
            "sqrtMdotW_rel_err_vs_eigh": rel(sq, exact),
            "oracle_lanczos_rel_err_vs_eigh": rel(ref, exact),
            "sqrtMdotW_rel_diff_vs_oracle_lanczos": rel(sq, ref),
        })
    config = {"workload": f"open-boundary RPY, N={N} uniform in a cube of side 24a (overlaps), "
                          f"a=1, eta=1/(6 pi), F,W~N(0,1); FP64 Mdot + sqrtMdotW tol 1e-6",
              "N": N}
    total = ms_mdot + ms_sqrt
    return result_line("Mdot + sqrtMdotW ms at N=1024 open FP64 tol 1e-6", total, "ms",
                       args.steps, args.warmup, total, False, "f64", config,
                       compute_roofline(pairs * fpp, ms_mdot, PEAK_FP64_TFLOPS,
                                        MEASURED_TRAFFIC_BYTES.get(
                                            ("c1", N, "ordered", tiling["targets_per_lane"],
                                             tiling["source_chunks"]))), extra, cpu)


# ---------------------------------------------------------------------------
# C2: single wall with torques, N = 65536


def bench_c2(args):
    g = Gpu()
    N = args.N or 65536
    pos, F, T, Wn = workloads.config_c2(N)
    pos_d, F_d, T_d = g.put(pos), g.put(F), g.put(T)
    Wn_d = g.put(Wn)
    U_d, W_d = g.zeros(3 * N), g.zeros(3 * N)
    fpp = flops_per_pair()["wall_6x6"]
    pairs = float(N) ** 2
    out = {}
    hs = {}
    for prec in ("fp32", "fp64"):
        log(f"{prec}-pair wall Mdot (F, T)")
        h = g.solver(pos_d, prec, wall=True, angular=True, tiling=args.tiling)
        # the drift's products follow the pair precision here, to state the
        # FP32 drift error (the library default is FP64 drift products)
        g.ffi.check(g.lib.lm_nbody_set_drift_precision(h, g.ffi.PAIR_PRECISION[prec]))
        hs[prec] = h
        steps = args.steps if prec == "fp32" else max(3, args.steps // 4)
        ms = g.time_events(lambda: g.mdot(h, F_d, T_d, U_d, W_d), steps, args.warmup)
        log(f"  {ms:.3f} ms/Mdot, {pairs / ms / 1e6:.1f} Gpair/s")
        blocks = {}
        for name, (f, t) in {"F": (F_d, None), "T": (None, T_d), "FT": (F_d, T_d)}.items():
            g.mdot(h, f, t, U_d, W_d)
            g.sync()
            blocks[name] = (U_d.cpu().numpy().reshape(-1, 3), W_d.cpu().numpy().reshape(-1, 3))
        rec = {"ms": ms, "blocks": blocks, "tiling": g.tiling(h), "evaluation": g.evaluation(h)}
        # thermal drift (RFD, delta 1e-3 a): both handles share the seed, so
        # they draw the same random displacement
        du, dw = g.zeros(3 * N), g.zeros(3 * N)
        g.ffi.check(g.lib.lm_nbody_div_m(h, du.data_ptr(), dw.data_ptr(), 1.0, g.ffi.LM_MEM_DEVICE))
        g.sync()
        rec["drift"] = np.concatenate([du.cpu().numpy(), dw.cpu().numpy()])
        if not args.no_lanczos:
            log(f"  {prec}-pair sqrtMdotW (6N, tol 1e-6)")
            ms_l, it_l = g.time_lanczos(h, Wn_d, 3 * N, 0, angular=True)
            su, sw = g.zeros(3 * N), g.zeros(3 * N)
            g.sqrt_noise(h, Wn_d, su, sw)
            g.sync()
            rec.update(sqrt_ms=ms_l, sqrt_it=it_l,
                       sqrt=np.concatenate([su.cpu().numpy(), sw.cpu().numpy()]))
        out[prec] = rec
    two_level = None
    if not args.no_lanczos:
        from libmobility_amd.precond import default_network
        net = default_network(True)
        h = hs["fp32"]
        g.ffi.check(g.lib.lm_nbody_set_coarse_level(h, COARSE_PPC))
        log(f"  fp32-pair sqrtMdotW (6N, tol 1e-6), coarse level of {COARSE_PPC} particles per "
            f"cell only")
        ms_c, it_c = g.time_lanczos(h, Wn_d, 3 * N, 1, angular=True)
        g.ffi.check(g.lib.lm_nbody_set_preconditioner(h, net._h))
        log(f"  fp32-pair sqrtMdotW (6N, tol 1e-6), two-level preconditioner (learned 6x6 blocks + "
            f"coarse level of {COARSE_PPC} particles per cell)")
        ms_t, it_t = g.time_lanczos(h, Wn_d, 3 * N, 1, angular=True)
        two_level = {"ms": ms_t, "iterations": it_t, "cells": g.coarse_cells(h),
                     "coarse_build_ms": g.coarse_build_ms(h), "coarse_only_ms": ms_c,
                     "coarse_only_iterations": it_c}
        g.ffi.check(g.lib.lm_nbody_set_coarse_level(h, 0))
        g.ffi.check(g.lib.lm_nbody_set_preconditioner(h, None))
    g.destroy(*hs.values())

    # blocks: U from F = UF F, W from F = WF F, U from T = UT T, W from T = WT T
    def split(b):
        return {"UF": b["F"][0], "WF": b["F"][1], "UT": b["T"][0], "WT": b["T"][1],
                "U": b["FT"][0], "W": b["FT"][1]}

    b32, b64 = split(out["fp32"]["blocks"]), split(out["fp64"]["blocks"])
    err32_vs_64 = {k: rel(b32[k], b64[k]) for k in b32}
    ms32, ms64 = out["fp32"]["ms"], out["fp64"]["ms"]
    extra = {
        "mdot_fp32pair_ms": ms32,
        "mdot_fp64pair_ms": ms64,
        "mdot_fp64pair_gpair_s": pairs / (ms64 * 1e-3) / 1e9,
        "mdot_fp64pair_valu_frac_of_fp64_peak": pairs * fpp / (ms64 * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
        "valu_frac_note": "algorithmic flops of the ordered-pair model (N^2 pairs x flops_per_pair); "
                          "the symmetric evaluation does each unordered pair once, so the fraction "
                          "can exceed the hardware's issued-flop fraction, and 1",
        "tiling_fp32pair": out["fp32"]["tiling"],
        "tiling_fp64pair": out["fp64"]["tiling"],
        "pair_evaluation_fp32pair": out["fp32"]["evaluation"],
        "pair_evaluation_fp64pair": out["fp64"]["evaluation"],
        "flops_per_pair": fpp,
        "rel_err_fp32pair_vs_fp64pair_all_rows": err32_vs_64,
        "divM_rel_diff_fp32drift_vs_fp64drift_delta1e-3": rel(out["fp32"]["drift"],
                                                               out["fp64"]["drift"]),
    }
    if not args.no_lanczos:
        extra.update({
            "sqrtMdotW_fp32pair_ms": out["fp32"]["sqrt_ms"],
            "sqrtMdotW_fp64pair_ms": out["fp64"]["sqrt_ms"],
            "sqrtMdotW_fp32pair_iterations": out["fp32"]["sqrt_it"],
            "sqrtMdotW_fp64pair_iterations": out["fp64"]["sqrt_it"],
            "sqrtMdotW_rel_diff_fp32pair_vs_fp64pair": rel(out["fp32"]["sqrt"], out["fp64"]["sqrt"]),
            "sqrtMdotW_two_level_fp32pair_ms": two_level["ms"],
            "sqrtMdotW_two_level_fp32pair_iterations": two_level["iterations"],
            "sqrtMdotW_two_level_cells": two_level["cells"],
            "sqrtMdotW_two_level_coarse_build_ms": two_level["coarse_build_ms"],
            "sqrtMdotW_two_level_speedup_vs_plain": out["fp32"]["sqrt_ms"] / two_level["ms"],
            "sqrtMdotW_coarse_only_fp32pair_ms": two_level["coarse_only_ms"],
            "sqrtMdotW_coarse_only_fp32pair_iterations": two_level["coarse_only_iterations"],
            "sqrtMdotW_learned_blocks_gain_over_coarse_only":
                two_level["coarse_only_ms"] / two_level["ms"],
            "sqrtMdotW_two_level_preconditioner": "learned 6x6 blocks (64-feature MLP, BF16 MFMA, "
                                                  "trained on the GPU) + coarse level",
        })
    config = {"workload": f"single-wall RPB/Swan-Brady, N={N}, x,y uniform in [0,256a), z uniform "
                          f"in [1.5a,20a], a=1, eta=1/(6 pi), F,T~N(0,1); 6x6 Mdot (UF,UT,WF,WT), "
                          f"FP32 pair terms with FP64 per-tile accumulation", "N": N}
    t32 = out["fp32"]["tiling"]
    key = ("c2", N, out["fp32"]["evaluation"], t32["targets_per_lane"], t32["source_chunks"])
    res = result_line("Wall 6x6 Mdot Gpair/s at N=65536 (FP32 pairs)", pairs / (ms32 * 1e-3) / 1e9,
                      "Gpair/s", args.steps, args.warmup, ms32, True, "f32", config,
                      counted_roofline(key, pairs * fpp, ms32), extra, None)
    if not args.no_cpu:
        rows = np.sort(np.random.default_rng(3).choice(N, size=min(N, args.oracle_rows),
                                                      replace=False))
        cpu, o = cpu_leg_c2(pos, F, T, rows)
        res["cpu_baseline"] = cpu
        extra["rel_err_fp64pair_vs_oracle_rows"] = {k: rel(b64[k][rows], o[k]) for k in o}
        extra["rel_err_fp32pair_vs_oracle_rows"] = {k: rel(b32[k][rows], o[k]) for k in o}
        extra["oracle_rows"] = int(rows.size)
    return res


def cpu_leg_c2(pos, F, T, rows):
    """CPU baseline of C2: the OpenMP FP64 oracle's wall 6x6 product on a
    sample of target rows, once for F alone and once for T alone (the blocks
    UF, WF and UT, WT).  Its outputs are also the checker for those rows."""
    import oracle
    N = pos.shape[0]
    log(f"CPU baseline: oracle rows, {rows.size} target rows x {N} sources, per block")
    t = time.perf_counter()
    uf, wf = oracle.mdot_rows(pos, ETA, 1.0, rows, F=F, angular=True, wall=True)
    ut, wt = oracle.mdot_rows(pos, ETA, 1.0, rows, T=T, angular=True, wall=True)
    dt = time.perf_counter() - t
    pairs = 2.0 * rows.size * N
    cpu = {
        "value": pairs / dt / 1e9, "unit": "Gpair/s", "cores": oracle.num_threads(),
        "kind": "port",
        "sample": f"FP64 single-wall 6x6 product (F alone, then T alone), {rows.size} random "
                  f"target rows x all {N} sources ({pairs:.3e} pair evaluations, {dt:.1f} s), "
                  f"OpenMP C oracle",
    }
    return cpu, dict(UF=uf, WF=wf, UT=ut, WT=wt, U=uf + ut, W=wf + wt)


# ---------------------------------------------------------------------------
# C4: 512 ragged systems, half open, half single wall


def bench_c4(args):
    g = Gpu()
    log("generating the 512-system ensemble")
    sizes, kinds, pos, F, Wn = workloads.config_c4()
    N = int(sizes.sum())
    off = np.concatenate([[0], np.cumsum(sizes)])
    pos_d, F_d, W_d = g.put(pos), g.put(F), g.put(Wn)
    U_d = g.zeros(3 * N)
    fp = flops_per_pair()
    sq = sizes.astype(np.float64) ** 2
    pairs = float(sq.sum())
    flops = float(sq[kinds == 0].sum() * fp["open_UF"] + sq[kinds == 1].sum() * fp["wall_UF"])

    log(f"ensemble: {len(sizes)} systems, N={N}, sum N^2={pairs:.3e}; one launch")
    he = g.solver(pos_d, "fp32", wall=True, tol=1e-6, sizes=sizes, kinds=kinds, persistent=1,
                  tiling=args.tiling)
    ms_e = g.time_events(lambda: g.mdot(he, F_d, None, U_d), args.steps, args.warmup)
    Ue = U_d.cpu().numpy()
    ev = g.evaluation(he)
    log(f"  {ms_e:.3f} ms per ensemble Mdot ({ev} evaluation), {pairs / ms_e / 1e6:.1f} Gpair/s")
    ho = g.solver(pos_d, "fp32", wall=True, tol=1e-6, sizes=sizes, kinds=kinds, persistent=1,
                  tiling=args.tiling)
    g.ffi.check(g.lib.lm_nbody_set_pair_symmetric(ho, 0))
    ms_o = g.time_events(lambda: g.mdot(ho, F_d, None, U_d), args.steps, args.warmup)
    Uo = U_d.cpu().numpy()
    t_o = g.tiling(ho)
    g.destroy(ho)
    log(f"  {ms_o:.3f} ms with the ordered evaluation (persistent grid)")
    # the same tiling as the persistent grid, so that the two sum in the same order
    hn = g.solver(pos_d, "fp32", wall=True, tol=1e-6, sizes=sizes, kinds=kinds, persistent=0,
                  tiling=(t_o["targets_per_lane"], t_o["source_chunks"]))
    g.ffi.check(g.lib.lm_nbody_set_pair_symmetric(hn, 0))
    ms_n = g.time_events(lambda: g.mdot(hn, F_d, None, U_d), args.steps, args.warmup)
    Un = U_d.cpu().numpy()
    g.destroy(hn)
    log(f"  {ms_n:.3f} ms with the ordered evaluation, non-persistent grid")

    log("per-system solves (one handle and one launch per system)")
    hs = [g.solver(pos_d[3 * a:3 * b], "fp32", wall=bool(k), tol=1e-6)
          for a, b, k in zip(off[:-1], off[1:], kinds)]
    Us = [U_d[3 * a:3 * b] for a, b in zip(off[:-1], off[1:])]
    Fs = [F_d[3 * a:3 * b] for a, b in zip(off[:-1], off[1:])]

    def per_system():
        for h, f, u in zip(hs, Fs, Us):
            g.mdot(h, f, None, u)

    ms_p = g.time_events(per_system, max(1, args.steps // 4), 1)
    Up = U_d.cpu().numpy()
    log(f"  {ms_p:.3f} ms for 512 per-system Mdots")
    extra = {
        "systems": int(len(sizes)), "N_total": N, "sum_N2": pairs,
        "pair_evaluation_ensemble": ev,
        "mdot_ensemble_ms": ms_e,
        "mdot_ensemble_ordered_persistent_ms": ms_o,
        "mdot_ensemble_ordered_plain_grid_ms": ms_n,
        "mdot_ensemble_symmetric_speedup_vs_ordered": ms_o / ms_e,
        "mdot_per_system_ms": ms_p,
        "mdot_speedup_vs_per_system": ms_p / ms_e,
        "mdot_rel_diff_ensemble_vs_per_system": rel(Ue, Up),
        "mdot_rel_diff_ensemble_vs_ordered": rel(Ue, Uo),
        "mdot_ordered_persistent_equals_plain_grid_bitwise": bool(np.array_equal(Uo, Un)),
        "tiling_ensemble": g.tiling(he),
        "tiling_ordered": t_o,
        "flops_per_launch": flops,
    }
    if not args.no_lanczos:
        log("sqrtMdotW (tol 1e-6): one Lanczos over the ensemble")
        ms_le, it_e = g.time_lanczos(he, W_d, 3 * N, args.lanczos_calls)
        Ule = g.zeros(3 * N)
        g.sqrt_noise(he, W_d, Ule)
        log("sqrtMdotW per system")
        Wns = [W_d[3 * a:3 * b] for a, b in zip(off[:-1], off[1:])]
        Ulp = g.zeros(3 * N)
        its = []
        for h, w, a, b in zip(hs, Wns, off[:-1], off[1:]):  # warm-up: workspaces, tiling
            g.sqrt_noise(h, w, Ulp[3 * a:3 * b])
        g.sync()
        Ulp.zero_()
        t = time.perf_counter()
        for h, w, a, b in zip(hs, Wns, off[:-1], off[1:]):
            g.sqrt_noise(h, w, Ulp[3 * a:3 * b])
            its.append(g.lib.lm_nbody_last_lanczos_iterations(h))
        g.sync()
        ms_lp = (time.perf_counter() - t) * 1e3
        log("sqrtMdotW: one recurrence per system, all systems in one launch per iteration")
        g.ffi.check(g.lib.lm_nbody_set_lanczos_per_system(he, 1))
        ms_lr, it_r = g.time_lanczos(he, W_d, 3 * N, args.lanczos_calls)
        Ulr = g.zeros(3 * N)
        g.sqrt_noise(he, W_d, Ulr)
        its_r = np.zeros(len(sizes), dtype=np.int32)
        g.ffi.check(g.lib.lm_nbody_last_lanczos_iterations_per_system(he, its_r.ctypes.data,
                                                                      len(sizes)))
        log(f"sqrtMdotW: one recurrence per system, two-level preconditioner (learned blocks + "
            f"coarse level of {COARSE_PPC} particles per cell, per system)")
        from libmobility_amd.precond import default_network
        net = default_network(False)
        g.ffi.check(g.lib.lm_nbody_set_preconditioner(he, net._h))
        g.ffi.check(g.lib.lm_nbody_set_coarse_level(he, COARSE_PPC))
        ms_lt, it_t = g.time_lanczos(he, W_d, 3 * N, args.lanczos_calls)
        its_t = np.zeros(len(sizes), dtype=np.int32)
        g.ffi.check(g.lib.lm_nbody_last_lanczos_iterations_per_system(he, its_t.ctypes.data,
                                                                      len(sizes)))
        cells_t, build_t = g.coarse_cells(he), g.coarse_build_ms(he)
        g.ffi.check(g.lib.lm_nbody_set_coarse_level(he, 0))
        g.ffi.check(g.lib.lm_nbody_set_preconditioner(he, None))
        up = Ulp.cpu().numpy()
        extra.update({
            "sqrtMdotW_per_system_two_level_ms": ms_lt,
            "sqrtMdotW_per_system_two_level_iterations_max": int(np.max(its_t)),
            "sqrtMdotW_per_system_two_level_iterations_mean": float(np.mean(its_t)),
            "sqrtMdotW_per_system_two_level_cells": cells_t,
            "sqrtMdotW_per_system_two_level_coarse_build_ms": build_t,
            "sqrtMdotW_per_system_two_level_speedup_vs_recurrences": ms_lr / ms_lt,
            "sqrtMdotW_ensemble_ms": ms_le,
            "sqrtMdotW_ensemble_iterations": it_e,
            "sqrtMdotW_per_system_ms": ms_lp,
            "sqrtMdotW_per_system_iterations_mean": float(np.mean(its)),
            "sqrtMdotW_per_system_iterations_max": int(np.max(its)),
            "sqrtMdotW_speedup_vs_per_system": ms_lp / ms_le,
            "sqrtMdotW_rel_diff_ensemble_vs_per_system": rel(Ule.cpu().numpy(), up),
            "sqrtMdotW_per_system_recurrences_ms": ms_lr,
            "sqrtMdotW_per_system_recurrences_iterations_max": it_r,
            "sqrtMdotW_per_system_recurrences_iterations_mean": float(np.mean(its_r)),
            "sqrtMdotW_per_system_recurrences_speedup_vs_per_system": ms_lp / ms_lr,
            "sqrtMdotW_per_system_recurrences_rel_diff_vs_per_system": rel(Ulr.cpu().numpy(), up),
        })
    g.destroy(he, *hs)
    config = {"workload": "512 independent suspensions, N_s log-uniform in [64, 8192] (seed 2024), "
                          "even systems open (RSA cube), odd systems single wall (RSA slab, z in "
                          "[1.5a, 10a]), phi uniform in [0.05, 0.25], a=1, eta=1/(6 pi), F~N(0,1); "
                          "FP32 pairs, FP64 accumulation; one persistent launch", "N": N}
    te = extra["tiling_ensemble"]
    key = ("c4", N, ev, te["targets_per_lane"], te["source_chunks"])
    return result_line("Ensemble Mdot Gpair/s, 512 ragged systems (FP32 pairs)",
                       pairs / (ms_e * 1e-3) / 1e9, "Gpair/s", args.steps, args.warmup, ms_e, True,
                       "f32", config, counted_roofline(key, flops, ms_e), extra, None)


# ---------------------------------------------------------------------------
# C5: single wall with torques, N = 1048576, hydrodynamicVelocities


def bench_c5(args):
    g = Gpu()
    N = args.N or 1048576
    pos, F, T, Lxy = workloads.config_c5(N)
    pos_d, F_d, T_d = g.put(pos), g.put(F), g.put(T)
    U_d, W_d = g.zeros(3 * N), g.zeros(3 * N)
    del pos, F, T
    fpp = flops_per_pair()["wall_6x6"]
    pairs = float(N) ** 2
    tiling = args.tiling or (2, 0)
    t_start = time.perf_counter()

    def progress(i, err):
        # one line per Lanczos iteration (a pair product of the 6N x 6N operator each)
        log(f"  Lanczos iteration {i}: error estimate {err:.3e}, "
            f"{time.perf_counter() - t_start:.0f} s")

    h32 = g.solver(pos_d, "fp32", wall=True, angular=True, delta=1e-4, tiling=tiling,
                   iteration_limit=1024, callback=progress)
    log("FP32-pair Mdot")
    ms32 = g.time_events(lambda: g.mdot(h32, F_d, T_d, U_d, W_d), 1, 0)
    ev32 = g.evaluation(h32)
    log(f"  {ms32:.1f} ms ({ev32} evaluation)")
    h64 = g.solver(pos_d, "fp64", wall=True, angular=True, delta=1e-4, tiling=tiling)
    log("FP64-pair Mdot")
    ms64 = g.time_events(lambda: g.mdot(h64, F_d, T_d, U_d, W_d), 1, 0)
    ev64 = g.evaluation(h64)
    log(f"  {ms64:.1f} ms ({ev64} evaluation)")
    extra = {"mdot_fp32pair_ms": ms32, "mdot_fp64pair_ms": ms64, "flops_per_pair": fpp,
             "pair_evaluation_fp32pair": ev32, "pair_evaluation_fp64pair": ev64,
             "tiling": {"targets_per_lane": tiling[0], "source_chunks": tiling[1]},
             "lanczos_iteration_limit": 1024}
    if not args.c5_plain:
        # thermal drift at delta = 1e-4 a on one random displacement: FP64 drift
        # products (the library default) against FP32 ones
        Wd = g.put(np.random.default_rng(13).normal(size=3 * N))
        drift = {}
        for prec in ("fp64", "fp32"):
            g.ffi.check(g.lib.lm_nbody_set_drift_precision(h32, g.ffi.PAIR_PRECISION[prec]))
            du, dw = g.zeros(3 * N), g.zeros(3 * N)
            g.sync()
            t = time.perf_counter()
            g.ffi.check(g.lib.lm_nbody_div_m_noise(h32, Wd.data_ptr(), du.data_ptr(), dw.data_ptr(),
                                                   1.0, g.ffi.LM_MEM_DEVICE))
            g.sync()
            drift[prec] = (time.perf_counter() - t, np.concatenate([du.cpu().numpy(),
                                                                    dw.cpu().numpy()]))
            log(f"RFD drift with {prec} products: {drift[prec][0]:.1f} s")
        g.ffi.check(g.lib.lm_nbody_set_drift_precision(h32, g.ffi.PAIR_PRECISION["fp64"]))
        extra.update({
            "divM_fp64drift_ms": drift["fp64"][0] * 1e3,
            "divM_fp32drift_ms": drift["fp32"][0] * 1e3,
            "divM_rel_err_fp32drift_vs_fp64drift_delta1e-4": rel(drift["fp32"][1],
                                                                 drift["fp64"][1]),
        })
        del Wd
        # FP64 plain Lanczos: cost per iteration over a bounded run (the full
        # FP64 plain step does not fit one call: ~200 iterations x the FP64 product)
        probe = 12
        g.ffi.check(g.lib.lm_nbody_set_lanczos_iteration_limit(h64, probe))
        Wn = g.put(np.random.default_rng(14).normal(size=6 * N))
        su, sw = g.zeros(3 * N), g.zeros(3 * N)
        g.sync()
        t = time.perf_counter()
        status = g.lib.lm_nbody_sqrt_mdot_w_noise(h64, Wn.data_ptr(), su.data_ptr(), sw.data_ptr(),
                                                  1.0, g.ffi.LM_MEM_DEVICE)
        g.sync()
        per_it64 = (time.perf_counter() - t) * 1e3 / probe
        log(f"FP64 plain Lanczos: {per_it64 / 1e3:.2f} s per iteration over {probe} iterations "
            f"(status {status}: stopped at the limit)")
        extra["sqrtMdotW_plain_fp64pair_ms_per_iteration"] = per_it64
        del Wn, su, sw
    g.destroy(h64)
    step_ms = None
    if not args.no_lanczos:
        if args.c5_plain:
            log("hydrodynamicVelocities (Mdot + sqrtMdotW tol 1e-6 + RFD drift), FP32 pairs, plain "
                "Lanczos")
        else:
            from libmobility_amd.precond import default_network
            net = default_network(True)
            g.ffi.check(g.lib.lm_nbody_set_preconditioner(h32, net._h))
            g.ffi.check(g.lib.lm_nbody_set_coarse_level(h32, COARSE_PPC))
            g.ffi.check(g.lib.lm_nbody_set_coarse_max_dim(h32, C5_COARSE_MAX_DIM))
            log("hydrodynamicVelocities (Mdot + sqrtMdotW tol 1e-6 + RFD drift), FP32 pairs, "
                "two-level preconditioner (learned 6x6 blocks + coarse level)")
        g.sync()
        t = t_start = time.perf_counter()
        g.ffi.check(g.lib.lm_nbody_hydrodynamic_velocities(
            h32, F_d.data_ptr(), 3 * N, T_d.data_ptr(), 3 * N, U_d.data_ptr(), W_d.data_ptr(),
            g.ffi.LM_MEM_DEVICE))
        g.sync()
        step_ms = (time.perf_counter() - t) * 1e3
        it = g.lib.lm_nbody_last_lanczos_iterations(h32)
        log(f"  {step_ms / 1e3:.1f} s, Lanczos iterations {it}")
        u = U_d.cpu().numpy()
        label = "plain" if args.c5_plain else "two_level"
        extra.update({
            f"hydrodynamicVelocities_fp32pair_{label}_ms": step_ms,
            "lanczos_iterations": it,
            "output_finite": bool(np.all(np.isfinite(u))),
        })
        if not args.c5_plain:
            extra["preconditioner"] = (f"learned 6x6 blocks + coarse level ({g.coarse_cells(h32)} "
                                       f"cells, 3C capped at {C5_COARSE_MAX_DIM}, built in "
                                       f"{g.coarse_build_ms(h32):.0f} ms)")
    g.destroy(h32)
    config = {"workload": f"single-wall RPB/Swan-Brady, N={N}, x,y uniform in [0,{Lxy:.0f}a), z "
                          f"uniform in [1.5a,30a], a=1, eta=1/(6 pi), F,T~N(0,1); 6x6 operator; "
                          f"hydrodynamicVelocities = Mdot + sqrtMdotW (prefactor 1, tol 1e-6) + "
                          f"RFD drift (delta 1e-4 a, FP64 drift products); FP32 pairs, FP64 "
                          f"accumulation, "
                          + ("plain Lanczos" if args.c5_plain else
                             "two-level preconditioner (learned 6x6 blocks + coarse level)"),
            "N": N}
    value = step_ms if step_ms is not None else ms32
    metric = ("hydrodynamicVelocities ms at N=1048576 single wall (FP32 pairs, "
              + ("plain Lanczos)" if args.c5_plain else "learned two-level preconditioner)"))
    return result_line(metric, value, "ms", 1, 1, value, False, "f32", config,
                       compute_roofline(pairs * fpp, ms32, PEAK_FP32_TFLOPS), extra, None)


def bench_c5_plain_fp64(args):
    """The C5 step on the plain FP64 path: FP64 pair terms in every product
    (Mdot, the plain Lanczos square root and the RFD drift)."""
    g = Gpu()
    N = args.N or 1048576
    pos, F, T, Lxy = workloads.config_c5(N)
    pos_d, F_d, T_d = g.put(pos), g.put(F), g.put(T)
    U_d, W_d = g.zeros(3 * N), g.zeros(3 * N)
    del pos, F, T
    fpp = flops_per_pair()["wall_6x6"]
    pairs = float(N) ** 2
    tiling = args.tiling or (2, 0)
    t_start = time.perf_counter()

    def progress(i, err):
        log(f"  Lanczos iteration {i}: error estimate {err:.3e}, "
            f"{time.perf_counter() - t_start:.0f} s")

    h64 = g.solver(pos_d, "fp64", wall=True, angular=True, delta=1e-4, tiling=tiling,
                   iteration_limit=1024, callback=progress)
    log("FP64-pair Mdot")
    ms64 = g.time_events(lambda: g.mdot(h64, F_d, T_d, U_d, W_d), 1, 0)
    log(f"  {ms64:.1f} ms ({g.evaluation(h64)} evaluation)")
    log("hydrodynamicVelocities (Mdot + sqrtMdotW tol 1e-6 + RFD drift), FP64 pairs, plain Lanczos")
    g.sync()
    t = t_start = time.perf_counter()
    g.ffi.check(g.lib.lm_nbody_hydrodynamic_velocities(
        h64, F_d.data_ptr(), 3 * N, T_d.data_ptr(), 3 * N, U_d.data_ptr(), W_d.data_ptr(),
        g.ffi.LM_MEM_DEVICE))
    g.sync()
    step_ms = (time.perf_counter() - t) * 1e3
    it = g.lib.lm_nbody_last_lanczos_iterations(h64)
    log(f"  {step_ms / 1e3:.1f} s, Lanczos iterations {it}")
    u = U_d.cpu().numpy()
    g.destroy(h64)
    extra = {"mdot_fp64pair_ms": ms64, "flops_per_pair": fpp,
             "hydrodynamicVelocities_fp64pair_plain_ms": step_ms, "lanczos_iterations": it,
             "output_finite": bool(np.all(np.isfinite(u))),
             "tiling": {"targets_per_lane": tiling[0], "source_chunks": tiling[1]}}
    config = {"workload": f"single-wall RPB/Swan-Brady, N={N}, x,y uniform in [0,{Lxy:.0f}a), z "
                          f"uniform in [1.5a,30a], a=1, eta=1/(6 pi), F,T~N(0,1); 6x6 operator; "
                          f"hydrodynamicVelocities = Mdot + sqrtMdotW (prefactor 1, tol 1e-6) + "
                          f"RFD drift (delta 1e-4 a); FP64 pairs everywhere, plain Lanczos",
              "N": N}
    return result_line("hydrodynamicVelocities ms at N=1048576 single wall (FP64 pairs, plain "
                       "Lanczos)", step_ms, "ms", 1, 1, step_ms, False, "f64", config,
                       compute_roofline(pairs * fpp, ms64, PEAK_FP64_TFLOPS), extra, None)


def copy_calibration():
    import torch
    n = 1 << 28  # doubles: 2 GiB read + 2 GiB written per copy
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8}), flush=True)


def main():
    global COARSE_PPC
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=("c3", "c1", "c2", "c4", "c5"), default="c3")
    ap.add_argument("--lanczos-calls", type=int, default=3)
    ap.add_argument("--N", type=int, default=None, help="override the particle count (c1, c2, c3, c5)")
    ap.add_argument("--tiling", default=None,
                    help="fix the pair tiling as TPL,CHUNKS instead of the autotuner")
    ap.add_argument("--oracle-rows", type=int, default=4096,
                    help="c2: target rows of the CPU baseline / checker")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lanczos", action="store_true")
    ap.add_argument("--no-multi", action="store_true",
                    help="c3: skip the several-vectors-per-pass leg")
    ap.add_argument("--no-bbt", action="store_true",
                    help="c3: skip the small-N ||BB^T - M|| check (thousands of small solves)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 2 GiB device-to-device copy (reference for the HBM byte counters)")
    ap.add_argument("--c5-plain", action="store_true",
                    help="c5: the step with plain Lanczos instead of the two-level preconditioner")
    ap.add_argument("--c5-plain-fp64", action="store_true",
                    help="c5: the step on the plain FP64 path (FP64 pairs everywhere, plain Lanczos)")
    ap.add_argument("--coarse-ppc", type=int, default=COARSE_PPC,
                    help="particles per cell of the preconditioner's coarse level")
    args = ap.parse_args()
    COARSE_PPC = args.coarse_ppc
    if args.gpus != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path.", file=sys.stderr)
        sys.exit(2)
    if args.tiling:
        args.tiling = tuple(int(x) for x in args.tiling.split(","))
    if args.copy_calibration:
        copy_calibration()
        return
    if args.workload == "c5" and args.c5_plain_fp64:
        res = bench_c5_plain_fp64(args)
    else:
        res = {"c3": bench_c3, "c1": bench_c1, "c2": bench_c2, "c4": bench_c4,
               "c5": bench_c5}[args.workload](args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
