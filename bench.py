"""Benchmark: NBody Mdot throughput and Lanczos sqrtMdotW time on one MI355X.

Workload (one GPU): open-boundary RPY, N = 131072 particles of radius a = 1,
random sequential addition at volume fraction 0.20 in a cube of side 140.0a,
viscosity 1/(6 pi), forces and Brownian noise W i.i.d. N(0,1).

A "step" is one Mdot over the whole system (N^2 ordered pairs, self pair
included) in FP32-pair / FP64-accumulate mode, inputs and outputs resident in
HBM.  K steps are timed with HIP events on the solver's stream.  The Lanczos
square root (tol 1e-6) is timed separately (median of a few calls) and
reported with its iteration count, for the plain FP32-pair and FP64-pair
operators.  The CPU baseline is the oracle's OpenMP FP64 matrix-free product
on a bounded sample of target rows, run on this host's cores.

Prints exactly one JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mdot Gpair/s + sqrtMdotW ms at N=131072 tol 1e-6: learned-prec vs plain"
PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 vector/matrix peak
PEAK_FP64_TFLOPS = 78.6


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(pos, F, eta, target_seconds=12.0):
    import oracle
    N = pos.shape[0]
    rows = 256
    t = time.perf_counter()
    oracle.mdot_open_uf_rows(pos, eta, 1.0, F, 0, rows)
    dt = time.perf_counter() - t
    rows = int(min(N, max(rows, rows * target_seconds / max(dt, 1e-6))))
    t = time.perf_counter()
    oracle.mdot_open_uf_rows(pos, eta, 1.0, F, 0, rows)
    dt = time.perf_counter() - t
    pairs = rows * N
    return {
        "value": pairs / dt / 1e9,
        "unit": "Gpair/s",
        "cores": oracle.num_threads(),
        "kind": "port",
        "sample": f"FP64 open-RPY UF product, target rows 0..{rows - 1} x all {N} sources "
                  f"({pairs:.3e} pairs, {dt:.1f} s), OpenMP C oracle",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--lanczos-calls", type=int, default=3)
    ap.add_argument("--N", type=int, default=131072)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lanczos", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 2 GiB device-to-device copy (reference for the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(sharding is described in DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import torch
    from libmobility_amd import _ffi
    from libmobility_amd.workloads import rsa_cube

    if args.copy_calibration:
        n = 1 << 28  # doubles: 2 GiB read + 2 GiB written per copy
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8}), flush=True)
        return

    lib = _ffi.load()
    dev = torch.device("cuda:0")
    N = args.N
    phi = 0.20
    eta = 1.0 / (6.0 * np.pi)
    log(f"generating RSA packing N={N} phi={phi}")
    pos, L = rsa_cube(N, phi, seed=2026)
    rng = np.random.default_rng(7)
    F = rng.normal(size=(N, 3))
    Wn = rng.normal(size=3 * N)

    pos_d = torch.from_numpy(pos.reshape(-1).copy()).to(dev)
    F_d = torch.from_numpy(F.reshape(-1).copy()).to(dev)
    W_d = torch.from_numpy(Wn).to(dev)
    U_d = torch.zeros(3 * N, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)

    def make_solver(prec, tol=1e-6):
        h = ctypes.c_void_p()
        _ffi.check(lib.lm_nbody_create(2, 2, 2, ctypes.byref(h)))
        _ffi.check(lib.lm_nbody_set_parameters(h, 3, -1, -1, 0, 0.0, 1e-3))
        _ffi.check(lib.lm_nbody_set_pair_precision(h, _ffi.PAIR_PRECISION[prec]))
        _ffi.check(lib.lm_nbody_set_stream(h, ctypes.c_void_p(stream.cuda_stream)))
        _ffi.check(lib.lm_nbody_initialize(h, eta, 1.0, 0, tol, 12345, _ffi.LANCZOS_CALLBACK(), None))
        _ffi.check(lib.lm_nbody_set_positions(h, pos_d.data_ptr(), 3 * N, _ffi.LM_MEM_DEVICE))
        _ffi.check(lib.lm_nbody_set_autotune(h, 1))
        return h

    def tiling(h):
        t, c = ctypes.c_int(), ctypes.c_int()
        _ffi.check(lib.lm_nbody_get_pair_tiling(h, ctypes.byref(t), ctypes.byref(c)))
        return {"targets_per_lane": t.value, "source_chunks": c.value}

    def time_mdot(h, steps, warmup):
        for _ in range(warmup):
            _ffi.check(lib.lm_nbody_mdot(h, F_d.data_ptr(), 3 * N, None, 0, U_d.data_ptr(), None,
                                         _ffi.LM_MEM_DEVICE))
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            _ffi.check(lib.lm_nbody_mdot(h, F_d.data_ptr(), 3 * N, None, 0, U_d.data_ptr(), None,
                                         _ffi.LM_MEM_DEVICE))
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps  # ms per Mdot

    def time_lanczos(h, calls):
        out = torch.zeros(3 * N, dtype=torch.float64, device=dev)
        times, iters = [], []
        for c in range(calls + 1):
            torch.cuda.synchronize()
            t = time.perf_counter()
            _ffi.check(lib.lm_nbody_sqrt_mdot_w_noise(h, W_d.data_ptr(), out.data_ptr(), None, 1.0,
                                                      _ffi.LM_MEM_DEVICE))
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t) * 1e3
            if c > 0:
                times.append(dt)
                iters.append(lib.lm_nbody_last_lanczos_iterations(h))
            log(f"  sqrtMdotW call {c}: {dt:.1f} ms, iterations {lib.lm_nbody_last_lanczos_iterations(h)}")
        return statistics.median(times), iters

    import json as _json
    flops = _json.load(open(os.path.join(ROOT, "libmobility_amd", "flop_model.json")))["flops_per_pair"]
    fpp = flops["open_UF"]
    pairs = float(N) * float(N)

    log("FP32-pair Mdot")
    h32 = make_solver("fp32")
    ms32 = time_mdot(h32, args.steps, args.warmup)
    gpair32 = pairs / (ms32 * 1e-3) / 1e9
    log(f"  {ms32:.3f} ms/Mdot, {gpair32:.1f} Gpair/s")
    log("FP64-pair Mdot")
    h64 = make_solver("fp64")
    ms64 = time_mdot(h64, max(3, args.steps // 4), 1)
    gpair64 = pairs / (ms64 * 1e-3) / 1e9
    log(f"  {ms64:.3f} ms/Mdot, {gpair64:.1f} Gpair/s")

    extra = {
        "tiling_fp32pair": tiling(h32),
        "tiling_fp64pair": tiling(h64),
        "mdot_fp32pair_ms": ms32,
        "mdot_fp64pair_ms": ms64,
        "mdot_fp64pair_gpair_s": gpair64,
        "mdot_fp64pair_valu_frac_of_fp64_peak": gpair64 * 1e9 * fpp / (PEAK_FP64_TFLOPS * 1e12),
        "flops_per_pair": fpp,
    }
    if not args.no_lanczos:
        log("plain Lanczos sqrtMdotW, FP32 pairs, tol 1e-6")
        ms_l32, it32 = time_lanczos(h32, args.lanczos_calls)
        extra["sqrtMdotW_plain_fp32pair_ms"] = ms_l32
        extra["sqrtMdotW_plain_fp32pair_iterations"] = it32
        log("plain Lanczos sqrtMdotW, FP64 pairs, tol 1e-6")
        ms_l64, it64 = time_lanczos(h64, max(1, args.lanczos_calls - 2))
        extra["sqrtMdotW_plain_fp64pair_ms"] = ms_l64
        extra["sqrtMdotW_plain_fp64pair_iterations"] = it64
        log("learned-preconditioned Lanczos sqrtMdotW, FP32 pairs, tol 1e-6")
        from libmobility_amd.precond import DEFAULT_WEIGHTS, LearnedPreconditioner
        net = LearnedPreconditioner(weights=DEFAULT_WEIGHTS)
        hl = make_solver("fp32")
        _ffi.check(lib.lm_nbody_set_preconditioner(hl, net._h))
        ms_ll, itl = time_lanczos(hl, args.lanczos_calls)
        lib.lm_nbody_destroy(hl)
        extra["sqrtMdotW_learned_fp32pair_ms"] = ms_ll
        extra["sqrtMdotW_learned_fp32pair_iterations"] = itl
        extra["sqrtMdotW_learned_speedup_vs_plain"] = ms_l32 / ms_ll
    for h in (h32, h64):
        lib.lm_nbody_destroy(h)

    achieved = gpair32 * 1e9 * fpp / 1e12
    result = {
        "metric": METRIC,
        "value": gpair32,
        "unit": "Gpair/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"open-boundary RPY Mdot (UF), N={N}, RSA packing phi={phi} in a cube of "
                        f"side {L:.1f}a, a=1, eta=1/(6 pi), F~N(0,1); FP32 pair terms with FP64 "
                        f"per-tile accumulation; sqrtMdotW tol 1e-6",
            "N": N,
            "pairs_per_step": pairs,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": None,
        },
        "extra": extra,
    }
    if not args.no_cpu:
        log("CPU baseline (oracle, OpenMP FP64)")
        result["cpu_baseline"] = cpu_baseline(pos, F, eta)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
