#!/usr/bin/env python3
"""Benchmark of the MI355X force engine on Orrery's headline direct workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu-baseline]

Workload (direct O(N^2), one GPU): a Plummer sphere of N = 262,144 equal-mass
particles (total mass 1, G = 1, scale radius a = 1, radii by inverse CDF
truncated at r < 10 a, Aarseth-rejection speeds, seed 1), Plummer softening
eps = 0.005, Yoshida-4 with dt = 5e-4. A step is one Yoshida-4 step on
device-resident FP64 state: three force evaluations (FP32 pairwise terms,
compensated accumulation) and the FP64 kick/drift updates around them.

Prints one JSON line. ``value`` is interactions per second over the K timed
steps, N(N-1) per evaluation, inputs resident in HBM. ``roofline`` is for the
dominant kernel (direct_tiles) from HIP events recorded on the engine's own
stream; ``cpu_baseline`` times the reference's threaded AVX2 FP64 kernel on
this host's cores over a bounded sample of targets of the same input.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

FLOP_PER_INTERACTION = 20          # the reference's convention (docs/performance.md:103-109)
FP32_PEAK_TFLOPS = 157.3           # MI355X published FP32 vector peak
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_threads() -> int:
    n = os.cpu_count() or 1
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    return max(1, min(16, n))


def cpu_baseline(state, eps: float, seconds: float = 10.0) -> dict:
    """Time the reference's AVX2 FP64 direct kernel (or the oracle's restatement)."""
    import ctypes
    import oracle
    from oracle import _ptr, _u64

    threads = cpu_threads()
    n = state.n
    ax, ay, az = np.zeros(n), np.zeros(n), np.zeros(n)
    ref = oracle.load_reference()
    if ref is not None:
        kind = "reference"

        def run(tb, te):
            ref.ref_direct_targets(_u64(n), _ptr(state.x), _ptr(state.y), _ptr(state.z),
                                   _ptr(state.m), ctypes.c_double(eps), ctypes.c_int(1),
                                   ctypes.c_uint(threads), _u64(tb), _u64(te),
                                   _ptr(ax), _ptr(ay), _ptr(az))
    else:
        kind = "port"
        lib = oracle.load()

        def run(tb, te):
            lib.orc_direct(_u64(n), _ptr(state.x), _ptr(state.y), _ptr(state.z), _ptr(state.m),
                           ctypes.c_double(eps * eps), ctypes.c_double(1.0), ctypes.c_int(1),
                           ctypes.c_uint(threads), _u64(tb), _u64(te), _ptr(ax), _ptr(ay), _ptr(az))

    probe = min(n, 256)
    run(0, probe)  # warm-up
    t0 = time.perf_counter()
    run(0, probe)
    dt = time.perf_counter() - t0
    rate = probe * (n - 1) / max(dt, 1e-9)
    targets = int(min(n, max(probe, seconds * rate / max(1, n - 1))))
    t0 = time.perf_counter()
    run(0, targets)
    dt = time.perf_counter() - t0
    inter = targets * (n - 1)
    return {"value": inter / dt, "unit": "interactions/s", "cores": threads, "kind": kind,
            "sample": f"{targets} of {n} targets x all {n} sources, FP64 AVX2 kernel with the "
                      f"work-stealing executor, {dt:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--error-targets", type=int, default=1024)
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this engine runs on exactly one GPU; multi-GPU figures are "
                                   "not produced (the direct sum would shard by target range "
                                   "with one all-gather of positions per evaluation)",
                          "n_gpus": args.gpus}), flush=True)
        return 2

    import oracle
    import orrery_amd as E
    from orrery_amd import _capi

    n, eps, dt = args.n, 0.005, 5e-4
    log(f"[bench] generating Plummer N={n} seed={args.seed}")
    state = oracle.plummer(n, args.seed, total_mass=1.0, scale_radius=1.0)

    ctx = E.Context(0)
    dev = E.describe(0)
    log(f"[bench] device {dev['name']} {dev['arch']} CUs={dev['compute_units']}")
    sys_ = E.DeviceSystem(ctx, n)
    sys_.upload(state.x, state.y, state.z, state.vx, state.vy, state.vz, state.m)
    cfg = E.StepConfig(integrator="yoshida4", solver="direct", softening=eps)

    # Force accuracy of one evaluation at the initial state vs the FP64 oracle.
    log("[bench] refresh + accuracy sample")
    sys_.refresh(cfg)
    acc = sys_.download()
    rng = np.random.default_rng(12345)
    tgt = np.sort(rng.choice(n, size=min(n, args.error_targets), replace=False)).astype(np.uint64)
    rx, ry, rz = oracle.reference_acceleration(state.x, state.y, state.z, state.m, eps * eps,
                                               targets=tgt, threads=cpu_threads())
    idx = tgt.astype(np.int64)
    rel = oracle.relative_errors(acc["ax"][idx], acc["ay"][idx], acc["az"][idx], rx, ry, rz)
    worst, rms = oracle.error_summary(rel)

    log(f"[bench] warmup {args.warmup} steps")
    sys_.step(cfg, dt, args.warmup)
    ctx.set_timing(True)
    ctx.reset_timings()
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    sys_.step(cfg, dt, args.steps)           # returns after a stream synchronise
    elapsed = time.perf_counter() - t0
    timings = ctx.timings()
    ctx.set_timing(False)

    evals = 3 * args.steps
    inter = evals * n * (n - 1)
    value = inter / elapsed
    k_ms, k_launches = timings.get("direct_kernel", (0.0, 0))
    kernel_ms = k_ms / max(1, k_launches)
    eval_ms = sum(timings.get(p, (0.0, 0))[0] for p in
                  ("direct_prepare", "direct_kernel", "direct_reduce", "bbox", "encode", "sort")) / max(1, evals)
    achieved = FLOP_PER_INTERACTION * n * (n - 1) / (kernel_ms * 1e-3) / 1e12

    result = {
        "metric": "direct interactions/s N=256k",
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / max(1, args.steps),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pairwise terms, fp64 state and update",
        "data": "synthetic Plummer sphere (seeded generator), device-resident",
        "config": {"workload": "direct O(N^2) Plummer, Yoshida-4, one GPU", "n": n, "softening": eps,
                   "timestep": dt, "integrator": "yoshida4", "evaluations_per_step": 3,
                   "seed": args.seed},
        "ms_per_evaluation": eval_ms,
        "kernel_ms_per_evaluation": kernel_ms,
        "force_error": {"rms_relative": rms, "max_relative": worst, "targets": int(len(tgt)),
                        "reference": "FP64 Neumaier direct sum"},
        "phases_ms": {k: v[0] for k, v in timings.items()},
        "roofline": {"bound": "fp32-valu", "kernel": "direct_tiles", "achieved": achieved,
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": None,
                     "note": "20 flop per interaction x N(N-1) per launch / mean launch time"},
    }
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline")
        result["cpu_baseline"] = cpu_baseline(state, eps)
    else:
        result["cpu_baseline"] = None
    sys_.close()
    ctx.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
