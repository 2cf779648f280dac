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


def measured_traffic(n: int):
    """HBM bytes per direct_tiles launch from the committed, calibrated PMC profile."""
    path = os.path.join(HERE, "profiles", "direct_tiles_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t.get("traffic_bytes") if f"N={n}," in t.get("workload", "") else None


def measure_ceilings(ctx) -> dict:
    """FP32 FMA, v_rsq_f32, HBM read and copy rates measured by the engine's probes."""
    ctx.probe("fma", 256)  # warm-up
    fma, _ = ctx.probe("fma", 4096)
    rsq, _ = ctx.probe("rsq", 1024)
    ctx.probe("read", 1 << 28)
    read, _ = ctx.probe("read", 1 << 31)
    copy, copy_ms = ctx.probe("copy", 1 << 30)
    return {"fp32_fma_TFLOPs": fma / 1e12, "rsq_Gops": rsq / 1e9, "hbm_read_GBs": read / 1e9,
            "hbm_copy_GBs": copy / 1e9, "copy_bytes": 2 * (1 << 30), "copy_ms": copy_ms}


def bench_barnes_hut(E, oracle, ctx, n: int, evals: int, error_targets: int, cpu: bool) -> dict:
    """Barnes-Hut on the two-disc collision: N particles, theta 0.5, leaf 16."""
    import ctypes

    eps, theta, cap = 0.01, 0.5, 16
    log(f"[bench] Barnes-Hut: generating two-disc collision N={n}")
    p = oracle.disc_collision(n, 1)
    d = [E.DeviceArray.from_host(ctx, a) for a in (p.x, p.y, p.z, p.m)]
    out = [E.DeviceArray(ctx, n) for _ in range(3)]
    rng = np.random.default_rng(2)
    tgt = np.sort(rng.choice(n, size=min(n, error_targets), replace=False)).astype(np.uint64)
    ref = oracle.reference_acceleration(p.x, p.y, p.z, p.m, eps * eps, targets=tgt, threads=cpu_threads())
    i = tgt.astype(np.int64)

    def run(quadrupole: bool):
        s = E.HipTreeSolver(E.TreeParameters(theta, cap, quadrupole=quadrupole), E.Softening(eps), context=ctx)
        s.evaluate_device(*(a.ptr for a in d), *(a.ptr for a in out), n)  # warm-up
        s.reset_interaction_count()
        ctx.set_timing(True)
        ctx.reset_timings()
        t0 = time.perf_counter()
        for _ in range(evals):
            s.evaluate_device(*(a.ptr for a in d), *(a.ptr for a in out), n)
        wall = (time.perf_counter() - t0) / evals
        phases = {k: v[0] / evals for k, v in ctx.timings().items()}
        ctx.set_timing(False)
        acc = [a.to_host() for a in out]
        err = oracle.error_summary(oracle.relative_errors(acc[0][i], acc[1][i], acc[2][i], *ref))
        return s, wall, phases, err

    mono_s, mono_wall, mono_phases, (mono_worst, mono_rms) = run(False)
    s, wall, phases, (worst, rms) = run(True)
    c = s.interaction_count()
    tree = s.tree_summary()
    M = tree["node_count"]
    build_phases = ("bbox", "encode", "sort", "gather", "build", "moments")
    build_ms = sum(phases.get(k, 0.0) for k in build_phases)
    # Compulsory bytes per evaluation of the streaming stages: 228 N + 64 M
    # (bbox 24N, encode 36N, sort 24N, gather 52N, build 24N + 32M, walk 28N + 32M,
    # scatter 40N); the walk's own stream lives in L2 / LDS.
    alg_bytes = 228 * n + 64 * M
    pp, pc = c.particle_particle // evals, c.particle_cell // evals
    out_d = {
        "workload": "Barnes-Hut, two exponential discs + Hernquist bulges, parabolic orbit",
        "n": n, "opening_angle": theta, "leaf_capacity": cap, "softening": eps, "quadrupole": True,
        "ms_per_evaluation": wall * 1e3,
        "monopole": {"ms_per_evaluation": mono_wall * 1e3, "phases_ms": mono_phases,
                     "force_error": {"rms_relative": mono_rms, "max_relative": mono_worst}},
        "phases_ms": phases,
        "interactions_per_evaluation": {"particle_particle": pp, "particle_cell": pc},
        "walk_interactions_per_s": (pp + pc) / (phases.get("walk", 0.0) * 1e-3) if phases.get("walk") else None,
        "tree": tree,
        "force_error": {"rms_relative": rms, "max_relative": worst, "targets": int(len(tgt)),
                        "reference": "FP64 Neumaier direct sum"},
        "streaming_stages": {"algorithmic_bytes": alg_bytes, "ms": build_ms,
                             "achieved_GBs": alg_bytes / (build_ms * 1e-3) / 1e9 if build_ms else None,
                             "peak_GBs": HBM_PEAK_GBS},
    }
    for a in d + out:
        a.release()
    if cpu:
        ref_lib = oracle.load_reference()
        sub = 262144
        q = oracle.disc_collision(sub, 1)
        ax, ay, az = np.zeros(sub), np.zeros(sub), np.zeros(sub)
        counts = np.zeros(2, dtype=np.uint64)
        from oracle import _ptr, _u64, _u64p
        t0 = time.perf_counter()
        if ref_lib is not None:
            kind = "reference"
            ref_lib.ref_barnes_hut(_u64(sub), _ptr(q.x), _ptr(q.y), _ptr(q.z), _ptr(q.m),
                                   ctypes.c_double(eps), ctypes.c_double(theta), _u64(cap), ctypes.c_int(0),
                                   ctypes.c_int(1), ctypes.c_uint(cpu_threads()), _ptr(ax), _ptr(ay), _ptr(az),
                                   _ptr(counts, _u64p))
        else:
            kind = "port"
            oracle.barnes_hut(q.x, q.y, q.z, q.m, eps * eps, theta, cap, threads=cpu_threads())
        dt = time.perf_counter() - t0
        out_d["cpu_baseline"] = {"value": dt * 1e3, "unit": "ms/evaluation", "cores": cpu_threads(),
                                 "kind": kind,
                                 "sample": f"N={sub} two-disc collision (same generator), theta 0.5, "
                                           f"leaf 16, the reference's threaded BarnesHutSolver"}
    return out_d


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--error-targets", type=int, default=1024)
    ap.add_argument("--bh-n", type=int, default=2097152)
    ap.add_argument("--bh-evals", type=int, default=5)
    ap.add_argument("--no-barnes-hut", action="store_true")
    ap.add_argument("--probes-only", action="store_true",
                    help="run only the machine-ceiling probes (copy probe: counter calibration)")
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

    if args.probes_only:
        ctx = E.Context(0)
        print(json.dumps({"ceilings": measure_ceilings(ctx)}), flush=True)
        ctx.close()
        return 0

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
                     "traffic": measured_traffic(n),
                     "note": "20 flop per interaction x N(N-1) per launch / mean launch time; "
                             "traffic = calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE bytes per launch "
                             "(profiles/direct_tiles_traffic.json)"},
    }
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline")
        result["cpu_baseline"] = cpu_baseline(state, eps)
    else:
        result["cpu_baseline"] = None
    sys_.close()
    if not args.no_barnes_hut:
        result["barnes_hut"] = bench_barnes_hut(E, oracle, ctx, args.bh_n, args.bh_evals,
                                                args.error_targets, not args.no_cpu_baseline)
    log("[bench] ceilings")
    result["ceilings"] = measure_ceilings(ctx)
    ctx.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
