#!/usr/bin/env python3
"""Benchmark of the MI355X force engine on Orrery's headline workloads.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu-baseline]

Headline (``value``), direct O(N^2) on one GPU: a Plummer sphere of
N = 262,144 equal-mass particles (total mass 1, G = 1, scale radius a = 1, radii
by inverse CDF truncated at r < 10 a, rejection-sampled speeds, seed 1, drawn
as the reference's sampler draws it), Plummer
softening eps = 0.005, Yoshida-4 with dt = 5e-4. A step is one Yoshida-4 step
on device-resident FP64 state: three force evaluations (FP32 pairwise terms,
compensated accumulation) and the FP64 kick/drift updates around them.
``value`` is interactions per second over the K timed steps, N(N-1) per
evaluation, inputs resident in HBM. ``roofline`` is for the dominant launches
(direct_pairs, which evaluates each far pair of tiles once for both tiles, and
direct_complement, the pairs inside each block of 8 tiles) from HIP events
recorded on the engine's own stream (direct_complement runs on a second
stream beside direct_pairs; the events bracket both).

Further legs in the same JSON line:
* ``direct_n65536``, ``direct_n524288`` -- the same direct stepping at
  N = 65,536 and 524,288 (the latter with its error on sampled targets);
* ``plummer4096_vv`` -- the plain CPU path's run (Plummer N = 4096, eps 0.01,
  velocity Verlet dt = 1e-3, 100 steps): time per step and energy drift on the
  host's reference kernel, beside the device-resident single-system path;
* ``barnes_hut`` -- two-disc collision at N = 2,097,152, theta 0.5, leaf 16;
* ``batched``    -- 512 Plummer spheres, N log-uniform in [64, 16384], eps 0.02,
  direct below 4096 particles and Barnes-Hut (theta 0.6) above, velocity Verlet
  dt = 1e-3, one batch step per step replayed as a hipGraph, against the same
  systems stepped one by one;
* ``served``     -- Poisson arrivals of run requests (N log-uniform in [1e3, 2e6],
  70/30 Plummer / two-disc collision, 100-1000 Yoshida-4 or velocity Verlet
  steps) at 0.8 of the single-system path's service rate, through the
  continuous-batching server: p50/p99 time to completion, throughput,
  per-request energy drift against the CPU reference, and beside them the same
  requests and arrivals served one at a time on the single-system path (every
  request run whole; first-in-first-out latencies);
* ``cpu_baseline`` and the ``force_error`` figures -- the CPU leg: the
  reference's threaded AVX2 FP64 kernels (or the oracle's restatement) timed on
  this host's cores under the reference's protocol (2 warm-ups, settling within
  5 %, 11 trials, median and quartiles; CPU model, AVX2/FMA and thread count
  stated) over bounded samples of the same inputs, and the FP64 Neumaier direct
  sums the GPU accelerations are checked against.

Inputs come from orrery_amd.workloads: Plummer spheres from the engine's
sampler, which draws the reference's mt19937_64 sequence in the reference's
order (src/initial_conditions/plummer.cpp), so a seed gives the reference's own
sample bit for bit; the two-disc model from a seeded numpy generator.
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

METRIC = "direct interactions/s N=256k; BH ms/eval N=2M; served p50/p99 latency"
FLOP_PER_INTERACTION = 20          # the reference's convention (docs/performance.md:103-109)
FP32_PEAK_TFLOPS = 157.3           # MI355X published FP32 vector peak
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_threads() -> int:
    """Every logical CPU in this process's affinity mask, unless OMP_NUM_THREADS
    names fewer: on a GPU host whose CPUs are shared between GPU jobs that
    variable states this job's share (it does on the MI355X hosts), and more
    threads than the share only contend with the neighbours' work."""
    n = os.cpu_count() or 1
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def cpu_description() -> dict:
    """CPU model, logical processors and the vector features the reference's
    AVX2 kernel needs (benchmarks/harness/machine_state.cpp:94-118 prints the same)."""
    model, flags = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                key, _, val = line.partition(":")
                key = key.strip()
                if key == "model name" and model == "unknown":
                    model = val.strip()
                elif key == "flags" and not flags:
                    flags = set(val.split())
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return {"model": model, "logical_processors": os.cpu_count() or 1, "affinity": affinity,
            "avx2": "avx2" in flags, "fma": "fma" in flags,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def run_trials(body, warmup: int = 2, settling_limit: int = 24, settled_within: float = 0.05,
               trials: int = 11, minimum_trial: float = 0.05) -> dict:
    """The reference's measurement protocol (benchmarks/harness/protocol.hpp:67-132,
    protocol.cpp, statistics.cpp): `warmup` discarded calls; one call sizes the
    trial (calls folded into a trial until it lasts `minimum_trial` seconds, at
    most 1000); trials are discarded until two consecutive ones agree within
    `settled_within` twice in a row (at most `settling_limit`); then `trials`
    timed trials, each recorded per call. Median and quartiles by nearest rank
    on the sorted trials; drift = the second half's median over the first's, - 1."""
    for _ in range(warmup):
        body()

    def timed(repeats):
        t0 = time.perf_counter()
        for _ in range(repeats):
            body()
        return time.perf_counter() - t0

    single = timed(1)
    repeats = 1 if minimum_trial <= 0 or single <= 0 or single >= minimum_trial else \
        int(min(minimum_trial / single, 1000))
    previous, agreements, settle = None, 0, 0
    for _ in range(settling_limit):
        current = timed(repeats)
        settle += 1
        if previous is not None:
            change = abs(current - previous) / previous if previous > 0 else 1.0
            agreements = agreements + 1 if change < settled_within else 0
            if agreements >= 2:
                break
        previous = current
    in_order = [timed(repeats) / repeats for _ in range(trials)]
    ordered = sorted(in_order)

    def quantile(f):
        return ordered[min(int(f * len(ordered)), len(ordered) - 1)]

    drift = 0.0
    if len(in_order) >= 4:
        half = len(in_order) // 2
        early, late = sorted(in_order[:half]), sorted(in_order[-half:])
        drift = late[len(late) // 2] / early[len(early) // 2] - 1.0
    med = quantile(0.5)
    return {"median_s": med, "q1_s": quantile(0.25), "q3_s": quantile(0.75),
            "relative_spread": (quantile(0.75) - quantile(0.25)) / med if med > 0 else 0.0,
            "fastest_s": ordered[0], "slowest_s": ordered[-1], "drift": drift,
            "trials": trials, "warmup": warmup, "settling_trials": settle,
            "calls_per_trial": repeats}


COOLDOWN_S = 0.75  # between configurations (protocol.hpp: cooldown)


class CpuLeg:
    """Everything this benchmark runs on the host CPU: the reference's own
    kernels (oracle/_ref, else the oracle's restatement) timed as the baseline
    under the reference's trial protocol, and the FP64 reference forces the GPU
    results are checked against. None of it is part of the measured GPU path."""

    def __init__(self):
        import oracle
        self.oracle = oracle
        self.ref = oracle.load_reference()
        self.threads = cpu_threads()
        self.machine = cpu_description()
        self.kind = "reference" if self.ref is not None else "port"
        log(f"[bench] CPU: {self.machine['model']}, {self.threads} threads of "
            f"{self.machine['affinity']} in the affinity mask, avx2 {self.machine['avx2']}, "
            f"fma {self.machine['fma']}")

    def stamp(self, stats: dict, value: float, unit: str, sample: str, **extra) -> dict:
        out = {"value": value, "unit": unit, "cores": self.threads, "kind": self.kind,
               "sample": sample, "cpu": self.machine, "protocol": stats}
        out.update(extra)
        return out

    def fp64_errors(self, state, eps: float, acc, targets) -> dict:
        o = self.oracle
        rx, ry, rz = o.reference_acceleration(state.x, state.y, state.z, state.m, eps * eps,
                                              targets=targets.astype(np.uint64), threads=self.threads)
        i = targets.astype(np.int64)
        worst, rms = o.error_summary(o.relative_errors(acc[0][i], acc[1][i], acc[2][i], rx, ry, rz))
        return {"rms_relative": rms, "max_relative": worst, "targets": int(len(targets)),
                "reference": "FP64 Neumaier direct sum"}

    def direct_baseline(self, state, eps: float, targets: int) -> dict:
        """Interactions/s of the AVX2 FP64 direct kernel on the work-stealing
        executor over `targets` targets (all N when targets >= N) x all sources."""
        import ctypes
        from oracle import _ptr, _u64
        n = state.n
        targets = min(n, targets)
        ax, ay, az = np.zeros(n), np.zeros(n), np.zeros(n)
        if self.ref is not None:
            def run():
                self.ref.ref_direct_targets(_u64(n), _ptr(state.x), _ptr(state.y), _ptr(state.z),
                                            _ptr(state.m), ctypes.c_double(eps), ctypes.c_int(1),
                                            ctypes.c_uint(self.threads), _u64(0), _u64(targets),
                                            _ptr(ax), _ptr(ay), _ptr(az))
        else:
            lib = self.oracle.load()

            def run():
                lib.orc_direct(_u64(n), _ptr(state.x), _ptr(state.y), _ptr(state.z), _ptr(state.m),
                               ctypes.c_double(eps * eps), ctypes.c_double(1.0), ctypes.c_int(1),
                               ctypes.c_uint(self.threads), _u64(0), _u64(targets), _ptr(ax),
                               _ptr(ay), _ptr(az))

        time.sleep(COOLDOWN_S)
        st = run_trials(run)
        inter = targets * (n - 1)
        what = "every target" if targets == n else f"{targets} of {n} targets"
        return self.stamp(st, inter / st["median_s"], "interactions/s",
                          f"{what} x all {n} sources per call, FP64 AVX2 kernel on the "
                          f"work-stealing executor; value = interactions per call / median call time",
                          interactions_per_call=inter,
                          iqr_interactions_per_s=[inter / st["q3_s"], inter / st["q1_s"]])

    def barnes_hut_baseline(self, state, eps: float, theta: float, cap: int, quadrupole: bool,
                            trials: int = 11, warmup: int = 2, settling_limit: int = 24) -> dict:
        import ctypes
        from oracle import _ptr, _u64, _u64p
        n = state.n
        ax, ay, az = np.zeros(n), np.zeros(n), np.zeros(n)
        counts = np.zeros(2, dtype=np.uint64)
        if self.ref is not None:
            def run():
                self.ref.ref_barnes_hut(_u64(n), _ptr(state.x), _ptr(state.y), _ptr(state.z),
                                        _ptr(state.m), ctypes.c_double(eps), ctypes.c_double(theta),
                                        _u64(cap), ctypes.c_int(1 if quadrupole else 0), ctypes.c_int(1),
                                        ctypes.c_uint(self.threads), _ptr(ax), _ptr(ay), _ptr(az),
                                        _ptr(counts, _u64p))
        else:
            def run():
                self.oracle.barnes_hut(state.x, state.y, state.z, state.m, eps * eps, theta, cap,
                                       quadrupole=quadrupole, threads=self.threads)

        time.sleep(COOLDOWN_S)
        st = run_trials(run, warmup=warmup, settling_limit=settling_limit, trials=trials)
        return self.stamp(st, st["median_s"] * 1e3, "ms/evaluation",
                          f"N={n} two-disc collision (the GPU leg's arrays), theta {theta}, leaf "
                          f"{cap}, {'quadrupole' if quadrupole else 'monopole'}, the reference's "
                          f"BarnesHutSolver on the work-stealing executor, AVX2 leaf kernel",
                          iqr_ms=[st["q1_s"] * 1e3, st["q3_s"] * 1e3],
                          interactions={"particle_particle": int(counts[0]),
                                        "particle_cell": int(counts[1])} if self.ref is not None else None)

    def plummer_vv(self, state, eps: float, dt: float, steps: int) -> dict:
        """The plain CPU run: `steps` velocity Verlet steps of the direct solver
        (AVX2 kernel, work-stealing executor) from `state`, timed per run under
        the protocol; FP64 energy drift of one run."""
        import ctypes
        from oracle import _ptr, _u64
        o = self.oracle
        n = state.n
        f = o.make_field("vv", "direct", eps2=eps * eps, threads=self.threads)
        start = o.Particles(*(a.copy() for a in state.arrays()))
        acc0 = o.refresh(start, f)
        work = o.Particles(*(a.copy() for a in state.arrays()))
        acc = [a.copy() for a in acc0]

        def reset():
            for a, b in zip(work.arrays(), start.arrays()):
                a[:] = b
            for a, b in zip(acc, acc0):
                a[:] = b

        if self.ref is not None:
            def run():
                reset()
                self.ref.ref_steps_threaded(_u64(n), *(_ptr(a) for a in work.arrays()[:6]),
                                            *(_ptr(a) for a in acc), _ptr(work.m), ctypes.c_int(0),
                                            ctypes.c_int(2), ctypes.c_double(eps), ctypes.c_double(0.5),
                                            _u64(16), ctypes.c_double(dt), _u64(steps),
                                            ctypes.c_uint(self.threads))
        else:
            def run():
                reset()
                o.steps(work, acc, f, dt, steps)

        time.sleep(COOLDOWN_S)
        st = run_trials(run)
        e0 = o.diagnostics(start, eps * eps)["total"]
        e1 = o.diagnostics(work, eps * eps)["total"]
        return self.stamp(st, st["median_s"] * 1e3 / steps, "ms/step",
                          f"{steps} velocity Verlet steps per call from the same initial state, "
                          f"direct FP64 AVX2 kernel on the work-stealing executor",
                          iqr_ms_per_step=[st["q1_s"] * 1e3 / steps, st["q3_s"] * 1e3 / steps],
                          energy_drift=(e1 - e0) / abs(e0))


def measured_traffic(n: int):
    """HBM bytes per evaluation of direct_pairs + direct_complement from the
    committed, calibrated PMC profile."""
    path = os.path.join(HERE, "profiles", "direct_pairs_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t.get("traffic_bytes") if f"N={n}," in t.get("workload", "") else None


def measure_ceilings(ctx) -> dict:
    """FP32 FMA, v_rsq_f32, HBM read and copy rates measured by the engine's probes."""
    # Each rate over tens of milliseconds: launches of a few milliseconds read
    # low while the clocks ramp.
    ctx.probe("fma", 256)  # warm-up
    fma, _ = ctx.probe("fma", 65536)
    rsq, _ = ctx.probe("rsq", 16384)
    mix, _ = ctx.probe("mix", 16384)
    ctx.probe("read", 1 << 28)
    read, _ = ctx.probe("read", 1 << 31)
    copy, copy_ms = ctx.probe("copy", 1 << 30)
    return {"fp32_fma_TFLOPs": fma / 1e12, "rsq_Gops": rsq / 1e9,
            "interaction_mix_per_s": mix, "hbm_read_GBs": read / 1e9,
            "hbm_copy_GBs": copy / 1e9, "copy_bytes": 2 * (1 << 30), "copy_ms": copy_ms}


def direct_accuracy(E, W, ctx, cpu, seeds=(1, 2, 3)) -> list:
    """Config 2's accuracy statement: force error of one evaluation against the
    FP64 direct sum at N = 65,536 (every target) and 262,144 (4096 targets), per seed."""
    eps, rows = 0.005, []
    rng = np.random.default_rng(77)
    for n, sample in ((65536, 65536), (262144, 4096)):
        for seed in seeds:
            state = W.plummer(n, seed)
            s = E.DeviceSystem(ctx, n)
            s.upload(*state.arrays())
            s.refresh(E.StepConfig(integrator="yoshida4", solver="direct", softening=eps))
            acc = s.download()
            s.close()
            tgt = np.arange(n) if sample >= n else np.sort(rng.choice(n, size=sample, replace=False))
            e = cpu.fp64_errors(state, eps, (acc["ax"], acc["ay"], acc["az"]), tgt)
            rows.append({"n": n, "seed": seed, "rms_relative": e["rms_relative"],
                         "max_relative": e["max_relative"], "targets": e["targets"]})
    return rows


def direct_rate(E, W, ctx, n: int, steps: int, seed: int = 1, cpu=None, targets: int = 1024) -> dict:
    """The same Yoshida-4 stepping of a device-resident Plummer sphere at
    another N, timed over `steps` steps after one warm-up step; with the CPU
    leg, the first evaluation's error on sampled targets against FP64."""
    eps, dt = 0.005, 5e-4
    state = W.plummer(n, seed)
    s = E.DeviceSystem(ctx, n)
    s.upload(*state.arrays())
    cfg = E.StepConfig(integrator="yoshida4", solver="direct", softening=eps)
    s.refresh(cfg)
    error = None
    if cpu is not None:
        a = s.download()
        rng = np.random.default_rng(seed + 100)
        tgt = np.sort(rng.choice(n, size=min(n, targets), replace=False))
        error = cpu.fp64_errors(state, eps, (a["ax"], a["ay"], a["az"]), tgt)
    s.step(cfg, dt, 1)
    ctx.set_timing(True)
    ctx.reset_timings()
    t0 = time.perf_counter()
    s.step(cfg, dt, steps)
    elapsed = time.perf_counter() - t0
    timings = ctx.timings()
    ctx.set_timing(False)
    s.close()
    evals = 3 * steps
    k_ms, k_launches = timings.get("direct_kernel", (0.0, 0))
    kernel_ms = k_ms / max(1, k_launches)
    fold_ms, fold_launches = timings.get("direct_reduce", (0.0, 0))
    return {"n": n, "steps": steps, "interactions_per_s": evals * n * (n - 1) / elapsed,
            "ms_per_evaluation": elapsed * 1e3 / evals, "kernel_ms_per_evaluation": kernel_ms,
            "fold_ms_per_evaluation": fold_ms / max(1, fold_launches),
            "kernel_frac_of_fp32_peak": FLOP_PER_INTERACTION * n * (n - 1) / (kernel_ms * 1e-3) / 1e12
            / FP32_PEAK_TFLOPS, "force_error": error}


def bench_plummer_vv(E, W, ctx, cpu) -> dict:
    """The plain CPU path's configuration: Plummer N = 4096 (seed 1), eps 0.01,
    velocity Verlet dt = 1e-3 for 100 steps. CPU: time per step under the
    reference's protocol and FP64 energy drift; beside it the device-resident
    single-system path on the same state (time per step over five runs of the
    100 steps, drift, and the first force evaluation against FP64)."""
    n, eps, dt, steps, runs = 4096, 0.01, 1e-3, 100, 5
    state = W.plummer(n, 1)
    cfg = E.StepConfig(integrator="velocity-verlet", solver="direct", softening=eps)
    s = E.DeviceSystem(ctx, n)
    times, drift, force = [], None, None
    for r in range(runs + 1):
        s.upload(*state.arrays())
        e0 = s.diagnostics(eps)["total"]
        s.refresh(cfg)
        if r == 1 and cpu is not None:
            a = s.download()
            force = cpu.fp64_errors(state, eps, (a["ax"], a["ay"], a["az"]), np.arange(n))
        t0 = time.perf_counter()
        s.step(cfg, dt, steps)
        elapsed = time.perf_counter() - t0
        e1 = s.diagnostics(eps)["total"]
        if r > 0:  # the first run warms up
            times.append(elapsed * 1e3 / steps)
            drift = (e1 - e0) / abs(e0)
    s.close()
    times.sort()
    out = {"workload": f"Plummer N={n} seed 1, eps {eps}, velocity Verlet dt {dt}, {steps} steps, "
                       f"direct", "n": n, "steps": steps,
           "gpu": {"ms_per_step": times[len(times) // 2], "runs": runs,
                   "range_ms_per_step": [times[0], times[-1]], "energy_drift": drift,
                   "force_error": force,
                   "path": "device-resident single-system path (orr_system), each step replayed as "
                           "one captured graph"}}
    if cpu is not None:
        log("[bench] CPU: Plummer N=4096 velocity Verlet, 100 steps per trial")
        out["cpu_baseline"] = cpu.plummer_vv(state, eps, dt, steps)
    return out


def bench_barnes_hut(E, W, ctx, n: int, evals: int, error_targets: int, cpu) -> dict:
    """Barnes-Hut on the two-disc collision: N particles, theta 0.5, leaf 16; the
    same settings at 262,144 particles as the cross-check size."""
    eps, theta, cap = 0.01, 0.5, 16
    log(f"[bench] Barnes-Hut: generating two-disc collision N={n}")
    p = W.disc_collision(n, 1)

    def run(p, quadrupole: bool):
        n = p.n
        d = [E.DeviceArray.from_host(ctx, a) for a in (p.x, p.y, p.z, p.m)]
        out = [E.DeviceArray(ctx, n) for _ in range(3)]
        rng = np.random.default_rng(2)
        tgt = np.sort(rng.choice(n, size=min(n, error_targets), replace=False))
        s = E.HipTreeSolver(E.TreeParameters(theta, cap, quadrupole=quadrupole), E.Softening(eps), context=ctx)
        # Warm-up: the first evaluation sizes the buffers, the second is
        # captured as a hipGraph; the timed ones replay it.
        for _ in range(2):
            s.evaluate_device(*(a.ptr for a in d), *(a.ptr for a in out), n)
        s.reset_interaction_count()
        t0 = time.perf_counter()
        for _ in range(evals):
            s.evaluate_device(*(a.ptr for a in d), *(a.ptr for a in out), n)
        wall = (time.perf_counter() - t0) / evals
        # Phase times from eager evaluations (events inside a graph are not
        # timed), launch gaps included.
        ctx.set_timing(True)
        ctx.reset_timings()
        for _ in range(evals):
            s.evaluate_device(*(a.ptr for a in d), *(a.ptr for a in out), n)
        phases = {k: v[0] / evals for k, v in ctx.timings().items()}
        ctx.set_timing(False)
        acc = [a.to_host() for a in out]
        err = cpu.fp64_errors(p, eps, acc, tgt) if cpu is not None else None
        for a in d + out:
            a.release()
        return s, wall, phases, err

    mono_s, mono_wall, mono_phases, mono_err = run(p, False)
    s, wall, phases, err = run(p, True)
    c = s.interaction_count()
    tree = s.tree_summary()
    M = tree["node_count"]
    build_phases = ("bbox", "encode", "sort", "gather", "build", "moments")
    # Streaming stages: their phase events in eager evaluations (host launch
    # gaps included, so an upper bound of their time in a replay).
    build_ms = sum(phases.get(k, 0.0) for k in build_phases)
    # Compulsory bytes per evaluation of the streaming stages: 228 N + 64 M
    # (bbox 24N, encode 36N, sort 24N, gather 52N, build 24N + 32M, walk 28N + 32M,
    # scatter 40N); the walk's own stream lives in L2 / LDS.
    alg_bytes = 228 * n + 64 * M
    pp, pc = c.particle_particle // (2 * evals), c.particle_cell // (2 * evals)
    out_d = {
        "workload": "Barnes-Hut, two exponential discs + Hernquist bulges, parabolic orbit",
        "n": n, "opening_angle": theta, "leaf_capacity": cap, "softening": eps, "quadrupole": True,
        "ms_per_evaluation": wall * 1e3,
        "monopole": {"ms_per_evaluation": mono_wall * 1e3, "phases_ms": mono_phases,
                     "force_error": mono_err},
        "phases_ms": phases,
        "interactions_per_evaluation": {"particle_particle": pp, "particle_cell": pc},
        "walk_interactions_per_s": (pp + pc) / (phases.get("walk", 0.0) * 1e-3) if phases.get("walk") else None,
        "walk_lane_efficiency": s.walk_lane_efficiency(),
        "tree": tree,
        "force_error": err,
        "streaming_stages": {"algorithmic_bytes": alg_bytes, "ms": build_ms,
                             "achieved_GBs": alg_bytes / (build_ms * 1e-3) / 1e9 if build_ms > 0 else None,
                             "peak_GBs": HBM_PEAK_GBS,
                             "replay_minus_walk_ms": wall * 1e3 - phases.get("walk", 0.0),
                             "note": "ms = the stages' phase events in eager evaluations, host "
                                     "launch gaps included; replay_minus_walk_ms = the replayed "
                                     "evaluation's wall time less the eager walk time (within the "
                                     "walk's run-to-run spread of about 0.05 ms)"},
    }
    small = W.disc_collision(262144, 1) if n != 262144 else p
    if n != 262144:
        # The cross-check size with the headline's settings (quadrupole).
        log("[bench] Barnes-Hut: N=262,144 cross-check")
        _, x_wall, x_phases, x_err = run(small, True)
        out_d["n262144"] = {"ms_per_evaluation": x_wall * 1e3, "phases_ms": x_phases,
                            "force_error": x_err}
    if cpu is not None:
        # The GPU headline's settings (quadrupole, theta 0.5, leaf 16) on the
        # same arrays; at 2M one evaluation takes seconds on the host, so that
        # size runs one warm-up and three timed trials (stated in "protocol").
        log("[bench] CPU: Barnes-Hut N=262,144 and the GPU leg's arrays")
        out_d["cpu_baseline"] = {
            "n262144": cpu.barnes_hut_baseline(small, eps, theta, cap, True),
            f"n{n}": cpu.barnes_hut_baseline(p, eps, theta, cap, True, trials=3, warmup=1,
                                              settling_limit=0),
        }
    return out_d


def bench_batched(E, W, ctx, count: int, steps: int, alone_steps: int, cpu) -> dict:
    """512 Plummer spheres of uneven N in one batch against the same systems one by one."""
    eps, dt, theta, below, seed = 0.02, 1e-3, 0.6, 4096, 4
    log(f"[bench] batched: generating {count} Plummer spheres")
    ns = W.batched_sizes(count, 64, 16384, seed)
    spheres = W.batched_spheres(count, 64, 16384, seed)
    systems = [E.BatchSystem(n, eps, dt) for n in ns]

    def fill(b):
        for k, s in enumerate(spheres):
            b.upload(k, *s.arrays())

    cfg = E.BatchConfig(integrator="velocity-verlet", opening_angle=theta, leaf_capacity=0,
                        tiles_per_item=0, direct_below=below, use_graph=True)
    b = E.DeviceBatch(ctx, systems, cfg)
    fill(b)
    t0 = time.perf_counter()
    info = b.autotune()
    tune_s = time.perf_counter() - t0
    log(f"[bench] batched: autotuned tiles_per_item={info['tiles_per_item']} "
        f"leaf_capacity={info['leaf_capacity']} in {tune_s:.2f} s")
    b.refresh()
    before = [b.diagnostics(k) for k in range(count)]
    b.step(2)  # the first step runs eagerly and captures the graph
    t0 = time.perf_counter()
    b.step(steps)
    wall = time.perf_counter() - t0
    info = b.info()
    tree_pairs = 0
    tree_cells = 0
    for k, n in enumerate(ns):
        if n >= below:
            pp, pc = b.counts(k)
            tree_pairs += pp
            tree_cells += pc
    per_eval = info["direct_pairs"] + tree_pairs + tree_cells
    after = [b.diagnostics(k) for k in range(count)]
    drift = [abs((a["total"] - e["total"]) / e["total"]) for a, e in zip(after, before)]

    # Accuracy of the last evaluation on a sample of systems, vs FP64.
    errors = None
    if cpu is not None:
        rng = np.random.default_rng(9)
        picks = sorted(set(int(k) for k in rng.choice(count, size=12, replace=False)))
        worst_d = rms_d = worst_t = rms_t = 0.0
        for k in picks:
            st = b.download(k)
            state = W.Workload(st["x"], st["y"], st["z"], st["vx"], st["vy"], st["vz"], spheres[k].m)
            tg = np.sort(rng.choice(ns[k], size=min(ns[k], 256), replace=False))
            e = cpu.fp64_errors(state, eps, (st["ax"], st["ay"], st["az"]), tg)
            if ns[k] < below:
                worst_d, rms_d = max(worst_d, e["max_relative"]), max(rms_d, e["rms_relative"])
            else:
                worst_t, rms_t = max(worst_t, e["max_relative"]), max(rms_t, e["rms_relative"])
        errors = {"systems_sampled": len(picks), "targets_per_system": 256,
                  "direct": {"max_rms_relative": rms_d, "max_relative": worst_d},
                  "barnes_hut": {"max_rms_relative": rms_t, "max_relative": worst_t},
                  "reference": "FP64 Neumaier direct sum"}
    b.close()

    # The same batch without the graph: per-launch host overhead included.
    cfg_eager = E.BatchConfig(integrator="velocity-verlet", opening_angle=theta,
                              leaf_capacity=info["leaf_capacity"],
                              tiles_per_item=info["tiles_per_item"], direct_below=below,
                              use_graph=False)
    b = E.DeviceBatch(ctx, systems, cfg_eager)
    fill(b)
    b.refresh()
    b.step(2)
    eager_steps = max(1, min(steps, 20))
    t0 = time.perf_counter()
    b.step(eager_steps)
    eager_ms = (time.perf_counter() - t0) * 1e3 / eager_steps
    b.close()

    # Each system alone: the plain single-system device path, one system at a time.
    alone_ms = None
    if alone_steps > 0:
        log(f"[bench] batched: stepping the {count} systems one by one ({alone_steps} steps each)")
        alone_ms = 0.0
    for k, s in enumerate(spheres if alone_steps > 0 else []):
        solver = "direct" if ns[k] < below else "barnes-hut"
        c = E.StepConfig(integrator="velocity-verlet", solver=solver, softening=eps,
                         opening_angle=theta, leaf_capacity=info["leaf_capacity"])
        one = E.DeviceSystem(ctx, ns[k])
        one.upload(*s.arrays())
        one.refresh(c)
        one.step(c, dt, 1)
        t0 = time.perf_counter()
        one.step(c, dt, alone_steps)
        alone_ms += (time.perf_counter() - t0) * 1e3 / alone_steps
        one.close()

    ms = wall * 1e3 / steps
    return {
        "workload": f"{count} Plummer spheres, N log-uniform in [64, 16384] (seed {seed}), "
                    f"direct below {below}, Barnes-Hut theta {theta} above, velocity Verlet",
        "systems": count, "particles": info["particles"],
        "direct_systems": info["direct_systems"], "tree_systems": info["tree_systems"],
        "softening": eps, "timestep": dt, "steps": steps,
        "autotuned": {"tiles_per_item": info["tiles_per_item"],
                      "leaf_capacity": info["leaf_capacity"], "seconds": tune_s},
        "ms_per_step": ms, "steps_per_s": 1e3 / ms,
        "interactions_per_step": {"direct_pairs": info["direct_pairs"],
                                  "tree_particle_particle": tree_pairs,
                                  "tree_particle_cell": tree_cells},
        "interactions_per_s": per_eval / (ms * 1e-3),
        "flop_rate_TFLOPs": FLOP_PER_INTERACTION * per_eval / (ms * 1e-3) / 1e12,
        "eager_ms_per_step": eager_ms,
        "one_by_one_ms_per_step": alone_ms,
        "speedup_vs_one_by_one": alone_ms / ms if alone_ms else None,
        "one_by_one": f"each system its own device-resident system (orr_system), stepped alone "
                      f"({alone_steps} steps each, after one warm-up step)",
        "energy_drift": {"max_abs_relative": max(drift), "median_abs_relative": float(np.median(drift)),
                         "steps": steps + 2},
        "force_error": errors,
    }


def make_request_state(spec):
    """One request's initial state (host only)."""
    kind, n, seed = spec
    from orrery_amd import workloads as W
    p = W.plummer(n, seed) if kind == "plummer" else W.disc_collision(n, seed)
    return p.arrays()


def bench_serving(E, W, ctx, count: int, rho: float, seed: int, cpu, use_graph: bool = True) -> dict:
    """The mixed request stream: continuous batching against one request at a
    time on the single-system path, on the same requests and arrival times.

    1. Each request alone, first in first out, on the device-resident
       single-system path (upload, start diagnostics, refresh, every step, end
       diagnostics, download), timed whole: its service time.
    2. Arrivals: a seeded unit-rate Poisson sequence scaled to the rate
       lambda = rho / mean(service), so the single-system server is busy a
       fraction rho of the time. The FIFO latencies follow from the arrival
       and service times exactly (one server, requests in arrival order).
    3. The same arrivals served in real time by the continuous-batching loop.
    """
    eps, dt, below = 0.01, 1e-3, 4096
    rng = np.random.default_rng(seed)
    sizes = [int(v) for v in np.exp(rng.uniform(np.log(1e3), np.log(2e6), count))]
    plummer = rng.random(count) < 0.7
    steps = [int(v) for v in rng.integers(100, 1001, count)]
    integ = ["yoshida4" if u < 0.5 else "velocity-verlet" for u in rng.random(count)]
    unit_arrivals = np.cumsum(rng.exponential(1.0, count))
    log(f"[bench] serving: generating {count} requests ({sum(sizes)} particles)")
    specs = [("plummer" if plummer[k] else "disc", n, 1000 + k) for k, n in enumerate(sizes)]
    work = [W.Workload(*make_request_state(spec)) for spec in specs]

    # 1. One at a time on the single-system path.
    log(f"[bench] serving: {count} requests one at a time on the single-system path")
    service, diag_s = [], []
    for k in range(count):
        cfg = E.StepConfig(integrator=integ[k], solver="direct" if sizes[k] < below else "barnes-hut",
                           softening=eps, opening_angle=0.5, leaf_capacity=16)
        t0 = time.perf_counter()
        s = E.DeviceSystem(ctx, sizes[k])
        s.upload(*work[k].arrays())
        t1 = time.perf_counter()
        s.diagnostics(eps)
        d0 = time.perf_counter() - t1
        s.refresh(cfg)
        s.step(cfg, dt, steps[k])
        t2 = time.perf_counter()
        s.diagnostics(eps)
        d1 = time.perf_counter() - t2
        s.download()
        s.close()
        service.append(time.perf_counter() - t0)
        diag_s.append(d0 + d1)
        log(f"[bench] serving: alone {k + 1}/{count} N={sizes[k]} {integ[k]} {steps[k]} steps "
            f"{service[-1]:.2f} s")
    service = np.array(service)
    rate = rho / float(service.mean())
    arrivals = unit_arrivals / rate
    free_at, fifo = 0.0, []
    for k in range(count):
        free_at = max(float(arrivals[k]), free_at) + float(service[k])
        fifo.append(free_at - float(arrivals[k]))
    fifo = np.array(fifo)
    big = int(np.argmax(sizes))

    # 3. Continuous batching, real time, same arrivals.
    server = E.Server(ctx, E.ServerConfig(round_ms=50.0, direct_below=below, alone_from=131072,
                                          opening_angle=0.5, leaf_capacity=16, tiles_per_item=16,
                                          use_graph=use_graph))
    log(f"[bench] serving: {count} requests through the continuous-batching loop, "
        f"{rate:.4f} requests/s (rho {rho})")
    ids = {}
    start = server.now()
    nxt = 0
    pending = 0
    last_log = time.perf_counter()
    while nxt < count or pending > 0:
        now = server.now() - start
        while nxt < count and arrivals[nxt] <= now:
            ids[nxt] = server.submit(work[nxt], integ[nxt], eps, dt, steps[nxt])
            nxt += 1
            pending += 1
        if pending == 0:
            time.sleep(max(0.0, arrivals[nxt] - (server.now() - start)))
            continue
        pending = server.round()
        if time.perf_counter() - last_log > 30:
            last_log = time.perf_counter()
            log(f"[bench] serving: {nxt} submitted, {pending} pending at {now:.0f} s")
    makespan = server.now() - start
    st = [server.status(ids[k]) for k in range(count)]
    stats = server.stats()
    server.close()
    latency = np.array([s["completed"] - s["submitted"] for s in st])
    queued = np.array([s["admitted"] - s["submitted"] for s in st])
    drift = [(s["end"]["total"] - s["start"]["total"]) / abs(s["start"]["total"]) for s in st]
    particle_steps = sum(n * k for n, k in zip(sizes, steps))
    large = np.array(sizes) >= 131072

    def pct(v):
        return {"p50": float(np.percentile(v, 50)), "p99": float(np.percentile(v, 99)),
                "max": float(v.max()), "min": float(v.min()), "mean": float(v.mean())}

    # Energy drift against the CPU reference (FP64 direct, same integrator and
    # initial state) for the requests the CPU can redo in a bounded time.
    cpu_drift = None
    if cpu is not None:
        o = cpu.oracle
        budget, spent, rows = 20.0, 0.0, []
        for k in sorted(range(count), key=lambda j: sizes[j] * steps[j]):
            if sizes[k] >= below:
                break
            cost = sizes[k] ** 2 * steps[k] * (3 if integ[k] == "yoshida4" else 1) / 2e10
            if spent + cost > budget:
                break
            t0 = time.perf_counter()
            f = o.make_field("yoshida4" if integ[k] == "yoshida4" else "vv", "direct", eps2=eps * eps,
                             threads=cpu.threads)
            ref = o.Particles(*(a.copy() for a in work[k].arrays()))
            e0 = o.diagnostics(ref, eps * eps)["total"]
            acc = o.refresh(ref, f)
            o.steps(ref, acc, f, dt, steps[k])
            e1 = o.diagnostics(ref, eps * eps)["total"]
            spent += time.perf_counter() - t0
            rows.append({"n": sizes[k], "steps": steps[k], "integrator": integ[k],
                         "device_drift": drift[k], "cpu_drift": (e1 - e0) / abs(e0)})
        cpu_drift = {"requests": rows, "cores": cpu.threads, "seconds": spent,
                     "reference": "oracle FP64 direct sum and integrator, same initial state"}
    return {
        "workload": f"{count} requests, Poisson arrivals at rho {rho} of the single-system path's "
                    f"measured service (lambda {rate:.4f}/s), N log-uniform in [1e3, 2e6], 70/30 "
                    f"Plummer / two-disc collision, 100-1000 steps, Yoshida-4 or velocity Verlet, "
                    f"eps {eps}, dt {dt}; direct below {below} particles, Barnes-Hut (theta 0.5, "
                    f"leaf 16) above; requests of 131,072 particles or more run alone, one at a time",
        "requests": count, "arrival_rate_per_s": rate, "utilisation_target": rho, "seed": seed,
        "large_requests": int(large.sum()),
        "particles_total": int(sum(sizes)), "particle_steps": int(particle_steps),
        "makespan_s": makespan, "requests_per_s": count / makespan,
        "particle_steps_per_s": particle_steps / makespan,
        "latency_s": pct(latency),
        "latency_small_s": pct(latency[~large]) if (~large).any() else None,
        "latency_large_s": pct(latency[large]) if large.any() else None,
        "queue_wait_s": {"p50": float(np.percentile(queued, 50)), "max": float(queued.max())},
        "single_system_fifo": {
            "latency_s": pct(fifo),
            "latency_small_s": pct(fifo[~large]) if (~large).any() else None,
            "latency_large_s": pct(fifo[large]) if large.any() else None,
            "makespan_s": float(max(float(arrivals[k]) + fifo[k] for k in range(count))),
            "busy_s": float(service.sum()),
            "method": "each request alone on the device-resident single-system path, every step "
                      "run and the whole request timed (upload, start diagnostics, refresh, steps, "
                      "end diagnostics, download); first-in-first-out latencies from those "
                      "service times and the same arrival times"},
        "diagnostics": {"largest_request_n": sizes[big],
                        "largest_request_diagnostics_s": diag_s[big],
                        "largest_request_service_s": float(service[big]),
                        "share_of_all_service": float(sum(diag_s) / service.sum()),
                        "note": "start + end FP64 diagnostics (O(N^2) pairwise potential) per request"},
        "graphs": use_graph,
        "rounds": stats["rounds"], "batch_builds": stats["batch_builds"],
        "busy_s": stats["busy_s"], "peak_resident_bytes": stats["peak_resident_bytes"],
        "memory_budget": stats["memory_budget"],
        "energy_drift": {"max_abs_relative": float(np.max(np.abs(drift))),
                         "median_abs_relative": float(np.median(np.abs(drift)))},
        "cpu_reference_drift": cpu_drift,
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--error-targets", type=int, default=1024)
    ap.add_argument("--bh-n", type=int, default=2097152)
    ap.add_argument("--bh-evals", type=int, default=5)
    ap.add_argument("--no-barnes-hut", action="store_true")
    ap.add_argument("--no-plummer-vv", action="store_true",
                    help="skip the N = 4096 velocity Verlet line (CPU path and device)")
    ap.add_argument("--batch-systems", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--alone-steps", type=int, default=3)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=24)
    ap.add_argument("--serve-rho", type=float, default=0.8,
                    help="arrival rate as a fraction of the single-system path's service rate")
    ap.add_argument("--serve-seed", type=int, default=5)
    ap.add_argument("--serve-no-graph", action="store_true",
                    help="served batches step eagerly instead of replaying a captured hipGraph")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-direct", action="store_true", help="skip the headline direct leg")
    ap.add_argument("--n65536-steps", type=int, default=20,
                    help="Yoshida-4 steps of the N = 65,536 direct leg (0 skips it)")
    ap.add_argument("--n524288-steps", type=int, default=2,
                    help="Yoshida-4 steps of the N = 524,288 direct leg (0 skips it)")
    ap.add_argument("--walk-targets", type=int, default=0,
                    help="targets per lane of the Barnes-Hut walk (1 or 2; 0: the engine's default)")
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

    import orrery_amd as E
    from orrery_amd import workloads as W

    if args.probes_only:
        ctx = E.Context(0)
        print(json.dumps({"ceilings": measure_ceilings(ctx)}), flush=True)
        ctx.close()
        return 0

    cpu = None if args.no_cpu_baseline else CpuLeg()
    ctx = E.Context(0)
    if args.walk_targets:
        ctx.set_tree_config(targets_per_lane=args.walk_targets)
    dev = E.describe(0)
    log(f"[bench] device {dev['name']} {dev['arch']} CUs={dev['compute_units']}")
    n, eps, dt = args.n, 0.005, 5e-4
    result = {
        "metric": METRIC, "value": None, "unit": "interactions/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 pairwise terms, fp64 state and update",
        "data": "synthetic Plummer sphere (seeded generator), device-resident",
        "config": {"workload": "direct O(N^2) Plummer, Yoshida-4, one GPU", "n": n, "softening": eps,
                   "timestep": dt, "integrator": "yoshida4", "evaluations_per_step": 3,
                   "seed": args.seed},
    }
    if not args.no_direct:
        log(f"[bench] generating Plummer N={n} seed={args.seed}")
        state = W.plummer(n, args.seed)
        sys_ = E.DeviceSystem(ctx, n)
        sys_.upload(*state.arrays())
        cfg = E.StepConfig(integrator="yoshida4", solver="direct", softening=eps)
        log("[bench] refresh")
        sys_.refresh(cfg)
        acc = sys_.download()
        log(f"[bench] warmup {args.warmup} steps")
        sys_.step(cfg, dt, args.warmup)
        # FP64 energy before and after the timed steps (outside the timed region).
        e_before = sys_.diagnostics(eps)
        ctx.set_timing(True)
        ctx.reset_timings()
        log(f"[bench] timing {args.steps} steps")
        t0 = time.perf_counter()
        sys_.step(cfg, dt, args.steps)           # returns after a stream synchronise
        elapsed = time.perf_counter() - t0
        timings = ctx.timings()
        ctx.set_timing(False)
        e_after = sys_.diagnostics(eps)
        sys_.close()
        energy = {"relative_drift": (e_after["total"] - e_before["total"]) / abs(e_before["total"]),
                  "steps": args.steps, "total_before": e_before["total"],
                  "momentum_after": [float(v) for v in e_after["momentum"]],
                  "note": "FP64 kinetic + pairwise potential over all pairs (softened), before and after "
                          "the timed Yoshida-4 steps"}

        evals = 3 * args.steps
        value = evals * n * (n - 1) / elapsed
        k_ms, k_launches = timings.get("direct_kernel", (0.0, 0))
        kernel_ms = k_ms / max(1, k_launches)
        # Wall time per evaluation (the Morton ordering is replayed as a graph
        # and untimed; the three direct launches carry the phase events).
        eval_ms = elapsed * 1e3 / max(1, evals)
        achieved = FLOP_PER_INTERACTION * n * (n - 1) / (kernel_ms * 1e-3) / 1e12
        # The issue-rate ceiling of the one-sided sequence: one interaction's
        # 12 VALU + 1 v_rsq_f32, no memory, sustained for about 0.1 s. The
        # pairs path spends 16 VALU + 1 v_rsq_f32 on a pair, i.e. on two
        # interactions, so it can pass this ceiling.
        ctx.probe("mix", 256)
        mix, _ = ctx.probe("mix", 65536)
        result.update({
            "value": value,
            "ms_per_step": elapsed * 1e3 / max(1, args.steps),
            "ms_per_evaluation": eval_ms,
            "kernel_ms_per_evaluation": kernel_ms,
            "energy": energy,
            "phases_ms": {k: v[0] for k, v in timings.items()},
            "roofline": {"bound": "fp32-valu", "kernel": "direct_pairs+direct_complement",
                         "achieved": achieved,
                         "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                         "traffic": measured_traffic(n),
                         "issue_ceiling": {"interactions_per_s": mix,
                                           "TFLOPs": FLOP_PER_INTERACTION * mix / 1e12,
                                           "frac": achieved * 1e12 / FLOP_PER_INTERACTION / mix},
                         "note": "20 flop per interaction x N(N-1) per evaluation (the reference's "
                                 "convention) / mean time per evaluation from the fork of "
                                 "direct_complement (second stream) to the join after direct_pairs "
                                 "(HIP events on the engine's stream); direct_pairs evaluates each far pair "
                                 "of tiles once for both tiles (16 VALU + 1 v_rsq_f32 per pair), so "
                                 "this counts interactions delivered, about twice the pair "
                                 "evaluations executed; issue_ceiling = the engine's mix probe, the "
                                 "one-sided sequence (12 VALU + 1 v_rsq_f32 per interaction, no "
                                 "memory, packed FP32 off); traffic = calibrated rocprofv3 "
                                 "FETCH_SIZE + WRITE_SIZE bytes per evaluation "
                                 "(profiles/direct_pairs_traffic.json)"},
        })
        if args.n65536_steps > 0:
            log("[bench] direct at N = 65,536")
            result["direct_n65536"] = direct_rate(E, W, ctx, 65536, args.n65536_steps)
        if args.n524288_steps > 0:
            log("[bench] direct at N = 524,288")
            result["direct_n524288"] = direct_rate(E, W, ctx, 524288, args.n524288_steps, cpu=cpu)
        if cpu is not None:
            log("[bench] CPU leg: FP64 reference forces and the AVX2 baseline")
            rng = np.random.default_rng(12345)
            tgt = np.sort(rng.choice(n, size=min(n, args.error_targets), replace=False))
            result["force_error"] = cpu.fp64_errors(state, eps, (acc["ax"], acc["ay"], acc["az"]), tgt)
            log("[bench] CPU: direct AVX2 baseline at N = 262,144 (sampled targets) and 65,536")
            result["cpu_baseline"] = cpu.direct_baseline(state, eps, targets=8192)
            if "direct_n65536" in result:
                result["direct_n65536"]["cpu_baseline"] = cpu.direct_baseline(W.plummer(65536, 1), eps,
                                                                              targets=65536)
            log("[bench] direct accuracy at N = 65,536 and 262,144, seeds 1-3")
            result["force_error_by_case"] = direct_accuracy(E, W, ctx, cpu)
        else:
            result["force_error"] = None
            result["cpu_baseline"] = None
    if not args.no_plummer_vv:
        log("[bench] Plummer N=4096, velocity Verlet, 100 steps")
        result["plummer4096_vv"] = bench_plummer_vv(E, W, ctx, cpu)
    if not args.no_barnes_hut:
        result["barnes_hut"] = bench_barnes_hut(E, W, ctx, args.bh_n, args.bh_evals,
                                                args.error_targets, cpu)
    if not args.no_batched:
        result["batched"] = bench_batched(E, W, ctx, args.batch_systems, args.batch_steps,
                                          args.alone_steps, cpu)
    if not args.no_serving:
        result["served"] = bench_serving(E, W, ctx, args.serve_requests, args.serve_rho,
                                         args.serve_seed, cpu, use_graph=not args.serve_no_graph)
    log("[bench] ceilings")
    result["ceilings"] = measure_ceilings(ctx)
    ctx.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
