"""The serving loop (include/orrery_hip.h orr_server_*).

Contract: requests that arrive while others run are admitted as memory frees
up, run concurrently in batches, and each result is the same alone as in the
mix (bit for bit: a system's arithmetic does not depend on its batch or on how
its steps are split into rounds). Energy drift is compared with the oracle's
FP64 integrator on the same initial state.
"""
import numpy as np
import pytest

import oracle
import orrery_amd as E
from orrery_amd import workloads as W

pytestmark = pytest.mark.gpu

EPS = 0.01


def requests():
    # (workload, integrator, steps): direct and Barnes-Hut sizes, both integrators.
    return [
        (W.plummer(1000, 1), "velocity-verlet", 30),
        (W.plummer(5000, 2), "yoshida4", 12),
        (W.disc_collision(4000, 3), "velocity-verlet", 25),
        (W.plummer(2500, 4), "yoshida4", 9),
        (W.disc_collision(9000, 5), "velocity-verlet", 17),
        (W.plummer(1200, 6), "velocity-verlet", 40),
    ]


def config(**kw):
    base = dict(round_ms=3.0, direct_below=4096, alone_from=8192, opening_angle=0.5,
                leaf_capacity=16, tiles_per_item=16)
    base.update(kw)
    return E.ServerConfig(**base)


def run_alone(ctx, state, integrator, steps):
    s = E.Server(ctx, config())
    rid = s.submit(state, integrator, EPS, 1e-3, steps)
    s.drain()
    out = s.result(rid)
    st = s.status(rid)
    s.close()
    return out, st


def test_alone_equals_mix_bitwise(ctx):
    reqs = requests()
    s = E.Server(ctx, config())
    ids = [s.submit(w, integ, EPS, 1e-3, steps) for w, integ, steps in reqs[:3]]
    s.round()
    s.round()
    # Arrivals while the first three are on the device.
    ids += [s.submit(w, integ, EPS, 1e-3, steps) for w, integ, steps in reqs[3:]]
    s.drain()
    stats = s.stats()
    assert stats["completed"] == len(reqs) and stats["active"] == 0 and stats["queued"] == 0
    assert stats["batch_builds"] >= 3 and stats["rounds"] >= 4
    mixed = [(s.result(i), s.status(i)) for i in ids]
    s.close()
    for (w, integ, steps), (got, st) in zip(reqs, mixed):
        assert st["state"] == "done" and st["steps_done"] == steps
        alone, st1 = run_alone(ctx, w, integ, steps)
        for key in ("x", "y", "z", "vx", "vy", "vz"):
            assert np.array_equal(got[key], alone[key]), (w.n, integ, key)
        assert st["end"]["total"] == st1["end"]["total"]


def test_matches_the_fp64_integrator_and_its_energy_drift(ctx):
    p = W.plummer(1500, 11)
    steps, dt = 60, 1e-3
    s = E.Server(ctx, config())
    rid = s.submit(p, "velocity-verlet", EPS, dt, steps)
    s.drain()
    got, st = s.result(rid), s.status(rid)
    s.close()
    f = oracle.make_field("vv", "direct", eps2=EPS * EPS)
    ref = oracle.Particles(*(a.copy() for a in p.arrays()))
    e0 = oracle.diagnostics(ref, EPS * EPS)["total"]
    acc = oracle.refresh(ref, f)
    oracle.steps(ref, acc, f, dt, steps)
    e1 = oracle.diagnostics(ref, EPS * EPS)["total"]
    assert st["start"]["total"] == pytest.approx(e0, rel=1e-14)
    assert np.max(np.abs(got["x"] - ref.x)) < 1e-9
    drift_gpu = (st["end"]["total"] - st["start"]["total"]) / abs(st["start"]["total"])
    drift_cpu = (e1 - e0) / abs(e0)
    assert abs(drift_gpu - drift_cpu) <= 0.05 * abs(drift_cpu) + 1e-10, (drift_gpu, drift_cpu)


def test_memory_budget_bounds_admission(ctx):
    reqs = requests()
    # Each request's charge (the batch bound of its system plus a rebuild
    # margin), read off a server with room for all; then a budget with room
    # for about two of them at a time.
    probe = E.Server(ctx, config())
    charges = [probe.status(probe.submit(w, integ, EPS, 1e-3, steps))["charged_bytes"]
               for w, integ, steps in reqs]
    probe.close()
    for (w, _, _), c in zip(reqs, charges):
        bound = E.DeviceBatch.bytes_bound([E.BatchSystem(w.n, EPS, 1e-3)],
                                          E.BatchConfig(opening_angle=0.5, leaf_capacity=16,
                                                        tiles_per_item=16, direct_below=4096))
        assert c > bound
    budget = sorted(charges)[-1] + sorted(charges)[-2]
    s = E.Server(ctx, config(memory_budget=budget))
    ids = [s.submit(w, integ, EPS, 1e-3, steps) for w, integ, steps in reqs]
    s.round()
    st = s.stats()
    assert st["queued"] > 0 and st["resident_bytes"] <= budget
    s.drain()
    st = s.stats()
    assert st["completed"] == len(reqs) and st["peak_resident_bytes"] <= budget
    assert st["resident_bytes"] == 0
    times = [s.status(i) for i in ids]
    assert all(t["state"] == "done" and t["admitted"] >= t["submitted"] for t in times)
    # Some request waited for memory to free up.
    assert max(t["admitted"] - t["submitted"] for t in times) > 0
    s.close()


def test_invalid_requests(ctx):
    s = E.Server(ctx, config(memory_budget=50 << 20))
    w = W.plummer(100, 1)
    with pytest.raises(E.solvers.C.EngineError) as e:
        s.submit(w, "velocity-verlet", EPS, 1e-3, 0)
    assert e.value.status == 1
    ragged = W.Workload(w.x[:50], w.y, w.z, w.vx, w.vy, w.vz, w.m)
    with pytest.raises(ValueError):
        s.submit(ragged, "velocity-verlet", EPS, 1e-3, 5)
    big = W.plummer(200000, 1)
    with pytest.raises(E.solvers.C.EngineError) as e:
        s.submit(big, "velocity-verlet", EPS, 1e-3, 5)
    assert e.value.status == 5
    rid = s.submit(w, "yoshida4", EPS, 1e-3, 3)
    with pytest.raises(E.solvers.C.EngineError):
        s.result(rid)
    s.drain()
    assert s.status(rid)["state"] == "done"
    s.result(rid)
    with pytest.raises(E.solvers.C.EngineError):
        s.result(rid)
    s.close()


def test_failed_rebuild_keeps_requests_and_is_retried(ctx):
    # The first two batch rebuilds fail after their uploads and copies (as an
    # allocation or copy failure would). The round reports the failure; the
    # requests and their host copies stay, later rounds rebuild, and every
    # result equals the undisturbed run's bit for bit.
    reqs = requests()[:4]
    clean = E.Server(ctx, config())
    ids = [clean.submit(w, integ, EPS, 1e-3, steps) for w, integ, steps in reqs]
    clean.drain()
    want = [clean.result(i) for i in ids]
    clean.close()

    s = E.Server(ctx, config(fail_rebuilds=2))
    ids = [s.submit(w, integ, EPS, 1e-3, steps) for w, integ, steps in reqs[:2]]
    failures = 0
    for _ in range(2):
        with pytest.raises(E.solvers.C.EngineError) as e:
            s.round()
        assert "injected" in str(e.value)
        failures += 1
    # Arrivals after the failures join the retried batches.
    ids += [s.submit(w, integ, EPS, 1e-3, steps) for w, integ, steps in reqs[2:]]
    s.drain()
    assert failures == 2
    for rid, (w, integ, steps), ref in zip(ids, reqs, want):
        st = s.status(rid)
        assert st["state"] == "done" and st["steps_done"] == steps
        got = s.result(rid)
        for key in ("x", "y", "z", "vx", "vy", "vz"):
            assert np.array_equal(got[key], ref[key]), (w.n, integ, key)
    assert s.stats()["resident_bytes"] == 0
    s.close()


def test_large_requests_run_one_at_a_time_in_arrival_order(ctx):
    # Requests of alone_from particles or more are admitted one at a time, in
    # arrival order: the second large request is admitted only after the first
    # has finished, while a small request submitted after both is admitted at
    # once and runs beside the first. Each result still equals the request
    # served alone.
    reqs = [
        (W.plummer(9000, 41), "velocity-verlet", 30),
        (W.disc_collision(10000, 42), "yoshida4", 12),
        (W.plummer(1500, 43), "velocity-verlet", 5),
    ]
    s = E.Server(ctx, config())
    ids = [s.submit(w, integ, EPS, 1e-3, steps) for w, integ, steps in reqs]
    s.round()
    st = [s.status(i) for i in ids]
    assert [x["state"] for x in st] == ["running", "queued", "running"]
    s.drain()
    st = [s.status(i) for i in ids]
    assert all(x["state"] == "done" for x in st)
    assert st[1]["admitted"] >= st[0]["completed"]
    assert st[2]["admitted"] < st[0]["completed"]
    got = [s.result(i) for i in ids]
    s.close()
    for (w, integ, steps), g in zip(reqs, got):
        alone, _ = run_alone(ctx, w, integ, steps)
        for key in ("x", "y", "z", "vx", "vy", "vz"):
            assert np.array_equal(g[key], alone[key]), (w.n, integ, key)
