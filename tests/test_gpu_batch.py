"""Batches of independent systems (include/orrery_hip.h orr_batch_*).

The reference steps one system at a time (sim/simulation.cpp:26-64), so the
oracle for a batch is the oracle for each of its systems: FP64 direct forces
per system, the FP64 integrator on the same seed, and -- the batching
contract -- that every system gives the same bits in a batch of its own as in
a mix of others.
"""
import numpy as np
import pytest

import oracle
import orrery_amd as E

pytestmark = pytest.mark.gpu

EPS = 0.02


def sizes(count, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return [int(v) for v in np.exp(rng.uniform(np.log(lo), np.log(hi), count))]


def make_batch(ctx, ns, config, seeds=None, eps=EPS, dt=1e-3, solver=-1):
    systems = [E.BatchSystem(n, eps, dt, solver=solver) for n in ns]
    b = E.DeviceBatch(ctx, systems, config)
    states = []
    for k, n in enumerate(ns):
        p = oracle.plummer(n, (seeds or {}).get(k, 100 + k)) if n > 0 else None
        if p is not None:
            b.upload(k, *p.arrays())
        states.append(p)
    return b, states


def test_direct_systems_match_fp64(ctx):
    ns = [1, 2, 63, 64, 65, 200, 0, 1000, 3000, 4095, 129]
    b, states = make_batch(ctx, ns, E.BatchConfig(use_graph=False))
    info = b.info()
    assert info["direct_systems"] == len(ns) and info["tree_systems"] == 0
    assert info["direct_pairs"] == sum(n * (n - 1) for n in ns)
    b.refresh()
    for k, n in enumerate(ns):
        if n == 0:
            continue
        got = b.download(k)
        p = states[k]
        ref = oracle.reference_acceleration(p.x, p.y, p.z, p.m, EPS * EPS)
        if n == 1:
            assert got["ax"][0] == 0 and got["ay"][0] == 0 and got["az"][0] == 0
            continue
        worst, rms = oracle.error_summary(oracle.relative_errors(got["ax"], got["ay"], got["az"], *ref))
        assert rms < 1e-6 and worst < 2e-5, (n, rms, worst)
        assert b.counts(k) == (n * (n - 1), 0)
    b.close()


def test_alone_equals_mix_bitwise(ctx):
    # Direct below 4096 particles, Barnes-Hut above (the batched-spheres setup),
    # fixed tile and leaf sizes; five graph-replayed velocity Verlet steps.
    ns = sizes(24, 64, 16384, 7)
    ns[3], ns[5] = 5000, 4096
    cfg = E.BatchConfig(opening_angle=0.6, leaf_capacity=16, tiles_per_item=8, use_graph=True)
    mix, _ = make_batch(ctx, ns, cfg)
    info = mix.info()
    assert info["tree_systems"] >= 2 and info["direct_systems"] >= 2
    mix.refresh()
    mix.step(5)
    picks = [k for k, n in enumerate(ns) if n < 4096][:3] + [k for k, n in enumerate(ns) if n >= 4096][:3]
    for k in picks:
        alone, _ = make_batch(ctx, [ns[k]], cfg, seeds={0: 100 + k})
        alone.refresh()
        alone.step(5)
        a, m = alone.download(0), mix.download(k)
        for key in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
            assert np.array_equal(a[key], m[key]), (k, ns[k], key)
        assert alone.counts(0) == mix.counts(k)
        alone.close()
    mix.close()


def test_tree_systems_match_single_solver_bitwise(ctx):
    ns = [5000, 300, 9000, 4100, 70, 12000]
    cfg = E.BatchConfig(opening_angle=0.5, leaf_capacity=16, quadrupole=True, use_graph=False)
    b, states = make_batch(ctx, ns, cfg)
    b.refresh()
    solver = E.HipTreeSolver(E.TreeParameters(0.5, 16, quadrupole=True), E.Softening(EPS), context=ctx)
    for k, n in enumerate(ns):
        if n < 4096:
            continue
        p = states[k]
        acc = (np.zeros(n), np.zeros(n), np.zeros(n))
        solver.reset_interaction_count()
        solver.evaluate((p.x, p.y, p.z), p.m, acc)
        got = b.download(k)
        assert np.array_equal(got["ax"], acc[0]) and np.array_equal(got["ay"], acc[1])
        assert np.array_equal(got["az"], acc[2])
        c = solver.interaction_count()
        assert b.counts(k) == (c.particle_particle, c.particle_cell)
    b.close()


def test_graph_replays_match_eager(ctx):
    # Yoshida-4 (three substeps per captured step); eager work between replays.
    ns = [700, 5000, 64, 2500, 6100]
    runs = {}
    for graph in (False, True):
        cfg = E.BatchConfig(integrator="yoshida4", opening_angle=0.6, leaf_capacity=16,
                            tiles_per_item=16, use_graph=graph)
        b, _ = make_batch(ctx, ns, cfg)
        b.refresh()
        frames = []
        for _ in range(4):
            b.step(2)
            frames.append([b.download(k) for k in range(len(ns))])
            b.diagnostics(1)
        assert b.info()["graph"] == (1 if graph else 0)
        runs[graph] = frames
        b.close()
    for f_eager, f_graph in zip(runs[False], runs[True]):
        for a, g in zip(f_eager, f_graph):
            for key in ("x", "vx", "ax"):
                assert np.array_equal(a[key], g[key])


@pytest.mark.parametrize("integrator", ["velocity-verlet", "yoshida4"])
def test_batch_steps_match_cpu_reference_path(ctx, integrator):
    ns = [2048, 512, 1500]
    eps, dt, steps = 0.01, 1e-3, 10
    b, states = make_batch(ctx, ns, E.BatchConfig(integrator=integrator), eps=eps, dt=dt)
    b.refresh()
    b.step(steps)
    for k, n in enumerate(ns):
        f = oracle.make_field(integrator="vv" if integrator == "velocity-verlet" else "yoshida4",
                              solver="direct", eps2=eps * eps)
        ref = oracle.Particles(*(a.copy() for a in states[k].arrays()))
        acc = oracle.refresh(ref, f)
        oracle.steps(ref, acc, f, dt, steps)
        st = b.download(k)
        assert np.max(np.abs(st["x"] - ref.x)) < 1e-9, k
        assert np.max(np.abs(st["vx"] - ref.vx)) < 1e-7, k
    b.close()


def test_per_system_timestep_and_diagnostics(ctx):
    ns = [800, 800]
    systems = [E.BatchSystem(800, 0.01, 1e-3), E.BatchSystem(800, 0.05, 5e-4)]
    b = E.DeviceBatch(ctx, systems, E.BatchConfig())
    p = oracle.plummer(800, 3)
    for k in range(2):
        b.upload(k, *p.arrays())
    for k, (eps, dt) in enumerate([(0.01, 1e-3), (0.05, 5e-4)]):
        d = b.diagnostics(k)
        r = oracle.diagnostics(p, eps * eps)
        assert abs(d["potential"] - r["potential"]) <= 1e-14 * abs(r["potential"])
    b.refresh()
    b.step(6)
    for k, (eps, dt) in enumerate([(0.01, 1e-3), (0.05, 5e-4)]):
        f = oracle.make_field(integrator="vv", solver="direct", eps2=eps * eps)
        ref = oracle.Particles(*(a.copy() for a in p.arrays()))
        acc = oracle.refresh(ref, f)
        oracle.steps(ref, acc, f, dt, 6)
        st = b.download(k)
        assert np.max(np.abs(st["x"] - ref.x)) < 1e-9, k
    b.close()


def test_copy_system_continues_identically(ctx):
    ns = [300, 5000, 1200]
    cfg = E.BatchConfig(leaf_capacity=16, tiles_per_item=16)
    a, _ = make_batch(ctx, ns, cfg)
    a.refresh()
    a.step(3)
    b = E.DeviceBatch(ctx, [E.BatchSystem(n, EPS, 1e-3) for n in [ns[2], 50, ns[1]]], cfg)
    b.copy_system(0, a, 2)
    b.copy_system(2, a, 1)
    b.upload(1, *oracle.plummer(50, 5).arrays())
    a.step(3)
    b.step(3)
    for src, dst in ((2, 0), (1, 2)):
        x, y = a.download(src), b.download(dst)
        assert np.array_equal(x["x"], y["x"]) and np.array_equal(x["vz"], y["vz"])
    a.close()
    b.close()


def test_autotune_picks_measured_sizes(ctx):
    ns = sizes(40, 64, 16384, 11)
    cfg = E.BatchConfig(opening_angle=0.6, leaf_capacity=0, tiles_per_item=0)
    b, states = make_batch(ctx, ns, cfg)
    info = b.autotune()
    assert info["tiles_per_item"] in (4, 8, 16, 32, 64)
    assert info["leaf_capacity"] in (8, 12, 16, 24, 32)
    k = min(range(len(ns)), key=lambda j: abs(ns[j] - 2000))
    p = states[k]
    got = b.download(k)
    ref = oracle.reference_acceleration(p.x, p.y, p.z, p.m, EPS * EPS)
    worst, rms = oracle.error_summary(oracle.relative_errors(got["ax"], got["ay"], got["az"], *ref))
    assert rms < 1e-6, rms
    b.close()


def test_invalid_batches_are_rejected(ctx):
    with pytest.raises(E.solvers.C.EngineError) as e:
        E.DeviceBatch(ctx, [E.BatchSystem(5000, 0.01, 1e-3), E.BatchSystem(6000, 0.02, 1e-3)],
                      E.BatchConfig())
    assert e.value.status == 1
    with pytest.raises(E.solvers.C.EngineError):
        E.DeviceBatch(ctx, [E.BatchSystem(10, 0.01, 1e-3, solver=3)], E.BatchConfig())
    b = E.DeviceBatch(ctx, [], E.BatchConfig())
    b.refresh()
    b.step(2)
    assert b.info()["particles"] == 0
    b.close()


def test_uploads_of_the_wrong_length_are_rejected(ctx):
    # The C ABI copies each system's own count of values from the pointer it is
    # given, so the host side must refuse shorter (over-read) or longer arrays.
    b = E.DeviceBatch(ctx, [E.BatchSystem(100, EPS, 1e-3), E.BatchSystem(5000, EPS, 1e-3)],
                      E.BatchConfig())
    small, large = oracle.plummer(100, 1), oracle.plummer(5000, 2)
    with pytest.raises(ValueError):
        b.upload(1, *small.arrays())
    with pytest.raises(ValueError):
        b.upload(0, *large.arrays())
    with pytest.raises(IndexError):
        b.upload(2, *small.arrays())
    b.close()
    s = E.DeviceSystem(ctx, 100)
    with pytest.raises(ValueError):
        s.upload(*large.arrays())
    s.close()
    with pytest.raises(ValueError):
        E.solvers.morton_order(ctx, np.zeros(10), np.zeros(9), np.zeros(10))


def test_system_order_does_not_change_results(ctx):
    # Every system is its own tree or direct sum: the same systems in another
    # order (sorted by size, as a caller may well hold them) give identical bits.
    ns = sizes(24, 64, 9000, 11)
    config = E.BatchConfig(opening_angle=0.6, leaf_capacity=16, tiles_per_item=8, direct_below=3000)
    order = sorted(range(len(ns)), key=lambda k: -ns[k])
    a, _ = make_batch(ctx, ns, config)
    b, _ = make_batch(ctx, [ns[k] for k in order], config, seeds={j: 100 + k for j, k in enumerate(order)})
    a.refresh()
    b.refresh()
    for j, k in enumerate(order):
        da, db = a.download(k), b.download(j)
        for key in ("ax", "ay", "az"):
            assert np.array_equal(da[key], db[key]), (k, key)
        if ns[k] >= 3000:
            assert a.counts(k) == b.counts(j)
    a.close()
    b.close()


def test_sort_forms_agree_bitwise(ctx):
    # 300 tree systems: the segmented Morton sort orders by code, then by two
    # digits of the system number. The one-sweep passes and the multi-kernel
    # passes give the same permutation, hence the same trees and bits.
    ns = sizes(300, 16, 2000, 21)
    config = E.BatchConfig(opening_angle=0.6, leaf_capacity=16, use_graph=False)
    results = []
    try:
        for mode in (1, 2):
            ctx.set_tree_config(sort_mode=mode)
            b, _ = make_batch(ctx, ns, config, solver=1)
            assert b.info()["tree_systems"] == 300
            b.refresh()
            results.append(([b.download(k) for k in range(len(ns))], [b.counts(k) for k in range(len(ns))]))
            b.close()
    finally:
        ctx.set_tree_config(sort_mode=1)
    (a, ca), (b, cb) = results
    for k in range(len(ns)):
        for key in ("ax", "ay", "az"):
            assert np.array_equal(a[k][key], b[k][key]), (k, key)
    assert ca == cb


def test_edge_systems_in_one_batch(ctx):
    # Zero softening (masked self terms), tiny systems forced onto the tree,
    # systems smaller than a leaf, and one direct system alone in its branch.
    ns = [1, 2, 7, 40, 3000]
    systems = [E.BatchSystem(1, 0.0, 1e-3, solver=0),      # single direct particle, eps = 0
               E.BatchSystem(2, 0.01, 1e-3, solver=1),     # two bodies on the tree
               E.BatchSystem(7, 0.01, 1e-3, solver=1),     # fewer than a leaf
               E.BatchSystem(40, 0.01, 1e-3, solver=1),
               E.BatchSystem(3000, 0.01, 1e-3, solver=1)]
    b = E.DeviceBatch(ctx, systems, E.BatchConfig(opening_angle=0.5, leaf_capacity=16))
    info = b.info()
    assert info["direct_systems"] == 1 and info["tree_systems"] == 4
    states = []
    for k, n in enumerate(ns):
        p = oracle.plummer(n, 50 + k) if n > 1 else oracle.Particles(*(np.array([v]) for v in (0.3, -0.2, 0.1, 0, 0, 0, 1.0)))
        b.upload(k, *p.arrays())
        states.append(p)
    b.refresh()
    got = b.download(0)
    assert got["ax"][0] == 0 and got["ay"][0] == 0 and got["az"][0] == 0
    for k in (1, 2):
        # Below a leaf's capacity the tree is one leaf: every pair directly.
        p = states[k]
        ref = oracle.reference_acceleration(p.x, p.y, p.z, p.m, 1e-4)
        g = b.download(k)
        worst, rms = oracle.error_summary(oracle.relative_errors(g["ax"], g["ay"], g["az"], *ref))
        assert worst < 1e-5, (ns[k], worst)
    solver = E.HipTreeSolver(E.TreeParameters(0.5, 16), E.Softening(0.01), context=ctx)
    for k in (3, 4):
        p = states[k]
        acc = (np.zeros(p.n), np.zeros(p.n), np.zeros(p.n))
        solver.evaluate((p.x, p.y, p.z), p.m, acc)
        g = b.download(k)
        assert np.array_equal(g["ax"], acc[0]) and np.array_equal(g["az"], acc[2]), ns[k]
    b.step(3)
    assert np.all(np.isfinite(b.download(2)["x"]))
    b.close()


@pytest.mark.parametrize("quadrupole,leaf,items", [(False, 16, 16), (True, 0, 0), (True, 8, 4)])
def test_device_bytes_within_bound(ctx, quadrupole, leaf, items):
    # orr_batch_bytes_bound is what the serving loop charges its budget with:
    # after refreshes, graph steps, autotuning and diagnostics the batch holds
    # no more device memory than it.
    ns = [300, 4095, 1, 5000, 20000, 90000, 2600, 64, 7000]
    cfg = E.BatchConfig(integrator="yoshida4", opening_angle=0.5, leaf_capacity=leaf,
                        quadrupole=quadrupole, tiles_per_item=items, use_graph=True)
    systems = [E.BatchSystem(n, EPS, 1e-3) for n in ns]
    bound = E.DeviceBatch.bytes_bound(systems, cfg)
    b, _ = make_batch(ctx, ns, cfg)
    b.refresh()
    if leaf == 0 or items == 0:
        b.autotune()
    b.step(3)
    for k in (0, 5, 6):
        b.diagnostics(k)
    held = b.info()["device_bytes"]
    print(f"held {held / 1e6:.1f} MB, bound {bound / 1e6:.1f} MB")
    assert 0 < held <= bound, (held, bound)
    # Bounds of disjoint sets of systems add.
    parts = sum(E.DeviceBatch.bytes_bound([s], cfg) for s in systems)
    assert bound <= parts
    b.close()
