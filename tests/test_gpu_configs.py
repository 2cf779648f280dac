"""The batched and served configurations at their configured workloads.

* Batched (512 Plummer spheres, N log-uniform in [64, 16384] with the
  benchmark's seed, eps = 0.02, direct below 4096 particles, Barnes-Hut
  theta = 0.6 above, velocity Verlet dt = 1e-3, autotuned tile and leaf sizes,
  hipGraph steps): every direct system against the FP64 oracle, every tree
  system bit for bit against the single-system HipTreeSolver at the tuned leaf
  size, and eight systems spread over the size range equal, bit for bit, to
  the same system stepped in a batch of its own.
* Served (continuous batching): a Barnes-Hut request of 1,048,576 particles
  (a two-disc collision, the serving stream's upper size class) among small
  requests of both integrators; each result equals, bit for bit, the same
  request served alone.
"""
import numpy as np
import pytest

import oracle
import orrery_amd as E
from orrery_amd import workloads as W

pytestmark = pytest.mark.gpu

EPS, DT, THETA, BELOW, SEED, COUNT = 0.02, 1e-3, 0.6, 4096, 4, 512


def batch_of(ctx, systems, spheres, cfg):
    b = E.DeviceBatch(ctx, [E.BatchSystem(s.n, EPS, DT) for s in spheres], cfg)
    for k, s in enumerate(spheres):
        b.upload(k, *s.arrays())
    return b


def test_config4_512_spheres(ctx):
    ns = W.batched_sizes(COUNT, 64, 16384, SEED)
    spheres = W.batched_spheres(COUNT, 64, 16384, SEED)
    assert [s.n for s in spheres] == ns
    cfg = E.BatchConfig(integrator="velocity-verlet", opening_angle=THETA, leaf_capacity=0,
                        tiles_per_item=0, direct_below=BELOW, use_graph=True)
    b = batch_of(ctx, ns, spheres, cfg)
    info = b.autotune()
    tiles, leaf = info["tiles_per_item"], info["leaf_capacity"]
    print(f"autotuned tiles_per_item={tiles} leaf_capacity={leaf}; "
          f"{info['direct_systems']} direct, {info['tree_systems']} tree systems")
    assert info["direct_systems"] + info["tree_systems"] == COUNT
    b.refresh()

    # Direct systems against the FP64 compensated sum, every particle.
    worst_all = 0.0
    for k, n in enumerate(ns):
        if n >= BELOW:
            continue
        got = b.download(k)
        p = spheres[k]
        ref = oracle.reference_acceleration(p.x, p.y, p.z, p.m, EPS * EPS)
        worst, rms = oracle.error_summary(oracle.relative_errors(got["ax"], got["ay"], got["az"], *ref))
        worst_all = max(worst_all, worst)
        assert rms <= 1e-6 and worst <= 1e-6, (k, n, rms, worst)
    print(f"direct systems: worst relative error {worst_all:.3e}")

    # Tree systems bit for bit against the single-system solver.
    solver = E.HipTreeSolver(E.TreeParameters(THETA, leaf), E.Softening(EPS), context=ctx)
    for k, n in enumerate(ns):
        if n < BELOW:
            continue
        p = spheres[k]
        acc = (np.zeros(n), np.zeros(n), np.zeros(n))
        solver.reset_interaction_count()
        solver.evaluate((p.x, p.y, p.z), p.m, acc)
        got = b.download(k)
        for c, key in enumerate(("ax", "ay", "az")):
            assert np.array_equal(got[key], acc[c]), (k, n, key)
        cnt = solver.interaction_count()
        assert b.counts(k) == (cnt.particle_particle, cnt.particle_cell), (k, n)

    # Alone equals mix: eight systems at the size range's quantiles, three
    # graph-replayed steps in the 512-system batch and alone.
    b.step(3)
    order = np.argsort(ns)
    picks = sorted({int(order[int(q * (COUNT - 1))]) for q in np.linspace(0, 1, 8)})
    fixed = E.BatchConfig(integrator="velocity-verlet", opening_angle=THETA, leaf_capacity=leaf,
                          tiles_per_item=tiles, direct_below=BELOW, use_graph=True)
    for k in picks:
        alone = batch_of(ctx, [ns[k]], [spheres[k]], fixed)
        alone.refresh()
        alone.step(3)
        a, m = alone.download(0), b.download(k)
        for key in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
            assert np.array_equal(a[key], m[key]), (k, ns[k], key)
        alone.close()
    b.close()


def test_config5_large_request_among_small(ctx):
    reqs = [
        (W.disc_collision(1048576, 31), "yoshida4", 2),
        (W.plummer(1000, 32), "velocity-verlet", 40),
        (W.plummer(5000, 33), "yoshida4", 12),
        (W.disc_collision(20000, 34), "velocity-verlet", 9),
        (W.plummer(150000, 35), "velocity-verlet", 3),
        (W.plummer(3000, 36), "yoshida4", 20),
    ]
    cfg = E.ServerConfig(round_ms=20.0, direct_below=BELOW, alone_from=131072, opening_angle=0.5,
                         leaf_capacity=16, tiles_per_item=16)

    def serve(batch):
        s = E.Server(ctx, cfg)
        ids = [s.submit(w, integ, 0.01, 1e-3, steps) for w, integ, steps in batch[:2]]
        s.round()
        ids += [s.submit(w, integ, 0.01, 1e-3, steps) for w, integ, steps in batch[2:]]
        s.drain()
        out = [(s.result(i), s.status(i)) for i in ids]
        s.close()
        return out

    mixed = serve(reqs)
    for (w, integ, steps), (got, st) in zip(reqs, mixed):
        assert st["state"] == "done" and st["steps_done"] == steps
        (alone, st1), = serve([(w, integ, steps)])
        for key in ("x", "y", "z", "vx", "vy", "vz"):
            assert np.array_equal(got[key], alone[key]), (w.n, integ, key)
        assert st["end"]["total"] == st1["end"]["total"]
