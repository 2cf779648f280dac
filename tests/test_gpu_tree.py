"""Barnes-Hut on the device against the oracle tree and the FP64 direct sum.

Mirrors tests/solvers/octree_test.cpp, barnes_hut_test.cpp, tree_walk_test.cpp
and cuda_tree_solver_test.cpp:
* structure parity: node count, leaf count, depth and the next / first / count
  arrays equal the reference tree's (the oracle is pinned bit for bit to the
  reference library), and masses, centres of mass and FP64 acceptance radii are
  equal bit for bit (the device sums in the reference's order without
  contraction);
* interaction counts equal, exactly, those of the oracle's mirror of the
  device's FP32 acceptance test on the same tree (the reference requires exact
  count equality only between walks with identical arithmetic,
  cuda_tree_solver_test.cpp:161-222);
* theta = 0 is the direct sum with pp = N(N-1), pc = 0 (barnes_hut_test.cpp:111-152);
* accuracy against the FP64 reference: RMS < 1e-2, worst < 1e-1 at theta = 0.5
  (cuda_tree_solver_test.cpp:392-412), error monotone in theta.
"""
import numpy as np
import pytest

import oracle
import orrery_amd as E

pytestmark = pytest.mark.gpu


def tree_eval(ctx, p, eps, theta, cap):
    s = E.HipTreeSolver(E.TreeParameters(theta, cap), E.Softening(eps), context=ctx)
    acc = [np.full(p.n, np.nan) for _ in range(3)]
    s.evaluate((p.x, p.y, p.z), p.m, acc)
    return acc, s


def oracle_tree(p, theta, cap):
    codes, index, origin, size = oracle.morton_order(p.x, p.y, p.z)
    i = index.astype(np.int64)
    return oracle.tree_build(p.x[i], p.y[i], p.z[i], p.m[i], codes, origin, size, theta, cap)


@pytest.mark.parametrize("gen,n,theta,cap", [("plummer", 5000, 0.5, 16), ("plummer", 3000, 0.6, 1),
                                             ("plummer", 20000, 0.5, 32), ("disc", 30000, 0.5, 16),
                                             ("plummer", 7, 0.5, 32), ("plummer", 2000, 0.0, 8)])
def test_structure_and_moments_bitwise(ctx, gen, n, theta, cap):
    p = oracle.plummer(n, 20260811) if gen == "plummer" else oracle.disc_collision(n, 2)
    _, s = tree_eval(ctx, p, 0.05, theta, cap)
    t = oracle_tree(p, theta, cap)
    got = s.nodes()
    summary = s.tree_summary()
    assert summary["node_count"] == len(t.mass)
    assert summary["leaf_count"] == t.leaves and summary["depth"] == t.depth
    assert np.array_equal(got["next"], t.next)
    assert np.array_equal(got["first"], t.first) and np.array_equal(got["count"], t.count)
    assert np.array_equal(got["mass"], t.mass)
    assert np.array_equal(got["com"], t.com)
    assert np.array_equal(got["acc_r2"], t.acc_r2)


@pytest.mark.parametrize("gen,n,cap", [("plummer", 5000, 16), ("disc", 20000, 16), ("plummer", 3000, 1)])
def test_quadrupole_moments_bitwise(ctx, gen, n, cap):
    # octree.cpp:75-84, 120-135, 165-184: leaf displaced moments, parallel-axis gather.
    p = oracle.plummer(n, 20260811) if gen == "plummer" else oracle.disc_collision(n, 4)
    s = E.HipTreeSolver(E.TreeParameters(0.5, cap, quadrupole=True), E.Softening(0.05), context=ctx)
    acc = [np.zeros(p.n) for _ in range(3)]
    s.evaluate((p.x, p.y, p.z), p.m, acc)
    codes, index, origin, size = oracle.morton_order(p.x, p.y, p.z)
    i = index.astype(np.int64)
    t = oracle.tree_build(p.x[i], p.y[i], p.z[i], p.m[i], codes, origin, size, 0.5, cap, quadrupole=True)
    assert np.array_equal(s.quadrupoles(), t.quad)


def test_quadrupole_accuracy_reference_setup(ctx):
    # docs/performance/tree_accuracy.csv setup: Plummer standard units, N = 16384,
    # eps = 0.05, leaf 32; the reference's quadrupole walk: RMS 9.76e-4 at theta 0.5.
    p = oracle.plummer(16384, 20260810, scale_radius=3 * np.pi / 16, cutoff=0.999)
    eps = 0.05
    ref = oracle.reference_acceleration(p.x, p.y, p.z, p.m, eps * eps)
    out = {}
    for quad in (False, True):
        s = E.HipTreeSolver(E.TreeParameters(0.5, 32, quadrupole=quad), E.Softening(eps), context=ctx)
        acc = [np.zeros(p.n) for _ in range(3)]
        s.evaluate((p.x, p.y, p.z), p.m, acc)
        out[quad] = oracle.error_summary(oracle.relative_errors(*acc, *ref))
        print(f"quadrupole={quad}: worst={out[quad][0]:.3e} rms={out[quad][1]:.3e}")
    assert out[True][1] < 1e-3 and out[True][1] < 0.7 * out[False][1]
    assert out[False][1] < 2.5e-3


def test_identical_points_stop_at_depth_21(ctx):
    n = 40
    x = np.full(n, 0.25)
    p = oracle.Particles(x, x.copy(), x.copy(), np.zeros(n), np.zeros(n), np.zeros(n), np.full(n, 1.0 / n))
    (ax, ay, az), s = tree_eval(ctx, p, 0.1, 0.5, 4)
    assert s.tree_summary()["depth"] == 21 and s.tree_summary()["leaf_count"] == 1
    assert np.all(ax == 0) and np.all(ay == 0) and np.all(az == 0)


def test_zero_masses_give_no_nan(ctx):
    # octree_test.cpp:434-451: a sized but unfilled container (zero masses at the
    # origin) builds a tree with zero masses and finite centres and radii.
    n = 16
    z = np.zeros(n)
    p = oracle.Particles(z, z.copy(), z.copy(), z.copy(), z.copy(), z.copy(), z.copy())
    s = E.HipTreeSolver(E.TreeParameters(0.5, 4, quadrupole=True), E.Softening(0.1), context=ctx)
    acc = [np.full(n, np.nan) for _ in range(3)]
    s.evaluate((p.x, p.y, p.z), p.m, acc)
    nodes = s.nodes()
    assert np.all(nodes["mass"] == 0)
    assert np.all(np.isfinite(nodes["com"])) and np.all(np.isfinite(nodes["acc_r2"]))
    assert np.all(np.isfinite(s.quadrupoles()))
    assert all(np.all(a == 0) for a in acc)


@pytest.mark.parametrize("quadrupole", [False, True])
def test_massless_particles_among_massive_ones(ctx, quadrupole):
    # Zero-mass cells take the mean of their particles (octree.cpp:75-84, 120-135):
    # every third particle massless, whole massless leaves included; moments
    # bitwise against the oracle, forces finite and equal to the massive set's.
    p = oracle.plummer(6000, 31)
    m = p.m.copy()
    m[::3] = 0.0
    m[p.x > 1.5] = 0.0
    q = oracle.Particles(p.x, p.y, p.z, p.vx, p.vy, p.vz, m)
    s = E.HipTreeSolver(E.TreeParameters(0.5, 8, quadrupole=quadrupole), E.Softening(0.02), context=ctx)
    acc = [np.zeros(q.n) for _ in range(3)]
    s.evaluate((q.x, q.y, q.z), q.m, acc)
    assert all(np.all(np.isfinite(a)) for a in acc)
    codes, index, origin, size = oracle.morton_order(q.x, q.y, q.z)
    i = index.astype(np.int64)
    t = oracle.tree_build(q.x[i], q.y[i], q.z[i], q.m[i], codes, origin, size, 0.5, 8, quadrupole=quadrupole)
    got = s.nodes()
    assert np.any(t.mass == 0)
    assert np.array_equal(got["mass"], t.mass) and np.array_equal(got["com"], t.com)
    assert np.array_equal(got["acc_r2"], t.acc_r2)
    if quadrupole:
        assert np.array_equal(s.quadrupoles(), t.quad)
    ref = oracle.reference_acceleration(q.x, q.y, q.z, q.m, 0.02 ** 2)
    worst, rms = oracle.error_summary(oracle.relative_errors(*acc, *ref))
    assert rms < 1e-2 and worst < 1e-1, (rms, worst)


@pytest.mark.parametrize("gen,n,theta,cap", [("plummer", 20000, 0.5, 16), ("plummer", 20000, 0.8, 32),
                                             ("disc", 50000, 0.5, 16), ("disc", 50000, 0.6, 16)])
def test_counts_equal_fp32_policy_mirror(ctx, gen, n, theta, cap):
    p = oracle.plummer(n, 20260811) if gen == "plummer" else oracle.disc_collision(n, 3)
    _, s = tree_eval(ctx, p, 0.01, theta, cap)
    c = s.interaction_count()
    pp, pc = oracle.barnes_hut_fp32_counts(p.x, p.y, p.z, p.m, theta, cap)
    assert (c.particle_particle, c.particle_cell) == (pp, pc)
    # and, on these inputs, the FP64 reference walk's counts as well
    _, _, _, (rpp, rpc) = oracle.barnes_hut(p.x, p.y, p.z, p.m, 1e-4, theta, cap)
    print(f"device pp={pp} pc={pc}; FP64 walk pp={rpp} pc={rpc}")
    assert abs(pp - rpp) <= 1e-4 * rpp and abs(pc - rpc) <= 1e-4 * max(1, rpc)


def test_theta_zero_is_direct_sum(ctx):
    p = oracle.plummer(3000, 4)
    eps = 0.02
    (ax, ay, az), s = tree_eval(ctx, p, eps, 0.0, 16)
    c = s.interaction_count()
    assert c.particle_particle == 3000 * 2999 and c.particle_cell == 0
    ref = oracle.reference_acceleration(p.x, p.y, p.z, p.m, eps * eps)
    worst, rms = oracle.error_summary(oracle.relative_errors(ax, ay, az, *ref))
    assert rms < 1e-6 and worst < 2e-5, (rms, worst)


def test_two_body_and_self(ctx):
    x = np.array([-1.0, 1.0])
    z = np.zeros(2)
    p = oracle.Particles(x, z, z.copy(), z.copy(), z.copy(), z.copy(), np.array([2.0, 3.0]))
    (ax, ay, az), s = tree_eval(ctx, p, 0.0, 0.5, 1)
    assert ax.tolist() == [0.75, -0.5] and ay.tolist() == [0.0, 0.0]
    one = oracle.Particles(np.array([2.0]), np.array([-3.0]), np.array([5.0]), *(np.zeros(1) for _ in range(3)),
                           np.array([4.0]))
    (ax, ay, az), _ = tree_eval(ctx, one, 0.0, 0.5, 16)
    assert (ax[0], ay[0], az[0]) == (0.0, 0.0, 0.0)


def test_empty_configuration(ctx):
    s = E.HipTreeSolver(E.TreeParameters(0.5, 16), E.Softening(0.1), context=ctx)
    e = np.zeros(0)
    s.evaluate((e, e, e), e, (e.copy(), e.copy(), e.copy()))
    assert s.interaction_count().evaluations == 1 and s.name() == "hip-tree"


def test_parameters_are_corrected(ctx):
    s = E.HipTreeSolver(E.TreeParameters(1.7, 0), E.Softening(0.1), context=ctx)
    assert s.parameters().opening_angle == 1.0 and s.parameters().leaf_capacity == 1


def test_accuracy_and_monotone_in_theta(ctx):
    p = oracle.plummer(20000, 20260811, scale_radius=3 * np.pi / 16, cutoff=0.999)
    eps = 0.05
    rng = np.random.default_rng(1)
    tgt = np.sort(rng.choice(p.n, 2048, replace=False)).astype(np.uint64)
    ref = oracle.reference_acceleration(p.x, p.y, p.z, p.m, eps * eps, targets=tgt)
    i = tgt.astype(np.int64)
    rms = []
    for theta in (0.3, 0.5, 0.7):
        (ax, ay, az), _ = tree_eval(ctx, p, eps, theta, 16)
        w, r = oracle.error_summary(oracle.relative_errors(ax[i], ay[i], az[i], *ref))
        print(f"theta={theta} rms={r:.3e} worst={w:.3e}")
        rms.append(r)
        if theta == 0.5:
            assert r < 1e-2 and w < 1e-1
    assert rms[0] < rms[1] < rms[2]


def test_momentum_residual(ctx):
    # cuda_tree_solver_test.cpp:535-568: sum m a small against sum m |a|.
    p = oracle.plummer(20000, 7)
    (ax, ay, az), _ = tree_eval(ctx, p, 0.02, 0.5, 16)
    scale = np.sum(p.m * np.sqrt(ax * ax + ay * ay + az * az))
    net = np.abs([np.sum(p.m * ax), np.sum(p.m * ay), np.sum(p.m * az)]) / scale
    assert np.all(net < 1e-3), net


def test_disc_collision_262144_sampled(ctx):
    # The reference's own bounds (cuda_tree_solver_test.cpp:392-412) without
    # quadrupoles, on the oracle's two-disc model.
    p = oracle.disc_collision(262144, 1)
    eps, theta, cap = 0.01, 0.5, 16
    (ax, ay, az), s = tree_eval(ctx, p, eps, theta, cap)
    rng = np.random.default_rng(2)
    tgt = np.sort(rng.choice(p.n, 1024, replace=False)).astype(np.uint64)
    ref = oracle.reference_acceleration(p.x, p.y, p.z, p.m, eps * eps, targets=tgt)
    i = tgt.astype(np.int64)
    w, r = oracle.error_summary(oracle.relative_errors(ax[i], ay[i], az[i], *ref))
    print(f"disc N=262144 theta=0.5 leaf 16: rms={r:.3e} worst={w:.3e} {s.tree_summary()}")
    assert r < 1e-2 and w < 1e-1


def disc_quadrupole_rms(ctx, n, targets, seed):
    from orrery_amd import workloads as W
    p = W.disc_collision(n, 1)
    eps, theta, cap = 0.01, 0.5, 16
    s = E.HipTreeSolver(E.TreeParameters(theta, cap, quadrupole=True), E.Softening(eps), context=ctx)
    acc = [np.full(n, np.nan) for _ in range(3)]
    s.evaluate((p.x, p.y, p.z), p.m, acc)
    rng = np.random.default_rng(seed)
    tgt = np.sort(rng.choice(n, targets, replace=False)).astype(np.uint64)
    ref = oracle.reference_acceleration(p.x, p.y, p.z, p.m, eps * eps, targets=tgt)
    i = tgt.astype(np.int64)
    w, r = oracle.error_summary(oracle.relative_errors(acc[0][i], acc[1][i], acc[2][i], *ref))
    print(f"two-disc N={n} theta=0.5 leaf 16 quadrupole, {targets} targets: rms={r:.3e} worst={w:.3e}")
    return r, w


@pytest.mark.parametrize("n,targets", [(262144, 4096), (2097152, 1024)])
def test_config3_rms_within_1e3(ctx, n, targets):
    # The Barnes-Hut configuration (two-disc collision, theta = 0.5, leaf 16,
    # quadrupoles, the benchmark's generator) against the FP64 direct sum on
    # sampled targets: RMS relative error <= 1e-3.
    r, w = disc_quadrupole_rms(ctx, n, targets, 5)
    assert r <= 1e-3 and w < 5e-2, (r, w)


def test_device_resident_tree_stepping(ctx):
    # Barnes-Hut as the field of device-resident Yoshida-4 steps vs the CPU path.
    p = oracle.plummer(3000, 12)
    eps, dt = 0.02, 1e-3
    sysd = E.DeviceSystem(ctx, p.n)
    sysd.upload(*p.arrays())
    cfg = E.StepConfig(integrator="yoshida4", solver="barnes-hut", softening=eps, opening_angle=0.5,
                       leaf_capacity=16)
    sysd.refresh(cfg)
    sysd.step(cfg, dt, 5)
    st = sysd.download()
    f = oracle.make_field("yoshida4", "barnes-hut", eps2=eps * eps, theta=0.5, leaf_capacity=16)
    ref = oracle.Particles(*(a.copy() for a in p.arrays()))
    acc = oracle.refresh(ref, f)
    oracle.steps(ref, acc, f, dt, 5)
    assert np.max(np.abs(st["x"] - ref.x)) < 1e-8
    sysd.close()


@pytest.mark.parametrize("quadrupole", [False, True])
def test_wave_shapes_are_bitwise(ctx, quadrupole):
    # A wave walking 128 targets (two per lane), 64 or 32 (half its lanes)
    # visits each target's nodes in the same DFS order with the same arithmetic;
    # counts and node visits are the same too. Also a batch of two systems
    # (the segmented walk) under each shape.
    p = oracle.disc_collision(30000, 4)
    q = oracle.plummer(7000, 9)
    results = []
    assert ctx.tree_config() == {"targets_per_lane": -1, "lanes": -1, "sort_mode": 1}  # the defaults
    try:
        for r, lanes in ((1, 64), (2, 64), (1, 32), (2, 32)):
            ctx.set_tree_config(targets_per_lane=r, lanes=lanes)
            ctx.set_tree_config()  # zeros keep
            assert ctx.tree_config() == {"targets_per_lane": r, "lanes": lanes, "sort_mode": 1}
            s = E.HipTreeSolver(E.TreeParameters(0.5, 16, quadrupole=quadrupole), E.Softening(0.01),
                                context=ctx)
            acc = (np.zeros(p.n), np.zeros(p.n), np.zeros(p.n))
            s.evaluate((p.x, p.y, p.z), p.m, acc)
            c = s.interaction_count()
            results.append((acc, c.particle_particle, c.particle_cell, s.node_visits()))
    finally:
        ctx.set_tree_config(targets_per_lane=-1, lanes=-1)
    a1, pp1, pc1, v1 = results[0]
    for a2, pp2, pc2, v2 in results[1:]:
        for k in range(3):
            assert np.array_equal(a1[k], a2[k])
        assert (pp1, pc1, v1) == (pp2, pc2, v2)
    # The segmented walk of a two-system batch, full and half lanes.
    batches = []
    try:
        for lanes in (64, 32):
            ctx.set_tree_config(lanes=lanes)
            cfg = E.BatchConfig(integrator="velocity-verlet", opening_angle=0.5, leaf_capacity=16,
                                quadrupole=quadrupole, tiles_per_item=16, direct_below=1000,
                                use_graph=False)
            b = E.DeviceBatch(ctx, [E.BatchSystem(p.n, 0.01, 1e-3), E.BatchSystem(q.n, 0.01, 1e-3)], cfg)
            b.upload(0, *p.arrays())
            b.upload(1, *q.arrays())
            b.refresh()
            batches.append([b.download(k) for k in range(2)] + [b.counts(0), b.counts(1)])
            b.close()
    finally:
        ctx.set_tree_config(lanes=-1)
    full, half = batches
    for k in range(2):
        for key in ("ax", "ay", "az"):
            assert np.array_equal(full[k][key], half[k][key]), (k, key)
    assert full[2:] == half[2:]


def test_replayed_evaluations_are_bitwise(ctx):
    # The second evaluation with unchanged arguments is captured as a hipGraph
    # and later ones replay it (phase timing on forces eager runs): every result
    # and count must be that of the eager first evaluation, and new positions in
    # the same device buffers must be honoured.
    p = oracle.disc_collision(30000, 6)
    s = E.HipTreeSolver(E.TreeParameters(0.5, 16, quadrupole=True), E.Softening(0.01), context=ctx)
    runs = []
    for k in range(5):
        if k == 3:
            ctx.set_timing(True)
            ctx.reset_timings()
        acc = (np.zeros(p.n), np.zeros(p.n), np.zeros(p.n))
        s.reset_interaction_count()
        s.evaluate((p.x, p.y, p.z), p.m, acc)
        c = s.interaction_count()
        runs.append((acc, c.particle_particle, c.particle_cell, s.tree_summary()))
    t = ctx.timings()
    ctx.set_timing(False)
    for acc, pp, pc, tree in runs[1:]:
        assert all(np.array_equal(acc[k], runs[0][0][k]) for k in range(3))
        assert (pp, pc, tree) == runs[0][1:]
    # Runs 3 and 4 ran eagerly (phase timing on) and timed the walk twice.
    assert t["walk"][1] == 2 and t["walk"][0] > 0
    # Moved particles, same buffers: the replay sees them.
    q = oracle.disc_collision(30000, 7)
    for _ in range(3):
        acc = (np.zeros(q.n), np.zeros(q.n), np.zeros(q.n))
        s.evaluate((q.x, q.y, q.z), q.m, acc)
    fresh = E.HipTreeSolver(E.TreeParameters(0.5, 16, quadrupole=True), E.Softening(0.01), context=E.Context(0))
    ref = (np.zeros(q.n), np.zeros(q.n), np.zeros(q.n))
    fresh.evaluate((q.x, q.y, q.z), q.m, ref)
    assert all(np.array_equal(acc[k], ref[k]) for k in range(3))
