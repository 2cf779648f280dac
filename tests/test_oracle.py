"""The CPU oracle against the reference's golden vectors and known answers.

These pin the checker before it is trusted: every function of the oracle is
checked against a fixture the reference holds, a known answer from the
reference's own tests, a closed form, or the reference library itself
(tests/test_oracle_vs_reference.py). None of them re-types the oracle's formula.
"""
import numpy as np
import pytest

import oracle
from tests.golden import otj

O = oracle


# ---- golden fixtures (web/tests/fixtures) -----------------------------------

def test_uniform_sphere_reproduces_agreement_frame0():
    # agreement.orrery: uniform sphere, N = 64, mass 1, radius 1, seed 20260813.
    t = otj.read("tests/golden/agreement.otj")
    u = O.uniform_sphere(64, 20260813)
    x0, y0, z0 = t["frames"][0]["positions"]
    assert np.array_equal(t["masses"], u.m)
    # Written by another platform's libm: equal to the last bits (agreement.test.ts:29-44).
    for a, b in ((x0, u.x), (y0, u.y), (z0, u.z)):
        assert np.max(np.abs(a - b)) <= 4e-16


def test_agreement_run_energy_and_positions():
    # 200 velocity-Verlet steps of direct summation, eps = 0.05, dt = 1e-3, FP64.
    t = otj.read("tests/golden/agreement.otj")
    rows = otj.read_csv("tests/golden/agreement.csv")
    x, y, z = (a.copy() for a in t["frames"][0]["positions"])
    n = t["n"]
    st = O.Particles(x, y, z, np.zeros(n), np.zeros(n), np.zeros(n), t["masses"].copy())
    eps2 = 0.05 * 0.05
    d0 = O.diagnostics(st, eps2)
    assert d0["total"] == pytest.approx(rows[0]["total_energy"], rel=1e-15, abs=0)
    f = O.make_field("vv", "direct", eps2=eps2, kernel=1)
    acc = O.refresh(st, f)
    O.steps(st, acc, f, t["timestep"], 200)
    d1 = O.diagnostics(st, eps2)
    # The energy error is a difference of nearly equal energies: the CSV's
    # drift 3.396e-9 must be reproduced to a few percent.
    drift = (d1["total"] - d0["total"]) / abs(d0["total"])
    assert drift == pytest.approx(rows[1]["relative_energy_error"], rel=1e-3)
    assert d1["total"] == pytest.approx(rows[1]["total_energy"], rel=1e-14)
    assert d1["kinetic"] == pytest.approx(rows[1]["kinetic_energy"], rel=1e-12)
    want = np.stack(t["frames"][1]["positions"])
    got = np.stack([st.x, st.y, st.z])
    assert np.max(np.abs(got - want)) < 1e-13


def test_kepler_fixture_bitwise():
    # examples/kepler.orrery: two bodies, direct, eps = 0, VV; frames every 20 steps.
    t = otj.read("tests/golden/kepler-double.otj")
    sep = 0.5
    speed = np.sqrt(1.0 * 2.0 * (1 + 0.5) / sep)
    st = O.Particles(np.array([-0.5 * sep, 0.5 * sep]), np.zeros(2), np.zeros(2),
                     np.zeros(2), np.array([-0.5 * speed, 0.5 * speed]), np.zeros(2),
                     np.array([1.0, 1.0]))
    f = O.make_field("vv", "direct", eps2=0.0, kernel=1)
    acc = O.refresh(st, f)
    for frame in t["frames"]:
        if frame["step"] > 0:
            O.steps(st, acc, f, t["timestep"], 20)
        want = frame["positions"]
        assert np.array_equal(st.x, want[0]) and np.array_equal(st.y, want[1]), frame["step"]


# ---- Morton known answers (tests/solvers/morton_test.cpp:45-182) -------------

def test_spread_bits_known_answers():
    lib = O.load()
    assert lib.orc_spread_bits(0) == 0
    assert lib.orc_spread_bits(1) == 1
    assert lib.orc_spread_bits(2) == 0b1000
    assert lib.orc_spread_bits(3) == 0b1001
    assert lib.orc_spread_bits(0b101) == 0b1000001
    assert lib.orc_spread_bits(1 << 20) == 1 << 60
    for bit in range(21):
        assert lib.orc_spread_bits(1 << bit) == 1 << (3 * bit)
    assert lib.orc_spread_bits(1 << 21) == 0


def test_interleave_order():
    f = O.load().orc_morton_interleave
    assert f(0, 0, 0) == 0
    assert f(1, 0, 0) == 0b100 and f(0, 1, 0) == 0b010 and f(0, 0, 1) == 0b001
    assert f(1, 1, 1) == 0b111
    assert f(2, 0, 0) == 0b100000 and f(2, 1, 1) == 0b100011


def test_octant_extraction():
    lib = O.load()
    g = (1 << 21) - 1
    corner = lib.orc_morton_interleave(g, 0, g)
    for level in range(1, 22):
        assert lib.orc_morton_octant(corner, level) == 0b101
    single = lib.orc_morton_interleave(1 << 18, 0, 0)
    assert [lib.orc_morton_octant(single, k) for k in (1, 2, 3, 4)] == [0, 0, 0b100, 0]
    assert lib.orc_morton_octant(corner, 0) == 0 and lib.orc_morton_octant(corner, 22) == 0


def test_bounding_cube_geometry():
    origin, size = O.bounding_cube(np.array([-2.0, 4.0]), np.array([0.0, 1.0]), np.array([1.0, 3.0]))
    assert size == 6.0
    assert (origin + size / 2).tolist() == [1.0, 0.5, 2.0]
    origin, size = O.bounding_cube(np.full(2, 3.0), np.full(2, 3.0), np.full(2, 3.0))
    assert size > 0 and (origin + size / 2).tolist() == [3.0, 3.0, 3.0]
    origin, size = O.bounding_cube(np.zeros(0), np.zeros(0), np.zeros(0))
    assert size == 0.0


def test_far_face_clamp():
    lib = O.load()
    g = (1 << 21) - 1
    codes = O.morton_codes(np.array([1.0, -1.0]), np.array([1.0, -1.0]), np.array([1.0, -1.0]),
                           np.zeros(3), 1.0)
    assert int(codes[0]) == lib.orc_morton_interleave(g, g, g)
    assert int(codes[1]) == 0


def test_ordering_is_a_permutation_with_ascending_codes():
    p = O.plummer(20000, 20260810, scale_radius=3 * np.pi / 16, cutoff=0.999)
    codes, index, origin, size = O.morton_order(p.x, p.y, p.z)
    assert np.array_equal(np.sort(index), np.arange(20000, dtype=np.uint64))
    assert np.all(codes[1:] >= codes[:-1])
    assert np.array_equal(codes, O.morton_codes(p.x, p.y, p.z, origin, size)[index.astype(np.int64)])


# ---- force law closed forms (tests/solvers/direct_solver_test.cpp) -----------

@pytest.mark.parametrize("kernel", [0, 1])
def test_two_body_exact(kernel):
    ax, ay, az = O.direct(np.array([-1.0, 1.0]), np.zeros(2), np.zeros(2), np.array([2.0, 3.0]),
                          0.0, kernel=kernel, threads=1)
    assert ax.tolist() == [0.75, -0.5] and ay.tolist() == [0.0, 0.0]
    rx, ry, rz = O.reference_acceleration(np.array([-1.0, 1.0]), np.zeros(2), np.zeros(2),
                                          np.array([2.0, 3.0]), 0.0)
    assert rx.tolist() == [0.75, -0.5]


def test_plummer_factor():
    ax, ay, az = O.reference_acceleration(np.array([0.0, 4.0]), np.zeros(2), np.zeros(2),
                                          np.array([1.0, 5.0]), 9.0)
    assert ax[0] == pytest.approx(5.0 * 4.0 / 25.0 ** 1.5, rel=4e-16)
    assert ay[0] == 0.0 and az[0] == 0.0


def test_self_term_is_zero():
    for e2 in (0.0, 0.01):
        a = O.direct(np.array([2.0]), np.array([-3.0]), np.array([5.0]), np.array([4.0]), e2)
        assert [v[0] for v in a] == [0.0, 0.0, 0.0]


def test_scalar_and_avx2_agree_with_fp64_reference():
    # kernel_accuracy_test.cpp:152-184 bounds: 200 sqrt(N) eps for FP64.
    p = O.plummer(2048, 20260810, scale_radius=3 * np.pi / 16, cutoff=0.999)
    e2 = 0.05 ** 2
    ref = O.reference_acceleration(p.x, p.y, p.z, p.m, e2)
    bound = 200 * np.sqrt(2048) * np.finfo(np.float64).eps
    for kernel in (0, 1):
        a = O.direct(p.x, p.y, p.z, p.m, e2, kernel=kernel)
        worst, _ = O.error_summary(O.relative_errors(*a, *ref))
        assert worst < bound


def test_threads_do_not_change_bits():
    p = O.plummer(1500, 7)
    one = O.direct(p.x, p.y, p.z, p.m, 1e-4, threads=1)
    many = O.direct(p.x, p.y, p.z, p.m, 1e-4, threads=7)
    assert all(np.array_equal(a, b) for a, b in zip(one, many))


def test_momentum_conservation_fp64():
    # conservation_test.cpp:55-80: sum m a vanishes to 1e-13 of the cancelled terms.
    p = O.plummer(128, 20260810)
    a = O.direct(p.x, p.y, p.z, p.m, 0.0025)
    scale = np.sum(p.m * np.sqrt(a[0] ** 2 + a[1] ** 2 + a[2] ** 2))
    for c in a:
        assert abs(np.sum(p.m * c)) < 1e-13 * scale


# ---- octree invariants (tests/solvers/octree_test.cpp) ------------------------

def _tree(p, theta=0.5, cap=16, quad=False):
    codes, index, origin, size = O.morton_order(p.x, p.y, p.z)
    i = index.astype(np.int64)
    sx, sy, sz, sm = p.x[i], p.y[i], p.z[i], p.m[i]
    return O.tree_build(sx, sy, sz, sm, codes, origin, size, theta, cap, quad), (sx, sy, sz, sm)


def test_leaves_partition_particles_once_and_escape_indices():
    p = O.plummer(5000, 20260810)
    t, (sx, sy, sz, sm) = _tree(p)
    M = len(t.mass)
    leaves = t.count > 0
    assert int(np.sum(t.count[leaves])) == 5000
    starts = np.sort(t.first[leaves])
    ends = starts + np.sort(t.count[leaves])[np.argsort(np.argsort(starts))]
    covered = np.zeros(5000, dtype=int)
    for f, c in zip(t.first[leaves], t.count[leaves]):
        covered[int(f):int(f + c)] += 1
    assert np.all(covered == 1)
    assert np.all(t.next[leaves] == np.arange(M)[leaves] + 1)
    assert np.all(t.next > np.arange(M)) and np.all(t.next <= M)
    assert t.next[0] == M


def test_node_moments_are_sums_of_own_particles():
    p = O.plummer(3000, 20260810)
    t, (sx, sy, sz, sm) = _tree(p)
    M = len(t.mass)
    # Particle range of every node from its DFS subtree.
    for node in range(0, M, 7):
        lo = node
        hi = int(t.next[node])
        sub_leaves = [k for k in range(lo, hi) if t.count[k] > 0]
        b = int(min(t.first[k] for k in sub_leaves))
        e = int(max(t.first[k] + t.count[k] for k in sub_leaves))
        mass = np.sum(sm[b:e])
        assert t.mass[node] == pytest.approx(mass, rel=1e-13)
        com = np.array([np.sum(sx[b:e] * sm[b:e]), np.sum(sy[b:e] * sm[b:e]), np.sum(sz[b:e] * sm[b:e])]) / mass
        assert np.allclose(t.com[node], com, rtol=0, atol=1e-12)


def test_theta_zero_opens_everything():
    p = O.plummer(500, 3)
    t, _ = _tree(p, theta=0.0)
    assert np.all(np.isinf(t.acc_r2))
    ax, ay, az, (pp, pc) = O.barnes_hut(p.x, p.y, p.z, p.m, 1e-4, 0.0, 16)
    assert pp == 500 * 499 and pc == 0
    d = O.direct(p.x, p.y, p.z, p.m, 1e-4)
    assert np.max(O.relative_errors(ax, ay, az, *d)) < 1e-13


def test_two_particle_leaf_com():
    # octree_test.cpp:98-125: masses 1 and 4 at 0 and 1 -> COM 4/5.
    x = np.array([0.0, 1.0])
    z = np.zeros(2)
    t, _ = _tree(O.Particles(x, z, z, z, z, z, np.array([1.0, 4.0])), cap=2)
    assert len(t.mass) == 1 and t.mass[0] == 5.0 and t.com[0][0] == 0.8


def test_identical_points_stop_at_depth_21():
    x = np.full(40, 0.25)
    z = np.zeros(40)
    t, _ = _tree(O.Particles(x, x.copy(), x.copy(), z, z, z, np.full(40, 1.0)), cap=4)
    assert t.depth == 21 and t.leaves == 1


def test_bh_error_monotone_in_theta():
    p = O.plummer(3000, 20260810, scale_radius=3 * np.pi / 16, cutoff=0.999)
    e2 = 0.05 ** 2
    ref = O.reference_acceleration(p.x, p.y, p.z, p.m, e2)
    rms = []
    for theta in (0.2, 0.5, 0.8):
        a = O.barnes_hut(p.x, p.y, p.z, p.m, e2, theta, 32)
        rms.append(O.error_summary(O.relative_errors(*a[:3], *ref))[1])
    assert rms[0] < rms[1] < rms[2]
    assert rms[1] < 1e-2


# ---- integrators: measured order (tests/integrators/convergence_test.cpp) -----

@pytest.mark.parametrize("integrator,order", [("vv", 2.0), ("yoshida4", 4.0), ("rk4", 4.0)])
def test_measured_order(integrator, order):
    # Circular two-body orbit, one period; error vs exact closure at dt and dt/2.
    def final_error(steps):
        st = O.Particles(np.array([-0.5, 0.5]), np.zeros(2), np.zeros(2), np.zeros(2),
                         np.array([-np.sqrt(0.5), np.sqrt(0.5)]) / 1.0, np.zeros(2), np.array([1.0, 1.0]))
        period = 2 * np.pi * np.sqrt(1.0 / 2.0)
        f = O.make_field(integrator, "direct", eps2=0.0, kernel=0)
        acc = O.refresh(st, f)
        O.steps(st, acc, f, period / steps, steps)
        return np.hypot(st.x[0] + 0.5, st.y[0])
    e1, e2 = final_error(200), final_error(400)
    measured = np.log2(e1 / e2)
    assert abs(measured - order) < 0.2, measured


def test_plummer_potential_closed_form():
    # plummer.cpp:130-136: U = -3 pi G M^2 / (32 a) for the untruncated sphere.
    p = O.plummer(4000, 1, cutoff=0.999)
    d = O.diagnostics(p, 0.0)
    assert d["potential"] == pytest.approx(-3 * np.pi / 32, rel=0.05)
    assert 2 * d["kinetic"] / abs(d["potential"]) == pytest.approx(1.0, abs=0.1)


def test_disc_collision_invariants():
    p = O.disc_collision(20000, 1)
    assert p.n == 20000
    assert np.sum(p.m) == pytest.approx(1.0, rel=1e-12)
    com = np.array([np.sum(p.m * p.x), np.sum(p.m * p.y), np.sum(p.m * p.z)])
    assert np.all(np.abs(com) < 1e-12)
    vcom = np.array([np.sum(p.m * p.vx), np.sum(p.m * p.vy), np.sum(p.m * p.vz)])
    assert np.all(np.abs(vcom) < 1e-12)
    # Two galaxies 20 apart along x, primary holding 3/4 of the particles.
    left = p.x < 0
    assert np.sum(left) == pytest.approx(15000, abs=2000)
