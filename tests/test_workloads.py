"""The synthetic workload generators.

The Plummer sampler is the reference's own seeded sample, so it is pinned bit
for bit: against the oracle's restatement of plummer.cpp / random.cpp and,
where it is built, against the reference library compiled from its sources.
The two-disc model is this project's own and draws from numpy's PCG64, so it is
checked by what fixes the model -- closed forms, invariants and distributions --
and against the oracle's C version of the same model by distribution.
"""
import numpy as np
import pytest

import oracle
from orrery_amd import workloads as W


def ks_two_sample(a, b):
    a, b = np.sort(a), np.sort(b)
    grid = np.concatenate([a, b])
    fa = np.searchsorted(a, grid, side="right") / len(a)
    fb = np.searchsorted(b, grid, side="right") / len(b)
    return np.max(np.abs(fa - fb))


def same_bits(p, q):
    return all(np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))
               for a, b in zip(p.arrays(), q.arrays()))


@pytest.mark.parametrize("n,seed", [(1, 1), (7, 2), (4096, 1), (65536, 1), (65536, 2), (65536, 3),
                                    (262144, 3)])
def test_plummer_is_the_reference_sample(n, seed):
    # mt19937_64(seed), top-53-bit uniforms, position direction before velocity
    # direction, compensated centre-of-mass shift: every bit as the reference.
    p = W.plummer(n, seed)
    assert same_bits(p, oracle.plummer(n, seed))
    ref = oracle.load_reference()
    if ref is not None:
        assert same_bits(p, oracle.plummer(n, seed, lib=ref))


def test_plummer_other_parameters_and_errors():
    p = W.plummer(1000, 9, total_mass=2.5, scale_radius=3 * np.pi / 16, cutoff=0.999)
    assert same_bits(p, oracle.plummer(1000, 9, total_mass=2.5, scale_radius=3 * np.pi / 16,
                                       cutoff=0.999))
    assert W.plummer(0, 1).n == 0
    from orrery_amd import _capi as C
    for kwargs in (dict(total_mass=0.0), dict(scale_radius=-1.0), dict(cutoff=1.0),
                   dict(cutoff=float("nan"))):
        with pytest.raises(C.EngineError):
            W.plummer(10, 1, **kwargs)


def test_plummer_invariants():
    p = W.plummer(20000, 3)
    assert p.n == 20000 and np.allclose(p.m, 1 / 20000)
    for pos, vel in ((p.x, p.vx), (p.y, p.vy), (p.z, p.vz)):
        assert abs(np.sum(pos * p.m)) < 1e-14 and abs(np.sum(vel * p.m)) < 1e-14
    r = np.sqrt(p.x ** 2 + p.y ** 2 + p.z ** 2)
    assert r.max() < 10.2  # truncated at r < 10 a before the centre-of-mass shift
    # Virial equilibrium: 2K/|U| = 1 within sampling noise.
    d = oracle.diagnostics(oracle.Particles(*p.arrays()), 1e-6)
    assert abs(2 * d["kinetic"] / abs(d["potential"]) - 1) < 0.05


def test_plummer_potential_energy_closed_form():
    # U = -3 pi G M^2 / (32 a) for the untruncated sphere (plummer.cpp:130-136);
    # truncation at 10 a lowers |U| by under 1 %.
    p = W.plummer(16384, 8)
    d = oracle.diagnostics(oracle.Particles(*p.arrays()), 1e-8)
    assert d["potential"] == pytest.approx(-3 * np.pi / 32, rel=0.03)


def test_disc_collision_model():
    n = 40000
    p = W.disc_collision(n, 2)
    assert p.n == n
    assert abs(p.m.sum() - 1.0) < 1e-12
    n1 = int(round(n * 0.75))
    assert np.allclose(p.m[:n1], 0.75 / n1) and np.allclose(p.m[n1:], 0.25 / (n - n1))
    for pos, vel in ((p.x, p.vx), (p.y, p.vy), (p.z, p.vz)):
        assert abs(np.sum(pos * p.m)) < 1e-12 and abs(np.sum(vel * p.m)) < 1e-12
    # Centres of mass 20 apart on the x axis, approaching on a parabolic orbit.
    c1 = np.array([np.average(a[:n1], weights=p.m[:n1]) for a in (p.x, p.y, p.z)])
    c2 = np.array([np.average(a[n1:], weights=p.m[n1:]) for a in (p.x, p.y, p.z)])
    assert np.allclose(c2 - c1, [20, 0, 0], atol=0.05)
    w1 = np.array([np.average(a[:n1], weights=p.m[:n1]) for a in (p.vx, p.vy, p.vz)])
    w2 = np.array([np.average(a[n1:], weights=p.m[n1:]) for a in (p.vx, p.vy, p.vz)])
    vrel = w2 - w1
    energy = 0.5 * vrel @ vrel - 1.0 / 20.0
    assert abs(energy) < 2e-3
    # The oracle's C version of the same model: radii from each centre alike.
    q = oracle.disc_collision(n, 2)
    for (lo, hi) in ((0, n1), (n1, n)):
        ra = np.sqrt((p.x[lo:hi] - p.x[lo:hi].mean()) ** 2 + (p.y[lo:hi] - p.y[lo:hi].mean()) ** 2)
        rb = np.sqrt((q.x[lo:hi] - q.x[lo:hi].mean()) ** 2 + (q.y[lo:hi] - q.y[lo:hi].mean()) ** 2)
        assert ks_two_sample(ra, rb) < 0.03


def test_batched_sizes_log_uniform():
    ns = W.batched_sizes(512, 64, 16384, 4)
    assert len(ns) == 512 and min(ns) >= 64 and max(ns) <= 16384
    logs = np.log(ns)
    # Log-uniform: mean of log N is the midpoint, and a quarter lie below 4 * 64.
    assert abs(logs.mean() - 0.5 * (np.log(64) + np.log(16384))) < 0.25
    assert 0.19 < np.mean(np.array(ns) < 256) < 0.31
    assert W.batched_sizes(512, 64, 16384, 4) == ns
