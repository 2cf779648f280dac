"""The oracle restatement against the reference's own CPU library, bit for bit.

oracle/_ref/libref.so is compiled by oracle/ref.mk from the reference sources
where they lie; these tests skip where it was not built (it never reaches a
machine without the reference checkout unless built beforehand).
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from oracle import _ptr, _u64, _u64p

REF = O.load_reference()
pytestmark = pytest.mark.skipif(REF is None, reason="reference library not built")


def same(a, b):
    return all(np.array_equal(u, v) for u, v in zip(a, b))


@pytest.mark.parametrize("n,seed", [(4096, 1), (65536, 2), (1000, 20260811)])
def test_plummer_bitwise(n, seed):
    assert same(O.plummer(n, seed).arrays(), O.plummer(n, seed, lib=REF).arrays())


def test_uniform_sphere_bitwise():
    assert same(O.uniform_sphere(200000, 5).arrays(), O.uniform_sphere(200000, 5, lib=REF).arrays())


@pytest.mark.parametrize("kernel", [0, 1])
def test_direct_bitwise(kernel):
    p = O.plummer(3001, 20260810)
    eps = 0.05
    a = O.direct(p.x, p.y, p.z, p.m, eps * eps, kernel=kernel, threads=4)
    r = [np.zeros(p.n) for _ in range(3)]
    REF.ref_direct(_u64(p.n), _ptr(p.x), _ptr(p.y), _ptr(p.z), _ptr(p.m), ctypes.c_double(eps),
                   ctypes.c_int(kernel), ctypes.c_uint(4), *(_ptr(v) for v in r))
    assert same(a, r)


def test_reference_kernel_bitwise():
    p = O.plummer(2000, 9)
    eps = 0.01
    a = O.reference_acceleration(p.x, p.y, p.z, p.m, eps * eps)
    r = [np.zeros(p.n) for _ in range(3)]
    REF.ref_reference_acceleration(_u64(p.n), _ptr(p.x), _ptr(p.y), _ptr(p.z), _ptr(p.m),
                                   ctypes.c_double(eps), None, _u64(0), *(_ptr(v) for v in r))
    assert same(a, r)


@pytest.mark.parametrize("gen", ["plummer", "disc"])
def test_morton_bitwise(gen):
    p = O.plummer(50000, 3) if gen == "plummer" else O.disc_collision(50000, 3)
    codes, index, origin, size = O.morton_order(p.x, p.y, p.z)
    rc, ri = np.zeros(p.n, dtype=np.uint64), np.zeros(p.n, dtype=np.uint64)
    ro, rs = np.zeros(3), ctypes.c_double()
    REF.ref_morton_order(_u64(p.n), _ptr(p.x), _ptr(p.y), _ptr(p.z), _ptr(ro), ctypes.byref(rs),
                         _ptr(rc, _u64p), _ptr(ri, _u64p))
    assert np.array_equal(codes, rc) and np.array_equal(index, ri)
    assert np.array_equal(origin, ro) and size == rs.value


@pytest.mark.parametrize("cap,theta,quad", [(32, 0.5, 0), (16, 0.5, 0), (1, 0.6, 0),
                                            (8, 0.0, 0), (16, 0.5, 1)])
def test_octree_bitwise(cap, theta, quad):
    p = O.plummer(6000, 20260810)
    codes, index, origin, size = O.morton_order(p.x, p.y, p.z)
    i = index.astype(np.int64)
    t = O.tree_build(p.x[i], p.y[i], p.z[i], p.m[i], codes, origin, size, theta, cap, bool(quad))
    M = 4 * p.n + 100
    f = [np.zeros(M) for _ in range(5)]
    u = [np.zeros(M, dtype=np.uint64) for _ in range(3)]
    q, ld = np.zeros((M, 6)), np.zeros(2, dtype=np.uint64)
    cnt = REF.ref_tree(_u64(p.n), _ptr(p.x), _ptr(p.y), _ptr(p.z), _ptr(p.m), ctypes.c_double(theta),
                       _u64(cap), ctypes.c_int(quad), ctypes.c_uint(4), _u64(M), *(_ptr(a) for a in f),
                       *(_ptr(a, _u64p) for a in u), _ptr(q), _ptr(ld, _u64p))
    k = len(t.mass)
    assert cnt == k and t.leaves == ld[0] and t.depth == ld[1]
    assert np.array_equal(t.mass, f[0][:k]) and np.array_equal(t.com[:, 0], f[1][:k])
    assert np.array_equal(t.com[:, 1], f[2][:k]) and np.array_equal(t.com[:, 2], f[3][:k])
    assert np.array_equal(t.acc_r2, f[4][:k])
    assert np.array_equal(t.next, u[0][:k]) and np.array_equal(t.first, u[1][:k])
    assert np.array_equal(t.count, u[2][:k])
    if quad:
        assert np.array_equal(t.quad, q[:k])


@pytest.mark.parametrize("quad", [0, 1])
def test_barnes_hut_bitwise(quad):
    p = O.plummer(5000, 20260810)
    eps = 0.05
    ax, ay, az, (pp, pc) = O.barnes_hut(p.x, p.y, p.z, p.m, eps * eps, 0.5, 32, quadrupole=bool(quad),
                                        threads=4)
    r = [np.zeros(p.n) for _ in range(3)]
    c = np.zeros(2, dtype=np.uint64)
    REF.ref_barnes_hut(_u64(p.n), _ptr(p.x), _ptr(p.y), _ptr(p.z), _ptr(p.m), ctypes.c_double(eps),
                       ctypes.c_double(0.5), _u64(32), ctypes.c_int(quad), ctypes.c_int(1),
                       ctypes.c_uint(4), *(_ptr(v) for v in r), _ptr(c, _u64p))
    assert same((ax, ay, az), r) and (pp, pc) == (int(c[0]), int(c[1]))


@pytest.mark.parametrize("integ,solver", [(0, 2), (1, 2), (0, 1), (1, 0), (2, 2), (2, 1)])
def test_integrators_bitwise(integ, solver):
    eps = 0.05
    s1, s2 = O.plummer(400, 3), O.plummer(400, 3)
    f = O.make_field(integrator=["vv", "yoshida4", "rk4"][integ],
                     solver="barnes-hut" if solver == 1 else "direct", eps2=eps * eps, theta=0.5,
                     leaf_capacity=16, kernel=1 if solver == 2 else 0)
    acc = O.refresh(s1, f)
    O.steps(s1, acc, f, 1e-3, 4)
    acc2 = O.refresh(s2, f)
    REF.ref_steps(_u64(400), *(_ptr(a) for a in (s2.x, s2.y, s2.z, s2.vx, s2.vy, s2.vz)),
                  *(_ptr(a) for a in acc2), _ptr(s2.m), ctypes.c_int(integ), ctypes.c_int(solver),
                  ctypes.c_double(eps), ctypes.c_double(0.5), _u64(16), ctypes.c_double(1e-3), _u64(4))
    assert same(s1.arrays(), s2.arrays()) and same(acc, acc2)


def test_diagnostics_bitwise():
    p = O.plummer(2500, 4)
    eps = 0.02
    d = O.diagnostics(p, eps * eps)
    out = np.zeros(8)
    REF.ref_diagnostics(_u64(p.n), *(_ptr(a) for a in p.arrays()), ctypes.c_double(eps), _ptr(out))
    assert d["kinetic"] == out[0] and d["potential"] == out[1]
    assert np.array_equal(d["momentum"], out[2:5]) and np.array_equal(d["angular_momentum"], out[5:8])
