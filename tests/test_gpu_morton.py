"""Bit-exact Morton keys and permutation on the device against the oracle.

The oracle restates solvers/morton.cpp:60-265 and is itself pinned bit for bit
to the reference library and to tests/solvers/morton_test.cpp's known answers
(tests/test_oracle_morton.py). Keys and the (code, index) order must be equal
element for element. Both forms of the device sort run every case: one-sweep
passes (one counting launch, then one scatter per pass that finds the blocks
before it by look-back; the default) and histogram, scan and scatter launches
per pass.
"""
import numpy as np
import pytest

import oracle
import orrery_amd as E

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[1, 2], ids=["one-sweep", "multi-kernel"])
def sort_mode(ctx, request):
    ctx.set_tree_config(sort_mode=request.param)
    yield request.param
    ctx.set_tree_config(sort_mode=1)


def check(ctx, x, y, z):
    codes, index, origin, size = E.morton_order(ctx, x, y, z)
    rc, ri, ro, rs = oracle.morton_order(x, y, z)
    assert size == rs
    assert np.array_equal(origin, ro)
    assert np.array_equal(codes, rc)
    assert np.array_equal(index.astype(np.uint64), ri)
    return codes, index


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 2047, 2048, 2049, 4095, 4096, 4097, 100000, 1000003])
def test_plummer_keys_and_permutation(ctx, sort_mode, n):
    p = oracle.plummer(n, 20260810)
    check(ctx, p.x, p.y, p.z)


def test_config_plummer_262144(ctx, sort_mode):
    p = oracle.plummer(262144, 1)
    codes, _ = check(ctx, p.x, p.y, p.z)
    assert np.all(codes[1:] >= codes[:-1])


def test_disc_collision_2m(ctx, sort_mode):
    p = oracle.disc_collision(2_097_152, 1)
    check(ctx, p.x, p.y, p.z)


def test_duplicates_keep_index_order(ctx, sort_mode):
    # Equal codes sort by index: the total order of morton.hpp:155-175.
    rng = np.random.default_rng(3)
    base = rng.normal(size=(50, 3))
    pts = base[rng.integers(0, 50, size=5000)]
    codes, index = check(ctx, pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy())
    same = codes[1:] == codes[:-1]
    assert np.all(index[1:][same] > index[:-1][same])


def test_degenerate_and_far_face(ctx, sort_mode):
    # No extent -> size 1 (morton.cpp:90-93); far-face clamp (morton_test.cpp:169-182).
    x = np.full(10, 3.0)
    check(ctx, x, x.copy(), x.copy())
    x = np.array([0.0, 1.0, 0.5, 1.0, 0.0])
    check(ctx, x, x[::-1].copy(), np.array([1.0, 1.0, 0.0, 0.0, 0.5]))


def test_sort_mode_is_checked(ctx):
    with pytest.raises(E.solvers.C.EngineError):
        ctx.set_tree_config(sort_mode=3)
    assert ctx.tree_config()["sort_mode"] == 1


def test_one_sweep_beyond_resident_blocks(ctx):
    # 8,388,608 keys are 4096 sort blocks, more than the device holds at once,
    # so the look-back's blocks wait on blocks that started earlier rather than
    # on blocks all resident together (the 2M case).
    assert ctx.tree_config()["sort_mode"] == 1
    rng = np.random.default_rng(17)
    n = 8_388_608
    x, y, z = (rng.standard_normal(n) for _ in range(3))
    check(ctx, x, y, z)
