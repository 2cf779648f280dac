"""orr_direct_eval_static, the direct sum that allocates nothing, against orr_direct_eval.

Both entry points are called through the C ABI on the same device arrays.
Two calls in a row, the second with fewer particles, so that the static
scratch still holds the first call's contents when the second reads it:

* N = 30,000: the one-sided kernel's split partials do not fit the 16 MiB
  scratch at the default split count, so the static path groups the same
  terms into fewer Kahan-summed splits. Held to the direct bound against
  orr_direct_eval and against the FP64 oracle: per particle
  |a - a_ref| / |a_ref| <= 1e-6. With the split count lowered by
  configuration until orr_direct_eval's partials fit too, bit for bit.
* N = 3,000: the default split count fits, so bit for bit.

Then the limits: one particle past the maximum is refused with
ORR_ERR_LENGTH, and the next call still works.
"""
import ctypes

import numpy as np
import pytest

import oracle
import orrery_amd as E
from orrery_amd import _capi as C
from orrery_amd import workloads as W

pytestmark = pytest.mark.gpu

BOUND = 1e-6


def call(ctx, entry, arrays, n, eps):
    x, y, z, m, ax, ay, az = arrays
    args = C.OrrDirectArgs()
    args.x, args.y, args.z, args.m = x.ptr, y.ptr, z.ptr, m.ptr
    args.ax, args.ay, args.az = ax.ptr, ay.ptr, az.ptr
    args.count = n
    args.softening_squared = eps * eps
    args.gravitational_constant = 1.0
    args.spatial_order = 1
    return getattr(C.load(), entry)(ctx.handle, ctypes.byref(args))


def device_arrays(ctx, p):
    return [E.DeviceArray.from_host(ctx, a) for a in (p.x, p.y, p.z, p.m)] + \
           [E.DeviceArray.from_host(ctx, np.full(p.n, np.nan)) for _ in range(3)]


def results(arrays):
    return [a.to_host() for a in arrays[4:]]


def relative(got, ref):
    num = np.sqrt(sum((g - r) ** 2 for g, r in zip(got, ref)))
    den = np.sqrt(sum(r ** 2 for r in ref))
    return num / den


def test_static_matches_allocating_entry_point(ctx):
    lib = C.load()
    max_count, scratch = ctypes.c_uint64(), ctypes.c_uint64()
    C.check(lib.orr_direct_static_limits(ctypes.byref(max_count), ctypes.byref(scratch)))
    assert max_count.value == 32768 and scratch.value == 16 << 20
    eps = 0.005
    default = ctx.direct_config()
    try:
        for n, seed in ((30000, 21), (3000, 22)):
            print(f"static direct: N={n} seed={seed}", flush=True)
            p = W.plummer(n, seed)
            arrays = device_arrays(ctx, p)
            C.check(call(ctx, "orr_direct_eval_static", arrays, n, eps))
            static = results(arrays)
            for a in arrays[4:]:
                a.upload(np.full(n, np.nan))
            C.check(call(ctx, "orr_direct_eval", arrays, n, eps))
            alloc = results(arrays)
            ref = oracle.reference_acceleration(p.x, p.y, p.z, p.m, eps * eps)
            assert not any(np.isnan(a).any() for a in static)
            assert relative(static, alloc).max() <= BOUND, (n, relative(static, alloc).max())
            assert relative(static, ref).max() <= BOUND, (n, relative(static, ref).max())
            if n == 3000:
                for s, a in zip(static, alloc):
                    assert np.array_equal(s, a), n
            else:
                # Fewer splits by configuration: orr_direct_eval's partials now
                # fit the scratch as well, and the two agree bit for bit.
                ctx.set_direct_config(block_target=256)
                for a in arrays[4:]:
                    a.upload(np.full(n, np.nan))
                C.check(call(ctx, "orr_direct_eval", arrays, n, eps))
                alloc_few = results(arrays)
                C.check(call(ctx, "orr_direct_eval_static", arrays, n, eps))
                static_few = results(arrays)
                for s, a in zip(static_few, alloc_few):
                    assert np.array_equal(s, a), n
                ctx.set_direct_config(block_target=default["block_target"])
            for a in arrays:
                a.release()

        # Past the maximum: refused before anything is launched.
        n = max_count.value + 1
        p = W.plummer(n, 23)
        arrays = device_arrays(ctx, p)
        status = call(ctx, "orr_direct_eval_static", arrays, n, eps)
        assert status == C.ORR_ERR_LENGTH, status
        assert "32768" in lib.orr_last_error().decode()
        # The refusal leaves the entry point usable.
        C.check(call(ctx, "orr_direct_eval_static", arrays, 2000, eps))
        got = [a[:2000] for a in results(arrays)]
        ref = oracle.reference_acceleration(p.x[:2000], p.y[:2000], p.z[:2000], p.m[:2000], eps * eps)
        assert relative(got, ref).max() <= BOUND
        for a in arrays:
            a.release()
    finally:
        ctx.set_direct_config(block_target=default["block_target"])
