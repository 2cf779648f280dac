"""Machine-ceiling probes (include/orrery_hip.h orr_probe), the rates the bench
prices its kernels against (the reference's benchmarks/cuda_probes.cu:15-117)."""
import ctypes
import math

import pytest

from orrery_amd import _capi as C

pytestmark = pytest.mark.gpu


def test_probe_rates_are_consistent(ctx):
    ctx.probe("fma", 256)  # the first launch pays for clocks ramping up
    fma, fma_ms = ctx.probe("fma", 4096)
    rsq, _ = ctx.probe("rsq", 1024)
    mix, _ = ctx.probe("mix", 1024)
    read, _ = ctx.probe("read", 1 << 28)
    copy, _ = ctx.probe("copy", 1 << 28)
    for rate in (fma, rsq, mix, read, copy):
        assert math.isfinite(rate) and rate > 0
    assert fma_ms > 0
    # One interaction is 12 FP32 VALU operations and one v_rsq_f32: its rate is
    # bounded by either part alone (an FMA is 2 flop).
    assert mix <= 1.05 * rsq and mix <= 1.05 * (fma / 2) / 12
    # Bandwidths within what HBM3E can carry (8 TB/s).
    assert read < 8.5e12 and copy < 8.5e12


def test_probe_arguments_are_checked(ctx):
    rate, ms = ctypes.c_double(), ctypes.c_double()
    lib = C.load()
    for kind, size in ((5, 16), (4, 0), (4, (1 << 24) + 1), (0, (1 << 24) + 1)):
        status = lib.orr_probe(ctx.handle, kind, size, ctypes.byref(rate), ctypes.byref(ms))
        assert status == C.ORR_ERR_INVALID_ARGUMENT, (kind, size)
