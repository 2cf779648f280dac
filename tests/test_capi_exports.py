"""The C ABI library loads on a machine without a GPU and exports every symbol
include/orrery_hip.h declares; no compute entry point is called here."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "orrery_hip.h")
LIB = os.path.join(ROOT, "orrery_amd", "liborrery_hip.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^ORR_API\s+[\w\s\*]+?\b(orr_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "orr_direct_eval" in names and "orr_tree_eval" in names
    assert len(names) >= 20


@pytest.mark.skipif(not os.path.exists(LIB), reason="engine library not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    # Nothing else leaks from the library.
    extra = sorted(n for n in exported if not n.startswith("orr_"))
    assert not extra, extra


@pytest.mark.skipif(not os.path.exists(LIB), reason="engine library not built")
def test_library_loads_and_reports_no_device_cleanly():
    import orrery_amd as E
    from orrery_amd import _capi
    lib = _capi.load()
    for name in declared():
        assert hasattr(lib, name)
    if E.device_count() == 0:
        h = ctypes.c_void_p()
        status = lib.orr_ctx_create(0, ctypes.byref(h))
        assert status == _capi.ORR_ERR_NO_DEVICE
        assert b"device" in lib.orr_last_error()
        assert E.HipDirectSolver.try_create(0.1) is None
        assert E.HipTreeSolver.try_create() is None
