"""The C++20 ForceSolver hosts (orrery_amd/host) over the C ABI.

* CPU: the hosts compile against the reference's own interface headers (the
  drop-in claim), and without a device try_create returns null
  (cuda_direct_solver_test.cpp:132-136).
* GPU: tests/cpp/test_host_solvers runs the reference's closed-form and
  agreement cases through HipDirectSolver / HipTreeSolver, once as built and
  once built with AddressSanitizer and UndefinedBehaviorSanitizer.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINARY = os.path.join(ROOT, "tests", "cpp", "test_host_solvers")
REFERENCE_INCLUDE = "/root/reference/include"


@pytest.mark.skipif(not os.path.isdir(REFERENCE_INCLUDE), reason="reference checkout absent")
def test_hosts_compile_against_reference_headers():
    cmd = ["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-Werror", "-I", REFERENCE_INCLUDE,
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "orrery_amd", "host", "include"),
           os.path.join(ROOT, "orrery_amd", "host", "hip_solvers.cpp"),
           os.path.join(ROOT, "orrery_amd", "host", "hip_assembly.cpp")]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


@pytest.mark.skipif(not os.path.exists(BINARY), reason="host test driver not built")
def test_try_create_is_null_without_device():
    import orrery_amd as E
    if E.device_count() != 0:
        pytest.skip("a device is present")
    res = subprocess.run([BINARY, "--expect-no-device"], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stdout + res.stderr


@pytest.mark.gpu
def test_host_solvers_on_device():
    assert os.path.exists(BINARY), "build first: python -m orrery_amd.build"
    res = subprocess.run([BINARY], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr


@pytest.mark.gpu
def test_host_solvers_on_device_under_asan_and_ubsan():
    # The driver and hosts built with AddressSanitizer and UndefinedBehaviorSanitizer
    # (host code only). Leaks inside the ROCm runtime's own libraries are
    # suppressed by library name (tests/cpp/lsan_rocm.supp); any other leak,
    # bad access or undefined behaviour fails the run.
    binary = BINARY + "_asan"
    assert os.path.exists(binary), "build first: python -m orrery_amd.build"
    supp = os.path.join(ROOT, "tests", "cpp", "lsan_rocm.supp")
    env = {**os.environ, "LSAN_OPTIONS": f"suppressions={supp}"}
    res = subprocess.run([binary], capture_output=True, text=True, timeout=300, env=env)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "all host-solver checks passed" in res.stdout
    assert "ERROR: AddressSanitizer" not in res.stderr and "runtime error" not in res.stderr, res.stderr
