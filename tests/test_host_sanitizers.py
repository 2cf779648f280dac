"""Host code under the compiler's sanitizers (g++ only; nothing here is GPU code).

* The C++ ForceSolver hosts (orrery_amd/host) and their test driver, built
  with AddressSanitizer and UndefinedBehaviorSanitizer over the engine's
  library, on the path that needs no device: try_create returning null, the
  assembly step's spellings, refusal and fallback note.
* The oracle (oracle/oracle.cpp, oracle_avx2.cpp), driven through every entry
  point by tests/cpp/oracle_sanitize.cpp, once with AddressSanitizer and
  UndefinedBehaviorSanitizer and once with ThreadSanitizer (its restated
  work-stealing executor runs the threaded cases).
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CXX = os.environ.get("CXX", "g++")
ENGINE = os.path.join(ROOT, "orrery_amd", "liborrery_hip.so")
ASAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]
ORACLE_FLAGS = ["-std=c++20", "-O1", "-g", "-pthread", "-ffp-contract=off", "-fno-builtin-sin",
                "-fno-builtin-cos", "-I", os.path.join(ROOT, "oracle")]


def build(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, " ".join(cmd) + "\n" + res.stderr


def run(binary, *args, env=None):
    res = subprocess.run([binary, *args], capture_output=True, text=True, timeout=600,
                         env={**os.environ, **(env or {})})
    assert res.returncode == 0, res.stdout + res.stderr
    assert "Sanitizer" not in res.stderr, res.stderr
    return res.stdout


@pytest.mark.skipif(not os.path.exists(ENGINE), reason="engine library not built")
def test_host_solvers_under_asan_and_ubsan(tmp_path):
    import orrery_amd as E
    if E.device_count() != 0:
        pytest.skip("a device is present: the no-device path is what runs here")
    binary = str(tmp_path / "host_asan")
    host = os.path.join(ROOT, "orrery_amd", "host")
    build([CXX, "-std=c++20", "-O1", "-g", *ASAN, "-Wall",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(host, "include"),
           "-I", os.path.join(host, "compat"),
           os.path.join(host, "hip_solvers.cpp"), os.path.join(host, "hip_assembly.cpp"),
           os.path.join(ROOT, "tests", "cpp", "test_host_solvers.cpp"),
           "-L", os.path.dirname(ENGINE), "-lorrery_hip", "-Wl,-rpath," + os.path.dirname(ENGINE),
           "-o", binary])
    assert "all host-solver checks passed" in run(binary, "--expect-no-device")


def oracle_binary(tmp_path, name, sanitize):
    objs = []
    for src, extra in (("oracle.cpp", []), ("oracle_avx2.cpp", ["-mavx2", "-mfma"])):
        obj = str(tmp_path / (name + "_" + src.replace(".cpp", ".o")))
        build([CXX, *ORACLE_FLAGS, *sanitize, *extra, "-c", os.path.join(ROOT, "oracle", src), "-o", obj])
        objs.append(obj)
    binary = str(tmp_path / name)
    build([CXX, *ORACLE_FLAGS, *sanitize, os.path.join(ROOT, "tests", "cpp", "oracle_sanitize.cpp"), *objs,
           "-o", binary])
    return binary


def test_oracle_under_asan_and_ubsan(tmp_path):
    assert "oracle sanitizer run passed" in run(oracle_binary(tmp_path, "oracle_asan", ASAN))


def test_oracle_under_tsan(tmp_path):
    binary = oracle_binary(tmp_path, "oracle_tsan", ["-fsanitize=thread"])
    assert "oracle sanitizer run passed" in run(binary, env={"TSAN_OPTIONS": "halt_on_error=1"})
