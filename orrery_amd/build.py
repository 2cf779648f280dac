"""Build the engine's shared library for gfx950 with hipcc.

    python -m orrery_amd.build          # -> orrery_amd/liborrery_hip.so

Every HIP translation unit is compiled separately (so one edit rebuilds one
object) and linked into one shared library whose only exports are the C ABI
of include/orrery_hip.h.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBRARY = os.path.join(HERE, "liborrery_hip.so")
SOURCES = ["context.hip", "capi.hip", "direct.hip", "morton.hip", "scan.hip", "tree.hip", "batch.hip", "serve.hip", "io.hip",
           "integrate.hip", "probes.hip"]
# Host-only translation units, compiled by the host compiler: the seeded
# initial-condition sampler must make the same libm calls, unfused, as the
# reference's CPU build.
HOST_SOURCES = ["sample.cpp"]
HOST_FLAGS = ["-std=c++20", "-O2", "-fPIC", "-fvisibility=hidden", "-Wall", "-ffp-contract=off",
              "-fno-builtin-sin", "-fno-builtin-cos"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops: no v_pk_*_f32 anywhere (scalar FP32 only).
# -fno-fast-math and no -ffast-math: IEEE semantics; contraction is decided per
# file with `#pragma clang fp contract`.
FLAGS = [
    "-std=c++20", "-O3", "-fPIC", f"--offload-arch={ARCH}", "-fno-gpu-rdc",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-result",
    "-munsafe-fp-atomics",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if src.endswith(".cpp"):
        cmd = [os.environ.get("CXX", "g++"), *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compiling {src} failed:\n{res.stderr}")
    return obj


ROOT = os.path.dirname(HERE)
HOST_LIBRARY = os.path.join(HERE, "liborrery_host.so")
HOST_TEST = os.path.join(ROOT, "tests", "cpp", "test_host_solvers")
HOST_TEST_ASAN = HOST_TEST + "_asan"
CXX = os.environ.get("CXX", "g++")
HOST_INCLUDES = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "host", "include"),
                 "-I", os.path.join(HERE, "host", "compat")]


def _run(cmd: list[str], what: str) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"{what} failed:\n{' '.join(cmd)}\n{res.stderr}")


def build_host(verbose: bool = False) -> None:
    """C++20 ForceSolver hosts over the C ABI, and their C++ test driver."""
    tmp = HOST_LIBRARY + ".tmp"
    _run([CXX, "-std=c++20", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra", *HOST_INCLUDES,
          os.path.join(HERE, "host", "hip_solvers.cpp"), os.path.join(HERE, "host", "hip_assembly.cpp"),
          "-L", HERE, "-lorrery_hip",
          "-Wl,-rpath,$ORIGIN", "-o", tmp], "host library")
    os.replace(tmp, HOST_LIBRARY)
    tmp = HOST_TEST + ".tmp"
    _run([CXX, "-std=c++20", "-O2", "-Wall", *HOST_INCLUDES,
          os.path.join(ROOT, "tests", "cpp", "test_host_solvers.cpp"), "-L", HERE, "-lorrery_host",
          "-lorrery_hip", "-Wl,-rpath,$ORIGIN/../../orrery_amd", "-o", tmp], "host test driver")
    os.replace(tmp, HOST_TEST)
    # The same driver and hosts under AddressSanitizer and UndefinedBehaviorSanitizer
    # (host code only; the engine library it links is the ordinary build).
    tmp = HOST_TEST_ASAN + ".tmp"
    _run([CXX, "-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
          "-fno-sanitize-recover=undefined", "-Wall", *HOST_INCLUDES,
          os.path.join(HERE, "host", "hip_solvers.cpp"), os.path.join(HERE, "host", "hip_assembly.cpp"),
          os.path.join(ROOT, "tests", "cpp", "test_host_solvers.cpp"), "-L", HERE, "-lorrery_hip",
          "-Wl,-rpath,$ORIGIN/../../orrery_amd", "-o", tmp], "sanitized host test driver")
    os.replace(tmp, HOST_TEST_ASAN)
    if verbose:
        print(f"built {HOST_LIBRARY}, {HOST_TEST} and {HOST_TEST_ASAN}")


REFERENCE = os.environ.get("ORRERY_REFERENCE", "/root/reference")
REFERENCE_CLI = os.path.join(HERE, "host", "_cli", "orrery")


def build_reference_cli(verbose: bool = False) -> None:
    """The reference's own simulator with reference.patch applied and the
    engine linked in (host/reference_cli.mk); only where the reference checkout
    exists, since it is compiled from the reference's sources."""
    if not os.path.isdir(os.path.join(REFERENCE, "src", "sim")):
        if verbose:
            print("reference checkout absent: reference simulator not built")
        return
    _run(["make", "-s", "-f", os.path.join(HERE, "host", "reference_cli.mk"), "-j8",
          f"REFERENCE={REFERENCE}"], "reference simulator")
    if verbose:
        print(f"built {REFERENCE_CLI}")


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=8) as pool:
        objs = list(pool.map(_compile, SOURCES + HOST_SOURCES))
    tmp = LIBRARY + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIBRARY)
    if verbose:
        print(f"built {LIBRARY}")
    build_host(verbose)
    build_reference_cli(verbose)
    return LIBRARY


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
