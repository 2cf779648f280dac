"""The FP64 close-pair term of the direct kernels, on the CPU.

orrery_amd/csrc/close_pair.hpp is plain FP64 arithmetic (the separation from
the positions' (hi, lo) float parts, a reciprocal square root seed and two
Newton steps), so the host compiler runs the same steps as the device from a
correctly rounded seed; the harness tests/cpp/test_close_pair.cpp checks the
term, split into two floats as the kernels add it, against long-double
arithmetic on 200,000 pairs whose separations reach down to the positions' own
rounding.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_close_pair_term_matches_long_double(tmp_path):
    exe = tmp_path / "test_close_pair"
    src = os.path.join(ROOT, "tests", "cpp", "test_close_pair.cpp")
    subprocess.run(["g++", "-std=c++20", "-O2", "-ffp-contract=off", "-Wall", "-Wno-unknown-pragmas",
                    src, "-o", str(exe)], check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
