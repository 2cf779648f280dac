"""The reference's own simulator driving the engine.

orrery_amd/host/reference_cli.mk builds the reference's `orrery` program
(apps/orrery.cpp) with orrery_amd/host/reference.patch applied: SolverKind
gains hip-direct and hip-tree (include/orrery/sim/configuration.hpp:56-61,
src/sim/configuration.cpp:31-33, 60-61, 101-112), and make_solver
(src/sim/assembly.cpp:246-272) hands them to the engine's assembly step. These
tests run that program on configuration files as a user would.

* CPU: the patch applies to the reference as it stands; the parser accepts and
  lists the two words and treats hip-tree as a tree solver; without a device a
  run refuses (allow_cpu_fallback = false) with the reference's wording, or
  prints the reference's note and runs the CPU equivalent, whose final state
  equals that of a run that asked for the CPU solver, bit for bit.
* GPU: hip-direct against the program's own FP64 direct solver and hip-tree
  against its Barnes-Hut, on the final state of a Plummer run; a run resumed
  from a checkpoint taken half way ends in the state of the uninterrupted run,
  bit for bit (the guarantee of docs/formats/checkpoint.md, asserted by
  tests/sim/simulation_test.cpp for the CPU solvers).
* GPU: the reference's example runs (examples/collision.orrery, 60,000
  particles under Barnes-Hut at theta 0.6, and examples/kepler.orrery, the
  two-body orbit under direct summation), fewer steps, on the engine and on
  the program's CPU solvers: final forces, energy error and time per step.
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINARY = os.path.join(ROOT, "orrery_amd", "host", "_cli", "orrery")
PATCH = os.path.join(ROOT, "orrery_amd", "host", "reference.patch")
REFERENCE = "/root/reference"
PATCHED_FILES = ["include/orrery/sim/configuration.hpp", "src/sim/configuration.cpp",
                 "src/sim/assembly.cpp", "src/solvers/CMakeLists.txt", "python/src/sim_bindings.cpp"]

needs_binary = pytest.mark.skipif(not os.path.exists(BINARY),
                                  reason="reference simulator not built (needs the reference checkout)")


def configuration(kind, count=1024, steps=10, integrator="velocity-verlet", extra=""):
    return f"""[run]
timestep = 0.001
steps    = {steps}
seed     = 20260811

[initial_conditions]
kind  = plummer
count = {count}

[solver]
kind      = {kind}
softening = 0.05
threads   = 8
{extra}
[integrator]
kind = {integrator}

[output]
checkpoint_path   = state.ock
checkpoint_stride = {steps}
"""


def run(args, cwd, expect=0):
    res = subprocess.run([BINARY, *args], cwd=cwd, capture_output=True, text=True, timeout=300)
    assert res.returncode == expect, res.stdout + res.stderr
    return res


def fnv1a64(data: bytes) -> int:
    h = 0xCBF29CE484222325
    for b in data:
        h = ((h ^ b) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def read_checkpoint(path):
    """docs/formats/checkpoint.md: header, configuration text, ten FP64 arrays
    (m, x, y, z, vx, vy, vz, ax, ay, az), FNV-1a of everything before it."""
    raw = open(path, "rb").read()
    assert raw[:8] == b"ORRERYCK"
    version, flags, step, n, length = struct.unpack_from("<IIQQI", raw, 8)
    assert version == 1 and flags == 0
    start = 36 + length
    end = start + 10 * n * 8
    if end <= 1 << 20:  # a byte-at-a-time hash; the program's own reader checks large files
        assert fnv1a64(raw[:end]) == struct.unpack_from("<Q", raw, end)[0]
    arrays = np.frombuffer(raw[start:end], dtype="<f8").reshape(10, n)
    names = ("m", "x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")
    return step, raw[36:start].decode(), dict(zip(names, arrays))


def simulate(tmp, name, kind, **kw):
    d = os.path.join(tmp, name)
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "run.orrery"), "w") as f:
        f.write(configuration(kind, **kw))
    res = run(["run", "run.orrery"], d)
    return res.stdout, read_checkpoint(os.path.join(d, "state.ock"))


def no_device():
    import orrery_amd as E
    return E.device_count() == 0


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="reference checkout absent")
def test_patch_applies_to_the_reference(tmp_path):
    for rel in PATCHED_FILES:
        os.makedirs(os.path.dirname(tmp_path / rel), exist_ok=True)
        shutil.copy(os.path.join(REFERENCE, rel), tmp_path / rel)
    with open(PATCH) as f:
        res = subprocess.run(["patch", "-p1", "--dry-run", "-d", str(tmp_path)], stdin=f,
                             capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    touched = sorted(line.split()[2] for line in res.stdout.splitlines() if line.startswith("checking file"))
    assert touched == sorted(PATCHED_FILES)
    # Patched, the two sim sources also compile in a build without the engine
    # (ORRERY_ENABLE_HIP undefined), the reference's default.
    with open(PATCH) as f:
        subprocess.run(["patch", "-p1", "-s", "-d", str(tmp_path)], stdin=f, check=True)
    for src in ("src/sim/assembly.cpp", "src/sim/configuration.cpp"):
        res = subprocess.run([os.environ.get("CXX", "g++"), "-std=c++20", "-fsyntax-only", "-Wall", "-Werror",
                              "-I", str(tmp_path / "include"), "-I", os.path.join(REFERENCE, "include"),
                              str(tmp_path / src)], capture_output=True, text=True)
        assert res.returncode == 0, src + res.stderr


@needs_binary
def test_parser_accepts_and_lists_the_hip_kinds(tmp_path):
    with open(tmp_path / "run.orrery", "w") as f:
        f.write(configuration("hip-tree"))
    shown = run(["show", "run.orrery"], tmp_path).stdout
    assert "kind = hip-tree" in shown
    # hip-tree reads the tree settings; for hip-direct an opening angle is a
    # problem with the configuration, as it is for direct.
    quiet = run(["show", "run.orrery", "--set", "solver.opening_angle=0.7"], tmp_path)
    assert "opening_angle" not in quiet.stderr
    loud = run(["show", "run.orrery", "--set", "solver.kind=hip-direct", "--set", "solver.opening_angle=0.7"],
               tmp_path, expect=1)
    assert "solver.opening_angle: is not used by a direct solver" in loud.stderr
    bad = run(["show", "run.orrery", "--set", "solver.kind=hip-trees"], tmp_path, expect=1)
    assert ("unknown solver.kind 'hip-trees'. The choices are direct, barnes-hut, sycl-direct, "
            "sycl-tree, hip-direct, hip-tree") in bad.stderr


@needs_binary
def test_refusal_without_a_device(tmp_path):
    if not no_device():
        pytest.skip("a device is present")
    with open(tmp_path / "run.orrery", "w") as f:
        f.write(configuration("hip-direct", extra="allow_cpu_fallback = false\n"))
    res = run(["run", "run.orrery"], tmp_path, expect=1)
    assert ("solver.kind is hip-direct and no usable device was found. Set "
            "solver.allow_cpu_fallback = true to run on the CPU instead") in res.stderr
    assert not os.path.exists(tmp_path / "state.ock")


@needs_binary
@pytest.mark.parametrize("kind,cpu", [("hip-tree", "barnes-hut"), ("hip-direct", "direct")])
def test_fallback_runs_the_cpu_equivalent(tmp_path, kind, cpu):
    if not no_device():
        pytest.skip("a device is present")
    out, (step, text, got) = simulate(tmp_path, "hip", kind)
    assert out.startswith(f"note: {kind} was asked for but no usable device was found, so this run "
                          f"uses {cpu} on the CPU\n")
    assert f"solver:      {cpu}," in out
    _, (step1, _, want) = simulate(tmp_path, "cpu", cpu)
    assert step == step1 == 10 and f"kind = {kind}" in text
    for key in want:
        assert np.array_equal(got[key], want[key]), key


# ---- on the device ---------------------------------------------------------

def relative(got, want):
    a = np.stack([got["ax"], got["ay"], got["az"]])
    b = np.stack([want["ax"], want["ay"], want["az"]])
    return np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)


@pytest.mark.gpu
def test_hip_direct_run_against_cpu_direct(tmp_path):
    assert os.path.exists(BINARY), "reference simulator not built"
    out, (_, _, got) = simulate(tmp_path, "hip", "hip-direct", count=4096)
    print(out)
    assert "solver:      hip-direct," in out and "note:" not in out
    assert "interactions: 184504320 particle-particle, 0 particle-cell" in out   # 11 x N(N-1)
    _, (_, _, want) = simulate(tmp_path, "cpu", "direct", count=4096)
    rel = relative(got, want)
    worst, rms = float(rel.max()), float(np.sqrt(np.mean(rel * rel)))
    print(f"hip-direct vs direct after 10 steps: worst {worst:.2e} rms {rms:.2e}")
    # The engine's direct tolerance, 1e-6 relative per particle; ten steps of
    # dt = 1e-3 move the states apart by far less than that changes a force.
    assert worst <= 1e-6 and rms <= 1e-7
    for key in ("x", "y", "z"):
        assert np.max(np.abs(got[key] - want[key])) <= 1e-10, key
    assert np.array_equal(got["m"], want["m"])


@pytest.mark.gpu
def test_hip_tree_run_against_cpu_tree_and_direct(tmp_path):
    assert os.path.exists(BINARY), "reference simulator not built"
    tree = "opening_angle = 0.5\nleaf_capacity = 16\nquadrupole = true\n"
    out, (_, _, got) = simulate(tmp_path, "hip", "hip-tree", count=4096, extra=tree)
    print(out)
    assert "solver:      hip-tree," in out and "note:" not in out
    _, (_, _, cpu) = simulate(tmp_path, "cpu", "barnes-hut", count=4096, extra=tree)
    vs_cpu = relative(got, cpu)
    # The physics: both trees against FP64 direct summation of the engine's
    # final positions, the north-star tree bound (RMS 1e-3 at theta 0.5).
    p = got
    ref = oracle.reference_acceleration(p["x"], p["y"], p["z"], p["m"], 0.05 * 0.05)
    exact = {"ax": ref[0], "ay": ref[1], "az": ref[2]}
    vs_exact = relative(got, exact)
    rms_cpu = float(np.sqrt(np.mean(vs_cpu * vs_cpu)))
    rms_exact = float(np.sqrt(np.mean(vs_exact * vs_exact)))
    print(f"hip-tree after 10 steps: vs barnes-hut rms {rms_cpu:.2e}, vs direct rms {rms_exact:.2e}")
    assert rms_exact <= 1e-3
    # Same tree, same acceptance rule: the two trees' forces differ by FP32
    # against FP64 arithmetic, far below the tree's own error.
    assert rms_cpu <= 1e-6
    for key in ("x", "y", "z"):
        assert np.max(np.abs(got[key] - cpu[key])) <= 1e-8, key


@pytest.mark.gpu
@pytest.mark.parametrize("kind,integrator", [("hip-tree", "yoshida4"), ("hip-direct", "velocity-verlet")])
def test_resumed_run_equals_uninterrupted_run(tmp_path, kind, integrator):
    assert os.path.exists(BINARY), "reference simulator not built"
    _, (step, _, whole) = simulate(tmp_path, "whole", kind, count=2048, steps=8, integrator=integrator)
    d = os.path.join(tmp_path, "split")
    os.makedirs(d)
    with open(os.path.join(d, "run.orrery"), "w") as f:
        f.write(configuration(kind, count=2048, steps=8, integrator=integrator).replace(
            "checkpoint_stride = 8", "checkpoint_stride = 4"))
    run(["run", "run.orrery", "--set", "run.steps=4"], d)
    assert read_checkpoint(os.path.join(d, "state.ock"))[0] == 4
    out = run(["resume", "state.ock", "--set", "run.steps=8"], d).stdout
    assert f"solver:      {kind}," in out
    step1, _, resumed = read_checkpoint(os.path.join(d, "state.ock"))
    assert step == step1 == 8
    for key in whole:
        assert np.array_equal(whole[key], resumed[key]), key


# The reference's own example runs (examples/collision.orrery and
# examples/kepler.orrery: their settings, fewer steps, no trajectory or
# diagnostics files), through the program on the engine and on the CPU.
COLLISION = """[run]
timestep = 0.005
steps    = {steps}
seed     = 20260812

[initial_conditions]
kind                  = galaxy-collision
count                 = 60000
total_mass            = 1.5
bulge_fraction        = 0.2
scale_length          = 1
scale_height          = 0.1
bulge_radius          = 0.2
mass_ratio            = 0.5
inclination           = 0.35
secondary_inclination = 1.2
separation            = 20
impact_parameter      = 2
approach_speed        = 0.8

[solver]
kind          = {kind}
softening     = 0.12
opening_angle = 0.6
threads       = 16

[integrator]
kind = velocity-verlet

[output]
checkpoint_path   = state.ock
checkpoint_stride = {steps}
"""

KEPLER = """[run]
timestep = 0.02221441469
steps    = {steps}
seed     = 1

[initial_conditions]
kind            = kepler
primary_mass    = 1
secondary_mass  = 1
semi_major_axis = 1
eccentricity    = 0.5

[solver]
kind    = {kind}
threads = 1

[integrator]
kind = velocity-verlet

[output]
checkpoint_path   = state.ock
checkpoint_stride = {steps}
"""


def run_text(tmp, name, text):
    d = os.path.join(tmp, name)
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "run.orrery"), "w") as f:
        f.write(text)
    out = run(["run", "run.orrery"], d).stdout
    per_step = float(out.split(" ms per step")[0].rsplit(", ", 1)[1])
    energy = float(out.split("relative energy error: ")[1].split()[0])
    return out, per_step, energy, read_checkpoint(os.path.join(d, "state.ock"))[2]


@pytest.mark.gpu
def test_collision_example_on_hip_tree(tmp_path):
    assert os.path.exists(BINARY), "reference simulator not built"
    steps = 40
    out, ms_gpu, de_gpu, got = run_text(tmp_path, "hip", COLLISION.format(steps=steps, kind="hip-tree"))
    assert "solver:      hip-tree," in out and "note:" not in out
    _, ms_cpu, de_cpu, cpu = run_text(tmp_path, "cpu", COLLISION.format(steps=steps, kind="barnes-hut"))
    rel = relative(got, cpu)
    rms = float(np.sqrt(np.mean(rel * rel)))
    print(f"collision example, 60000 particles, {steps} steps: hip-tree {ms_gpu:.2f} ms/step, "
          f"barnes-hut (16 threads) {ms_cpu:.2f} ms/step; energy error {de_gpu:.4e} vs {de_cpu:.4e}; "
          f"final forces RMS {rms:.2e} apart")
    # Both trees at theta 0.6 against each other: FP32 against FP64 terms
    # (measured RMS 5.6e-6 after 40 steps).
    assert rms <= 2e-5
    assert abs(de_gpu - de_cpu) <= 1e-3 * abs(de_cpu) + 1e-9


@pytest.mark.gpu
def test_kepler_example_on_hip_direct(tmp_path):
    assert os.path.exists(BINARY), "reference simulator not built"
    steps = 2000
    out, ms_gpu, de_gpu, got = run_text(tmp_path, "hip", KEPLER.format(steps=steps, kind="hip-direct"))
    assert "solver:      hip-direct," in out and "note:" not in out
    _, ms_cpu, de_cpu, cpu = run_text(tmp_path, "cpu", KEPLER.format(steps=steps, kind="direct"))
    dx = max(float(np.max(np.abs(got[k] - cpu[k]))) for k in ("x", "y", "z"))
    print(f"kepler example, {steps} steps (10 orbits): hip-direct {ms_gpu:.3f} ms/step, "
          f"direct {ms_cpu:.4f} ms/step; energy error {de_gpu:.6e} vs {de_cpu:.6e}; positions {dx:.2e} apart")
    # The two particles span their tile, so their pair is not a close pair and
    # its term is FP32 (relative error a few 1e-8): over 10 orbits (200 steps each) of
    # e = 0.5 that moves the phase by about 1.7e-5 and the energy error, which
    # the integrator's truncation sets at 1.5e-3, by a few parts in 1e4.
    assert dx <= 1e-4
    assert abs(de_gpu - de_cpu) <= 1e-3 * abs(de_cpu)
