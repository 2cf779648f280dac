"""The reference's own Python extension driving the engine.

orrery_amd/host/reference_cli.mk also builds the reference's pybind11 module
(python/src, imported as `_orrery`) over the same patched library:
SolverKind gains hip_direct / hip_tree (python/src/sim_bindings.cpp:62-68 with
reference.patch), and `assemble` reaches the engine through the patched
make_solver (src/sim/assembly.cpp:246-272). These tests use it as a notebook
would.

* CPU: the kinds parse from their spellings and from a configuration document;
  without a device `assemble` writes the reference's note to
  sys.stderr (sim_bindings.cpp:40-54) and runs the CPU equivalent, whose steps
  equal those of a run that asked for it, bit for bit, or raises when
  allow_cpu_fallback is false.
* GPU: `assemble` builds hip-direct / hip-tree; ten steps against the same run
  on the module's own CPU solvers, read through the NumPy views of the state.
"""
import contextlib
import io
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODULE_DIR = os.path.join(ROOT, "orrery_amd", "host", "_cli")


def module():
    if not os.path.isdir(MODULE_DIR) or not any(
            name.startswith("_orrery") and name.endswith(".so") for name in os.listdir(MODULE_DIR)):
        return None
    if MODULE_DIR not in sys.path:
        sys.path.insert(0, MODULE_DIR)
    import _orrery
    return _orrery


o = module()
needs_module = pytest.mark.skipif(o is None, reason="reference extension not built (needs the reference checkout)")

DOCUMENT = """[run]
timestep = 0.001
steps    = 10
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
kind = velocity-verlet
"""


def configuration(kind, count=1024, extra=""):
    return o.parse_configuration(DOCUMENT.format(kind=kind, count=count, extra=extra))


def assemble(c):
    err = io.StringIO()
    with contextlib.redirect_stderr(err):
        sim = o.assemble(c)
    return sim, err.getvalue()


def state(sim):
    p = sim.particles
    names = ("position_x", "position_y", "position_z", "velocity_x", "velocity_y", "velocity_z",
             "acceleration_x", "acceleration_y", "acceleration_z", "mass")
    return {k: np.array(getattr(p, k)) for k in names}


def no_device():
    import orrery_amd as E
    return E.device_count() == 0


@needs_module
def test_kinds_parse():
    assert o.parse_solver_kind("hip-tree") == o.SolverKind.hip_tree
    assert o.parse_solver_kind("hip-direct") == o.SolverKind.hip_direct
    assert o.SolverKind.hip_tree.name == "hip_tree" and o.SolverKind.hip_direct.name == "hip_direct"
    c = configuration("hip-tree", extra="opening_angle = 0.6\nquadrupole = true\n")
    assert c.solver.kind == o.SolverKind.hip_tree
    assert c.solver.opening_angle == 0.6 and c.solver.quadrupole


@needs_module
@pytest.mark.parametrize("kind,cpu", [("hip-tree", "barnes-hut"), ("hip-direct", "direct")])
def test_fallback_without_a_device(kind, cpu):
    if not no_device():
        pytest.skip("a device is present")
    sim, note = assemble(configuration(kind))
    assert note == (f"note: {kind} was asked for but no usable device was found, so this run uses "
                    f"{cpu} on the CPU\n")
    assert sim.solver_name == cpu
    ref, _ = assemble(configuration(cpu))
    sim.run(10)
    ref.run(10)
    got, want = state(sim), state(ref)
    for key in want:
        assert np.array_equal(got[key], want[key]), key


@needs_module
def test_refusal_without_a_device():
    if not no_device():
        pytest.skip("a device is present")
    with pytest.raises(RuntimeError, match="solver.kind is hip-tree and no usable device was found"):
        assemble(configuration("hip-tree", extra="allow_cpu_fallback = false\n"))


def relative(got, want):
    a = np.stack([got["acceleration_x"], got["acceleration_y"], got["acceleration_z"]])
    b = np.stack([want["acceleration_x"], want["acceleration_y"], want["acceleration_z"]])
    return np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,cpu,extra", [
    ("hip-direct", "direct", ""),
    ("hip-tree", "barnes-hut", "opening_angle = 0.5\nleaf_capacity = 16\nquadrupole = true\n"),
], ids=["hip-direct", "hip-tree"])
def test_engine_runs_from_python(kind, cpu, extra):
    assert o is not None, "reference extension not built"
    sim, note = assemble(configuration(kind, count=4096, extra=extra))
    assert note == "" and sim.solver_name == kind
    ref, _ = assemble(configuration(cpu, count=4096, extra=extra))
    sim.run(10)
    ref.run(10)
    assert sim.step_index == ref.step_index == 10
    got, want = state(sim), state(ref)
    rel = relative(got, want)
    worst, rms = float(rel.max()), float(np.sqrt(np.mean(rel * rel)))
    count = sim.interaction_count
    print(f"{kind} from Python after 10 steps: worst {worst:.2e} rms {rms:.2e} against {cpu}; "
          f"{count.particle_particle} pp, {count.particle_cell} pc")
    # hip-direct: the engine's direct tolerance (1e-6 per particle); hip-tree:
    # the same tree in FP32 against FP64 arithmetic (RMS 1.6e-7 measured
    # through the command-line program on the same settings).
    if kind == "hip-direct":
        assert worst <= 1e-6 and rms <= 1e-7
        assert count.particle_particle == 11 * 4096 * 4095
    else:
        assert rms <= 1e-6
    assert np.array_equal(got["mass"], want["mass"])
    assert np.max(np.abs(got["position_x"] - want["position_x"])) <= 1e-8
