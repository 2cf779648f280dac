"""Checkpoint and trajectory files from device-resident state.

Format parity is checked with the reference's own reader and writer
(sim/checkpoint.cpp, sim/trajectory.cpp, built into oracle/_ref), and the
trajectory also with tests/golden/otj.py, which reads the reference's committed
fixtures. The resume contract (docs/formats/checkpoint.md: a resumed run is a
copy of the stored state) is checked bit for bit on the device.
"""
import os

import numpy as np
import pytest

import oracle
import orrery_amd as E
from tests.golden import otj

pytestmark = pytest.mark.gpu

NAMES = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")
needs_ref = pytest.mark.skipif(oracle.load_reference() is None,
                               reason="the reference library (oracle/_ref) is not built here")


def stepped(ctx, n=3000, seed=2, steps=5, integrator="yoshida4", solver="barnes-hut"):
    p = oracle.plummer(n, seed)
    s = E.DeviceSystem(ctx, n)
    s.upload(*p.arrays())
    cfg = E.StepConfig(integrator=integrator, solver=solver, softening=0.02, leaf_capacity=16)
    s.refresh(cfg)
    s.step(cfg, 1e-3, steps)
    return s, p, cfg


@pytest.mark.parametrize("integrator", ["velocity-verlet", "yoshida4", "runge-kutta4"])
@pytest.mark.parametrize("solver", ["direct", "barnes-hut"])
def test_resume_from_checkpoint_is_bitwise(ctx, tmp_path, integrator, solver):
    # simulation_test.cpp:225-318: for every integrator and solver, a run resumed
    # from its checkpoint continues bit for bit (the stored accelerations are
    # current, so the resumed run skips the refresh).
    a, p, cfg = stepped(ctx, integrator=integrator, solver=solver)
    path = str(tmp_path / "run.ock")
    a.checkpoint(path, 5, "[run]\ntimestep = 0.001\n")
    b, step, text = E.DeviceSystem.from_checkpoint(ctx, path)
    assert step == 5 and b.count == a.count and text == "[run]\ntimestep = 0.001\n"
    assert not os.path.exists(path + ".partial")
    a.step(cfg, 1e-3, 4)
    b.step(cfg, 1e-3, 4)
    sa, sb = a.download(), b.download()
    for k in NAMES:
        assert np.array_equal(sa[k], sb[k]), k
    a.close()
    b.close()


@needs_ref
def test_reference_reads_engine_checkpoint(ctx, tmp_path):
    a, p, _ = stepped(ctx, n=1500, solver="direct", integrator="velocity-verlet")
    # A configuration the reference itself wrote, so its parser accepts it.
    seed_path = str(tmp_path / "seed.ock")
    assert oracle.ref_checkpoint_write(seed_path, 0, [np.zeros(1)] * 10) == 0
    text = oracle.ref_checkpoint_read(seed_path, 1)["configuration"]
    path = str(tmp_path / "engine.ock")
    a.checkpoint(path, 123, text)
    r = oracle.ref_checkpoint_read(path, a.count)
    assert "refused" not in r, r
    st = a.download()
    assert r["step"] == 123 and r["count"] == a.count and r["configuration"] == text
    assert np.array_equal(r["m"], p.m)
    for k in NAMES:
        assert np.array_equal(r[k], st[k]), k
    a.close()


@needs_ref
def test_engine_reads_reference_checkpoint(ctx, tmp_path):
    p = oracle.plummer(777, 4)
    acc = oracle.refresh(p, oracle.make_field("vv", "direct", eps2=1e-4))
    path = str(tmp_path / "ref.ock")
    arrays = [p.m, p.x, p.y, p.z, p.vx, p.vy, p.vz, *acc]
    assert oracle.ref_checkpoint_write(path, 42, arrays, "[run]\ntimestep = 0.002\n") == 0
    s, step, text = E.DeviceSystem.from_checkpoint(ctx, path)
    assert step == 42 and s.count == 777 and "timestep = 0.002" in text
    st = s.download()
    for k, ref in zip(NAMES, [p.x, p.y, p.z, p.vx, p.vy, p.vz, *acc]):
        assert np.array_equal(st[k], ref), k
    s.close()


def test_damaged_checkpoints_are_refused(ctx, tmp_path):
    a, _, _ = stepped(ctx, n=200, steps=1)
    path = str(tmp_path / "good.ock")
    a.checkpoint(path, 1)
    raw = open(path, "rb").read()
    cases = {
        "magic": b"ORRERYXX" + raw[8:],
        "version": raw[:8] + (2).to_bytes(4, "little") + raw[12:],
        "flags": raw[:12] + (4).to_bytes(4, "little") + raw[16:],
        "single": raw[:12] + (1).to_bytes(4, "little") + raw[16:],
        "short": raw[:-9],
        "checksum": raw[:200] + bytes([raw[200] ^ 1]) + raw[201:],
    }
    for name, data in cases.items():
        bad = str(tmp_path / f"{name}.ock")
        open(bad, "wb").write(data)
        with pytest.raises(E.solvers.C.EngineError) as e:
            E.DeviceSystem.from_checkpoint(ctx, bad)
        assert e.value.status == 1, name
    a.close()


def test_trajectory_frames(ctx, tmp_path):
    s, p, cfg = stepped(ctx, n=900, steps=0, integrator="velocity-verlet", solver="direct")
    path = str(tmp_path / "run.otj")
    t = E.Trajectory(s, path, 1e-3, velocities=True)
    states = []
    for k in range(3):
        t.append(s, 10 * k)
        states.append(s.download())
        s.step(cfg, 1e-3, 10)
    t.close()
    # The committed reader of the reference's fixtures parses it ...
    got = otj.read(path)
    assert got["n"] == 900 and got["timestep"] == 1e-3
    assert np.array_equal(got["masses"], p.m)
    assert [f["step"] for f in got["frames"]] == [0, 10, 20]
    for f, st in zip(got["frames"], states):
        assert f["time"] == f["step"] * 1e-3
        assert np.array_equal(np.stack(f["positions"]), np.stack([st["x"], st["y"], st["z"]]))
    # ... and so does the reference's own TrajectoryReader.
    r = oracle.ref_trajectory_read(path, 900, 8)
    if r is not None:
        assert "refused" not in r and r["velocities"] and len(r["frames"]) == 3
        for f, st in zip(r["frames"], states):
            want = np.stack([st[k] for k in ("x", "y", "z", "vx", "vy", "vz")])
            assert np.array_equal(f["values"], want)
    s.close()
