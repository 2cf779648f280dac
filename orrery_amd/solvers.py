"""Host classes of the MI355X engine, mirroring the reference's solver seam.

``HipDirectSolver`` and ``HipTreeSolver`` implement the ForceSolver contract of
include/orrery/solvers/force_solver.hpp:41-87 and
include/orrery/integrators/acceleration_field.hpp:37-70 over the C ABI:

* ``evaluate(positions, masses, accelerations)`` takes three equal-length SoA
  component arrays for positions (the Vec3Span of core/vec3_span.hpp:33-74), the
  masses, and three arrays it *writes* (never accumulates);
* ``name()``, ``softening()``, ``interaction_count()``, ``reset_interaction_count()``;
* ``try_create(...)`` returns None when there is no usable device, as
  CudaDirectSolver::try_create does (cuda_direct_solver.cpp:140-164).

Host arrays go through device staging (the reference CUDA solvers' H2D, kernel,
D2H). ``evaluate_device`` takes device pointers instead, for state that already
lives in HBM.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import _capi as C


@dataclass(frozen=True)
class Softening:
    """Plummer softening held as its square (core/softening.hpp:50-78)."""

    length_: float = 0.0

    def squared(self) -> float:
        return float(self.length_) * float(self.length_)

    def length(self) -> float:
        return math.sqrt(self.squared())


@dataclass
class InteractionCount:
    """solvers/interaction_count.hpp:37-76."""

    evaluations: int = 0
    particle_particle: int = 0
    particle_cell: int = 0


@dataclass(frozen=True)
class TreeParameters:
    """solvers/octree.hpp:147-199; corrected as octree.cpp:189-194 does."""

    opening_angle: float = 0.5
    leaf_capacity: int = 32
    quadrupole: bool = False

    def corrected(self) -> "TreeParameters":
        return TreeParameters(min(max(float(self.opening_angle), 0.0), 1.0),
                              max(int(self.leaf_capacity), 1), bool(self.quadrupole))


def device_count() -> int:
    n = ctypes.c_int32()
    C.check(C.load().orr_device_count(ctypes.byref(n)))
    return int(n.value)


def describe(device: int = 0) -> dict:
    d = C.OrrDeviceDesc()
    C.check(C.load().orr_device_describe(device, ctypes.byref(d)))
    return dict(name=d.name.decode(), arch=d.arch.decode(), compute_units=d.compute_units,
                wavefront_size=d.wavefront_size, clock_khz=d.clock_khz,
                global_memory_bytes=d.global_memory_bytes, l2_bytes=d.l2_bytes)


class Context:
    """One device, one stream, the engine's scratch (orr_ctx)."""

    def __init__(self, device: int = 0):
        self._lib = C.load()
        h = ctypes.c_void_p()
        C.check(self._lib.orr_ctx_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if getattr(self, "handle", None):
            C.check(self._lib.orr_ctx_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- device memory ---------------------------------------------------
    def alloc(self, nbytes: int) -> int:
        p = ctypes.c_void_p()
        C.check(self._lib.orr_malloc(self.handle, int(nbytes), ctypes.byref(p)))
        return int(p.value or 0)

    def free(self, ptr: int) -> None:
        C.check(self._lib.orr_free(self.handle, ctypes.c_void_p(ptr)))

    def upload(self, dst: int, src: np.ndarray) -> None:
        src = np.ascontiguousarray(src)
        C.check(self._lib.orr_upload(self.handle, ctypes.c_void_p(dst),
                                     src.ctypes.data_as(ctypes.c_void_p), src.nbytes))

    def download(self, dst: np.ndarray, src: int) -> None:
        assert dst.flags.c_contiguous
        C.check(self._lib.orr_download(self.handle, dst.ctypes.data_as(ctypes.c_void_p),
                                       ctypes.c_void_p(src), dst.nbytes))

    # -- direct kernel configuration (autotuner knob) -----------------------
    def set_direct_config(self, targets_per_lane: int = 0, refine: int = -1, block_target: int = 0,
                          pair_mode: int = 0, equal_masses: int = 0, small_order: int = 0) -> None:
        c = C.OrrDirectConfig(targets_per_lane, refine, block_target, pair_mode, equal_masses, small_order)
        C.check(self._lib.orr_ctx_set_direct_config(self.handle, ctypes.byref(c)))

    def direct_config(self) -> dict:
        c = C.OrrDirectConfig()
        C.check(self._lib.orr_ctx_get_direct_config(self.handle, ctypes.byref(c)))
        return dict(targets_per_lane=c.targets_per_lane, refine=c.refine, block_target=c.block_target,
                    pair_mode=c.pair_mode, equal_masses=c.equal_masses, small_order=c.small_order)

    def set_tree_config(self, targets_per_lane: int = 0, lanes: int = 0, sort_mode: int = 0) -> None:
        c = C.OrrTreeConfig(targets_per_lane, lanes, sort_mode)
        C.check(self._lib.orr_ctx_set_tree_config(self.handle, ctypes.byref(c)))

    def tree_config(self) -> dict:
        c = C.OrrTreeConfig()
        C.check(self._lib.orr_ctx_get_tree_config(self.handle, ctypes.byref(c)))
        return dict(targets_per_lane=c.targets_per_lane, lanes=c.lanes, sort_mode=c.sort_mode)

    # -- machine ceilings (benchmarks/cuda_probes.cu equivalents) -----------
    PROBES = {"fma": 0, "rsq": 1, "read": 2, "copy": 3, "mix": 4}

    def probe(self, kind: str, size: int) -> tuple[float, float]:
        """(rate, milliseconds): flop/s (fma), rsq/s (rsq), interactions/s (mix: one
        interaction's 12 VALU + 1 v_rsq_f32 without memory) or bytes/s (read, copy)."""
        rate, ms = ctypes.c_double(), ctypes.c_double()
        C.check(self._lib.orr_probe(self.handle, self.PROBES[kind], int(size), ctypes.byref(rate),
                                    ctypes.byref(ms)))
        return float(rate.value), float(ms.value)

    # -- timings -----------------------------------------------------------
    def set_timing(self, enabled: bool) -> None:
        C.check(self._lib.orr_ctx_set_timing(self.handle, 1 if enabled else 0))

    def timings(self) -> dict:
        t = C.OrrTimings()
        C.check(self._lib.orr_ctx_timings(self.handle, ctypes.byref(t)))
        return {C.PHASE_NAMES[k]: (float(t.milliseconds[k]), int(t.launches[k]))
                for k in range(C.ORR_PHASES) if t.launches[k]}

    def reset_timings(self) -> None:
        C.check(self._lib.orr_ctx_reset_timings(self.handle))

    def graph_launches(self) -> int:
        """Evaluations replayed as one captured hipGraph since the last reset."""
        t = C.OrrTimings()
        C.check(self._lib.orr_ctx_timings(self.handle, ctypes.byref(t)))
        return int(t.graph_launches)


class DeviceArray:
    """A 1-D device allocation of a numpy dtype (backend/cuda_memory.hpp's CudaArray)."""

    def __init__(self, ctx: Context, count: int, dtype=np.float64):
        self.ctx = ctx
        self.count = int(count)
        self.dtype = np.dtype(dtype)
        self.ptr = ctx.alloc(max(1, self.count) * self.dtype.itemsize)

    @classmethod
    def from_host(cls, ctx: Context, host: np.ndarray) -> "DeviceArray":
        a = cls(ctx, host.shape[0], host.dtype)
        if a.count:
            ctx.upload(a.ptr, host)
        return a

    def to_host(self, out: np.ndarray | None = None) -> np.ndarray:
        if out is None:
            out = np.empty(self.count, dtype=self.dtype)
        if self.count:
            self.ctx.download(out, self.ptr)
        return out

    def upload(self, host: np.ndarray) -> None:
        assert host.shape[0] == self.count and host.dtype == self.dtype
        if self.count:
            self.ctx.upload(self.ptr, host)

    def release(self) -> None:
        if self.ptr:
            self.ctx.free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class _Staging:
    """Grow-only device staging for host-array evaluations (cuda_direct_solver.cpp:78-137)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.capacity = 0
        self.arrays: list[DeviceArray] = []

    def reserve(self, n: int) -> list[DeviceArray]:
        if n > self.capacity:
            for a in self.arrays:
                a.release()
            cap = max(n, 1)
            self.arrays = [DeviceArray(self.ctx, cap) for _ in range(7)]
            self.capacity = cap
        return self.arrays


def _check_spans(positions, masses, accelerations):
    x, y, z = positions
    ax, ay, az = accelerations
    n = len(masses)
    for a in (x, y, z, ax, ay, az):
        if len(a) != n:
            raise ValueError("positions, masses and accelerations must have equal length")
    return n


class HipDirectSolver:
    """Direct O(N^2) summation on the GPU; name ``hip-direct``.

    Drop-in for CudaDirectSolver (include/orrery/solvers/cuda_direct_solver.hpp).
    """

    def __init__(self, softening: Softening | float = Softening(), context: Context | None = None,
                 spatial_order: bool = True, gravitational_constant: float = 1.0):
        self._softening = softening if isinstance(softening, Softening) else Softening(softening)
        self.ctx = context or Context(0)
        self.spatial_order = spatial_order
        self.G = float(gravitational_constant)
        self._count = InteractionCount()
        self._staging = _Staging(self.ctx)

    @classmethod
    def try_create(cls, softening: Softening | float = Softening(), **kw):
        try:
            if device_count() == 0:
                return None
            return cls(softening, **kw)
        except C.NoDeviceError:
            return None

    def name(self) -> str:
        return "hip-direct"

    def softening(self) -> Softening:
        return self._softening

    def interaction_count(self) -> InteractionCount:
        return InteractionCount(**vars(self._count))

    def reset_interaction_count(self) -> None:
        self._count = InteractionCount()

    def timings(self) -> dict:
        return self.ctx.timings()

    def _args(self, ptrs, n) -> C.OrrDirectArgs:
        a = C.OrrDirectArgs()
        a.x, a.y, a.z, a.m, a.ax, a.ay, a.az = ptrs
        a.count = n
        a.softening_squared = self._softening.squared()
        a.gravitational_constant = self.G
        a.spatial_order = 1 if self.spatial_order else 0
        return a

    def evaluate_device(self, x: int, y: int, z: int, m: int, ax: int, ay: int, az: int, n: int):
        """Evaluate on device pointers (state already resident in HBM)."""
        args = self._args((x, y, z, m, ax, ay, az), n)
        C.check(C.load().orr_direct_eval(self.ctx.handle, ctypes.byref(args)))
        self._tally(n)

    def evaluate(self, positions, masses, accelerations) -> None:
        n = _check_spans(positions, masses, accelerations)
        if n == 0:
            self._tally(0)
            return
        d = self._staging.reserve(n)
        for dev, host in zip(d[:4], (*positions, masses)):
            self.ctx.upload(dev.ptr, np.ascontiguousarray(host, dtype=np.float64))
        args = self._args(tuple(a.ptr for a in d), n)
        C.check(C.load().orr_direct_eval(self.ctx.handle, ctypes.byref(args)))
        for dev, host in zip(d[4:], accelerations):
            out = np.empty(n)
            self.ctx.download(out, dev.ptr)
            host[:] = out
        self._tally(n)

    def _tally(self, n: int) -> None:
        # direct_solver.cpp:75-79: one evaluation, N(N-1) ordered pairs.
        self._count.evaluations += 1
        if n > 1:
            self._count.particle_particle += n * (n - 1)


class HipTreeSolver:
    """Barnes-Hut with on-device ordering, build and walk; name ``hip-tree``.

    Drop-in for CudaTreeSolver (include/orrery/solvers/cuda_tree_solver.hpp).
    """

    def __init__(self, parameters: TreeParameters = TreeParameters(),
                 softening: Softening | float = Softening(), context: Context | None = None,
                 gravitational_constant: float = 1.0):
        self._parameters = parameters.corrected()
        self._softening = softening if isinstance(softening, Softening) else Softening(softening)
        self.ctx = context or Context(0)
        self.G = float(gravitational_constant)
        self._count = InteractionCount()
        self._visits = 0
        self._wave_steps = 0
        self._stats = C.OrrTreeStats()
        self._staging = _Staging(self.ctx)

    @classmethod
    def try_create(cls, parameters: TreeParameters = TreeParameters(),
                   softening: Softening | float = Softening(), **kw):
        try:
            if device_count() == 0:
                return None
            return cls(parameters, softening, **kw)
        except C.NoDeviceError:
            return None

    def name(self) -> str:
        return "hip-tree"

    def softening(self) -> Softening:
        return self._softening

    def parameters(self) -> TreeParameters:
        return self._parameters

    def interaction_count(self) -> InteractionCount:
        return InteractionCount(**vars(self._count))

    def reset_interaction_count(self) -> None:
        self._count = InteractionCount()
        self._visits = 0
        self._wave_steps = 0

    def node_visits(self) -> int:
        return self._visits

    def walk_lane_efficiency(self) -> float:
        """Node visits per lane-iteration of the walk (1.0: no lane ever idles)."""
        return self._visits / (64.0 * self._wave_steps) if self._wave_steps else 0.0

    def tree_summary(self) -> dict:
        s = self._stats
        return dict(node_count=int(s.node_count), leaf_count=int(s.leaf_count), depth=int(s.depth))

    def timings(self) -> dict:
        return self.ctx.timings()

    def _args(self, ptrs, n) -> C.OrrTreeArgs:
        a = C.OrrTreeArgs()
        a.x, a.y, a.z, a.m, a.ax, a.ay, a.az = ptrs
        a.count = n
        a.softening_squared = self._softening.squared()
        a.gravitational_constant = self.G
        a.opening_angle = self._parameters.opening_angle
        a.leaf_capacity = self._parameters.leaf_capacity
        a.quadrupole = 1 if self._parameters.quadrupole else 0
        return a

    def evaluate_device(self, x: int, y: int, z: int, m: int, ax: int, ay: int, az: int, n: int):
        args = self._args((x, y, z, m, ax, ay, az), n)
        C.check(C.load().orr_tree_eval(self.ctx.handle, ctypes.byref(args), ctypes.byref(self._stats)))
        self._tally()

    def evaluate(self, positions, masses, accelerations) -> None:
        n = _check_spans(positions, masses, accelerations)
        if n == 0:
            self._stats = C.OrrTreeStats()
            self._count.evaluations += 1
            return
        d = self._staging.reserve(n)
        for dev, host in zip(d[:4], (*positions, masses)):
            self.ctx.upload(dev.ptr, np.ascontiguousarray(host, dtype=np.float64))
        args = self._args(tuple(a.ptr for a in d), n)
        C.check(C.load().orr_tree_eval(self.ctx.handle, ctypes.byref(args), ctypes.byref(self._stats)))
        for dev, host in zip(d[4:], accelerations):
            out = np.empty(n)
            self.ctx.download(out, dev.ptr)
            host[:] = out
        self._tally()

    def _tally(self) -> None:
        self._count.evaluations += 1
        self._count.particle_particle += int(self._stats.particle_particle)
        self._count.particle_cell += int(self._stats.particle_cell)
        self._visits += int(self._stats.node_visits)
        self._wave_steps += int(self._stats.wave_steps)

    def nodes(self) -> dict:
        """Nodes of the last evaluation in DFS order (TreeNode fields)."""
        M = int(self._stats.node_count)
        mass, cx, cy, cz, acc = (np.zeros(M) for _ in range(5))
        nxt, first, count = (np.zeros(M, dtype=np.uint64) for _ in range(3))
        ptrs = [a.ctypes.data_as(ctypes.c_void_p) for a in (mass, cx, cy, cz, acc, nxt, first, count)]
        C.check(C.load().orr_tree_nodes(self.ctx.handle, *ptrs))
        return dict(mass=mass, com=np.stack([cx, cy, cz], axis=1), acc_r2=acc, next=nxt,
                    first=first, count=count)


def _tree_quadrupoles(self) -> np.ndarray:
    """Quadrupole moments of the last evaluation, shape (node_count, 6)."""
    M = int(self._stats.node_count)
    out = np.zeros((M, 6))
    C.check(C.load().orr_tree_quadrupoles(self.ctx.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


HipTreeSolver.quadrupoles = _tree_quadrupoles


def morton_order(ctx: Context, x: np.ndarray, y: np.ndarray, z: np.ndarray):
    """Device Morton ordering of host positions: (codes, index, origin, size)."""
    n = len(x)
    if len(y) != n or len(z) != n:
        raise ValueError(f"x, y, z hold {n}, {len(y)}, {len(z)} values; they must match")
    dx, dy, dz = (DeviceArray.from_host(ctx, np.ascontiguousarray(a, dtype=np.float64)) for a in (x, y, z))
    codes = DeviceArray(ctx, n, np.uint64)
    index = DeviceArray(ctx, n, np.uint32)
    cube = np.zeros(4)
    C.check(C.load().orr_morton_order(ctx.handle, dx.ptr, dy.ptr, dz.ptr, n, codes.ptr, index.ptr,
                                      cube.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    out = codes.to_host(), index.to_host(), cube[:3].copy(), float(cube[3])
    for a in (dx, dy, dz, codes, index):
        a.release()
    return out


INTEGRATORS = {"velocity-verlet": 0, "vv": 0, "yoshida4": 1, "runge-kutta4": 2, "rk4": 2}
SOLVERS = {"direct": 0, "hip-direct": 0, "barnes-hut": 1, "hip-tree": 1}


@dataclass
class StepConfig:
    integrator: str = "velocity-verlet"
    solver: str = "direct"
    softening: float = 0.0
    gravitational_constant: float = 1.0
    opening_angle: float = 0.5
    leaf_capacity: int = 32
    quadrupole: bool = False

    def native(self) -> C.OrrStepConfig:
        c = C.OrrStepConfig()
        c.integrator = INTEGRATORS[self.integrator]
        c.solver = SOLVERS[self.solver]
        c.softening_squared = float(self.softening) * float(self.softening)
        c.gravitational_constant = self.gravitational_constant
        c.opening_angle = self.opening_angle
        c.leaf_capacity = self.leaf_capacity
        c.quadrupole = 1 if self.quadrupole else 0
        return c


class DeviceSystem:
    """Particle state resident in HBM across steps (FP64 x, v, a, m)."""

    def __init__(self, ctx: Context, count: int):
        self.ctx = ctx
        self.count = int(count)
        h = ctypes.c_void_p()
        C.check(C.load().orr_system_create(ctx.handle, self.count, ctypes.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            C.check(C.load().orr_system_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _p(a, count=None, name="array"):
        """FP64 contiguous copy of a; with count, exactly count values or ValueError (the
        C ABI copies count values from the pointer, so a shorter buffer would be over-read)."""
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64)
        if count is not None and a.shape != (count,):
            raise ValueError(f"{name}: expected {count} values, got shape {a.shape}")
        return a

    @staticmethod
    def _checked(count, x, y, z, vx, vy, vz, m):
        names = ("x", "y", "z", "vx", "vy", "vz", "m")
        return [DeviceSystem._p(a, count, n) for a, n in zip((x, y, z, vx, vy, vz, m), names)]

    def upload(self, x, y, z, vx, vy, vz, m) -> None:
        arrs = self._checked(int(self.count), x, y, z, vx, vy, vz, m)
        ptrs = [None if a is None else a.ctypes.data_as(ctypes.c_void_p) for a in arrs]
        C.check(C.load().orr_system_upload(self.handle, *ptrs))

    def download(self):
        out = [np.zeros(self.count) for _ in range(9)]
        C.check(C.load().orr_system_download(self.handle, *(a.ctypes.data_as(ctypes.c_void_p) for a in out)))
        return dict(x=out[0], y=out[1], z=out[2], vx=out[3], vy=out[4], vz=out[5],
                    ax=out[6], ay=out[7], az=out[8])

    def refresh(self, config: StepConfig) -> None:
        c = config.native()
        C.check(C.load().orr_system_refresh(self.handle, ctypes.byref(c)))

    def step(self, config: StepConfig, timestep: float, steps: int = 1) -> None:
        c = config.native()
        C.check(C.load().orr_system_step(self.handle, ctypes.byref(c), float(timestep), int(steps)))

    def checkpoint(self, path: str, step: int, configuration: str = "") -> None:
        """Write the reference's checkpoint format (sim/checkpoint.cpp) from device state."""
        text = configuration.encode()
        C.check(C.load().orr_checkpoint_write(self.handle, os.fsencode(path), int(step), text, len(text)))

    @classmethod
    def from_checkpoint(cls, ctx: Context, path: str):
        """(system, step, configuration text) from a checkpoint file."""
        h = ctypes.c_void_p()
        step = ctypes.c_uint64()
        length = ctypes.c_uint64()
        buf = ctypes.create_string_buffer(1 << 20)
        C.check(C.load().orr_checkpoint_read(ctx.handle, os.fsencode(path), ctypes.byref(h),
                                             ctypes.byref(step), buf, len(buf), ctypes.byref(length)))
        s = cls.__new__(cls)
        s.ctx = ctx
        s.handle = h
        count = ctypes.c_uint64()
        C.check(C.load().orr_system_count(h, ctypes.byref(count)))
        s.count = count.value
        return s, step.value, buf.value.decode()

    def diagnostics(self, softening: float, gravitational_constant: float = 1.0) -> dict:
        out = np.zeros(8)
        C.check(C.load().orr_system_diagnostics(self.handle, float(softening) * float(softening),
                                                float(gravitational_constant),
                                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return dict(kinetic=out[0], potential=out[1], total=out[0] + out[1],
                    momentum=out[2:5].copy(), angular_momentum=out[5:8].copy())


@dataclass
class BatchSystem:
    """One system of a batch: size, solver (-1 by size, 0 direct, 1 Barnes-Hut),
    softening length, G and timestep."""
    count: int
    softening: float
    timestep: float
    solver: int = -1
    gravitational_constant: float = 1.0

    def native(self) -> C.OrrBatchSystem:
        return C.OrrBatchSystem(int(self.count), int(self.solver), 0,
                                float(self.softening) * float(self.softening),
                                float(self.gravitational_constant), float(self.timestep))


@dataclass
class BatchConfig:
    """Batch-wide settings; 0 for leaf_capacity / tiles_per_item leaves them to
    DeviceBatch.autotune (16 until then)."""
    integrator: str = "velocity-verlet"
    opening_angle: float = 0.5
    leaf_capacity: int = 16
    quadrupole: bool = False
    tiles_per_item: int = 16
    direct_below: int = 4096
    use_graph: bool = True

    def native(self) -> C.OrrBatchConfig:
        return C.OrrBatchConfig(INTEGRATORS[self.integrator], int(bool(self.quadrupole)),
                                float(self.opening_angle), int(self.leaf_capacity),
                                int(self.tiles_per_item), int(self.direct_below),
                                int(bool(self.use_graph)))


class DeviceBatch:
    """Many independent systems resident in HBM, stepped together (one launch
    sequence per step for all of them, optionally replayed as a hipGraph)."""

    def __init__(self, ctx: Context, systems, config: BatchConfig):
        self.ctx = ctx
        self.systems = list(systems)
        arr = (C.OrrBatchSystem * max(1, len(self.systems)))(*(s.native() for s in self.systems))
        cfg = config.native()
        h = ctypes.c_void_p()
        C.check(C.load().orr_batch_create(ctx.handle, arr, len(self.systems), ctypes.byref(cfg),
                                          ctypes.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            C.check(C.load().orr_batch_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        out = C.OrrBatchInfo()
        C.check(C.load().orr_batch_info_get(self.handle, ctypes.byref(out)))
        return {name: getattr(out, name) for name, _ in C.OrrBatchInfo._fields_}

    @staticmethod
    def bytes_bound(systems, config: BatchConfig) -> int:
        """Upper bound of the device bytes a batch of these systems holds (orr_batch_bytes_bound)."""
        systems = list(systems)
        arr = (C.OrrBatchSystem * max(1, len(systems)))(*(s.native() for s in systems))
        cfg = config.native()
        out = ctypes.c_uint64()
        C.check(C.load().orr_batch_bytes_bound(arr, len(systems), ctypes.byref(cfg), ctypes.byref(out)))
        return int(out.value)

    def upload(self, system: int, x, y, z, vx, vy, vz, m) -> None:
        if not 0 <= int(system) < len(self.systems):
            raise IndexError(f"system {system} of a batch of {len(self.systems)}")
        arrs = DeviceSystem._checked(int(self.systems[system].count), x, y, z, vx, vy, vz, m)
        ptrs = [None if a is None else a.ctypes.data_as(ctypes.c_void_p) for a in arrs]
        C.check(C.load().orr_batch_upload(self.handle, int(system), *ptrs))

    def download(self, system: int) -> dict:
        n = int(self.systems[system].count)
        out = [np.zeros(n) for _ in range(9)]
        C.check(C.load().orr_batch_download(self.handle, int(system),
                                            *(a.ctypes.data_as(ctypes.c_void_p) for a in out)))
        return dict(x=out[0], y=out[1], z=out[2], vx=out[3], vy=out[4], vz=out[5],
                    ax=out[6], ay=out[7], az=out[8])

    def copy_system(self, system: int, source: "DeviceBatch", source_system: int) -> None:
        C.check(C.load().orr_batch_copy_system(self.handle, int(system), source.handle,
                                               int(source_system)))

    def refresh(self) -> None:
        C.check(C.load().orr_batch_refresh(self.handle))

    def step(self, steps: int = 1) -> None:
        C.check(C.load().orr_batch_step(self.handle, int(steps)))

    def autotune(self) -> dict:
        C.check(C.load().orr_batch_autotune(self.handle))
        return self.info()

    def diagnostics(self, system: int) -> dict:
        out = np.zeros(8)
        C.check(C.load().orr_batch_diagnostics(self.handle, int(system),
                                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return dict(kinetic=out[0], potential=out[1], total=out[0] + out[1],
                    momentum=out[2:5].copy(), angular_momentum=out[5:8].copy())

    def counts(self, system: int) -> tuple[int, int]:
        out = (ctypes.c_uint64 * 2)()
        C.check(C.load().orr_batch_counts(self.handle, int(system), out))
        return int(out[0]), int(out[1])


@dataclass
class ServerConfig:
    """Serving-loop settings (0 / None: the engine's defaults)."""
    memory_budget: int = 0
    round_ms: float = 50.0
    direct_below: int = 4096
    alone_from: int = 131072
    opening_angle: float = 0.5
    leaf_capacity: int = 16
    tiles_per_item: int = 16
    quadrupole: bool = False
    use_graph: bool = True
    fail_rebuilds: int = 0  # testing aid: the first n batch rebuilds fail

    def native(self) -> C.OrrServerConfig:
        return C.OrrServerConfig(int(self.memory_budget), float(self.round_ms), int(self.direct_below),
                                 int(self.alone_from), float(self.opening_angle),
                                 int(self.leaf_capacity), int(self.tiles_per_item),
                                 int(bool(self.quadrupole)), int(bool(self.use_graph)),
                                 int(self.fail_rebuilds), 0)


class Server:
    """Continuous-batching request loop on one device (orr_server_*)."""

    STATES = {0: "queued", 1: "running", 2: "done"}

    def __init__(self, ctx: Context, config: ServerConfig | None = None):
        self.ctx = ctx
        cfg = (config or ServerConfig()).native()
        h = ctypes.c_void_p()
        C.check(C.load().orr_server_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self.counts = {}

    def close(self):
        if getattr(self, "handle", None):
            C.check(C.load().orr_server_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def now(self) -> float:
        t = ctypes.c_double()
        C.check(C.load().orr_server_now(self.handle, ctypes.byref(t)))
        return t.value

    def submit(self, state, integrator: str, softening: float, timestep: float, steps: int,
               gravitational_constant: float = 1.0) -> int:
        raw = state.arrays()
        arrs = DeviceSystem._checked(len(raw[6]), *raw)
        req = C.OrrRequest(len(arrs[6]), INTEGRATORS[integrator], 0, float(softening) ** 2,
                           float(gravitational_constant), float(timestep), int(steps),
                           *(a.ctypes.data_as(ctypes.c_void_p) for a in arrs))
        rid = ctypes.c_uint64()
        C.check(C.load().orr_server_submit(self.handle, ctypes.byref(req), ctypes.byref(rid)))
        self.counts[rid.value] = len(arrs[6])
        return rid.value

    def round(self) -> int:
        pending = ctypes.c_uint32()
        C.check(C.load().orr_server_round(self.handle, ctypes.byref(pending)))
        return pending.value

    def status(self, rid: int) -> dict:
        st = C.OrrRequestStatus()
        C.check(C.load().orr_server_status(self.handle, int(rid), ctypes.byref(st)))

        def diag(a):
            return dict(kinetic=a[0], potential=a[1], total=a[0] + a[1], momentum=list(a[2:5]),
                        angular_momentum=list(a[5:8]))
        return dict(state=self.STATES[st.state], steps_done=st.steps_done, submitted=st.submitted_s,
                    admitted=st.admitted_s, completed=st.completed_s,
                    start=diag(list(st.diagnostics_start)), end=diag(list(st.diagnostics_end)),
                    charged_bytes=st.charged_bytes)

    def result(self, rid: int) -> dict:
        n = self.counts[rid]
        out = [np.zeros(n) for _ in range(6)]
        C.check(C.load().orr_server_result(self.handle, int(rid),
                                           *(a.ctypes.data_as(ctypes.c_void_p) for a in out)))
        return dict(zip(("x", "y", "z", "vx", "vy", "vz"), out))

    def stats(self) -> dict:
        st = C.OrrServerStats()
        C.check(C.load().orr_server_stats_get(self.handle, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in C.OrrServerStats._fields_}

    def drain(self) -> None:
        while self.round() > 0:
            pass


class Trajectory:
    """The reference's trajectory format (sim/trajectory.cpp) written from device state."""

    def __init__(self, system: DeviceSystem, path: str, timestep: float, velocities: bool = False):
        h = ctypes.c_void_p()
        C.check(C.load().orr_trajectory_open(system.handle, os.fsencode(path), float(timestep),
                                             1 if velocities else 0, ctypes.byref(h)))
        self.handle = h

    def append(self, system: DeviceSystem, step: int) -> None:
        C.check(C.load().orr_trajectory_append(self.handle, system.handle, int(step)))

    def close(self) -> None:
        if getattr(self, "handle", None):
            C.check(C.load().orr_trajectory_close(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
