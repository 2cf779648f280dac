"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline
leg, as the checker. The product package ``orrery_amd`` never imports it.

``load()`` returns the oracle restatement (oracle/liboracle.so);
``load_reference()`` returns the reference's own CPU library compiled from its
sources (oracle/_ref/libref.so) or None where it was not built.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_d = ctypes.POINTER(ctypes.c_double)
_f = ctypes.POINTER(ctypes.c_float)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u64 = ctypes.c_uint64


def _ptr(a, kind=_d):
    if a is None:
        return None
    return a.ctypes.data_as(kind)


class OrcDiscParams(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_uint64),
        ("mass_ratio", ctypes.c_double),
        ("scale_length", ctypes.c_double * 2),
        ("scale_height", ctypes.c_double),
        ("truncation", ctypes.c_double),
        ("bulge_fraction", ctypes.c_double),
        ("bulge_scale", ctypes.c_double),
        ("dispersion", ctypes.c_double),
        ("pericentre", ctypes.c_double),
        ("separation", ctypes.c_double),
        ("inclined_fraction", ctypes.c_double),
        ("inclination", ctypes.c_double),
        ("total_mass", ctypes.c_double),
        ("seed", ctypes.c_uint64),
    ]


class OrcField(ctypes.Structure):
    _fields_ = [
        ("integrator", ctypes.c_int),
        ("solver", ctypes.c_int),
        ("softening_squared", ctypes.c_double),
        ("G", ctypes.c_double),
        ("opening_angle", ctypes.c_double),
        ("leaf_capacity", ctypes.c_uint64),
        ("threads", ctypes.c_uint),
    ]


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (and the reference library where its sources exist)."""
    out = subprocess.DEVNULL if quiet else None
    subprocess.run(["make", "-s", "-C", HERE], check=True, stdout=out)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-j8", "-C", HERE, "ref"], check=True, stdout=out)


def load():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        lib = ctypes.CDLL(path)
        lib.orc_spread_bits.restype = ctypes.c_uint64
        lib.orc_spread_bits.argtypes = [ctypes.c_uint32]
        lib.orc_morton_interleave.restype = ctypes.c_uint64
        lib.orc_morton_interleave.argtypes = [ctypes.c_uint32] * 3
        lib.orc_morton_octant.restype = ctypes.c_uint
        lib.orc_morton_octant.argtypes = [ctypes.c_uint64, ctypes.c_uint]
        lib.orc_tree_build.restype = ctypes.c_void_p
        lib.orc_tree_node_count.restype = ctypes.c_uint64
        lib.orc_tree_node_count.argtypes = [ctypes.c_void_p]
        lib.orc_tree_leaf_count.restype = ctypes.c_uint64
        lib.orc_tree_leaf_count.argtypes = [ctypes.c_void_p]
        lib.orc_tree_depth.restype = ctypes.c_uint
        lib.orc_tree_depth.argtypes = [ctypes.c_void_p]
        lib.orc_tree_nodes.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 9
        lib.orc_tree_free.argtypes = [ctypes.c_void_p]
        _LIB = lib
    return _LIB


def load_reference():
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            return None
        lib = ctypes.CDLL(path)
        lib.ref_tree.restype = ctypes.c_int64
        _REF = lib
    return _REF


# ---------------------------------------------------------------------------
# Thin numpy wrappers.


@dataclass
class Particles:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    m: np.ndarray

    @property
    def n(self) -> int:
        return int(self.m.shape[0])

    def arrays(self):
        return (self.x, self.y, self.z, self.vx, self.vy, self.vz, self.m)


def _empty(n):
    return Particles(*(np.zeros(n) for _ in range(7)))


def plummer(n, seed, total_mass=1.0, scale_radius=1.0, cutoff=None, lib=None,
            velocity_direction_first=False):
    """Plummer sphere; cutoff defaults to truncation at r < 10 a.
    velocity_direction_first: the draw order of a GCC build of plummer.cpp:122
    (the oracle's restatement only)."""
    if cutoff is None:
        cutoff = (100.0 / 101.0) ** 1.5
    p = _empty(n)
    lib = lib or load()
    if velocity_direction_first:
        rc = lib.orc_plummer_ordered(_u64(n), ctypes.c_double(total_mass), ctypes.c_double(scale_radius),
                                     ctypes.c_double(cutoff), _u64(seed), ctypes.c_int(1),
                                     *(_ptr(a) for a in p.arrays()))
        assert rc == 0
        return p
    fn = lib.orc_plummer if hasattr(lib, "orc_plummer") else lib.ref_plummer
    rc = fn(_u64(n), ctypes.c_double(total_mass), ctypes.c_double(scale_radius),
            ctypes.c_double(cutoff), _u64(seed), *(_ptr(a) for a in p.arrays()))
    assert rc == 0
    return p


def uniform_sphere(n, seed, total_mass=1.0, radius=1.0, lib=None):
    p = _empty(n)
    lib = lib or load()
    fn = lib.orc_uniform_sphere if hasattr(lib, "orc_uniform_sphere") else lib.ref_uniform_sphere
    rc = fn(_u64(n), ctypes.c_double(total_mass), ctypes.c_double(radius), _u64(seed),
            *(_ptr(a) for a in p.arrays()))
    assert rc == 0
    return p


DISC_DEFAULTS = dict(mass_ratio=3.0, scale_length=(1.0, 0.7), scale_height=0.1, truncation=5.0,
                     bulge_fraction=0.2, bulge_scale=0.2, dispersion=0.1, pericentre=2.0,
                     separation=20.0, inclined_fraction=0.3, inclination=np.deg2rad(30.0),
                     total_mass=1.0)


def disc_collision(n, seed, **overrides):
    """Two exponential discs with Hernquist bulges on a parabolic orbit."""
    cfg = dict(DISC_DEFAULTS)
    cfg.update(overrides)
    prm = OrcDiscParams()
    prm.count = n
    prm.seed = seed
    for k, v in cfg.items():
        if k == "scale_length":
            prm.scale_length[0], prm.scale_length[1] = v
        else:
            setattr(prm, k, v)
    p = _empty(n)
    rc = load().orc_disc_collision(ctypes.byref(prm), *(_ptr(a) for a in p.arrays()))
    assert rc == 0
    return p


def reference_acceleration(x, y, z, m, eps2, G=1.0, targets=None, threads=0):
    n = len(m)
    t = None if targets is None else np.ascontiguousarray(targets, dtype=np.uint64)
    k = n if t is None else len(t)
    ax, ay, az = np.zeros(k), np.zeros(k), np.zeros(k)
    rc = load().orc_reference_acceleration(_u64(n), _ptr(x), _ptr(y), _ptr(z), _ptr(m),
                                           ctypes.c_double(eps2), ctypes.c_double(G),
                                           _ptr(t, _u64p), _u64(k), _ptr(ax), _ptr(ay), _ptr(az),
                                           ctypes.c_uint(threads))
    assert rc == 0
    return ax, ay, az


def direct(x, y, z, m, eps2, G=1.0, kernel=1, threads=0, target_range=None):
    n = len(m)
    tb, te = (0, n) if target_range is None else target_range
    ax, ay, az = np.zeros(n), np.zeros(n), np.zeros(n)
    rc = load().orc_direct(_u64(n), _ptr(x), _ptr(y), _ptr(z), _ptr(m), ctypes.c_double(eps2),
                           ctypes.c_double(G), ctypes.c_int(kernel), ctypes.c_uint(threads),
                           _u64(tb), _u64(te), _ptr(ax), _ptr(ay), _ptr(az))
    assert rc == 0
    return ax, ay, az


def bounding_cube(x, y, z):
    origin = np.zeros(3)
    size = ctypes.c_double()
    load().orc_bounding_cube(_u64(len(x)), _ptr(x), _ptr(y), _ptr(z), _ptr(origin), ctypes.byref(size))
    return origin, size.value


def morton_order(x, y, z):
    n = len(x)
    origin = np.zeros(3)
    size = ctypes.c_double()
    codes = np.zeros(n, dtype=np.uint64)
    index = np.zeros(n, dtype=np.uint64)
    load().orc_morton_order(_u64(n), _ptr(x), _ptr(y), _ptr(z), _ptr(origin), ctypes.byref(size),
                            _ptr(codes, _u64p), _ptr(index, _u64p))
    return codes, index, origin, size.value


def morton_codes(x, y, z, origin, size):
    n = len(x)
    codes = np.zeros(n, dtype=np.uint64)
    load().orc_morton_codes(_u64(n), _ptr(x), _ptr(y), _ptr(z),
                            _ptr(np.ascontiguousarray(origin, dtype=np.float64)),
                            ctypes.c_double(size), _ptr(codes, _u64p))
    return codes


@dataclass
class Tree:
    mass: np.ndarray
    com: np.ndarray          # (M, 3)
    acc_r2: np.ndarray
    next: np.ndarray
    first: np.ndarray
    count: np.ndarray
    quad: np.ndarray | None
    leaves: int
    depth: int


def tree_build(sx, sy, sz, sm, codes, origin, size, theta, leaf_capacity, quadrupole=False):
    lib = load()
    h = lib.orc_tree_build(_u64(len(sm)), _ptr(sx), _ptr(sy), _ptr(sz), _ptr(sm),
                           _ptr(codes, _u64p), _ptr(np.ascontiguousarray(origin, dtype=np.float64)),
                           ctypes.c_double(size), ctypes.c_double(theta), _u64(leaf_capacity),
                           ctypes.c_int(1 if quadrupole else 0))
    try:
        M = lib.orc_tree_node_count(h)
        mass, cx, cy, cz, acc = (np.zeros(M) for _ in range(5))
        nxt, first, count = (np.zeros(M, dtype=np.uint64) for _ in range(3))
        quad = np.zeros((M, 6)) if quadrupole else None
        lib.orc_tree_nodes(h, _ptr(mass), _ptr(cx), _ptr(cy), _ptr(cz), _ptr(acc),
                           _ptr(nxt, _u64p), _ptr(first, _u64p), _ptr(count, _u64p),
                           _ptr(quad) if quad is not None else None)
        return Tree(mass, np.stack([cx, cy, cz], axis=1), acc, nxt, first, count, quad,
                    int(lib.orc_tree_leaf_count(h)), int(lib.orc_tree_depth(h)))
    finally:
        lib.orc_tree_free(h)


def barnes_hut(x, y, z, m, eps2, theta, leaf_capacity, G=1.0, quadrupole=False, kernel=1, threads=0):
    n = len(m)
    ax, ay, az = np.zeros(n), np.zeros(n), np.zeros(n)
    counts = np.zeros(2, dtype=np.uint64)
    rc = load().orc_barnes_hut(_u64(n), _ptr(x), _ptr(y), _ptr(z), _ptr(m), ctypes.c_double(eps2),
                               ctypes.c_double(G), ctypes.c_double(theta), _u64(leaf_capacity),
                               ctypes.c_int(1 if quadrupole else 0), ctypes.c_int(kernel),
                               ctypes.c_uint(threads), _ptr(ax), _ptr(ay), _ptr(az),
                               _ptr(counts, _u64p))
    assert rc == 0
    return ax, ay, az, (int(counts[0]), int(counts[1]))


def barnes_hut_fp32_counts(x, y, z, m, theta, leaf_capacity, threads=0):
    """(pp, pc) of the device walk's FP32 acceptance policy on the reference tree."""
    counts = np.zeros(2, dtype=np.uint64)
    rc = load().orc_barnes_hut_fp32_counts(_u64(len(m)), _ptr(x), _ptr(y), _ptr(z), _ptr(m),
                                           ctypes.c_double(theta), _u64(leaf_capacity),
                                           ctypes.c_uint(threads), _ptr(counts, _u64p))
    assert rc == 0
    return int(counts[0]), int(counts[1])


def make_field(integrator="vv", solver="direct", eps2=0.0, G=1.0, theta=0.5, leaf_capacity=32,
               threads=0, kernel=1):
    f = OrcField()
    f.integrator = {"vv": 0, "yoshida4": 1, "rk4": 2}[integrator]
    f.solver = {"direct": 2 if kernel == 1 else 0, "barnes-hut": 1}[solver]
    f.softening_squared = eps2
    f.G = G
    f.opening_angle = theta
    f.leaf_capacity = leaf_capacity
    f.threads = threads
    return f


def refresh(state: Particles, field: OrcField):
    n = state.n
    a = [np.zeros(n) for _ in range(3)]
    load().orc_refresh(_u64(n), _ptr(state.x), _ptr(state.y), _ptr(state.z), _ptr(state.m),
                       ctypes.byref(field), *(_ptr(v) for v in a))
    return a


def steps(state: Particles, acc, field: OrcField, dt, count):
    """Advance `state` in place by `count` steps; `acc` (3 arrays) in/out."""
    rc = load().orc_steps(_u64(state.n), _ptr(state.x), _ptr(state.y), _ptr(state.z),
                          _ptr(state.vx), _ptr(state.vy), _ptr(state.vz),
                          _ptr(acc[0]), _ptr(acc[1]), _ptr(acc[2]), _ptr(state.m),
                          ctypes.byref(field), ctypes.c_double(dt), _u64(count))
    assert rc == 0


def diagnostics(state: Particles, eps2, G=1.0):
    out = np.zeros(8)
    load().orc_diagnostics(_u64(state.n), *(_ptr(a) for a in state.arrays()),
                           ctypes.c_double(eps2), ctypes.c_double(G), _ptr(out))
    return dict(kinetic=out[0], potential=out[1], total=out[0] + out[1],
                momentum=out[2:5].copy(), angular_momentum=out[5:8].copy())


def relative_errors(ax, ay, az, rx, ry, rz):
    """Per-particle |a - a_ref| / |a_ref| (cuda_direct_solver_test.cpp:86-111)."""
    mag = np.sqrt(rx * rx + ry * ry + rz * rz)
    err = np.sqrt((ax - rx) ** 2 + (ay - ry) ** 2 + (az - rz) ** 2)
    ok = mag > 0
    rel = np.zeros_like(mag)
    rel[ok] = err[ok] / mag[ok]
    return rel


def error_summary(rel):
    """(worst, rms) with rms over all particles, as the reference's tests define it."""
    return float(rel.max()) if rel.size else 0.0, float(np.sqrt(np.sum(rel * rel) / max(1, rel.size)))


# ---------------------------------------------------------------------------
# The reference's own checkpoint and trajectory code (oracle/_ref), for format
# parity. Each returns None when the reference library is not built.


def ref_checkpoint_read(path, capacity):
    R = load_reference()
    if R is None:
        return None
    step, count = ctypes.c_uint64(), ctypes.c_uint64()
    arrays = [np.zeros(capacity) for _ in range(10)]
    ptrs = (ctypes.c_void_p * 10)(*[a.ctypes.data for a in arrays])
    text = ctypes.create_string_buffer(1 << 16)
    R.ref_checkpoint_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64,
                                      ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p, ctypes.c_uint64]
    rc = R.ref_checkpoint_read(os.fsencode(path), ctypes.byref(step), ctypes.byref(count), capacity,
                               ptrs, text, len(text))
    if rc != 0:
        return {"refused": rc}
    n = count.value
    names = ("m", "x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")
    out = {k: a[:n].copy() for k, a in zip(names, arrays)}
    out.update(step=step.value, count=n, configuration=text.value.decode())
    return out


def ref_checkpoint_write(path, step, arrays, text=""):
    """arrays in the order m, x, y, z, vx, vy, vz, ax, ay, az."""
    R = load_reference()
    if R is None:
        return None
    arrays = [np.ascontiguousarray(a, dtype=np.float64) for a in arrays]
    ptrs = (ctypes.c_void_p * 10)(*[a.ctypes.data for a in arrays])
    R.ref_checkpoint_write.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint64,
                                       ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p]
    return R.ref_checkpoint_write(os.fsencode(path), int(step), len(arrays[0]), ptrs, text.encode())


def ref_trajectory_read(path, capacity, max_frames):
    R = load_reference()
    if R is None:
        return None
    count, dt, vel = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_int()
    masses = np.zeros(capacity)
    steps = np.zeros(max_frames, dtype=np.uint64)
    times = np.zeros(max_frames)
    values = np.zeros(max_frames * 6 * capacity)
    R.ref_trajectory_read.restype = ctypes.c_long
    R.ref_trajectory_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                                      ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    frames = R.ref_trajectory_read(os.fsencode(path), ctypes.byref(count), ctypes.byref(dt),
                                   ctypes.byref(vel), masses.ctypes.data, capacity, max_frames,
                                   steps.ctypes.data, times.ctypes.data, values.ctypes.data)
    if frames < 0:
        return {"refused": True}
    n = count.value
    out = {"count": n, "timestep": dt.value, "velocities": bool(vel.value), "masses": masses[:n].copy(),
           "frames": []}
    for f in range(frames):
        v = values[f * 6 * n: (f + 1) * 6 * n].reshape(6, n)  # frames packed at 6 n
        out["frames"].append({"step": int(steps[f]), "time": float(times[f]), "values": v.copy()})
    return out
