"""ctypes binding of liborrery_hip.so (the C ABI of include/orrery_hip.h).

The library is built in-tree by ``python -m orrery_amd.build``. Loading fails
loudly when it is missing: there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "liborrery_hip.so")

ORR_OK = 0
ORR_ERR_INVALID_ARGUMENT = 1
ORR_ERR_NO_DEVICE = 2
ORR_ERR_RUNTIME = 3
ORR_ERR_LENGTH = 4
ORR_ERR_OUT_OF_MEMORY = 5
ORR_PHASES = 16

PHASE_NAMES = ["direct_prepare", "direct_kernel", "direct_reduce", "bbox", "encode", "sort",
               "gather", "build", "moments", "walk", "integrate", "energy", "phase12",
               "phase13", "phase14", "phase15"]

_d = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


class OrrDeviceDesc(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("arch", ctypes.c_char * 64),
                ("compute_units", ctypes.c_int32), ("wavefront_size", ctypes.c_int32),
                ("max_threads_per_block", ctypes.c_int32), ("clock_khz", ctypes.c_int32),
                ("global_memory_bytes", ctypes.c_uint64), ("lds_per_block_bytes", ctypes.c_uint64),
                ("l2_bytes", ctypes.c_uint64)]


class OrrDirectArgs(ctypes.Structure):
    _fields_ = [("x", _vp), ("y", _vp), ("z", _vp), ("m", _vp), ("ax", _vp), ("ay", _vp),
                ("az", _vp), ("count", ctypes.c_uint64), ("softening_squared", ctypes.c_double),
                ("gravitational_constant", ctypes.c_double), ("spatial_order", ctypes.c_int32)]


class OrrDirectConfig(ctypes.Structure):
    _fields_ = [("targets_per_lane", ctypes.c_int32), ("refine", ctypes.c_int32),
                ("block_target", ctypes.c_uint32), ("pair_mode", ctypes.c_int32),
                ("equal_masses", ctypes.c_int32), ("small_order", ctypes.c_int32)]


class OrrTreeArgs(ctypes.Structure):
    _fields_ = [("x", _vp), ("y", _vp), ("z", _vp), ("m", _vp), ("ax", _vp), ("ay", _vp),
                ("az", _vp), ("count", ctypes.c_uint64), ("softening_squared", ctypes.c_double),
                ("gravitational_constant", ctypes.c_double), ("opening_angle", ctypes.c_double),
                ("leaf_capacity", ctypes.c_uint64), ("quadrupole", ctypes.c_int32)]


class OrrTreeStats(ctypes.Structure):
    _fields_ = [("particle_particle", ctypes.c_uint64), ("particle_cell", ctypes.c_uint64),
                ("node_visits", ctypes.c_uint64), ("node_count", ctypes.c_uint64),
                ("leaf_count", ctypes.c_uint64), ("depth", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("wave_steps", ctypes.c_uint64)]


class OrrTimings(ctypes.Structure):
    _fields_ = [("milliseconds", ctypes.c_double * ORR_PHASES),
                ("launches", ctypes.c_uint64 * ORR_PHASES),
                ("graph_launches", ctypes.c_uint64), ("reserved", ctypes.c_uint64 * 3)]


class OrrStepConfig(ctypes.Structure):
    _fields_ = [("integrator", ctypes.c_int32), ("solver", ctypes.c_int32),
                ("softening_squared", ctypes.c_double), ("gravitational_constant", ctypes.c_double),
                ("opening_angle", ctypes.c_double), ("leaf_capacity", ctypes.c_uint64),
                ("quadrupole", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class OrrBatchSystem(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint64), ("solver", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("softening_squared", ctypes.c_double), ("gravitational_constant", ctypes.c_double),
                ("timestep", ctypes.c_double)]


class OrrBatchConfig(ctypes.Structure):
    _fields_ = [("integrator", ctypes.c_int32), ("quadrupole", ctypes.c_int32),
                ("opening_angle", ctypes.c_double), ("leaf_capacity", ctypes.c_uint32),
                ("tiles_per_item", ctypes.c_uint32), ("direct_below", ctypes.c_uint32),
                ("use_graph", ctypes.c_int32)]


class OrrBatchInfo(ctypes.Structure):
    _fields_ = [("systems", ctypes.c_uint32), ("direct_systems", ctypes.c_uint32),
                ("tree_systems", ctypes.c_uint32), ("tiles_per_item", ctypes.c_uint32),
                ("leaf_capacity", ctypes.c_uint32), ("direct_tiles", ctypes.c_uint32),
                ("direct_items", ctypes.c_uint32), ("graph", ctypes.c_int32),
                ("particles", ctypes.c_uint64), ("direct_particles", ctypes.c_uint64),
                ("tree_particles", ctypes.c_uint64), ("direct_pairs", ctypes.c_uint64),
                ("device_bytes", ctypes.c_uint64)]


class OrrServerConfig(ctypes.Structure):
    _fields_ = [("memory_budget", ctypes.c_uint64), ("round_ms", ctypes.c_double),
                ("direct_below", ctypes.c_uint32), ("alone_from", ctypes.c_uint32),
                ("opening_angle", ctypes.c_double), ("leaf_capacity", ctypes.c_uint32),
                ("tiles_per_item", ctypes.c_uint32), ("quadrupole", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("fail_rebuilds", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class OrrRequest(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint64), ("integrator", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("softening_squared", ctypes.c_double), ("gravitational_constant", ctypes.c_double),
                ("timestep", ctypes.c_double), ("steps", ctypes.c_uint64),
                ("x", _vp), ("y", _vp), ("z", _vp), ("vx", _vp), ("vy", _vp), ("vz", _vp), ("m", _vp)]


class OrrRequestStatus(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("reserved", ctypes.c_int32), ("steps_done", ctypes.c_uint64),
                ("submitted_s", ctypes.c_double), ("admitted_s", ctypes.c_double),
                ("completed_s", ctypes.c_double), ("diagnostics_start", ctypes.c_double * 8),
                ("diagnostics_end", ctypes.c_double * 8), ("charged_bytes", ctypes.c_uint64)]


class OrrServerStats(ctypes.Structure):
    _fields_ = [("rounds", ctypes.c_uint64), ("batch_builds", ctypes.c_uint64),
                ("particle_steps", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("resident_bytes", ctypes.c_uint64), ("peak_resident_bytes", ctypes.c_uint64),
                ("memory_budget", ctypes.c_uint64), ("active", ctypes.c_uint32),
                ("queued", ctypes.c_uint32), ("busy_s", ctypes.c_double)]


class OrrTreeConfig(ctypes.Structure):
    _fields_ = [("targets_per_lane", ctypes.c_int32), ("lanes", ctypes.c_int32), ("sort_mode", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


# (name, argtypes) of every entry point the header declares.
SIGNATURES = {
    "orr_device_count": [ctypes.POINTER(ctypes.c_int32)],
    "orr_device_describe": [ctypes.c_int32, ctypes.POINTER(OrrDeviceDesc)],
    "orr_last_error": [],
    "orr_ctx_create": [ctypes.c_int32, ctypes.POINTER(_vp)],
    "orr_ctx_destroy": [_vp],
    "orr_malloc": [_vp, ctypes.c_uint64, ctypes.POINTER(_vp)],
    "orr_free": [_vp, _vp],
    "orr_upload": [_vp, _vp, _vp, ctypes.c_uint64],
    "orr_download": [_vp, _vp, _vp, ctypes.c_uint64],
    "orr_memset": [_vp, _vp, ctypes.c_int32, ctypes.c_uint64],
    "orr_direct_eval": [_vp, ctypes.POINTER(OrrDirectArgs)],
    "orr_direct_eval_static": [_vp, ctypes.POINTER(OrrDirectArgs)],
    "orr_direct_static_limits": [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)],
    "orr_ctx_set_direct_config": [_vp, ctypes.POINTER(OrrDirectConfig)],
    "orr_ctx_get_direct_config": [_vp, ctypes.POINTER(OrrDirectConfig)],
    "orr_tree_eval": [_vp, ctypes.POINTER(OrrTreeArgs), ctypes.POINTER(OrrTreeStats)],
    "orr_ctx_set_tree_config": [_vp, ctypes.POINTER(OrrTreeConfig)],
    "orr_ctx_get_tree_config": [_vp, ctypes.POINTER(OrrTreeConfig)],
    "orr_morton_order": [_vp, _vp, _vp, _vp, ctypes.c_uint64, _vp, _vp, _d],
    "orr_tree_nodes": [_vp] + [_vp] * 8,
    "orr_probe": [_vp, ctypes.c_int32, ctypes.c_uint64, _d, _d],
    "orr_tree_quadrupoles": [_vp, _d],
    "orr_ctx_set_timing": [_vp, ctypes.c_int32],
    "orr_ctx_timings": [_vp, ctypes.POINTER(OrrTimings)],
    "orr_ctx_reset_timings": [_vp],
    "orr_system_create": [_vp, ctypes.c_uint64, ctypes.POINTER(_vp)],
    "orr_system_destroy": [_vp],
    "orr_system_count": [_vp, ctypes.POINTER(ctypes.c_uint64)],
    "orr_system_upload": [_vp] + [_vp] * 7,
    "orr_system_download": [_vp] + [_vp] * 9,
    "orr_system_refresh": [_vp, ctypes.POINTER(OrrStepConfig)],
    "orr_system_step": [_vp, ctypes.POINTER(OrrStepConfig), ctypes.c_double, ctypes.c_uint64],
    "orr_system_diagnostics": [_vp, ctypes.c_double, ctypes.c_double, _d],
    "orr_checkpoint_write": [_vp, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_char_p, ctypes.c_uint64],
    "orr_checkpoint_read": [_vp, ctypes.c_char_p, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_uint64),
                            ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)],
    "orr_trajectory_open": [_vp, ctypes.c_char_p, ctypes.c_double, ctypes.c_int32, ctypes.POINTER(_vp)],
    "orr_trajectory_append": [_vp, _vp, ctypes.c_uint64],
    "orr_trajectory_close": [_vp],
    "orr_batch_create": [_vp, ctypes.POINTER(OrrBatchSystem), ctypes.c_uint32,
                         ctypes.POINTER(OrrBatchConfig), ctypes.POINTER(_vp)],
    "orr_batch_destroy": [_vp],
    "orr_batch_info_get": [_vp, ctypes.POINTER(OrrBatchInfo)],
    "orr_batch_bytes_bound": [_vp, ctypes.c_uint32, _vp, ctypes.POINTER(ctypes.c_uint64)],
    "orr_batch_upload": [_vp, ctypes.c_uint32] + [_vp] * 7,
    "orr_batch_download": [_vp, ctypes.c_uint32] + [_vp] * 9,
    "orr_batch_copy_system": [_vp, ctypes.c_uint32, _vp, ctypes.c_uint32],
    "orr_batch_refresh": [_vp],
    "orr_batch_step": [_vp, ctypes.c_uint64],
    "orr_batch_autotune": [_vp],
    "orr_batch_diagnostics": [_vp, ctypes.c_uint32, _d],
    "orr_batch_counts": [_vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)],
    "orr_server_create": [_vp, ctypes.POINTER(OrrServerConfig), ctypes.POINTER(_vp)],
    "orr_server_destroy": [_vp],
    "orr_server_now": [_vp, _d],
    "orr_server_submit": [_vp, ctypes.POINTER(OrrRequest), ctypes.POINTER(ctypes.c_uint64)],
    "orr_server_round": [_vp, ctypes.POINTER(ctypes.c_uint32)],
    "orr_server_status": [_vp, ctypes.c_uint64, ctypes.POINTER(OrrRequestStatus)],
    "orr_server_result": [_vp, ctypes.c_uint64] + [_vp] * 6,
    "orr_server_stats_get": [_vp, ctypes.POINTER(OrrServerStats)],
    "orr_plummer_sample": [ctypes.c_uint64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                           ctypes.c_uint64] + [_d] * 7,
}

_LIB = None


class EngineError(RuntimeError):
    """A non-zero status from the engine (the reference's CudaError, cuda_memory.hpp:129-153)."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


class NoDeviceError(EngineError):
    pass


def library_path() -> str:
    return LIBRARY


def load():
    """Load the engine library; raise if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBRARY):
        raise RuntimeError(
            f"{LIBRARY} is not built; run `python -m orrery_amd.build` (there is no CPU fallback)")
    lib = ctypes.CDLL(LIBRARY)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_char_p if name == "orr_last_error" else ctypes.c_int
    _LIB = lib
    return lib


def check(status: int) -> None:
    if status == ORR_OK:
        return
    msg = load().orr_last_error().decode(errors="replace")
    if status == ORR_ERR_NO_DEVICE:
        raise NoDeviceError(status, msg)
    raise EngineError(status, msg)
