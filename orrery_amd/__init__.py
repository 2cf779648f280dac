"""MI355X-native force engine for Orrery: direct summation and Barnes-Hut on gfx950.

The compute path is liborrery_hip.so (hand-written HIP kernels behind the C ABI
of include/orrery_hip.h). This package is its Python host; it has no CPU
fallback and fails loudly when the library is not built.
"""
from .solvers import (BatchConfig, BatchSystem, Context, DeviceArray, DeviceBatch, DeviceSystem,
                      HipDirectSolver, HipTreeSolver, InteractionCount, Server, ServerConfig,
                      Softening, StepConfig, Trajectory, TreeParameters, describe, device_count,
                      morton_order)

__all__ = ["BatchConfig", "BatchSystem", "Context", "DeviceArray", "DeviceBatch", "DeviceSystem",
           "HipDirectSolver", "HipTreeSolver", "InteractionCount", "Server", "ServerConfig",
           "Softening", "StepConfig", "Trajectory", "TreeParameters", "describe", "device_count",
           "morton_order"]
