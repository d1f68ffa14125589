"""MI355X-native multi-layer XPBD substep (drop-in for the layerSolve path of
AlexandreMercierAubin/MultiLayerSolver2024).

The GPU work is in libmlxpbd.so (HIP, gfx950), reached through the C ABI of
include/mlxpbd.h; this package holds the ctypes binding, scene builders and
the host-side mirror of the reference's integrator interface.
"""
from . import scenes  # noqa: F401

__all__ = ["scenes", "mesh", "integrator"]
