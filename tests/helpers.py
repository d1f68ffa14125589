"""Shared scene/integrator setups for the tests."""
import numpy as np

from multilayersolver2024_amd import scenes
from multilayersolver2024_amd.integrator import PlaneContactFinder, XpbdLayer2D, XpbdLayer3D
from multilayersolver2024_amd.mesh import Mesh


def beam_case(nx=8, ny=4, nz=4, layers=(98.4375, 87.5, 0), iterations=(2, 4, 14), seed=0,
              residuals=True, pct=0.0, z0=0.05):
    sc = scenes.beam3d(nx, ny, nz, seed=seed, z0=z0)
    m = Mesh(sc)
    it = XpbdLayer3D()
    it.Gravity = -9.81
    it.setLayers(list(layers))
    it.iterations = list(iterations)
    it.computeResiduals = residuals
    it.StopAtImprovementPercent = pct
    return m, it, PlaneContactFinder([0, 0, 1], [0, 0, 0])


def cantilever_case():
    m = Mesh(scenes.cantilever2d())
    it = XpbdLayer2D()
    it.Gravity = -9.81
    it.iterations = [10]
    return m, it


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
