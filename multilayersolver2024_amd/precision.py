"""Precision study: FP32 bulk projection against an all-FP64 run of the
same scene on the GPU.

The scene is the 655,360-tet beam (128 x 32 x 32 hex cells, 5 tets per
cube, per-element E log-uniform in [1e4, 1e7] Pa and nu uniform in
[0.30, 0.45] from the seed), both end faces pinned, the far end cap turned
about the beam axis at 1 rad/s (scripted pins, `Simulation.set_pinned_targets`;
the reference freezes pinned positions at frame 1, so the twist is an
extension).  In the FP32 run the element projection, positions and
lambdas are FP32; residual and energy reductions, rigid-body sums,
contacts and the layer maps stay FP64.  Per substep the study reports the
relative L2 error of positions and lambdas against the FP64 run, the
relative error of the elastic (StVK) energy, and both runs' iteration
counts.

    python -m multilayersolver2024_amd.precision --steps 300 --method plain

prints one JSON object per substep and a summary line.
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import _lib as L
from . import scenes
from .integrator import PlaneContactFinder, Simulation, XpbdLayer3D
from .mesh import Mesh


def stvk_energy(mesh, scene, p):
    """Sum over tets of vol * (mu |E|_F^2 + lam/2 tr(E)^2), E = (F^T F - I)/2,
    F = [p1-p0, p2-p0, p3-p0] DmInv (FP64 on the host)."""
    X = np.asarray(p, dtype=np.float64).reshape(-1, 3)
    T = mesh.T0
    Ds = np.stack([X[T[:, 1]] - X[T[:, 0]], X[T[:, 2]] - X[T[:, 0]], X[T[:, 3]] - X[T[:, 0]]],
                  axis=2)  # nT x 3 (rows xyz) x 3 (edge)
    D = mesh.DmInv.reshape(-1, 3, 3).transpose(0, 2, 1)  # pages are column-major
    F = np.einsum("nrk,nkc->nrc", Ds, D)
    E = 0.5 * (np.einsum("nrj,nrk->njk", F, F) - np.eye(3))
    tr = np.trace(E, axis1=1, axis2=2)
    W = mesh.elV * (scene["mu"] * np.einsum("njk,njk->n", E, E) + 0.5 * scene["lam"] * tr * tr)
    return float(W.sum())


def integrator_for(method):
    it = XpbdLayer3D()
    it.Gravity = -9.81
    if method == "plain":
        it.setLayers([0])
        it.iterations = [20]
    else:
        it.setLayers([98.4375, 87.5, 0])
        it.iterations = [2, 4, 14]
    it.computeResiduals = True
    it.StopAtImprovementPercent = 99.99
    it.hangingStop = 1000
    return it


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def run(nx, ny, nz, steps, method, seed=0, omega=1.0, h=1.0 / 60.0, out=sys.stdout):
    import copy

    sc = scenes.beam3d(nx, ny, nz, seed=seed, z0=0.5, pin_far_end=True)
    mesh = Mesh(sc)
    plane = PlaneContactFinder([0, 0, 1], [0, 0, 0])
    it = integrator_for(method)
    s64 = Simulation(mesh, it, h, plane, precision=L.MLX_FP64)
    s32 = Simulation(mesh, copy.deepcopy(it), h, plane, precision=L.MLX_FP32)
    rows = []
    try:
        for k in range(steps):
            if k >= 1:
                tgt = scenes.twist_pinned_targets(mesh, omega * k * h)
                s64.set_pinned_targets(tgt)
                s32.set_pinned_targets(tgt)
            r64 = s64.step()
            r32 = s32.step()
            p64, v64, l64, _ = s64.download()
            p32, v32, l32, _ = s32.download()
            e64 = stvk_energy(mesh, sc, p64)
            e32 = stvk_energy(mesh, sc, p32)
            row = {"substep": k + 1, "pos_rel": rel(p32, p64), "vel_rel": rel(v32, v64),
                   "lambda_rel": rel(l32, l64),
                   "energy_rel": abs(e32 - e64) / max(abs(e64), 1e-300), "energy_fp64": e64,
                   "iters_fp64": list(r64.itersPerLayer), "iters_fp32": list(r32.itersPerLayer),
                   "finite": bool(np.isfinite(p32).all() and np.isfinite(p64).all())}
            rows.append(row)
            print(json.dumps(row), file=out, flush=True)
    finally:
        s64.close()
        s32.close()
    summ = {"summary": True, "scene": f"beam {nx}x{ny}x{nz} hex x 5 tets ({mesh.nT} tets, "
                                      f"{mesh.nV} vertices), both ends pinned, far end cap "
                                      f"turned at {omega} rad/s",
            "method": method, "steps": steps,
            "pos_rel_max": max(r["pos_rel"] for r in rows),
            "lambda_rel_max": max(r["lambda_rel"] for r in rows),
            "energy_rel_max": max(r["energy_rel"] for r in rows),
            "pos_rel_final": rows[-1]["pos_rel"], "energy_rel_final": rows[-1]["energy_rel"],
            "iters_equal_fraction": float(np.mean([r["iters_fp64"] == r["iters_fp32"]
                                                   for r in rows])),
            "all_finite": all(r["finite"] for r in rows)}
    print(json.dumps(summ), file=out, flush=True)
    return rows, summ


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=128)
    ap.add_argument("--ny", type=int, default=32)
    ap.add_argument("--nz", type=int, default=32)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--method", choices=["plain", "multilayer"], default="plain")
    ap.add_argument("--omega", type=float, default=1.0)
    args = ap.parse_args(argv)
    t0 = time.perf_counter()
    run(args.nx, args.ny, args.nz, args.steps, args.method, omega=args.omega)
    print(json.dumps({"wall_s": time.perf_counter() - t0}), flush=True)


if __name__ == "__main__":
    main()
