#!/usr/bin/env python3
"""Benchmark of the multi-layer XPBD substep on one MI355X.

Workload: the 81,920-tet box beam (64 x 16 x 16 hex cells x 5 tets, 0.02 m
cells, x = 0 face pinned, per-element E log-uniform in [1e4, 1e7] Pa and nu
uniform in [0.30, 0.45] from PCG32 seed 0, rho = 1000, alpha1 = 0.02),
h = 1/60 s, gravity -9.81 on z, ground plane z = 0, layers
[98.4375, 87.5, 0] (two coarse layers with 1/64 and 1/8 of the elements
elastic) with iterations [2, 4, 14], and the relative-residual stop
(StopAtImprovementPercent = 99.99, i.e. residual <= 1e-4 x the first one,
hangingStop 1000), useRigidContacts off (see make_integrator).  Bulk
projection in FP32, rigid layer and residual reductions in FP64.  The five
warmup substeps are the free fall (one iteration each: the first residual
is under the 1e-8 floor); the timed ones start at touchdown.

A step is one integrate() substep: layer maps, layerSolve to the stop rule,
velocity update.  value = tet projections / s over the timed steps (every
elastic element update in every iteration); ms_per_step = wall ms/substep.
The roofline entry is for the element-sweep kernel: over --profile-steps
substeps that follow the timed ones, one HIP event pair (no system fence)
brackets the back-to-back color launches of every iteration on the mesh's
stream (direct launches; the timed substeps replay the same kernels from
captured hipGraphs), and the average launch duration is that time over the
number of launches.  `phases` splits the timed substeps' host wall time into
contacts + layer maps and layerSolve.  cpu_baseline is the FP64 CPU oracle
(a restatement of the reference path, one core) on a bounded sample.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from multilayersolver2024_amd import _lib as L  # noqa: E402
from multilayersolver2024_amd import scenes  # noqa: E402
from multilayersolver2024_amd.integrator import (PlaneContactFinder, Simulation,  # noqa: E402
                                                 XpbdLayer3D)
from multilayersolver2024_amd.mesh import Mesh  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of one tet projection in FP32: 4 vertex ids 16 + x read 48
# + x write 48 + x_old 48 + 1/m 16 + DmInv 36 + vol 4 + material
# (alpha_d, alpha_o, alpha_s, alpha1) 16 + lambda read 24 + lambda write 24.
BYTES_PER_TET = {"fp32": 280, "fp64": 544}
# Algorithmic bytes of one element's residual evaluation in FP32: 4 vertex
# ids 16 + x read 48 + DmInv 36 + vol 4 + (alpha_d, alpha_o, alpha_s) 12 +
# lambda read 24 (FP64: 16 + 96 + 72 + 8 + 24 + 48).
RESIDUAL_BYTES_PER_TET = {"fp32": 140, "fp64": 264}
# HBM-side bytes per dispatch on the default workloads, from the rocprofv3
# FETCH_SIZE / WRITE_SIZE runs (KiB per dispatch; k_sweep3d averaged over
# 8,200 dispatches of the 81,920-tet beam with 64-thread workgroups,
# k_batch_plain over the 256-beam substeps at 1,024 threads), scaled by what
# the same counters read for a 1 GiB device copy (FETCH_SIZE reads half the
# bytes fetched, WRITE_SIZE all).
TRAFFIC_KIB = {"k_sweep3d<float>": {"fetch": 197.73, "fetch_scale": 2.0,
                                    "write": 96.19, "write_scale": 1.0},
               "k_batch_plain<float,1024>": {"fetch": 15336905.12, "fetch_scale": 2.0,
                                             "write": 3388201.03, "write_scale": 1.0}}


def traffic_bytes(kernel):
    tk = TRAFFIC_KIB.get(kernel)
    return (1024 * (tk["fetch"] * tk["fetch_scale"] + tk["write"] * tk["write_scale"])
            if tk else None)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_integrator(stop_rule=True, iterations=(2, 4, 14), hanging_stop=1000, rigid_contacts=False):
    """Integrator of the plain workload.  useRigidContacts is off: the
    reference's 3D rigid contact (penaltyConstraint.projectRigid with
    rigidGeneralizedInverseMass, xpbdLayer3D.m:445-484) does not split the
    body's mass over its simultaneous contacts, so when the 98.4 %-rigid
    coarse body lands on ~1,100 ground contacts its position correction is
    ~550x too large and the substep diverges to NaN two substeps after
    touchdown (the oracle and the GPU agree on it); a NaN run would iterate
    to hangingStop and time garbage.  Contacts are then resolved in the
    elastic layer only, as useRigidContacts = false does in the reference."""
    it = XpbdLayer3D()
    it.Gravity = -9.81
    it.setLayers([98.4375, 87.5, 0])
    it.iterations = list(iterations)
    it.computeResiduals = True
    if stop_rule:
        it.StopAtImprovementPercent = 99.99
    it.hangingStop = hanging_stop
    it.useRigidContacts = rigid_contacts
    return it


def bench_batched(h, plane, steps, warmup, precision, nscenes=256, graph=True):
    """Batched form: nscenes beams (tet counts log-uniform in [2000, 60000],
    random aspect ratios, per-scene E log-uniform in [1e4, 1e7]) integrated
    in one persistent launch per substep, plain method (one all-elastic
    layer, 20 iterations, the same 1e-4 relative-residual stop per scene).
    The block size and the scene split (target tets per workgroup; 0 = one
    workgroup per scene) are chosen by timing each candidate pair on the
    same substep."""
    from multilayersolver2024_amd.batch import BatchSimulation, merge_meshes, random_beams

    merged = merge_meshes([Mesh(s) for s in random_beams(nscenes, seed=0)])
    it = XpbdLayer3D()
    it.Gravity = -9.81
    it.iterations = [20]
    it.computeResiduals = True
    it.StopAtImprovementPercent = 99.99
    it.hangingStop = 1000
    prec = L.MLX_FP32 if precision == "fp32" else L.MLX_FP64
    b = BatchSimulation(merged, it, h, plane, precision=prec, block=512)
    # warmup and autotune launch the kernels directly: the kernel time per
    # candidate comes from its own events
    b.set_graph(0)
    for _ in range(warmup):
        b.step()
    # block-size autotune: every candidate runs the same substep from the
    # post-warmup state
    p0, v0, _ = b.download()
    frame0 = b.frame
    tune = {}
    for blk in (256, 512, 768, 1024):
        for split in (0, 8000, 25000):
            b.set_block(blk)
            b.set_split(split)
            b.set_state(p0, v0)
            b.frame = frame0
            proj, kms = b.step()
            tune[(blk, split)] = proj / (kms / 1e3)
            log(f"batched block {blk} split {split}: {kms:.2f} ms kernel, "
                f"{tune[(blk, split)]:.3e} tet-proj/s")
    best = max(tune, key=tune.get)
    b.set_block(best[0])
    b.set_split(best[1])
    b.set_state(p0, v0)
    b.frame = frame0
    # the timed substeps replay the captured graph of the whole substep; the
    # first replayable substep captures it and is not timed
    b.set_graph(int(graph))
    b.step()
    p1, v1, _ = b.download()
    frame1 = b.frame
    L.check("mlx_synchronize", L.lib().mlx_synchronize())
    t0 = time.perf_counter()
    proj, iters, iters_max = 0, [], []
    for _ in range(steps):
        pr, _ = b.step()
        proj += pr
        iters.append(float(np.mean(b.iters)))
        iters_max.append(int(np.max(b.iters)))
    L.check("mlx_synchronize", L.lib().mlx_synchronize())
    wall = time.perf_counter() - t0
    p_end, v_end, _ = b.download()
    # kernel duration for the roofline: the same substeps again from the
    # same state with the kernels launched directly, the solve bracketed by
    # the batch's own events
    b.set_graph(0)
    b.set_state(p1, v1)
    b.frame = frame1
    kproj, kms, ksteps = 0, 0.0, steps
    for _ in range(ksteps):
        pr, k = b.step()
        kproj += pr
        kms += k
    b.close()
    # one element-iteration of the batched kernel = a projection and a
    # residual evaluation of that element (every element is evaluated every
    # iteration)
    ftype = "float" if precision == "fp32" else "double"
    bpt = BYTES_PER_TET[precision] + RESIDUAL_BYTES_PER_TET[precision]
    achieved = kproj * bpt / (kms / 1e3) / 1e9
    return {
        "value": proj / wall, "unit": "tet-proj/s", "ms_per_step": 1e3 * wall / steps,
        "steps": steps, "warmup": warmup,
        "workload": f"{nscenes} independent beams, {merged.nT} tets total "
                    f"({min(m.nT for m in merged.meshes)}..{max(m.nT for m in merged.meshes)} per "
                    f"scene), plain method, 20 iterations + rel. residual 1e-4 stop per scene, "
                    f"ground contact, h=1/60, one solve launch per substep",
        "block_size": best[0], "split_tets_per_workgroup": best[1],
        "autotune_tet_proj_per_s": {f"{k[0]}/{k[1]}": v for k, v in tune.items()},
        "mean_iterations_per_scene_substep": float(np.mean(iters)),
        "max_iterations_per_scene_substep": int(np.max(iters_max)),
        "finite": bool(np.isfinite(p_end).all() and np.isfinite(v_end).all()),
        "roofline": {"bound": "hbm", "kernel": f"k_batch_plain<{ftype},{best[0]}>"
                     if best[1] == 0 else f"k_batch_split<{ftype},{best[0]}>",
                     "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_bytes(f"k_batch_plain<{ftype},{best[0]}>")
                     if best[1] == 0 and nscenes == 256 else None,
                     "bytes_per_unit": bpt, "unit_of_work": "element-iteration (projection + "
                     "residual evaluation)", "avg_launch_ms": kms / ksteps,
                     "measured_on": "the timed substeps rerun from the same state with the "
                                    "kernels launched directly"},
        "graph": "substeps replay one captured hipGraph (contacts, prologue, solve, velocity)"
                 if graph else "substeps launched directly",
    }


def cpu_baseline(mesh, h, plane, budget_s=12.0, max_steps=4):
    """The oracle (FP64, one core) on the same scene with the iteration counts
    fixed at [2, 4, 14] (the stop rule off), substeps 1.. until budget_s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.build()
    osim = O.OracleSim(mesh, make_integrator(stop_rule=False), h, plane)
    proj, t, steps = 0, 0.0, 0
    while steps < max_steps and (t < budget_s or steps == 0):
        t0 = time.perf_counter()
        r = osim.step()
        t += time.perf_counter() - t0
        proj += r["projections"]
        steps += 1
    return dict(value=proj / t, unit="tet-proj/s", cores=1, kind="port",
                sample=f"{steps} substeps of the same beam from rest, iterations fixed at "
                       f"[2,4,14] (no stop rule), residual monitor on; "
                       f"{proj} tet projections in {t:.2f} s; {1e3 * t / steps:.1f} ms/substep")


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--ny", type=int, default=16)
    ap.add_argument("--nz", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--hanging-stop", type=int, default=1000,
                    help="iteration cap of the stop rule (reference default 1000)")
    ap.add_argument("--graphs", type=int, default=1,
                    help="replay iteration chunks as hipGraphs (1) or launch directly (0)")
    ap.add_argument("--persistent", type=int, default=0,
                    help="run layers without rigid elements in the persistent cooperative loop")
    ap.add_argument("--persist-wg", type=int, default=0,
                    help="workgroups of the persistent iteration loop (0 = from the mesh size)")
    ap.add_argument("--persist-local", type=int, default=1,
                    help="persistent loop on one XCD's workgroups with L2-keeping barriers")
    ap.add_argument("--sweep-block", type=int, default=64, choices=[64, 128, 256],
                    help="threads per workgroup of the per-color sweep launches")
    ap.add_argument("--sweep-wg", type=int, default=0,
                    help="sparse rigid layers sweep all colors in one workgroup (1) or launch "
                         "per color (0)")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--batch-graph", type=int, default=1, choices=[0, 1],
                    help="timed batched substeps replay a captured graph (1) or launch directly (0, for kernel-trace profiling)")
    ap.add_argument("--batch-warmup", type=int, default=6)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only; --gpus must be 1 (multi-GPU sharding of the "
              "scene is a design note, not a built path)", file=sys.stderr)
        sys.exit(2)

    h = 1.0 / 60.0
    sc = scenes.beam3d(args.nx, args.ny, args.nz, seed=0)
    mesh = Mesh(sc)
    plane = PlaneContactFinder([0, 0, 1], [0, 0, 0])
    prec = L.MLX_FP32 if args.precision == "fp32" else L.MLX_FP64
    log(f"plain: {mesh.nT} tets, {mesh.nV} vertices, {mesh.numColors} colors")
    sim = Simulation(mesh, make_integrator(hanging_stop=args.hanging_stop), h, plane,
                     precision=prec)
    lib = L.lib()
    sim.dmesh.set_option(L.MLX_OPT_GRAPHS, args.graphs)
    sim.dmesh.set_option(L.MLX_OPT_PERSISTENT, args.persistent)
    sim.dmesh.set_option(L.MLX_OPT_PERSIST_WG, args.persist_wg)
    sim.dmesh.set_option(L.MLX_OPT_PERSIST_LOCAL, args.persist_local)
    sim.dmesh.set_option(L.MLX_OPT_SWEEP_BLOCK, args.sweep_block)
    sim.dmesh.set_option(L.MLX_OPT_SWEEP_WG, args.sweep_wg)
    ncol, _ = sim.dmesh.info()
    for _ in range(args.warmup):
        sim.step()
    log(f"warmup done ({args.warmup} substeps)")
    L.check("mlx_synchronize", lib.mlx_synchronize())
    ph = [C.c_double() for _ in range(3)]
    nsub = C.c_int64()
    L.check("mlx_profile_phases", lib.mlx_profile_phases(sim.dmesh.handle, *[C.byref(x) for x in ph],
                                                         C.byref(nsub)))
    t0 = time.perf_counter()
    proj, iters, contacts = 0, [], 0
    for _ in range(args.steps):
        r = sim.step()
        proj += r.elasticProjections
        iters.append(sum(r.itersPerLayer))
    L.check("mlx_synchronize", lib.mlx_synchronize())
    t1 = time.perf_counter()
    L.check("mlx_profile_phases", lib.mlx_profile_phases(sim.dmesh.handle, *[C.byref(x) for x in ph],
                                                         C.byref(nsub)))
    phases = {"layer_maps_ms_per_step": ph[0].value / max(nsub.value, 1),
              "layer_solve_ms_per_step": ph[1].value / max(nsub.value, 1),
              "integrate_ms_per_step": ph[2].value / max(nsub.value, 1)}
    # Kernel timing window: the next substeps.  The persistent loop (layers
    # without rigid elements) is timed with one HIP event pair per launch;
    # the per-color sweeps of the rigid layers with one pair around each
    # iteration's back-to-back color launches.
    log(f"timed: {args.steps} substeps in {t1 - t0:.3f} s")
    L.check("mlx_profile_enable", lib.mlx_profile_enable(sim.dmesh.handle, 1))
    for _ in range(args.profile_steps):
        sim.step()
    ms, pms = C.c_double(), C.c_double()
    launches, elems = C.c_int64(), C.c_int64()
    plaunches, pelems, piters = C.c_int64(), C.c_int64(), C.c_int64()
    L.check("mlx_profile_read", lib.mlx_profile_read(sim.dmesh.handle, C.byref(ms),
                                                     C.byref(launches), C.byref(elems)))
    L.check("mlx_profile_read_persistent", lib.mlx_profile_read_persistent(
        sim.dmesh.handle, C.byref(pms), C.byref(plaunches), C.byref(pelems), C.byref(piters)))
    L.check("mlx_profile_enable", lib.mlx_profile_enable(sim.dmesh.handle, 0))
    pfin, vfin, _, contacts = sim.download()
    finite = bool(np.isfinite(pfin).all() and np.isfinite(vfin).all())
    sim.close()
    wall = t1 - t0
    bpt = BYTES_PER_TET[args.precision]
    ftype = "float" if args.precision == "fp32" else "double"
    sweep_s = ms.value / 1e3
    sweeps = {"kernel": f"k_sweep3d<{ftype}>", "ms": ms.value, "launches": launches.value,
              "elements": elems.value, "bytes_per_unit": bpt,
              "avg_launch_us": 1e3 * ms.value / max(launches.value, 1),
              "achieved_GBs": (elems.value * bpt) / sweep_s / 1e9 if sweep_s > 0 else 0.0}
    if plaunches.value > 0:
        # one unit = one element in one iteration: its projection plus its
        # residual evaluation (the monitor runs every iteration)
        bpu = bpt + RESIDUAL_BYTES_PER_TET[args.precision]
        pk_s = pms.value / 1e3
        achieved = pelems.value * bpu / pk_s / 1e9
        roof = {"kernel": f"k_persist<{ftype},3>", "bytes_per_unit": bpu,
                "avg_launch_us": 1e3 * pms.value / plaunches.value, "launches": plaunches.value,
                "elements_per_launch": pelems.value / plaunches.value,
                "iterations_per_launch": piters.value / plaunches.value,
                "per_color_sweeps": sweeps}
    else:
        achieved = sweeps["achieved_GBs"]
        roof = {"kernel": sweeps["kernel"], "bytes_per_unit": bpt,
                "avg_launch_us": sweeps["avg_launch_us"], "launches": launches.value,
                "elements_per_launch": elems.value / max(launches.value, 1)}
    traffic = (traffic_bytes(roof["kernel"]) if (args.nx, args.ny, args.nz) == (64, 16, 16)
               and args.sweep_block == 64 else None)
    line = {
        "metric": "ms/substep & tet-proj/s to rel. residual 1e-4: plain vs learned+batched, 1 GPU",
        "value": proj / wall,
        "unit": "tet-proj/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"plain multi-layer XPBD, box beam {args.nx}x{args.ny}x{args.nz} hex x 5 tets "
                        f"({mesh.nT} tets, {mesh.nV} vertices), StVK-Voigt, layers [98.4375,87.5,0] "
                        f"iterations [2,4,14], rel. residual 1e-4 stop, ground contact, h=1/60",
            "form": "plain",
            "precision": args.precision,
            "graphs": bool(args.graphs),
            "sweep_block": args.sweep_block, "sweep_wg": bool(args.sweep_wg),
            "persistent": bool(args.persistent),
            "persist_local": bool(args.persist_local) if args.persistent else None,
            "colors": ncol,
            "mean_iterations_per_substep": float(np.mean(iters)),
            "tet_projections": int(proj),
            "contacts_last_substep": int(contacts),
            "finite": finite,
            "rigid_contacts": False,
            "phases": phases,
        },
        "roofline": dict({
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
        }, **roof),
    }
    if not args.no_batched:
        log("batched leg")
        line["batched"] = bench_batched(h, plane, args.batch_steps, args.batch_warmup,
                                        args.precision, graph=bool(args.batch_graph))
    if not args.no_cpu_baseline:
        log("cpu baseline")
        line["cpu_baseline"] = cpu_baseline(mesh, h, plane)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
