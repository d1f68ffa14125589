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
After the warmup the window runs once untimed (capturing the iteration
graphs it needs) and is then timed from the same state; the other legs
restart from that state too: `monitor_excluded` (the window with each
substep's iteration counts fixed to the stop rule's and the residual monitor
off, the time the reference reports), `fp64` (the same window in FP64) with
`fp32_error_vs_fp64`, and `learned` (the committed network attached);
`neohookean_beam` runs the same beam with NeoHookean tetrahedra over a
short window after the same warmup.
The roofline entry is for the element-sweep kernel: over --profile-steps
substeps that follow the timed ones, one HIP event pair (no system fence)
brackets the back-to-back color launches of every iteration on the mesh's
stream (direct launches; the timed substeps replay the same kernels from
captured hipGraphs), and the average launch duration is that time over the
number of launches: kernel time plus the dispatch gap to the next launch
(a pair around every launch adds more than the gap).  `phases` splits the timed substeps' host wall time into
contacts + layer maps and layerSolve.  cpu_baseline is the FP64 CPU oracle
(a C restatement of the reference path) on one substep from the window-start
state with the same stop rule, median of 5 runs, on one core and with OpenMP
over each color's elements on every core the process may use.
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
# HBM-side traffic from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs (KiB),
# scaled by what the same counters read for a 1 GiB device copy (FETCH_SIZE
# reads half the bytes fetched, WRITE_SIZE all).  k_sweep3d: per launched
# thread slot (one slot per slot of the color, elastic or not), averaged over
# the 12,872 sweep dispatches of the 81,920-tet beam run (64-thread
# workgroups); a launch's traffic is that times its slots.  k_batch_plain:
# per dispatch (one per substep of the 256 beams, 1,024 threads).
TRAFFIC_KIB = {"k_sweep3d<float>": {"fetch": 114.394 / 1024, "fetch_scale": 2.0,
                                    "write": 55.865 / 1024, "write_scale": 1.0,
                                    "per": "thread slot"},
               "k_batch_plain<float,1024>": {"fetch": 15498759.3, "fetch_scale": 2.0,
                                             "write": 3375125.8, "write_scale": 1.0,
                                             "per": "dispatch"},
               # the three bucket dispatches (512 threads each) of one substep
               "k_batch_layered<float,512>": {"fetch": 92102241.8, "fetch_scale": 2.0,
                                              "write": 14835576.8, "write_scale": 1.0,
                                              "per": "substep"}}
# Where the constants above come from.
TRAFFIC_SOURCE = ("rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of this bench (--graphs 0; "
                  "the sweep from the plain leg over 12,872 dispatches, the plain batched kernel "
                  "from the plain batched leg with --batch-legs plain --batch-splits 0, 6 "
                  "dispatches, the cluster kernel from --batch-legs clusters --batch-graph 0, the "
                  "4 timed substeps' 12 bucket dispatches), scaled by the same counters of a 1 GiB "
                  "device copy (FETCH x2.0, WRITE x1.0); commands and figures in "
                  "profiles/hbm_traffic_*.md")
# Digest of the kernel sources (multilayersolver2024_amd/csrc) the counters
# above were read on.  When the sources differ, the constants may no longer
# describe the kernels that run: the bench then reports no traffic figure
# and says why, instead of a stale one.
TRAFFIC_SOURCES_DIGEST = "20578410e67ea1e8"
CSRC = os.path.join(ROOT, "multilayersolver2024_amd", "csrc")


def kernel_sources_digest():
    """SHA-256 (first 16 hex digits) over the names and bytes of the kernel
    sources, in name order."""
    import hashlib
    h = hashlib.sha256()
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".h", ".hip", ".cpp")):
            h.update(name.encode())
            with open(os.path.join(CSRC, name), "rb") as f:
                h.update(f.read())
    return h.hexdigest()[:16]


def traffic_current():
    return kernel_sources_digest() == TRAFFIC_SOURCES_DIGEST


TRAFFIC_STALE = ("null: the kernel sources differ from those the counters were read on "
                 "(kernel_sources_digest() != TRAFFIC_SOURCES_DIGEST); rerun the FETCH_SIZE and "
                 "WRITE_SIZE passes of profiles/hbm_traffic_*.md")


def traffic_bytes(kernel, units=1.0):
    """HBM bytes of one launch of `kernel` covering `units` (thread slots
    for the sweep, 1 for a batched dispatch); None when the kernel has no
    counter reading or the sources changed since it was taken."""
    tk = TRAFFIC_KIB.get(kernel)
    if tk is None or not traffic_current():
        return None
    return 1024 * units * (tk["fetch"] * tk["fetch_scale"] + tk["write"] * tk["write_scale"])


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


def bench_batched(h, plane, steps, warmup, precision, nscenes=256, graph=True,
                  splits=(0, 8000, 25000)):
    """Batched form: nscenes beams (tet counts log-uniform in [2000, 60000],
    random aspect ratios, per-scene E log-uniform in [1e4, 1e7]) integrated
    in one persistent launch per substep, plain method (one all-elastic
    layer, 20 iterations, the same 1e-4 relative-residual stop per scene).
    The block size and the scene split (target tets per workgroup; 0 = one
    workgroup per scene) are chosen by timing each candidate pair on the
    same substep (split candidates other than 0 are cooperative launches)."""
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
        for split in splits:
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
                     "traffic_source": None if traffic_current() else TRAFFIC_STALE,
                     "bytes_per_unit": bpt, "unit_of_work": "element-iteration (projection + "
                     "residual evaluation)", "avg_launch_ms": kms / ksteps,
                     "measured_on": "the timed substeps rerun from the same state with the "
                                    "kernels launched directly"},
        "graph": "substeps replay one captured hipGraph (contacts, prologue, solve, velocity)"
                 if graph else "substeps launched directly",
    }


def host_info():
    """CPU model and the cores this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return model, cores


def oracle_from_snapshot(O, mesh, h, plane, snap, integ):
    """The oracle's integrator positioned at a GPU snapshot: positions,
    velocities, cache.oldp, pinnedx (the pinned DOFs' positions) and
    runUntil, so its next substep is the one the GPU runs from there."""
    osim = O.OracleSim(mesh, integ, h, plane)
    osim.p = snap["p"].copy()
    osim.v = snap["v"].copy()
    osim.oldp = snap["oldp"].copy()
    osim.pinnedx = snap["p"][np.asarray(mesh.pinnedDOFs) - 1].copy()
    osim.frame = int(snap["frame"])
    osim.s.runUntil = float(snap["runUntil"])
    return osim


def bench_batched_clusters(h, plane, steps, warmup, precision, nscenes=256, graph=True,
                           model_path=None):
    """The batched workload as specified: nscenes beams (tet counts
    log-uniform in [2000, 60000], random aspect ratios, per-scene E
    log-uniform in [1e4, 1e7]), odd scenes NeoHookean, each with its own
    ground contact and 1..3 rigid clusters (slabs across the beam) that form
    the coarse layer; per scene the two-layer layerSolve (iterations [2, 18],
    the 1e-4 relative-residual stop, hangingStop 1000) in one launch per
    color bucket, the block size per bucket chosen by the library's
    autotuner (mlx_batch_autotune), the substep replayed as a captured
    hipGraph."""
    from multilayersolver2024_amd.batch import BatchSimulation, cluster_workload, merge_meshes

    scs, masks = cluster_workload(nscenes, seed=0)
    meshes = [Mesh(s) for s in scs]
    merged = merge_meshes(meshes)
    it = XpbdLayer3D()
    it.Gravity = -9.81
    it.setLayers([50, 0])
    it.iterations = [2, 18]
    it.computeResiduals = True
    it.StopAtImprovementPercent = 99.99
    it.hangingStop = 1000
    it.useRigidContacts = False
    prec = L.MLX_FP32 if precision == "fp32" else L.MLX_FP64
    b = BatchSimulation(merged, it, h, plane, precision=prec, block=256)
    b.set_rigid(np.concatenate(masks))
    b.set_graph(0)
    for _ in range(warmup):
        b.step()
    tuned = b.autotune()
    log(f"batched clusters: autotune {tuned}")
    p0, v0, _ = b.download()
    frame0 = b.frame
    b.set_graph(int(graph))
    b.step()  # captures the graph
    p1, v1, _ = b.download()
    frame1 = b.frame
    L.check("mlx_synchronize", L.lib().mlx_synchronize())
    t0 = time.perf_counter()
    proj, its = 0, []
    for _ in range(steps):
        pr, _ = b.step()
        proj += pr
        its.append(b.iters.copy())
    L.check("mlx_synchronize", L.lib().mlx_synchronize())
    wall = time.perf_counter() - t0
    p_end, v_end, _ = b.download()
    # kernel time: the same substeps from the same state, replayed with one
    # HIP event pair around each graph (the buckets' launches run
    # concurrently inside it; launched directly from their streams they
    # partly serialise)
    b.set_state(p1, v1)
    b.frame = frame1
    kproj, kms, kres = 0, 0.0, 0
    nts = np.diff(merged.elemOff)
    for _ in range(steps):
        pr, k = b.step()
        kproj += pr
        kms += k
        kres += int(np.dot(b.iters, nts))
    its = np.array(its)
    nh = np.array([m.neohookean.any() for m in meshes])
    # the learned form on the same workload: the committed network attached
    # (mlx_batch_set_learned), the same window from the same state (one
    # untimed substep captures the graph and seeds the multipliers the
    # network reads), each scene's warm start judged by its equation residual
    learned = None
    if model_path and os.path.exists(model_path):
        from multilayersolver2024_amd import learn
        net = learn.Mlp.load(model_path)
        b.set_learned(net, 2.0)
        b.set_state(p0, v0)
        b.frame = frame0
        b.step()
        b.learn_stats()
        L.check("mlx_synchronize", L.lib().mlx_synchronize())
        t0 = time.perf_counter()
        lproj, lits = 0, []
        for _ in range(steps):
            pr, _ = b.step()
            lproj += pr
            lits.append(b.iters.copy())
        L.check("mlx_synchronize", L.lib().mlx_synchronize())
        lwall = time.perf_counter() - t0
        kept, rej = b.learn_stats()
        pl, vl, _ = b.download()
        lits = np.array(lits)
        b.set_learned(None)
        net.close()
        learned = {
            "value": lproj / lwall, "unit": "tet-proj/s", "ms_per_step": 1e3 * lwall / steps,
            "steps": steps,
            "mean_iterations_stvk_scene_substep": float(lits[:, ~nh].mean()),
            "mean_iterations_neohookean_scene_substep": float(lits[:, nh].mean()),
            "warm_starts_kept": kept, "warm_starts_rejected": rej,
            "finite": bool(np.isfinite(pl).all() and np.isfinite(vl).all()),
            "model": os.path.relpath(model_path, ROOT),
            "what": "the batched workload with the learned warm start of every scene's "
                    "all-elastic layer (mlx_batch_set_learned), judged per scene; same window "
                    "length from the post-warmup state, graph replays"}
        log(f"learned batched: {learned['ms_per_step']:.1f} ms/substep, kept {kept}, "
            f"rejected {rej}")
    b.close()
    bpt = BYTES_PER_TET[precision]
    rbytes = RESIDUAL_BYTES_PER_TET[precision]
    achieved = (kproj * bpt + kres * rbytes) / (kms / 1e3) / 1e9
    ftype = "float" if precision == "fp32" else "double"
    return {
        "value": proj / wall, "unit": "tet-proj/s", "ms_per_step": 1e3 * wall / steps,
        "steps": steps, "warmup": warmup,
        "workload": f"{nscenes} independent beams, {merged.nT} tets total "
                    f"({min(m.nT for m in meshes)}..{max(m.nT for m in meshes)} per scene), "
                    f"half NeoHookean, 1..3 rigid clusters per scene (coarse layer), layers "
                    f"[clusters, 0] iterations [2, 18], rel. residual 1e-4 stop, hangingStop 1000, "
                    f"ground contact, h=1/60",
        "autotune": [{"threads": t[0], "kernel_ms": t[1], "scenes": t[2]} for t in tuned],
        "mean_iterations_stvk_scene_substep": float(its[:, ~nh].mean()),
        "mean_iterations_neohookean_scene_substep": float(its[:, nh].mean()),
        "rigid_bodies": int(sum(int(np.asarray(mk).any()) for mk in masks)),
        "finite": bool(np.isfinite(p_end).all() and np.isfinite(v_end).all()),
        "roofline": {"bound": "hbm", "kernel": f"k_batch_layered<{ftype},*> (per bucket)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (traffic_bytes("k_batch_layered<float,512>")
                                 if precision == "fp32" and all(t[0] == 512 for t in tuned)
                                 else None),
                     "traffic_per": "substep (the buckets' dispatches)",
                     "traffic_source": None if traffic_current() else TRAFFIC_STALE,
                     "bytes_per_unit": f"{bpt} per projection + {rbytes} per residual evaluation",
                     "avg_launch_ms": kms / steps,
                     "measured_on": "the timed substeps rerun from the same state, one HIP event "
                                    "pair around each graph replay (the buckets' solves and the "
                                    "substep's small kernels)"},
        "note": "the NeoHookean pair does not converge at h = 1/60 (its scenes run to "
                "hangingStop every substep); the StVK-Voigt scenes stop by the relative rule",
        "learned": learned,
    }


def cpu_baseline(mesh, h, plane, snap, runs=5, threads=None, label="one substep"):
    """The oracle (FP64 C restatement of the reference path) timed on the
    host from the GPU's timed-window start state: one substep with the same
    integrator and stop rule, run `runs` times from that state after one
    untimed run, median reported; serially and with OpenMP over the elements
    of each color on every core this process may use.  Returns the baseline
    and the oracle's one-substep positions from that state."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.build()
    model, ncores = host_info()
    threads = threads or ncores
    legs, p_one = {}, None
    for nt in (1, threads):
        O.lib().ora_set_threads(nt)
        try:
            times, proj, iters = [], None, None
            for k in range(runs + 1):
                osim = oracle_from_snapshot(O, mesh, h, plane, snap, make_integrator())
                t0 = time.perf_counter()
                r = osim.step()
                dt = time.perf_counter() - t0
                if k > 0:
                    times.append(dt)
                proj, iters = r["projections"], sum(r["iters"])
                if p_one is None:
                    p_one = osim.p.copy()
            med = float(np.median(times))
            legs[nt] = dict(value=proj / med, ms_per_substep=1e3 * med, tet_projections=proj,
                            iterations=iters, runs_ms=[1e3 * t for t in times])
            log(f"cpu baseline, {nt} thread(s): {1e3 * med:.0f} ms/substep median, "
                f"{iters} iterations")
        finally:
            O.lib().ora_set_threads(1)
    one, many = legs[1], legs[threads]
    return dict(
        value=many["value"], unit="tet-proj/s", cores=threads, kind="port",
        sample=(f"{label}, run from the GPU's state before it (contacts {snap['contacts']}, "
                f"same layers, iterations and rel. residual 1e-4 stop: "
                f"{many['iterations']} iterations, {many['tet_projections']} tet projections), "
                f"median of {runs} runs after 1 untimed run; OpenMP over the elements of each "
                f"color on {threads} threads, the rest of the substep serial; FP64"),
        ms_per_substep=many["ms_per_substep"], runs_ms=many["runs_ms"],
        single_thread=dict(value=one["value"], ms_per_substep=one["ms_per_substep"],
                           runs_ms=one["runs_ms"], cores=1),
        nproc=ncores, cpu_model=model,
        what="CPU restatement of the reference path (oracle/mlx_oracle.c), not Matlab"), p_one


def run_window(sim, steps, settings_per_step=None):
    """`steps` substeps, wall-timed between device synchronisations; with
    settings_per_step, a callback that adjusts the settings before each."""
    lib = L.lib()
    ph = [C.c_double() for _ in range(3)]
    nsub = C.c_int64()
    L.check("mlx_synchronize", lib.mlx_synchronize())
    L.check("mlx_profile_phases", lib.mlx_profile_phases(sim.dmesh.handle, *[C.byref(x) for x in ph],
                                                         C.byref(nsub)))
    t0 = time.perf_counter()
    proj, iters, per_layer = 0, [], []
    for k in range(steps):
        if settings_per_step is not None:
            settings_per_step(k)
        r = sim.step()
        proj += r.elasticProjections
        iters.append(sum(r.itersPerLayer))
        per_layer.append(list(r.itersPerLayer))
    L.check("mlx_synchronize", lib.mlx_synchronize())
    wall = time.perf_counter() - t0
    L.check("mlx_profile_phases", lib.mlx_profile_phases(sim.dmesh.handle, *[C.byref(x) for x in ph],
                                                         C.byref(nsub)))
    n = max(nsub.value, 1)
    return dict(wall=wall, proj=proj, iters=iters, per_layer=per_layer,
                value=proj / wall, ms_per_step=1e3 * wall / steps,
                mean_iterations=float(np.mean(iters)),
                phases={"layer_maps_ms_per_step": ph[0].value / n,
                        "layer_solve_ms_per_step": ph[1].value / n,
                        "integrate_ms_per_step": ph[2].value / n})


def timed_from(sim, snap, steps, settings_per_step=None, record=None):
    """The window from `snap`: one untimed pass (it captures the iteration
    graphs the window needs; with `record`, a list, the state before each of
    its substeps is appended), then the timed pass from the same state."""
    sim.restore(snap)
    if record is None:
        run_window(sim, steps, settings_per_step)
    else:
        for k in range(steps):
            if settings_per_step is not None:
                settings_per_step(k)
            s_ = sim.snapshot()
            s_["contacts"] = sim.download()[3]
            record.append(s_)
            sim.step()
    sim.restore(snap)
    return run_window(sim, steps, settings_per_step)


def one_substep(sim, snap):
    sim.restore(snap)
    sim.step()
    return sim.download()[0]


def set_options(sim, args):
    sim.dmesh.set_option(L.MLX_OPT_GRAPHS, args.graphs)
    sim.dmesh.set_option(L.MLX_OPT_PERSISTENT, args.persistent)
    sim.dmesh.set_option(L.MLX_OPT_PERSIST_WG, args.persist_wg)
    sim.dmesh.set_option(L.MLX_OPT_PERSIST_LOCAL, args.persist_local)
    sim.dmesh.set_option(L.MLX_OPT_SWEEP_BLOCK, args.sweep_block)
    sim.dmesh.set_option(L.MLX_OPT_SWEEP_WG, args.sweep_wg)
    sim.dmesh.set_option(L.MLX_OPT_FLOW, args.flow)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


LEARNED_MODEL = os.path.join(ROOT, "multilayersolver2024_amd", "learned_warmstart.npz")


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
    ap.add_argument("--no-extra-legs", action="store_true",
                    help="skip the FP64, monitor-excluded and learned legs")
    ap.add_argument("--cpu-runs", type=int, default=5)
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
    ap.add_argument("--flow", type=int, default=0, choices=[0, 1],
                    help="layers without rigid elements sweep their colors in the dataflow kernel "
                         "(k_flow: 64-slot chunks wait for the chunks that last updated their "
                         "vertices instead of one launch per color)")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--batch-graph", type=int, default=1, choices=[0, 1],
                    help="timed batched substeps replay a captured graph (1) or launch directly (0, for kernel-trace profiling)")
    ap.add_argument("--batch-warmup", type=int, default=6)
    ap.add_argument("--batch-legs", choices=["both", "clusters", "plain"], default="both",
                    help="which batched legs run: the cluster workload, the plain StVK one, or both")
    ap.add_argument("--batch-splits", default="0,8000,25000",
                    help="split candidates (tets per workgroup) of the plain batched autotune; "
                         "0 = one workgroup per scene")
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
    set_options(sim, args)
    ncol, sim_buckets = sim.dmesh.info()
    for _ in range(args.warmup):
        sim.step()
    snap0 = sim.snapshot()
    snap0["contacts"] = sim.download()[3]
    log(f"warmup done ({args.warmup} substeps, {snap0['contacts']} contacts)")
    window_states = []
    w = timed_from(sim, snap0, args.steps, record=window_states)
    snap_end = sim.snapshot()
    snap_end["contacts"] = sim.download()[3]
    # the window's representative substep (all layers run, iteration count
    # nearest the window's median): the CPU baseline times that one
    med = float(np.median(w["iters"]))
    cand = [k for k in range(args.steps) if len(w["per_layer"][k]) == 3] or list(range(args.steps))
    k_rep = min(cand, key=lambda k: (abs(w["iters"][k] - med), k))
    snap_rep = window_states[k_rep]
    snap_rep["contacts"] = (window_states[k_rep + 1]["contacts"] if k_rep + 1 < args.steps
                            else snap_end["contacts"])
    log(f"timed: {args.steps} substeps, {w['ms_per_step']:.3f} ms/substep, "
        f"{w['mean_iterations']:.2f} iterations")
    # Kernel timing window: the next substeps.  The persistent loop (layers
    # without rigid elements) is timed with one HIP event pair per launch;
    # the per-color sweeps with one pair around each iteration's
    # back-to-back color launches.
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

    extra = {}
    p32_one = p32_rep = None
    if not args.no_extra_legs:
        # monitor-excluded time (the reference's sumNonMonitoredTime,
        # xpbdLayer3D.m:742): the same window with each substep's iteration
        # counts fixed to the ones the stop rule chose and the residual
        # monitor off -- the same projections in the same order
        st = sim._settings
        if all(len(pl) == len(w["per_layer"][0]) for pl in w["per_layer"]) and \
                len(w["per_layer"][0]) == st.nLayers:
            keep = (st.computeResiduals, st.stopAtImprovementPercent)

            def fixed(k):
                st.computeResiduals = 0
                st.stopAtImprovementPercent = 0.0
                st.runUntilResSmallerThan = -1.0
                for l, n in enumerate(w["per_layer"][k]):
                    st.iterations[l] = n

            nm = timed_from(sim, snap0, args.steps, fixed)
            same = bool(np.array_equal(sim.download()[0], snap_end["p"]))
            st.computeResiduals, st.stopAtImprovementPercent = keep
            for l, n in enumerate((2, 4, 14)):
                st.iterations[l] = n
            extra["monitor_excluded"] = dict(
                ms_per_step=nm["ms_per_step"], value=nm["value"], unit="tet-proj/s",
                same_positions_as_timed=same,
                what="the timed window again with each substep's iteration counts fixed to the "
                     "stop rule's and the residual monitor off (the reference times the solve "
                     "without its monitor, xpbdLayer3D.m:742)")
            log(f"monitor excluded: {nm['ms_per_step']:.3f} ms/substep (same positions: {same})")
        p32_one = one_substep(sim, snap0)
        p32_rep = one_substep(sim, snap_rep)
        # FP64 plain line and the FP32 error against it
        sim64 = Simulation(mesh, make_integrator(hanging_stop=args.hanging_stop), h, plane,
                           precision=L.MLX_FP64)
        set_options(sim64, args)
        w64 = timed_from(sim64, snap0, args.steps)
        p64_end = sim64.download()[0]
        p64_one = one_substep(sim64, snap0)
        sim64.close()
        extra["fp64"] = dict(value=w64["value"], unit="tet-proj/s", ms_per_step=w64["ms_per_step"],
                             mean_iterations_per_substep=w64["mean_iterations"],
                             phases=w64["phases"], dtype="f64")
        extra["fp32_error_vs_fp64"] = dict(
            one_substep=rel(p32_one, p64_one), after_window=rel(snap_end["p"], p64_end),
            what="relative 2-norm error of the FP32 positions against the FP64 GPU path (itself "
                 "equal to the FP64 oracle to 1e-9 in the parity tests): one substep from the "
                 "window-start state, and after the timed window")
        log(f"fp64: {w64['ms_per_step']:.3f} ms/substep; fp32 error one substep "
            f"{extra['fp32_error_vs_fp64']['one_substep']:.2e}, after window "
            f"{extra['fp32_error_vs_fp64']['after_window']:.2e}")
        # learned form: the committed network attached, same window
        if os.path.exists(LEARNED_MODEL) and args.precision == "fp32":
            from multilayersolver2024_amd import learn
            net = learn.Mlp.load(LEARNED_MODEL)
            simL = Simulation(mesh, make_integrator(hanging_stop=args.hanging_stop), h, plane,
                              precision=prec)
            set_options(simL, args)
            learn.attach(simL, net)
            simL.restore(snap0)
            run_window(simL, args.steps)
            simL.restore(snap0)
            learn.stats(simL)
            wl = run_window(simL, args.steps)
            kept, rej = learn.stats(simL)
            pl_end = simL.download()[0]
            simL.close()
            net.close()
            extra["learned"] = dict(
                value=wl["value"], unit="tet-proj/s", ms_per_step=wl["ms_per_step"],
                mean_iterations_per_substep=wl["mean_iterations"],
                warm_starts_kept=kept, warm_starts_rejected=rej,
                position_difference_vs_plain=rel(pl_end, snap_end["p"]),
                model=os.path.relpath(LEARNED_MODEL, ROOT),
                what="the same window with the learned warm start and key factors "
                     "(mlx_mesh_set_learned, MLX_LEARN_WARMSTART | MLX_LEARN_KEYS)")
            log(f"learned: {wl['ms_per_step']:.3f} ms/substep, {wl['mean_iterations']:.2f} "
                f"iterations, kept {kept} rejected {rej}")
        # the beam with NeoHookean tetrahedra (per-element E and nu from
        # the same seed), same layers and stop rule, a short window after
        # the same warmup
        scN = scenes.beam3d(args.nx, args.ny, args.nz, seed=0, neohookean=True)
        simN = Simulation(Mesh(scN), make_integrator(hanging_stop=args.hanging_stop), h, plane,
                          precision=prec)
        set_options(simN, args)
        for _ in range(args.warmup):
            simN.step()
        nN = min(args.steps, 5)
        wN = timed_from(simN, simN.snapshot(), nN)
        pN = simN.download()[0]
        simN.close()
        extra["neohookean_beam"] = dict(
            value=wN["value"], unit="tet-proj/s", ms_per_step=wN["ms_per_step"], steps=nN,
            mean_iterations_per_substep=wN["mean_iterations"],
            iterations_per_layer=wN["per_layer"], finite=bool(np.isfinite(pN).all()),
            what="the same beam, layers and stop rule with NeoHookean tetrahedra (the "
                 "NeoHookean material of the 3D beam configuration), a short window after the "
                 "same warmup")
        log(f"neohookean beam: {wN['ms_per_step']:.1f} ms/substep, "
            f"{wN['mean_iterations']:.1f} iterations")
    sim.close()
    wall = w["wall"]
    bpt = BYTES_PER_TET[args.precision]
    ftype = "float" if args.precision == "fp32" else "double"
    sweep_s = ms.value / 1e3
    sweeps = {"kernel": f"k_sweep3d<{ftype}>", "ms": ms.value, "launches": launches.value,
              "elements": elems.value, "bytes_per_unit": bpt,
              "avg_launch_us": 1e3 * ms.value / max(launches.value, 1),
              "achieved_GBs": (elems.value * bpt) / sweep_s / 1e9 if sweep_s > 0 else 0.0}
    if plaunches.value > 0 and args.flow:
        # k_flow: one launch = one iteration's sweep of the all-elastic
        # layer (its colors, contacts and pins; the residual follows in
        # k_residual_tail); one unit = one tet projection
        pk_s = pms.value / 1e3
        achieved = pelems.value * bpt / pk_s / 1e9
        roof = {"kernel": f"k_flow<{ftype},0>", "bytes_per_unit": bpt,
                "avg_launch_us": 1e3 * pms.value / plaunches.value, "launches": plaunches.value,
                "elements_per_launch": pelems.value / plaunches.value,
                "per_color_sweeps": sweeps}
    elif plaunches.value > 0:
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
    # thread slots per sweep launch: the colors' slot counts rounded up to
    # the workgroup, averaged over the colors (the all-elastic layer, which
    # makes most of the launches, launches every color)
    bucket = np.asarray(sim_buckets, dtype=np.int64)
    slots = float(np.mean(-(-bucket // args.sweep_block) * args.sweep_block))
    traffic = (traffic_bytes(roof["kernel"], slots) if roof["kernel"] == "k_sweep3d<float>"
               and args.sweep_block == 64 else None)
    roof["thread_slots_per_launch"] = slots
    line = {
        "metric": "ms/substep & tet-proj/s to rel. residual 1e-4: plain vs learned+batched, 1 GPU",
        "value": w["value"],
        "unit": "tet-proj/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": w["ms_per_step"],
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
            "flow": bool(args.flow),
            "persist_local": bool(args.persist_local) if args.persistent else None,
            "colors": ncol,
            "mean_iterations_per_substep": w["mean_iterations"],
            "tet_projections": int(w["proj"]),
            "contacts_at_window_start": int(snap0["contacts"]),
            "contacts_last_substep": int(contacts),
            "finite": finite,
            "rigid_contacts": False,
            "phases": w["phases"],
            "timing": "after the warmup substeps, one untimed pass over the window (iteration "
                      "graphs captured), then the window timed again from the same state",
        },
        "roofline": dict({
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": (TRAFFIC_SOURCE if traffic is not None
                               else None if traffic_current() else TRAFFIC_STALE),
        }, **roof),
    }
    line.update(extra)
    if not args.no_batched and args.batch_legs in ("both", "clusters"):
        log("batched leg (rigid clusters, NeoHookean half)")
        line["batched"] = bench_batched_clusters(
            h, plane, args.batch_steps, args.batch_warmup, args.precision,
            graph=bool(args.batch_graph),
            model_path=None if args.no_extra_legs else LEARNED_MODEL)
        log(f"batched: {line['batched']['ms_per_step']:.1f} ms/substep, "
            f"{line['batched']['value']:.3e} tet-proj/s")
        line["learned_batched"] = line["batched"].pop("learned")
    if not args.no_batched and args.batch_legs in ("both", "plain"):
        log("batched leg (plain method, StVK-Voigt)")
        splits = tuple(int(v) for v in args.batch_splits.split(",") if v.strip())
        line["batched_plain_stvk"] = bench_batched(h, plane, args.batch_steps, args.batch_warmup,
                                                   args.precision, graph=bool(args.batch_graph),
                                                   splits=splits)
    if not args.no_cpu_baseline:
        log("cpu baseline")
        cb, p_oracle_one = cpu_baseline(mesh, h, plane, snap_rep, runs=args.cpu_runs,
                                        label=f"substep {k_rep + 1} of the timed window "
                                              f"({w['iters'][k_rep]} GPU iterations; window "
                                              f"median {med:g})")
        line["cpu_baseline"] = cb
        if p32_rep is not None:
            line["fp32_error_vs_fp64"]["representative_substep_vs_oracle"] = rel(p32_rep,
                                                                                 p_oracle_one)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
