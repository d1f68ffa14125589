"""Many independent scenes integrated in one persistent launch.

``merge_meshes`` concatenates precomputed scenes the way Mesh3D.mergeMesh
does (3d/@Mesh3D/Mesh3D.m:300-402): vertex ids offset per scene and each
scene keeping its own graph coloring.  ``BatchSimulation`` drives
``mlx_batch_integrate``: every scene runs its own plain-method layerSolve
(own residual and stop rule) inside one workgroup of a single kernel launch.
``random_beams`` builds the batched workload: beams with log-uniform tet
counts, random aspect ratios and per-scene materials from PCG32.
"""
import ctypes as C
import math

import numpy as np

from . import _lib as L
from . import scenes
from .mesh import Mesh


class MergedMesh:
    """Concatenation of 3D Mesh objects (duck-types Mesh.desc())."""

    def __init__(self, meshes):
        self.meshes = meshes
        self.dim = 3
        self.nc = 6
        self.vertOff = np.zeros(len(meshes) + 1, dtype=np.int64)
        self.elemOff = np.zeros(len(meshes) + 1, dtype=np.int64)
        for i, m in enumerate(meshes):
            self.vertOff[i + 1] = self.vertOff[i] + m.nV
            self.elemOff[i + 1] = self.elemOff[i] + m.nT
        self.nV, self.nT = int(self.vertOff[-1]), int(self.elemOff[-1])
        self.T0 = np.concatenate([m.T0 + int(self.vertOff[i]) for i, m in enumerate(meshes)])
        self.t = L.fortran(self.T0 + 1, np.int32)
        self.p = np.concatenate([m.p for m in meshes])
        self.DmInv = np.concatenate([m.DmInv for m in meshes])
        self.alpha = np.concatenate([m.alpha for m in meshes])
        self.beta = np.concatenate([m.beta for m in meshes])
        self.elV = np.concatenate([m.elV for m in meshes])
        self.mass = np.concatenate([m.mass for m in meshes])
        self.elementColor = L.i32(np.concatenate([m.elementColor for m in meshes]))
        self.numColors = max(m.numColors for m in meshes)
        self.pinnedDOFs = L.i32(np.concatenate(
            [m.pinnedDOFs + 3 * int(self.vertOff[i]) for i, m in enumerate(meshes)]))
        self.forcedElastic = L.u8(np.concatenate([m.forcedElastic for m in meshes]))
        self.neohookean = L.u8(np.concatenate([m.neohookean for m in meshes]))
        self.elMu = L.f64(np.concatenate([m.elMu for m in meshes]))
        self.elLambda = L.f64(np.concatenate([m.elLambda for m in meshes]))

    def desc(self):
        return Mesh.desc(self)


def merge_meshes(meshes):
    return MergedMesh(meshes)


class BatchSimulation:
    def __init__(self, merged, integrator, h, contact=None, precision=L.MLX_FP32, block=512,
                 device=0, split=0):
        lib = L.lib()
        L.check("mlx_init", lib.mlx_init(device))
        self.merged = merged
        self.h = float(h)
        self._desc = merged.desc()
        b = C.c_void_p()
        L.check("mlx_batch_create", lib.mlx_batch_create(
            C.byref(self._desc), len(merged.meshes), L.ptr(merged.vertOff, C.c_int64),
            L.ptr(merged.elemOff, C.c_int64), precision, C.byref(b)))
        self.handle = b
        self.set_block(block)
        self.set_split(split)
        v0 = np.zeros_like(merged.p)
        L.check("mlx_batch_set_state", lib.mlx_batch_set_state(b, L.ptr(merged.p), L.ptr(v0)))
        self.frame = 0
        self._settings = integrator.settings()
        self._plane = contact.plane() if contact is not None else None
        ns = len(merged.meshes)
        self.iters = np.zeros(ns, dtype=np.int32)
        self.lastres = np.zeros(ns)
        self.contacts = np.zeros(ns, dtype=np.int32)

    def set_state(self, p, v):
        """Positions and velocities of every scene (merged vertex order)."""
        p = L.f64(p)
        v = L.f64(v)
        L.check("mlx_batch_set_state", L.lib().mlx_batch_set_state(self.handle, L.ptr(p), L.ptr(v)))

    def set_split(self, tets_per_workgroup):
        """Workgroups per scene = round(tets / tets_per_workgroup); 0: one."""
        L.check("mlx_batch_set_split", L.lib().mlx_batch_set_split(self.handle,
                                                                   int(tets_per_workgroup)))

    def set_lds(self, enable):
        """Keep positions of scenes that fit in LDS there (default on)."""
        L.check("mlx_batch_set_lds", L.lib().mlx_batch_set_lds(self.handle, int(bool(enable))))

    def set_graph(self, enable):
        """Replay a captured hipGraph of the whole substep after frame 1
        (default on); off launches every kernel directly.  The kernel time
        returned by step() is measured only for direct launches."""
        L.check("mlx_batch_set_graph", L.lib().mlx_batch_set_graph(self.handle, int(bool(enable))))

    def set_block(self, threads):
        L.check("mlx_batch_set_block", L.lib().mlx_batch_set_block(self.handle, int(threads)))

    def autotune(self):
        """Per-bucket block sizes from one timed substep per candidate (the
        next substep's state, restored afterwards).  Returns a list of
        (threads, kernel ms, scenes) per bucket."""
        blk = np.zeros(3, np.int32)
        ms = np.zeros(3, np.float32)
        cnt = np.zeros(3, np.int32)
        pl = C.byref(self._plane) if self._plane is not None else None
        L.check("mlx_batch_autotune", L.lib().mlx_batch_autotune(
            self.handle, C.byref(self._settings), self.h, self.frame + 1, pl,
            L.ptr(blk, C.c_int32), L.ptr(ms, C.c_float), L.ptr(cnt, C.c_int32)))
        return [(int(a), float(b_), int(c)) for a, b_, c in zip(blk, ms, cnt)]

    def set_rigid(self, rigid):
        """Rigid clusters of the coarse layer (merged element order; None
        returns to the plain method).  The integrator then needs two layers,
        [x, 0] with x != 0, and iterations for both."""
        r = None if rigid is None else L.u8(np.asarray(rigid, dtype=bool))
        L.check("mlx_batch_set_rigid", L.lib().mlx_batch_set_rigid(
            self.handle, L.ptr(r, C.c_uint8) if r is not None else None))

    def set_learned(self, mlp, kappa_max=2.0):
        """Attach a network (learn.Mlp) for the learned warm start of every
        scene's all-elastic layer (rigid clusters needed); None detaches.
        The network must outlive the attachment."""
        L.check("mlx_batch_set_learned", L.lib().mlx_batch_set_learned(
            self.handle, mlp.handle if mlp is not None else None, float(kappa_max)))

    def learn_stats(self):
        """(warm starts kept, rejected) over all scenes since the last call."""
        k, r = np.zeros(1, np.int64), np.zeros(1, np.int64)
        L.check("mlx_batch_learn_stats", L.lib().mlx_batch_learn_stats(
            self.handle, L.ptr(k, C.c_int64), L.ptr(r, C.c_int64)))
        return int(k[0]), int(r[0])

    def step(self):
        self.frame += 1
        out = L.BatchOut()
        out.iterations = L.ptr(self.iters, C.c_int32)
        out.lastResidual = L.ptr(self.lastres)
        out.contacts = L.ptr(self.contacts, C.c_int32)
        pl = C.byref(self._plane) if self._plane is not None else None
        L.check("mlx_batch_integrate", L.lib().mlx_batch_integrate(
            self.handle, C.byref(self._settings), self.h, self.frame, pl, C.byref(out)))
        return int(out.elasticProjections), float(out.kernelMs)

    def download(self):
        m = self.merged
        p = np.zeros(3 * m.nV)
        v = np.zeros(3 * m.nV)
        lam = np.zeros(6 * m.nT)
        L.check("mlx_batch_download", L.lib().mlx_batch_download(
            self.handle, L.ptr(p), L.ptr(v), L.ptr(lam)))
        return p, v, lam.reshape(6, m.nT).T.copy()

    def close(self):
        if self.handle:
            L.lib().mlx_batch_destroy(self.handle)
            self.handle = None


def rigid_components(mesh, rigid):
    """Body id per element (0 elastic, 1-based) of the rigid elements less
    the pinned ones: elements sharing a vertex form a body, bodies numbered
    by ascending smallest element (breadth-first search; the numbering of
    mexUnionFindByVertice)."""
    T = np.asarray(mesh.T0)
    rig = np.asarray(rigid, dtype=bool) & ~np.asarray(mesh.forcedElastic, dtype=bool)
    by_vert = {}
    for e in np.nonzero(rig)[0]:
        for v in T[e]:
            by_vert.setdefault(int(v), []).append(int(e))
    body = np.zeros(len(T), dtype=np.int32)
    nb = 0
    for e in np.nonzero(rig)[0]:
        if body[e]:
            continue
        nb += 1
        body[e] = nb
        queue = [int(e)]
        while queue:
            f = queue.pop()
            for v in T[f]:
                for g in by_vert[int(v)]:
                    if not body[g]:
                        body[g] = nb
                        queue.append(g)
    return body


def cluster_mask(scene, count):
    """Elements of `count` (1..3) rigid slabs across a beam from beam3d: the
    beam's cells along x (after the pinned first cell) are cut into `count`
    segments and each segment's interior cells (a one-cell elastic gap on
    both sides, so slabs share no vertex) are rigid, all five tets of each."""
    a = int(round(np.ptp(np.asarray(scene["p"])[:, 0]) / 0.02))
    nT = len(scene["t"])
    per_slab = nT // a  # tets per x cell (5 b c)
    mask = np.zeros(nT, dtype=bool)
    free = a - 1
    if free < 1:
        return mask
    count = max(1, min(count, (free + 1) // 2))
    seg = free / count
    for j in range(count):
        lo = 1 + int(round(j * seg))
        hi = 1 + int(round((j + 1) * seg))
        lo_i, hi_i = (lo + 1, hi - 1) if hi - lo >= 3 else (lo, lo + 1)
        for i in range(lo_i, max(hi_i, lo_i + 1)):
            if i < a:
                mask[i * per_slab:(i + 1) * per_slab] = True
    return mask


def random_beams(n=256, seed=0, tets=(2000, 60000), cell=0.02, z0=0.05,
                 E_range=(1e4, 1e7), nu_range=(0.30, 0.45), neohookean_odd=False):
    """n beam scenes: tet count log-uniform in `tets` (snapped to 5 a b c
    cubes with aspect ratios drawn from the seed), per-scene E log-uniform,
    nu uniform; each pinned at x = 0 over the ground plane z = 0.  With
    neohookean_odd, odd-numbered scenes are NeoHookean."""
    rng = scenes.Pcg32(seed, stream=77)
    out = []
    lo, hi = math.log10(tets[0]), math.log10(tets[1])
    for i in range(n):
        target = 10 ** (lo + (hi - lo) * rng.next_double())
        r1 = 1.0 + 3.0 * rng.next_double()
        r2 = 1.0 + rng.next_double()
        cubes = target / 5.0
        c = max(1, int(round((cubes / (r1 * r2 * r2)) ** (1.0 / 3.0))))
        b = max(1, int(round(c * r2)))
        a = max(1, int(round(cubes / (b * c))))
        E = 10 ** (math.log10(E_range[0]) + (math.log10(E_range[1]) - math.log10(E_range[0]))
                   * rng.next_double())
        nu = nu_range[0] + (nu_range[1] - nu_range[0]) * rng.next_double()
        out.append(scenes.beam3d(a, b, c, cell=cell, z0=z0, E=E, nu=nu,
                                 neohookean=bool(neohookean_odd and i % 2 == 1)))
    return out


def cluster_workload(n=256, seed=0, tets=(2000, 60000)):
    """The batched workload with rigid clusters: random_beams with odd
    scenes NeoHookean, and 1..3 rigid slabs per scene (count from the seed).
    Returns (scenes, per-scene masks)."""
    scs = random_beams(n, seed=seed, tets=tets, neohookean_odd=True)
    rng = scenes.Pcg32(seed, stream=91)
    masks = [cluster_mask(sc, 1 + int(3 * rng.next_double()) % 3) for sc in scs]
    return scs, masks
