"""Learned warm start of the 3D substep: the per-element network behind
``mlx_mlp_*`` / ``mlx_mesh_set_learned``, its training data from
plain-solver trajectories, training on the device and the saved model.

The reference has no learned component: every substep restarts the
multipliers at zero (3d/integrator/xpbdLayer3D.m:666) and orders the
coarse layers by the strain rate of the substep that just ended
(xpbdLayer3D.m:601-619, mexEdiffNorm3D.cpp:25-61).  The network (21 -> 64
-> 64 -> 2, SiLU, BF16 weights and activations, FP32 accumulation on the
matrix cores) predicts per element

* ``kappa``: the warm start lambda_0 = kappa * lambda_prev of the finest
  layer (target: the least-squares factor <lambda_n, lambda_{n-1}> /
  |lambda_{n-1}|^2 of the plain solver's own multipliers), and
* ``c``: the log10 factor from the last substep's strain-rate key to the
  coming substep's (target: log10 of the ratio of consecutive keys).

The library keeps a warm start only when it does not raise the residual of
the equations the projection solves (C + alpha~ lambda + damping per
constraint, r phi per contact) over the cold start's, so a poor model costs
time, never accuracy; the stop rule is the reference's, unchanged.

Run ``python -m multilayersolver2024_amd.learn --out FILE`` on a GPU to
generate the data, train and save a model.
"""
import argparse
import ctypes as C
import json
import math
import sys
import time

import numpy as np

from . import _lib as L

N_IN, N_HID, N_OUT = 21, 64, 2
FEATURES = ([f"F{r}{c}" for r in range(3) for c in range(3)]
            + [f"dF{r}{c}" for r in range(3) for c in range(3)]
            + ["log10_lambda_rel", "log10_E", "nu"])
OUTPUTS = ["kappa", "key_log10_factor"]
PARAM_SHAPES = {"W1": (N_HID, N_IN), "b1": (N_HID,), "W2": (N_HID, N_HID), "b2": (N_HID,),
                "W3": (N_OUT, N_HID), "b3": (N_OUT,)}
NORM_SHAPES = {"fmean": (N_IN,), "fistd": (N_IN,), "oscale": (N_OUT,), "oshift": (N_OUT,)}
GRAD_ORDER = ["W1", "b1", "W2", "b2", "W3", "b3"]  # mlx_mlp_grad layout


def init_params(seed=0):
    """Glorot-uniform weights, zero biases, identity normalisation."""
    rng = np.random.default_rng(seed)
    p = {}
    for k, shp in PARAM_SHAPES.items():
        if k.startswith("W"):
            lim = math.sqrt(6.0 / (shp[0] + shp[1]))
            p[k] = rng.uniform(-lim, lim, shp).astype(np.float32)
        else:
            p[k] = np.zeros(shp, np.float32)
    p["fmean"] = np.zeros(N_IN, np.float32)
    p["fistd"] = np.ones(N_IN, np.float32)
    p["oscale"] = np.ones(N_OUT, np.float32)
    p["oshift"] = np.zeros(N_OUT, np.float32)
    return p


def constant_params(kappa, c=0.0):
    """A network whose outputs are the constants (kappa, c) for every input
    (zero weights; the output map carries the constants)."""
    p = {k: np.zeros(s, np.float32) for k, s in PARAM_SHAPES.items()}
    p.update(fmean=np.zeros(N_IN, np.float32), fistd=np.ones(N_IN, np.float32),
             oscale=np.ones(N_OUT, np.float32),
             oshift=np.array([kappa, c], np.float32))
    return p


class Mlp:
    """A network on the current device (mlx_mlp)."""

    def __init__(self, params=None):
        h = C.c_void_p()
        L.check("mlx_mlp_create", L.lib().mlx_mlp_create(C.byref(h)))
        self.handle = h
        self.norm = {k: np.zeros(s, np.float32) for k, s in NORM_SHAPES.items()}
        self.set(params if params is not None else init_params())

    def set(self, p):
        a = {k: L.f32(np.asarray(p[k]).reshape(s)) for k, s in {**PARAM_SHAPES, **NORM_SHAPES}.items()}
        L.check("mlx_mlp_set", L.lib().mlx_mlp_set(
            self.handle, *(L.ptr(a[k], C.c_float) for k in ("W1", "b1", "W2", "b2", "W3", "b3",
                                                             "fmean", "fistd", "oscale",
                                                             "oshift"))))
        for k in NORM_SHAPES:
            self.norm[k] = a[k].copy()

    def get(self):
        out = {k: np.zeros(s, np.float32) for k, s in PARAM_SHAPES.items()}
        L.check("mlx_mlp_get", L.lib().mlx_mlp_get(
            self.handle, *(L.ptr(out[k], C.c_float) for k in ("W1", "b1", "W2", "b2", "W3",
                                                               "b3"))))
        out.update({k: v.copy() for k, v in self.norm.items()})
        return out

    def infer(self, X):
        X = L.f32(X).reshape(-1, N_IN)
        Y = np.zeros((len(X), N_OUT), np.float32)
        L.check("mlx_mlp_infer", L.lib().mlx_mlp_infer(self.handle, L.ptr(X, C.c_float), len(X),
                                                       L.ptr(Y, C.c_float)))
        return Y

    def grad(self, X, T, W=None):
        """Gradient of sum_s w_s |y_s - t_s|^2 / sum_s w_s (dict like
        PARAM_SHAPES) and the loss."""
        X = L.f32(X).reshape(-1, N_IN)
        T = L.f32(T).reshape(-1, N_OUT)
        Wa = None if W is None else L.f32(W)
        n = sum(int(np.prod(PARAM_SHAPES[k])) for k in GRAD_ORDER)
        g = np.zeros(n, np.float32)
        loss = np.zeros(1)
        L.check("mlx_mlp_grad", L.lib().mlx_mlp_grad(
            self.handle, L.ptr(X, C.c_float), L.ptr(T, C.c_float),
            L.ptr(Wa, C.c_float) if Wa is not None else None, len(X), L.ptr(g, C.c_float),
            L.ptr(loss)))
        out, o = {}, 0
        for k in GRAD_ORDER:
            sz = int(np.prod(PARAM_SHAPES[k]))
            out[k] = g[o:o + sz].reshape(PARAM_SHAPES[k])
            o += sz
        return out, float(loss[0])

    def train(self, X, T, W=None, epochs=10, batch=4096, lr=1e-3, seed=0):
        X = L.f32(X).reshape(-1, N_IN)
        T = L.f32(T).reshape(-1, N_OUT)
        Wa = None if W is None else L.f32(W)
        losses = np.zeros(max(epochs, 1), np.float32)
        L.check("mlx_mlp_train", L.lib().mlx_mlp_train(
            self.handle, L.ptr(X, C.c_float), L.ptr(T, C.c_float),
            L.ptr(Wa, C.c_float) if Wa is not None else None, len(X), int(epochs), int(batch),
            float(lr), int(seed), L.ptr(losses, C.c_float)))
        return losses[:epochs]

    def save(self, path, meta=None):
        p = self.get()
        np.savez(path, **p, meta=np.array(json.dumps(meta or {})))

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            p = {k: z[k] for k in list(PARAM_SHAPES) + list(NORM_SHAPES)}
        return cls(p)

    def close(self):
        if self.handle:
            L.lib().mlx_mlp_destroy(self.handle)
            self.handle = None


def attach(sim, mlp, flags=L.MLX_LEARN_WARMSTART | L.MLX_LEARN_KEYS, kappa_max=2.0):
    """Attach a network to a Simulation's mesh (mlp None detaches)."""
    L.check("mlx_mesh_set_learned", L.lib().mlx_mesh_set_learned(
        sim.dmesh.handle, mlp.handle if mlp is not None else None, int(flags), float(kappa_max)))


def stats(sim):
    """(warm starts kept, rejected) since the last call."""
    k, r = np.zeros(1, np.int64), np.zeros(1, np.int64)
    L.check("mlx_learn_stats", L.lib().mlx_learn_stats(sim.dmesh.handle, L.ptr(k, C.c_int64),
                                                        L.ptr(r, C.c_int64)))
    return int(k[0]), int(r[0])


def outputs(sim):
    """The last substep's lambda_0 (nT x 6) and key factors (nT)."""
    nT = sim.mesh.nT
    lam0, ks = np.zeros(6 * nT), np.zeros(nT)
    L.check("mlx_learn_outputs", L.lib().mlx_learn_outputs(sim.dmesh.handle, L.ptr(lam0),
                                                            L.ptr(ks)))
    return lam0.reshape(nT, 6), ks


def features(sim):
    """The 21 features of every element (nT x 21) as the next substep's
    inference sees them."""
    f = np.zeros((sim.mesh.nT, N_IN), np.float32)
    L.check("mlx_learn_features", L.lib().mlx_learn_features(sim.state, sim.h,
                                                             L.ptr(f, C.c_float)))
    return f


# ---------------------------------------------------------------------------
# training data
# ---------------------------------------------------------------------------

def rest_inverse(P0, T):
    """Dm^-1 per element from the rest positions."""
    R = P0.reshape(-1, 3)
    return np.linalg.inv(np.stack([R[T[:, k]] - R[T[:, 0]] for k in (1, 2, 3)], axis=2))


def green_strain(Dinv, T, x):
    """E = (F^T F - I) / 2 per element, F = Ds Dm^-1."""
    X = x.reshape(-1, 3)
    F = np.stack([X[T[:, k]] - X[T[:, 0]] for k in (1, 2, 3)], axis=2) @ Dinv
    return 0.5 * (np.swapaxes(F, 1, 2) @ F)


def strain_rate_keys(Ea, Eb, h):
    """Per-element ||E_a - E_b||_F^2 / h^2: the strain-rate measure the
    layers are sorted by (mexEdiffNorm3D.cpp:25-61)."""
    return np.sum((Ea - Eb) ** 2, axis=(1, 2)) / (h * h)


def kappa_target(lam_new, lam_prev, kappa_max=2.0):
    """Least-squares factor of lambda_new on lambda_prev per element, clipped
    to [0, kappa_max]; weight 1 where lambda_prev is nonzero, else 0."""
    den = np.sum(lam_prev * lam_prev, axis=1)
    num = np.sum(lam_new * lam_prev, axis=1)
    ok = den > 0
    k = np.where(ok, num / np.where(ok, den, 1.0), 0.0)
    return np.clip(k, 0.0, kappa_max), ok.astype(np.float32)


def key_target(ed_next, ed_now):
    """log10 of the ratio of consecutive strain-rate keys, clipped to +-3
    (keys floored at 1e-12 of the substep's largest)."""
    fl = 1e-12 * max(float(ed_now.max()), float(ed_next.max()), 1e-300)
    return np.clip(np.log10((ed_next + fl) / (ed_now + fl)), -3.0, 3.0)


def beam_family(seed, rng):
    """A beam of the 81,920-tet workload's family: same generator, the
    material draw from `seed`, dimensions and drop height varied."""
    from . import scenes
    nx, ny, nz = [(32, 8, 8), (48, 12, 12), (64, 16, 16), (40, 8, 12)][int(rng.integers(4))]
    z0 = float(rng.uniform(0.01, 0.25))
    return scenes.beam3d(nx, ny, nz, cell=0.02, z0=z0, seed=seed)


def collect(seeds, substeps, per_substep, precision=L.MLX_FP32, log=print, base_rng=0,
            graphs=True):
    """Plain-solver trajectories (layers [98.4375, 87.5, 0], iterations
    [2, 4, 14], relative-residual stop 1e-4, ground contact, h = 1/60) of
    beams of the given seeds.  Returns X (n x 21 features), Y (n x 2
    targets) and W (n weights), per_substep elements sampled per substep.
    graphs=False launches every kernel directly instead of replaying the
    iteration graphs (same results; for kernel profilers)."""
    from .integrator import PlaneContactFinder, Simulation, XpbdLayer3D
    from .mesh import Mesh

    h = 1.0 / 60.0
    rng = np.random.default_rng(base_rng)
    Xs, Ys, Ws = [], [], []
    probe = Mlp(constant_params(0.0))
    for seed in seeds:
        sc = beam_family(seed, rng)
        mesh = Mesh(sc)
        it = XpbdLayer3D()
        it.Gravity = -9.81
        it.setLayers([98.4375, 87.5, 0])
        it.iterations = [2, 4, 14]
        it.computeResiduals = True
        it.StopAtImprovementPercent = 99.99
        it.hangingStop = 1000
        it.useRigidContacts = False
        sim = Simulation(mesh, it, h, contact=PlaneContactFinder([0, 0, 1], [0, 0, 0]),
                         precision=precision)
        sim.dmesh.set_option(L.MLX_OPT_GRAPHS, int(graphs))
        attach(sim, probe, flags=0)
        T = mesh.T0
        Dinv = rest_inverse(mesh.p, T)
        t0 = time.time()
        n_before = sum(len(x) for x in Xs)
        p_now, _, lam_prev, _ = sim.download()
        E_prev = E_now = green_strain(Dinv, T, p_now)
        for n in range(substeps):
            X = features(sim) if n >= 1 else None
            sim.step()
            p_next, _, lam_new, _ = sim.download()
            E_next = green_strain(Dinv, T, p_next)
            if n >= 1:
                kap, w = kappa_target(lam_new, lam_prev)
                c = key_target(strain_rate_keys(E_next, E_now, h),
                               strain_rate_keys(E_now, E_prev, h))
                idx = rng.choice(mesh.nT, size=min(per_substep, mesh.nT), replace=False)
                ok = np.isfinite(X[idx]).all(axis=1) & np.isfinite(kap[idx]) & np.isfinite(c[idx])
                idx = idx[ok]
                Xs.append(X[idx])
                Ys.append(np.stack([kap[idx], c[idx]], axis=1).astype(np.float32))
                Ws.append(w[idx])
            lam_prev = lam_new
            E_prev, E_now = E_now, E_next
        log(f"seed {seed}: {sc['name']} z0={sc['p'][:, 2].min():.3f}, {substeps} substeps, "
            f"{sum(len(x) for x in Xs) - n_before} samples, {time.time() - t0:.1f} s")
        sim.close()
    probe.close()
    return np.concatenate(Xs), np.concatenate(Ys), np.concatenate(Ws)


def fit(X, Y, W, epochs=20, batch=4096, lr=2e-3, seed=0, log=print):
    """A network with the dataset's feature/target normalisation, trained
    on the device; returns (Mlp, epoch losses)."""
    p = init_params(seed)
    w = W > 0
    mean = X[w].mean(axis=0)
    std = X[w].std(axis=0)
    p["fmean"] = mean.astype(np.float32)
    p["fistd"] = (1.0 / np.where(std > 1e-12, std, 1.0)).astype(np.float32)
    ym, ys = Y[w].mean(axis=0), Y[w].std(axis=0)
    p["oshift"] = ym.astype(np.float32)
    p["oscale"] = np.where(ys > 1e-6, ys, 1.0).astype(np.float32)
    net = Mlp(p)
    losses = net.train(X, Y, W, epochs=epochs, batch=batch, lr=lr, seed=seed)
    for e, l_ in enumerate(losses):
        log(f"epoch {e}: loss {l_:.5f}")
    return net, losses


def evaluate(model, nx=64, ny=16, nz=16, seed=0, substeps=60, warmup=30,
             flags=L.MLX_LEARN_WARMSTART | L.MLX_LEARN_KEYS, precision=L.MLX_FP32, log=print):
    """Plain against learned on one beam (the 81,920-tet workload by
    default, seed 0, which the training data never uses): both runs take
    `warmup` plain substeps from rest, then `substeps` substeps timed, the
    learned run with the model attached.  Returns per-run mean iterations
    and ms per substep, and the warm starts kept / rejected."""
    from . import scenes
    from .integrator import PlaneContactFinder, Simulation, XpbdLayer3D
    from .mesh import Mesh

    h = 1.0 / 60.0
    mesh = Mesh(scenes.beam3d(nx, ny, nz, seed=seed))
    out = {}
    for name in ("plain", "learned"):
        it = XpbdLayer3D()
        it.Gravity = -9.81
        it.setLayers([98.4375, 87.5, 0])
        it.iterations = [2, 4, 14]
        it.computeResiduals = True
        it.StopAtImprovementPercent = 99.99
        it.hangingStop = 1000
        it.useRigidContacts = False
        sim = Simulation(mesh, it, h, contact=PlaneContactFinder([0, 0, 1], [0, 0, 0]),
                         precision=precision)
        try:
            for _ in range(warmup):
                sim.step()
            if name == "learned":
                attach(sim, model, flags)
                stats(sim)
            iters, ms, finals = [], [], []
            for _ in range(substeps):
                t0 = time.perf_counter()
                r = sim.step()
                ms.append(1e3 * (time.perf_counter() - t0))
                iters.append(sum(r.itersPerLayer))
                finals.append(float(r.residualArray[-1]))
            kept, rej = stats(sim) if name == "learned" else (0, 0)
            out[name] = dict(mean_iterations=float(np.mean(iters)),
                             median_ms=float(np.median(ms)), mean_ms=float(np.mean(ms)),
                             mean_final_residual=float(np.mean(finals)), kept=kept,
                             rejected=rej, iterations=iters)
            log(f"{name}: {np.mean(iters):.2f} iterations/substep, "
                f"{np.median(ms):.2f} ms median, kept {kept} rejected {rej}")
        finally:
            sim.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--scenes", type=int, default=24)
    ap.add_argument("--substeps", type=int, default=200)
    ap.add_argument("--per-substep", type=int, default=1024)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--graphs", type=int, default=1, choices=[0, 1],
                    help="replay iteration graphs while collecting (1) or launch directly (0)")
    ap.add_argument("--eval", type=int, default=0,
                    help="then compare plain and learned over this many substeps of the "
                         "81,920-tet beam")
    args = ap.parse_args(argv)
    lg = lambda *a: print("[learn]", *a, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.time()
    X, Y, W = collect(range(1, args.scenes + 1), args.substeps, args.per_substep, log=lg,
                      graphs=bool(args.graphs))
    lg(f"dataset: {len(X)} samples, {args.scenes * args.substeps} substeps, "
       f"{time.time() - t0:.1f} s")
    t1 = time.time()
    net, losses = fit(X, Y, W, epochs=args.epochs, batch=args.batch, lr=args.lr, log=lg)
    # the trivial predictor (the target mean) for reference
    w = W > 0
    base = float(np.mean(np.sum((Y[w] - Y[w].mean(axis=0)) ** 2, axis=1)))
    meta = dict(scenes=args.scenes, substeps=args.substeps, per_substep=args.per_substep,
                samples=int(len(X)), epochs=args.epochs, batch=args.batch, lr=args.lr,
                final_loss=float(losses[-1]), mean_predictor_loss=base,
                train_seconds=time.time() - t1, data_seconds=t1 - t0)
    if args.eval:
        ev = evaluate(net, substeps=args.eval, log=lg)
        meta["eval"] = {k: {a: b for a, b in v.items() if a != "iterations"} for k, v in ev.items()}
    net.save(args.out, meta)
    lg(json.dumps(meta))
    net.close()


if __name__ == "__main__":
    main()
