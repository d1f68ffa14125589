"""The learned warm start on the GPU: the BF16 matrix-core network against
numpy (forward and gradient, with and without BF16 rounding emulated), a
training run that lowers the loss, and the learned substep against the
oracle's restatement of it (ora_learn_*): a constant-zero model gives the
plain substep bit for bit, an overshooting model is rejected and restored
bit for bit, and a kappa = 1 model with key factors matches the oracle in
FP64 with the same keep / reject decisions."""
import numpy as np
import pytest

import oracle as O
from helpers import beam_case, rel
from multilayersolver2024_amd import _lib as L
from multilayersolver2024_amd import learn
from multilayersolver2024_amd.integrator import Simulation

pytestmark = pytest.mark.gpu

H = 1 / 60


def _say(*a):
    print("[learn]", *a, flush=True)


def bf16(a):
    """Round float32 to bfloat16 (nearest, ties to even), returned as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


def silu(z):
    return z / (1 + np.exp(-z))


def dsilu(z):
    s = 1 / (1 + np.exp(-z))
    return s * (1 + z * (1 - s))


def forward(p, X, emulate=True):
    """The network in numpy; with emulate, operands and stored activations
    rounded to bf16 where the device stores them so."""
    if emulate:
        q, ft = bf16, np.float32
        Xn = q(((X.astype(np.float32) - p["fmean"]) * p["fistd"]).astype(np.float32))
    else:
        q, ft = (lambda a: np.asarray(a, np.float64)), np.float64
        Xn = (X.astype(np.float64) - p["fmean"]) * p["fistd"]
    W1, W2, W3 = (q(p[k]).astype(ft) for k in ("W1", "W2", "W3"))
    z1 = Xn @ W1.T + p["b1"]
    h1 = q(silu(z1)).astype(ft)
    z2 = h1 @ W2.T + p["b2"]
    h2 = q(silu(z2)).astype(ft)
    raw = h2 @ W3.T + p["b3"]
    return dict(Xn=Xn, z1=z1, h1=h1, z2=z2, h2=h2, raw=raw, y=raw * p["oscale"] + p["oshift"],
                W1=W1, W2=W2, W3=W3)


def backward(p, X, T, W, emulate=True):
    q = bf16 if emulate else (lambda a: a)
    a = forward(p, X, emulate)
    wsum = W.sum()
    e = a["y"] - T
    loss = float(np.sum(W[:, None] * e * e) / wsum)
    dY = 2 * W[:, None] * e * p["oscale"] / wsum
    dYq = q(dY.astype(np.float32)) if emulate else dY
    dZ2 = q((dYq @ a["W3"]) * dsilu(a["z2"]))
    dZ1 = q((dZ2 @ a["W2"]) * dsilu(a["z1"]))
    g = dict(W1=dZ1.T @ a["Xn"], b1=dZ1.sum(0), W2=dZ2.T @ a["h1"], b2=dZ2.sum(0),
             W3=dYq.T @ a["h2"], b3=dY.sum(0))
    return g, loss


def _params(seed):
    p = learn.init_params(seed)
    rng = np.random.default_rng(seed)
    for k in ("b1", "b2", "b3"):
        p[k] = (0.1 * rng.standard_normal(p[k].shape)).astype(np.float32)
    p["fmean"] = rng.standard_normal(21).astype(np.float32)
    p["fistd"] = (0.5 + rng.random(21)).astype(np.float32)
    p["oscale"] = np.array([0.7, 1.9], np.float32)
    p["oshift"] = np.array([1.0, -0.3], np.float32)
    return p


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((n, 21)) * 1.5 + 0.2).astype(np.float32)
    return X


def test_numpy_backprop_matches_finite_differences():
    # pins the test's own reference gradient (float64, no rounding)
    p = {k: v.astype(np.float64) for k, v in _params(3).items()}
    X = _data(40, 3).astype(np.float64)
    rng = np.random.default_rng(9)
    T = rng.standard_normal((40, 2))
    W = rng.random(40)
    g, _ = backward(p, X, T, W, emulate=False)
    for k, idx in (("W1", (5, 7)), ("b1", (11,)), ("W2", (3, 60)), ("b2", (0,)), ("W3", (1, 9)),
                   ("b3", (0,))):
        eps = 1e-6
        pp, pm = {a: b.copy() for a, b in p.items()}, {a: b.copy() for a, b in p.items()}
        pp[k][idx] += eps
        pm[k][idx] -= eps
        fd = (backward(pp, X, T, W, False)[1] - backward(pm, X, T, W, False)[1]) / (2 * eps)
        assert g[k][idx] == pytest.approx(fd, rel=1e-5, abs=1e-9), k


def test_mlp_inference_matches_numpy():
    p = _params(1)
    X = _data(1000, 1)  # not a multiple of the 16-row tile
    net = learn.Mlp(p)
    try:
        Y = net.infer(X)
    finally:
        net.close()
    em = forward(p, X, True)["y"]
    ex = forward(p, X, False)["y"]
    d_em = np.abs(Y - em).max() / np.abs(em).max()
    d_ex = np.abs(Y - ex).max() / np.abs(ex).max()
    _say("infer vs bf16-emulated", d_em, "vs fp64", d_ex)
    assert d_em < 2e-3
    assert d_ex < 3e-2
    assert np.isfinite(Y).all()


def test_mlp_gradient_matches_numpy():
    p = _params(2)
    X = _data(777, 2)
    rng = np.random.default_rng(22)
    T = rng.standard_normal((777, 2)).astype(np.float32)
    W = rng.random(777).astype(np.float32)
    net = learn.Mlp(p)
    try:
        g, loss = net.grad(X, T, W)
    finally:
        net.close()
    ge, le = backward(p, X, T, W, True)
    gx, lx = backward({k: v.astype(np.float64) for k, v in p.items()}, X.astype(np.float64),
                      T.astype(np.float64), W.astype(np.float64), False)
    assert loss == pytest.approx(le, rel=1e-4)
    assert loss == pytest.approx(lx, rel=1e-2)
    for k in learn.GRAD_ORDER:
        r_em, r_ex = rel(g[k], ge[k]), rel(g[k], gx[k])
        _say(k, "vs emulated", r_em, "vs fp64", r_ex)
        assert r_em < 2e-3, k
        assert r_ex < 5e-2, k


def test_mlp_training_lowers_the_loss():
    rng = np.random.default_rng(5)
    X = _data(8192, 5)
    T = np.stack([np.tanh(X[:, 0] + 0.5 * X[:, 1]), 0.3 * X[:, 2] - 0.1 * X[:, 3] ** 2],
                 axis=1).astype(np.float32)
    p = learn.init_params(0)
    p["fmean"], p["fistd"] = X.mean(0), (1 / X.std(0)).astype(np.float32)
    p["oshift"], p["oscale"] = T.mean(0), T.std(0).astype(np.float32)
    net = learn.Mlp(p)
    try:
        base = float(np.mean(np.sum((T - T.mean(0)) ** 2, axis=1)))
        losses = net.train(X, T, None, epochs=30, batch=512, lr=3e-3, seed=1)
        Y = net.infer(X)
    finally:
        net.close()
    _say("train loss", losses[0], "->", losses[-1], "mean predictor", base)
    assert losses[-1] < 0.1 * base
    assert np.mean(np.sum((Y - T) ** 2, axis=1)) < 0.15 * base
    assert np.all(np.isfinite(losses))


def _run(model_params, flags, substeps, precision=L.MLX_FP64, kappa_max=2.0, nx=16,
         attach_at=0):
    m, it, contact = beam_case(nx, 4, 4, pct=99.99, z0=0.004)
    it.useRigidContacts = False
    it.hangingStop = 300
    sim = Simulation(m, it, H, contact=contact, precision=precision)
    net = None
    outs, st = [], []
    try:
        for n in range(substeps):
            if n == attach_at and model_params is not None:
                net = learn.Mlp(model_params)
                learn.attach(sim, net, flags, kappa_max)
            r = sim.step()
            lo = learn.outputs(sim) if net is not None else None
            outs.append((r, lo, sim.download()))
            if net is not None:
                st.append(learn.stats(sim))  # (kept, rejected) of this substep
    finally:
        sim.close()
        if net is not None:
            net.close()
    return m, it, contact, outs, st


def test_zero_model_is_the_plain_substep():
    both = L.MLX_LEARN_WARMSTART | L.MLX_LEARN_KEYS
    _, _, _, plain, _ = _run(None, 0, 8)
    _, _, _, zero, st = _run(learn.constant_params(0.0, 0.0), both, 8)
    for (rp, _, sp), (rz, _, sz) in zip(plain, zero):
        assert rp.itersPerLayer == rz.itersPerLayer
        for a, b in zip(sp[:3], sz[:3]):
            assert np.array_equal(a, b)
    _say("zero model decisions", st)
    # the warm start runs (and is kept: it changes nothing) whenever the last
    # layer runs (a coarse layer may stop the solve)
    assert st == [(int(r.layersRun == 3), 0) for r, _, _ in zero]


def test_overshooting_model_is_rejected_and_restored():
    """A model attached once the beam rests loaded on the ground, whose warm
    start is 1e4 times the last multipliers: every warm start is rejected
    and the run stays the plain run bit for bit."""
    _, _, _, plain, _ = _run(None, 0, 16)
    _, _, _, big, st = _run(learn.constant_params(1e4, 0.0), L.MLX_LEARN_WARMSTART, 16,
                            kappa_max=1e5, attach_at=10)
    _say("overshoot decisions", st)
    ran = [r for r, _, _ in big[10:]]
    assert st == [(0, int(r.layersRun == 3)) for r in ran] and sum(r.layersRun == 3 for r in ran) > 0
    for (rp, _, sp), (rb, _, sb) in zip(plain, big):
        assert rp.itersPerLayer == rb.itersPerLayer
        for a, b in zip(sp[:3], sb[:3]):
            assert np.array_equal(a, b)


def _random_model(seed):
    """A random network with outputs spread per element: kappa around 0.8,
    c around 0 (so the key order changes)."""
    p = learn.init_params(seed)
    p["fmean"] = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1] + [0] * 9 + [-6, 5.5, 0.375], np.float32)
    p["fistd"] = np.array([20.0] * 9 + [2.0] * 9 + [0.3, 1.2, 20.0], np.float32)
    p["oscale"] = np.array([0.3, 1.0], np.float32)
    p["oshift"] = np.array([0.8, 0.0], np.float32)
    return p


def test_features_match_numpy():
    """The 21 features the solver feeds the network, against numpy from the
    state: F, (F - F_prev)/h, log10(|lambda|/(vol h^2 E) + 1e-12), log10 E, nu."""
    m, it, contact = beam_case(16, 4, 4, pct=99.99, z0=0.004)
    it.useRigidContacts = False
    sim = Simulation(m, it, H, contact=contact)
    net = learn.Mlp(learn.constant_params(0.0))
    try:
        learn.attach(sim, net, 0)
        for _ in range(3):
            p_before = sim.download()[0]
            sim.step()
        p, _, lam, _ = sim.download()
        f = learn.features(sim)
    finally:
        sim.close()
        net.close()
    sc = m.scene
    P0 = np.asarray(sc["p"]).reshape(-1, 3)
    T = m.T0
    Dm = np.stack([P0[T[:, k]] - P0[T[:, 0]] for k in (1, 2, 3)], axis=2)
    Dinv = np.linalg.inv(Dm)

    def F(x):
        X = x.reshape(-1, 3)
        return np.stack([X[T[:, k]] - X[T[:, 0]] for k in (1, 2, 3)], axis=2) @ Dinv

    Fn, Fo = F(p), F(p_before)
    vol = np.abs(np.linalg.det(Dm)) / 6
    E, nu = np.asarray(sc["E"]), np.asarray(sc["nu"])
    want = np.concatenate([Fn.reshape(-1, 9), ((Fn - Fo) / H).reshape(-1, 9),
                           np.log10(np.linalg.norm(lam, axis=1) / (vol * H * H * E) + 1e-12)[:, None],
                           np.log10(E)[:, None], nu[:, None]], axis=1)
    err = np.abs(f - want) / (np.abs(want) + 1e-3)
    _say("features max rel err per column", np.round(err.max(0), 7))
    assert err[:, :18].max() < 1e-5
    assert err[:, 18].max() < 1e-4
    assert err[:, 19:].max() < 1e-5


@pytest.mark.parametrize("case", ["kappa1", "kappa0.8_c0.5", "random"])
def test_learned_substeps_match_the_oracle(case):
    """FP64: each learned substep of the device (warm start and key factors)
    against the oracle run from the device's state at the start of that
    substep (positions, velocities, cache.oldp, runUntil) and fed the
    device's lambda_0 and key factors: same iterations, same keep / reject
    decision, positions and multipliers.  (Starting each substep from the
    same state keeps the comparison off the strain-rate keys' near-ties,
    which a last-bit difference in the positions would flip.)"""
    import copy
    both = L.MLX_LEARN_WARMSTART | L.MLX_LEARN_KEYS
    model = {"kappa1": learn.constant_params(1.0, 0.0),
             "kappa0.8_c0.5": learn.constant_params(0.8, 0.5),
             "random": _random_model(7)}[case]
    m, it, contact = beam_case(16, 4, 4, pct=99.99, z0=0.004)
    it.useRigidContacts = False
    it.hangingStop = 300
    sim = Simulation(m, copy.deepcopy(it), H, contact=contact)
    osim = O.OracleSim(m, copy.deepcopy(it), H, contact=contact)
    net = learn.Mlp(model)
    decisions = []
    try:
        learn.attach(sim, net, both)
        p_prev = None
        for n in range(14):
            p, v, _, _ = sim.download()
            osim.p, osim.v = p.copy(), v.copy()
            if p_prev is not None:
                osim.oldp = p_prev.copy()
            osim.s.runUntil = sim._settings.runUntilResSmallerThan
            r = sim.step()
            lam0, ks = learn.outputs(sim)
            st = learn.stats(sim)
            o = osim.step(lam0=lam0, key_scale=ks)
            pg, _, lam, _ = sim.download()
            e_p, e_l = rel(pg, osim.p), rel(lam, osim.lambdas())
            _say(f"{case} substep {n + 1}: kept {o['learnKept']} dev {st} res {o['learnRes']} "
                 f"iters {r.itersPerLayer}/{o['iters']} pos {e_p:.2e} lam {e_l:.2e}")
            assert (o["learnKept"] >= 0) == (r.layersRun == 3)
            assert st == ((int(o["learnKept"] == 1), int(o["learnKept"] == 0))
                          if r.layersRun == 3 else (0, 0))
            assert r.itersPerLayer == o["iters"]
            assert e_p < 1e-10 and e_l < 1e-8
            decisions.append(o["learnKept"])
            p_prev = p
    finally:
        sim.close()
        net.close()
    assert decisions.count(1) + decisions.count(0) > 0
    if case == "random":
        assert np.ptp(np.log10(ks)) > 0.1 and np.ptp(lam0) > 0
