"""The oracle's NeoHookean pair on tetrahedra (ora_neohookean_*3d): the 2D
reference's deviatoric/volumetric constraints (mexNeoHookeanDConstraint.cpp,
mexNeoHookeanHConstraint.cpp) with F = Ds DmInv of a tetrahedron.  The
reference's 3D integrator projects StVK-Voigt only, so nothing in the
reference pins these functions; the checks below come from the mathematics:
finite differences of ||F||_F and det F, closed forms under scaling and
rotation, the XPBD update rule against finite-difference gradients, and the
rest state as a fixed point at the equilibrium multipliers.  The last test
records the convergence behaviour at the beam's step h = 1/60: the pair
does not converge there (the NeoHookean half of the batched workload runs to
hangingStop), and does with small steps."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from multilayersolver2024_amd import scenes
from multilayersolver2024_amd.integrator import XpbdLayer3D
from multilayersolver2024_amd.mesh import Mesh

P = O.P
X0 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]]) * 0.02


def _dminv(X):
    return np.ascontiguousarray(np.linalg.inv((X[1:] - X[0]).T).reshape(-1, order="F"))


def _vals(D, x):
    L = O.lib()
    cd, J = np.zeros(1), np.zeros(1)
    L.ora_neohookean_grad3d(P(D), P(np.ascontiguousarray(x)), None, None, P(cd), P(J))
    return cd[0], J[0]


def _grads(D, x):
    L = O.lib()
    gD, gH = np.zeros(12), np.zeros(12)
    L.ora_neohookean_grad3d(P(D), P(np.ascontiguousarray(x)), P(gD), P(gH), None, None)
    return gD, gH


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gradients_match_central_differences(seed, oracle_lib):
    rng = np.random.default_rng(seed)
    D = _dminv(X0)
    x = (X0 + 0.005 * rng.standard_normal(X0.shape)).reshape(-1)
    gD, gH = _grads(D, x)
    eps = 1e-7
    for q in range(12):
        xp, xm = x.copy(), x.copy()
        xp[q] += eps
        xm[q] -= eps
        fd = (np.array(_vals(D, xp)) - np.array(_vals(D, xm))) / (2 * eps)
        assert abs(fd[0] - gD[q]) <= 1e-7 * np.abs(gD).max()
        assert abs(fd[1] - gH[q]) <= 1e-7 * np.abs(gH).max()


def test_scaling_and_rotation_closed_forms(oracle_lib):
    D = _dminv(X0)
    rng = np.random.default_rng(3)
    Y = X0 + 0.004 * rng.standard_normal(X0.shape)
    cd0, J0 = _vals(D, Y.reshape(-1))
    F = (Y[1:] - Y[0]).T @ np.linalg.inv((X0[1:] - X0[0]).T)
    assert cd0 == pytest.approx(np.linalg.norm(F), rel=1e-14)
    assert J0 == pytest.approx(np.linalg.det(F), rel=1e-13)
    for s in (0.5, 1.0, 1.7):
        cd, J = _vals(D, (X0[0] + s * (X0 - X0[0])).reshape(-1))
        assert cd == pytest.approx(s * math.sqrt(3), rel=1e-14)
        assert J == pytest.approx(s ** 3, rel=1e-14)
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] = -Q[:, 0]
    cd, J = _vals(D, (Y @ Q.T + 0.3).reshape(-1))
    assert cd == pytest.approx(cd0, rel=1e-13)
    assert J == pytest.approx(J0, rel=1e-12)


@pytest.mark.parametrize("which", ["D", "H"])
def test_update_rule_against_finite_difference_gradients(which, oracle_lib):
    """dlambda (sum_a w_a |grad_a C|^2 + alpha~) = -(C + alpha~ lambda) and
    dx = w grad C dlambda, with grad C from finite differences."""
    L = O.lib()
    rng = np.random.default_rng(7)
    D = _dminv(X0)
    x = (X0 + 0.004 * rng.standard_normal(X0.shape)).reshape(-1)
    h, mu, lam, vol, lam0 = 1 / 60, 3.0e4, 9.0e4, 0.02 ** 3 / 6, -2.5e-9
    mass = 1000 * vol / 4 * (1 + rng.random(12))
    dx, dl, Cc = np.zeros(12), np.zeros(1), np.zeros(1)
    eps = 1e-7
    g = np.zeros(12)
    k = 0 if which == "D" else 1
    for q in range(12):
        xp, xm = x.copy(), x.copy()
        xp[q] += eps
        xm[q] -= eps
        g[q] = (_vals(D, xp)[k] - _vals(D, xm)[k]) / (2 * eps)
    cd, J = _vals(D, x)
    if which == "D":
        L.ora_neohookean_d3d(C.c_double(h), P(x), P(D), C.c_double(mu), C.c_double(vol),
                             C.c_double(lam0), P(mass), P(dx), P(dl), P(Cc))
        Cv, at = cd, 1 / (vol * h * h * mu)
    else:
        gamma = 1 + mu / lam
        L.ora_neohookean_h3d(C.c_double(h), P(x), P(D), C.c_double(lam), C.c_double(vol),
                             C.c_double(lam0), P(mass), C.c_double(gamma), P(dx), P(dl), P(Cc))
        Cv, at = J - gamma, 1 / (vol * h * h * lam)
    assert Cc[0] == pytest.approx(Cv, rel=1e-14)
    s = np.sum(g * g / mass) + at
    assert dl[0] * s == pytest.approx(-(Cv + at * lam0), rel=1e-7)
    np.testing.assert_allclose(dx, g / mass * dl[0], rtol=1e-6, atol=1e-9 * np.abs(dx).max())


def test_rest_state_is_a_fixed_point_at_the_equilibrium_multipliers(oracle_lib):
    """At F = I, grad C_D = I/sqrt(3) g_a and grad C_H = I g_a, so the forces
    balance for lambda_D = -sqrt(3) vol h^2 mu, lambda_H = vol h^2 mu
    (gamma = 1 + mu/lambda), where both updates vanish."""
    L = O.lib()
    D = _dminv(X0)
    x = np.ascontiguousarray(X0.reshape(-1))
    h, mu, lam, vol = 1 / 60, 3.7e4, 8.6e4, 0.02 ** 3 / 6
    mass = np.full(12, 1000 * vol / 4)
    dx, dl = np.zeros(12), np.zeros(1)
    L.ora_neohookean_d3d(C.c_double(h), P(x), P(D), C.c_double(mu), C.c_double(vol),
                         C.c_double(-math.sqrt(3) * vol * h * h * mu), P(mass), P(dx), P(dl), None)
    assert abs(dl[0]) < 1e-12 * math.sqrt(3) * vol * h * h * mu
    assert np.abs(dx).max() < 1e-14 * 0.02
    L.ora_neohookean_h3d(C.c_double(h), P(x), P(D), C.c_double(lam), C.c_double(vol),
                         C.c_double(vol * h * h * mu), P(mass), C.c_double(1 + mu / lam),
                         P(dx), P(dl), None)
    assert abs(dl[0]) < 1e-12 * vol * h * h * mu
    assert np.abs(dx).max() < 1e-14 * 0.02


def test_sweep_sums_both_displacements_from_the_same_positions(oracle_lib):
    """One elastic tet of a cube: one sweep moves its vertices by dx_D + dx_H,
    both evaluated at the positions the color started from, and adds both
    multipliers."""
    L = O.lib()
    sc = scenes.beam3d(1, 1, 1, seed=2, neohookean=True)
    m = Mesh(sc)
    om = O.OracleMesh(m)
    rng = np.random.default_rng(5)
    xi = m.p + 0.002 * rng.standard_normal(m.p.shape)
    lam = np.zeros(6 * m.nT)
    first = int(np.nonzero(m.elementColor == 1)[0][0])
    isEl = np.zeros(m.nT, dtype=np.uint8)
    isEl[first] = 1
    x = np.ascontiguousarray(xi.reshape(-1, 3)[m.T0[first]].reshape(-1))
    ms = np.ascontiguousarray(om.mass.reshape(-1, 3)[m.T0[first]].reshape(-1))
    D = np.ascontiguousarray(om.DmInv[9 * first:9 * first + 9])
    h = 1 / 60
    mu, la = m.elMu[first], m.elLambda[first]
    dxD, dxH, dlD, dlH = np.zeros(12), np.zeros(12), np.zeros(1), np.zeros(1)
    L.ora_neohookean_d3d(C.c_double(h), P(x), P(D), C.c_double(mu), C.c_double(om.vol[first]),
                         C.c_double(0.0), P(ms), P(dxD), P(dlD), None)
    L.ora_neohookean_h3d(C.c_double(h), P(x), P(D), C.c_double(la), C.c_double(om.vol[first]),
                         C.c_double(0.0), P(ms), C.c_double(1 + mu / la), P(dxH), P(dlH), None)
    before = xi.copy()
    L.ora_neohookean_constraints3d(C.byref(om.s), P(isEl, C.c_uint8), C.c_double(h), P(lam), P(xi))
    moved = xi.reshape(-1, 3)[m.T0[first]].reshape(-1)
    np.testing.assert_array_equal(moved, (x + dxD) + dxH)
    assert lam[6 * first] == dlD[0] and lam[6 * first + 1] == dlH[0]
    assert not np.array_equal(before, xi)


def _residual_trace(h, iterations):
    m = Mesh(scenes.beam3d(8, 2, 2, seed=0, neohookean=True))
    it = XpbdLayer3D()
    it.Gravity = -9.81
    it.setLayers([0])
    it.iterations = [iterations]
    it.computeResiduals = True
    it.hangingStop = 100000
    sim = O.OracleSim(m, it, h, None, residual_cap=iterations + 8)
    return sim.step()["residuals"]


def test_pair_converges_only_with_small_steps(oracle_lib):
    """With lambda restarted at zero every substep, as the reference's
    integrators do, ||F|| = sqrt(3) at rest puts every substep's first
    iterate far from equilibrium: at h = 1/6000 the residual drops by 1e-4
    within 20 iterations, at h = 1/60 (the beam workload's step) it is still
    above 1 % of the first after 400."""
    r = _residual_trace(1 / 6000, 30)
    assert r[20] < 1e-4 * r[1]
    r = _residual_trace(1 / 60, 400)
    assert r[400] > 1e-2 * r[1]


def test_neohookean_mask_reaches_the_mesh_descriptor():
    """The NeoHookean mask and the Lame parameters travel to mlx_mesh_create
    (single meshes and merged batches); StVK-only meshes pass NULL."""
    import ctypes as C
    from multilayersolver2024_amd.batch import MergedMesh
    a = Mesh(scenes.beam3d(2, 1, 1, seed=0, neohookean=True))
    b = Mesh(scenes.beam3d(2, 1, 1, seed=1))
    d = a.desc()
    assert bool(d.neohookean) and bool(d.mu) and bool(d.lambda_)
    assert d.neohookean[0] == 1 and d.mu[3] == a.elMu[3] and d.lambda_[3] == a.elLambda[3]
    assert not bool(b.desc().neohookean)
    mm = MergedMesh([b, a])
    dm = mm.desc()
    assert dm.neohookean[0] == 0 and dm.neohookean[b.nT] == 1
    assert dm.mu[b.nT + 2] == a.elMu[2]
