"""Pins of the oracle's rigid layer (restriction, step, prolongation,
rigid-elastic boundary and contacts) and of its residual, by closed forms and
invariants."""
import copy
import ctypes as C

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle as O
from multilayersolver2024_amd import scenes
from multilayersolver2024_amd.integrator import XpbdLayer3D
from multilayersolver2024_amd.mesh import Mesh

P = O.P


def _bodies(nb):
    return dict(M=np.zeros(nb), COM=np.zeros(3 * nb), COMDOT=np.zeros(3 * nb), R=np.zeros(9 * nb),
                I0=np.zeros(9 * nb), I=np.zeros(9 * nb), omega=np.zeros(3 * nb))


def _restrict(p, v, mass, vb, nb):
    b = _bodies(nb)
    ri = np.zeros_like(p)
    O.lib().ora_rigid_properties3d(len(p) // 3, nb, P(vb, C.c_int32), P(p), P(v), P(mass),
                                   P(b["M"]), P(b["COM"]), P(b["COMDOT"]), P(b["R"]), P(b["I0"]),
                                   P(b["I"]), P(b["omega"]), P(ri))
    return b, ri


def test_restriction_against_numpy_sums():
    rng = np.random.default_rng(20)
    nV = 40
    p = rng.normal(size=3 * nV)
    v = rng.normal(size=3 * nV)
    mass = np.repeat(rng.uniform(0.5, 2.0, size=nV), 3)
    vb = rng.integers(0, 4, size=nV).astype(np.int32)  # bodies 1..3, 0 = elastic
    b, ri = _restrict(p, v, mass, vb, 3)
    X, V, m = p.reshape(-1, 3), v.reshape(-1, 3), mass[0::3]
    for k in range(3):
        sel = vb == k + 1
        M = m[sel].sum()
        com = (m[sel, None] * X[sel]).sum(0) / M
        cdot = (m[sel, None] * V[sel]).sum(0) / M
        r = X[sel] - com
        I = sum(mi * (ri_ @ ri_ * np.eye(3) - np.outer(ri_, ri_)) for mi, ri_ in zip(m[sel], r))
        Lm = sum(mi * np.cross(ri_, vi - cdot) for mi, ri_, vi in zip(m[sel], r, V[sel]))
        assert b["M"][k] == pytest.approx(M, rel=1e-14)
        np.testing.assert_allclose(b["COM"][3 * k:3 * k + 3], com, rtol=1e-13)
        np.testing.assert_allclose(b["COMDOT"][3 * k:3 * k + 3], cdot, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(b["I0"][9 * k:9 * k + 9].reshape(3, 3), I, rtol=1e-12)
        np.testing.assert_allclose(I @ b["omega"][3 * k:3 * k + 3], Lm, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(ri.reshape(-1, 3)[sel], r, atol=1e-14)
        assert np.array_equal(b["R"][9 * k:9 * k + 9].reshape(3, 3), np.eye(3))


def test_free_rigid_body_flight_closed_form():
    """All elements rigid (layers = [100]), no pins, no gravity, no contacts:
    one substep moves the mesh as x -> c + h v0 + Rot(h w)(x - c)
    (restriction, rigidSymplecticImperativeStep and rotateAllVerts)."""
    sc = scenes.beam3d(2, 2, 2, cell=0.1, z0=1.0)
    sc["pinned"] = np.zeros(0, dtype=np.int32)
    m = Mesh(sc)
    it = XpbdLayer3D()
    it.Gravity = 0.0
    it.setLayers([100])
    it.iterations = [1]
    X = sc["p"]
    mv = m.mass[0::3]
    c = (mv[:, None] * X).sum(0) / mv.sum()
    v0 = np.array([0.3, -0.1, 0.2])
    w = np.array([1.5, -2.0, 0.7])
    V = v0 + np.cross(w, X - c)
    h = 0.01
    sim = O.OracleSim(m, it, h, None, v0=V.reshape(-1))
    sim.step()
    want = c + h * v0 + (X - c) @ Rotation.from_rotvec(h * w).as_matrix().T
    np.testing.assert_allclose(sim.p.reshape(-1, 3), want, atol=1e-13)


def test_step_keeps_rotation_orthonormal():
    rng = np.random.default_rng(21)
    R = np.eye(3).reshape(-1, order="F").copy()
    COM = np.zeros(3)
    I0 = np.diag([1.0, 2.0, 3.0]).reshape(-1).copy()
    I = I0.copy()
    for k in range(50):
        w = rng.normal(size=3) * 5
        O.lib().ora_rigid_step3d(C.c_double(0.01), P(COM), P(np.zeros(3)), P(R), P(w), P(I0), P(I))
    Rm = R.reshape(3, 3, order="F")
    np.testing.assert_allclose(Rm @ Rm.T, np.eye(3), atol=1e-13)
    np.testing.assert_allclose(I.reshape(3, 3, order="F"), Rm @ I0.reshape(3, 3) @ Rm.T, atol=1e-12)


def _one_body_setup(rng, nV=10):
    p = rng.normal(size=3 * nV)
    mass = np.repeat(rng.uniform(0.5, 2.0, size=nV), 3)
    vb = np.ones(nV, dtype=np.int32)
    b, ri = _restrict(p, np.zeros_like(p), mass, vb, 1)
    return p, mass, vb, b, ri


def test_boundary_satisfied_constraint_leaves_body_unchanged():
    rng = np.random.default_rng(22)
    p, mass, vb, b, ri = _one_body_setup(rng)
    xi = p.copy()  # every vertex exactly at R ri + COM
    bverts = np.array([1, 4, 7], dtype=np.int32)
    lamB = np.zeros(3)
    COM0, R0 = b["COM"].copy(), b["R"].copy()
    O.lib().ora_boundary3d(1, P(b["M"]), P(b["COM"]), P(b["R"]), P(b["I0"]), P(b["I"]), 3,
                           P(bverts, C.c_int32), P(vb, C.c_int32), P(ri), P(mass), C.c_double(1e-8),
                           P(lamB), P(xi), C.c_double(0.01))
    assert np.array_equal(b["COM"], COM0) and np.array_equal(b["R"], R0)
    assert np.all(lamB == 0)


def test_boundary_single_vertex_closed_form():
    """One boundary vertex displaced by d from its rigid position, alpha~ = 0:
    dlambda = -|d| / (1/m + w), w = 1/M + (r x n).I^-1(r x n), COM moves by
    dlambda n / M, and the vertex snaps to the updated rigid position."""
    rng = np.random.default_rng(23)
    p, mass, vb, b, ri = _one_body_setup(rng)
    q = 3
    d = np.array([0.01, -0.02, 0.005])
    xi = p.copy()
    xi[3 * q:3 * q + 3] -= d  # rigid position - elastic position = d
    lamB = np.zeros(1)
    COM0 = b["COM"].copy()
    I = b["I"].reshape(3, 3, order="F").copy()
    r = ri[3 * q:3 * q + 3]
    n = d / np.linalg.norm(d)
    ax = np.cross(r, n)
    w = 1 / b["M"][0] + ax @ np.linalg.solve(I, ax)
    dl = -np.linalg.norm(d) / (1 / mass[3 * q + 2] + w)
    O.lib().ora_boundary3d(1, P(b["M"]), P(b["COM"]), P(b["R"]), P(b["I0"]), P(b["I"]), 1,
                           P(np.array([q], dtype=np.int32), C.c_int32), P(vb, C.c_int32), P(ri),
                           P(mass), C.c_double(0.0), P(lamB), P(xi), C.c_double(0.01))
    assert lamB[0] == pytest.approx(dl, rel=1e-12)
    np.testing.assert_allclose(b["COM"] - COM0, dl * n / b["M"][0], rtol=1e-10, atol=1e-16)
    R = b["R"].reshape(3, 3, order="F")
    np.testing.assert_allclose(xi[3 * q:3 * q + 3], R @ r + b["COM"], atol=1e-14)
    # the rotation is exp(h I^-1 (r x p / h))
    wv = np.linalg.solve(I, np.cross(r, dl * n / 0.01))
    np.testing.assert_allclose(R, Rotation.from_rotvec(0.01 * wv).as_matrix(), atol=1e-13)


def test_elastic_contact_closed_form():
    """penaltyConstraint.projectElastic with lambda = 0: one vertex at depth
    phi < 0 moves by -r phi m^-1 / (2 m^-1 + alpha/h^2) n."""
    xi = np.array([0.1, 0.2, -0.03])
    mass = np.full(3, 2.0)
    vert = np.array([0], dtype=np.int32)
    nrm = np.array([0.0, 0.0, 1.0])
    pts = np.array([0.1, 0.2, 0.0])
    comp = np.array([1e-4])
    oc = O.OraContacts()
    oc.K, oc.vertex, oc.normal, oc.point, oc.compliance = 1, P(vert, C.c_int32), P(nrm), P(pts), P(comp)
    cl = np.zeros(1)
    b = _bodies(1)
    h, res = 0.01, 0.5
    O.lib().ora_contacts3d(0, P(b["M"]), P(b["COM"]), P(b["R"]), P(b["I0"]), P(b["I"]), C.byref(oc),
                           P(np.zeros(1, dtype=np.int32), C.c_int32), P(np.zeros(3)), P(mass), P(cl),
                           P(xi), C.c_double(h), C.c_double(res), 1)
    mi = 0.5
    dx = -res * (-0.03) * mi / (2 * mi + 1e-4 / h ** 2)
    assert xi[2] == pytest.approx(-0.03 + dx, rel=1e-14)
    assert xi[0] == 0.1 and xi[1] == 0.2


def test_residual3d_closed_form_with_contact_expansion():
    """simpleEvaluateAllConstraintResiduals: element part |C| - min(|C|,
    |alpha~ lambda|) and the contact part with the K x 1 x K expansion of
    contactCompliance(1x1xK) .* contactLambdas(:)."""
    sc = scenes.beam3d(1, 1, 1, cell=0.1)
    m = Mesh(sc)
    om = O.OracleMesh(m)
    rng = np.random.default_rng(24)
    x = m.p + 0.003 * rng.normal(size=m.p.shape)
    lam = 1e-4 * rng.normal(size=6 * m.nT)
    h = 0.02
    verts = np.array([0, 3], dtype=np.int32)
    nrm = np.array([0, 0, 1.0, 0, 0, 1.0])
    pts = np.array([x[0], x[1], x[2] + 0.01, x[9], x[10], x[11] + 0.02])
    comp = np.array([1e-4, 3e-4])
    cl = np.array([2e-3, -1e-3])
    oc = O.OraContacts()
    oc.K, oc.vertex, oc.normal, oc.point, oc.compliance = 2, P(verts, C.c_int32), P(nrm), P(pts), P(comp)
    r = O.lib().ora_residual3d(C.byref(om.s), C.c_double(h), P(lam), P(x), P(x), C.byref(oc), P(cl),
                               C.c_double(0.5))
    acc = 0.0
    X = x.reshape(-1, 3)
    for e in range(m.nT):
        T = m.T0[e]
        Dm = np.stack([sc["p"][T[k]] - sc["p"][T[0]] for k in (1, 2, 3)], axis=1)
        F = np.stack([X[T[k]] - X[T[0]] for k in (1, 2, 3)], axis=1) @ np.linalg.inv(Dm)
        E = 0.5 * (F.T @ F - np.eye(3))
        Cv = m.elV[e] * np.array([E[0, 0], E[1, 1], E[2, 2], E[0, 1], E[0, 2], E[1, 2]])
        A = m.alpha[36 * e:36 * e + 36].reshape(6, 6, order="F") / h ** 2
        corr = np.abs(A @ lam[6 * e:6 * e + 6])
        acc += np.sum((np.abs(Cv) - np.minimum(np.abs(Cv), corr)) ** 2)
    for i in range(2):
        phi = min(nrm[3 * i:3 * i + 3] @ (X[verts[i]] - pts[3 * i:3 * i + 3]), 0.0)
        aC = abs(0.5 * phi)
        for j in range(2):
            corr = abs(comp[j] / h ** 2 * cl[i])
            acc += (aC - min(aC, corr)) ** 2
    assert r == pytest.approx(np.sqrt(acc), rel=1e-10)
