"""Pins of the oracle's 2D rigid layer (xpbdLayer2D.m:296-444, 606-738).

* bodies from connectivity against numpy sums;
* the boundary projection against the textbook XPBD point constraint
  between a particle and a body point (mass splitting over the body's
  boundary vertices), its zero-distance guard, and the direction of the
  pose change; the angular update is half the textbook impulse response
  (rigidBoundaryConstraint2D.m:33, the reference's factor 0.5);
* rigid contacts: no push without penetration, a push along the normal
  with it, the 2 w denominator (penaltyConstraint2D.m:51);
* the whole multi-layer substep against the independent vectorised numpy
  restatement (tests/numpy_xpbd.py) on a cantilever in ground contact.
"""
import copy
import ctypes as C

import numpy as np
import pytest

import oracle as O
from helpers import rel
from multilayersolver2024_amd import scenes
from multilayersolver2024_amd.integrator import PlaneContactFinder, XpbdLayer2D
from multilayersolver2024_amd.mesh import Mesh
from numpy_xpbd import NumpyXpbd2DLayers

P = O.P


def _bodies(p, vb, mass, nb):
    nV = len(vb)
    M, COM, I, ri = np.zeros(nb), np.zeros(2 * nb), np.zeros(nb), np.zeros(2 * nV)
    O.lib().ora_rigid_properties2d(nV, nb, P(vb, C.c_int32), P(p), P(mass), P(M), P(COM), P(I),
                                   P(ri))
    return M, COM, I, ri


def test_rigid_properties2d_match_numpy_sums():
    rng = np.random.default_rng(3)
    nV, nb = 40, 3
    p = rng.normal(size=2 * nV)
    m = np.repeat(rng.uniform(0.5, 2.0, nV), 2)
    vb = rng.integers(0, nb + 1, nV).astype(np.int32)
    M, COM, I, ri = _bodies(p, vb, m, nb)
    X = p.reshape(-1, 2)
    for k in range(nb):
        sel = vb == k + 1
        mk = m[0::2][sel]
        com = (X[sel] * mk[:, None]).sum(0) / mk.sum()
        assert M[k] == pytest.approx(mk.sum(), rel=1e-15)
        assert np.allclose(COM[2 * k:2 * k + 2], com, rtol=1e-14, atol=1e-15)
        r = X[sel] - com
        assert I[k] == pytest.approx((mk * (r ** 2).sum(1)).sum(), rel=1e-13)
        assert np.allclose(ri.reshape(-1, 2)[sel], r, atol=1e-15)
    assert np.all(ri.reshape(-1, 2)[vb == 0] == 0)


def _boundary(M, COM, ang, I, bverts, vb, ri, mass, alpha, lamB, xi, h):
    O.lib().ora_boundary2d(len(M), P(M), P(COM), P(ang), P(I), len(bverts),
                           P(np.asarray(bverts, np.int32), C.c_int32), P(vb, C.c_int32), P(ri),
                           P(mass), C.c_double(alpha), P(lamB), P(xi), C.c_double(h))


def _square_body(extra=1):
    """Four unit-mass particles on the unit square (body 1) plus `extra`
    elastic particles of mass 2 (body 0)."""
    pts = [[0, 0], [1, 0], [1, 1], [0, 1]] + [[2, 0.5]] * extra
    p = np.array(pts, dtype=np.float64).reshape(-1)
    mass = np.repeat([1.0, 1, 1, 1] + [2.0] * extra, 2)
    vb = np.array([1, 1, 1, 1] + [0] * extra, dtype=np.int32)
    M, COM, I, ri = _bodies(p, vb, mass, 1)
    return p, mass, vb, M, COM, I, ri


def test_boundary2d_zero_distance_is_a_no_op():
    p, mass, vb, M, COM, I, ri = _square_body()
    ang = np.zeros(1)
    lamB = np.zeros(1)
    xi = p.copy()
    COM0 = COM.copy()
    _boundary(M, COM, ang, I, [1], vb, ri, mass, 0.0, lamB, xi, 0.01)
    assert np.all(np.isfinite(xi)) and lamB[0] == 0
    assert np.array_equal(COM, COM0) and ang[0] == 0


def test_boundary2d_through_com_is_the_textbook_point_constraint():
    """Boundary vertices at r = 0 have no rotational term: with alpha = 0 the
    XPBD point constraint gives dlambda = -C / (w_particle + w_body/nb) and
    the body moves by dlambda n / M per vertex (Tonge et al. mass splitting).
    Two coincident boundary vertices (nb = 2) must move the body by twice the
    single-vertex impulse computed with the halved body weight."""
    for nb in (1, 2):
        p = np.array([0, 0, 2, 0, 1, 0] + [1, 0] * nb, dtype=np.float64)
        # vertices 0,1,2 rigid (COM at (1,0)), nb extra rigid copies of the COM
        vb = np.array([1, 1, 1] + [1] * nb, dtype=np.int32)
        mass = np.repeat([1.0, 1.0, 2.0] + [1.0] * nb, 2)
        M, COM, I, ri = _bodies(p, vb, mass, 1)
        bverts = list(range(3, 3 + nb))
        xi = p.copy()
        delta = 0.1
        for v in bverts:
            xi[2 * v + 1] += delta  # elastic side displaced +y
        ang, lamB = np.zeros(1), np.zeros(nb)
        COM0 = COM.copy()
        _boundary(M, COM, ang, I, bverts, vb, ri, mass, 0.0, lamB, xi, 0.01)
        dl = -delta / (1.0 / 1.0 + (1.0 / M[0]) / nb)
        assert lamB == pytest.approx(np.full(nb, dl), rel=1e-14)
        # n = (rigid - elastic)/C = (0, -1): body pulled toward +y
        assert COM[1] - COM0[1] == pytest.approx(nb * (-dl) / M[0], rel=1e-13)
        assert COM[0] == COM0[0]
        assert ang[0] == 0.0


def test_boundary2d_off_center_turns_the_body_with_half_the_impulse():
    p, mass, vb, M, COM, I, ri = _square_body(extra=0)
    # boundary vertex 1 at (1,0): r = (0.5, -0.5) from COM (0.5, 0.5)
    xi = p.copy()
    xi[3] += 0.05  # elastic side above the rigid point
    ang, lamB = np.zeros(1), np.zeros(1)
    COM0, h = COM.copy(), 0.01
    _boundary(M, COM, ang, I, [1], vb, ri, mass, 0.0, lamB, xi, h)
    r = np.array([0.5, -0.5])
    n = np.array([0.0, -1.0])
    crn = r[0] * n[1] - r[1] * n[0]
    w = (1.0 / M[0]) / 1 + crn ** 2 / I[0]
    dl = -0.05 / (1.0 / mass[3] + w)
    pimp = dl * n
    assert lamB[0] == pytest.approx(dl, rel=1e-14)
    assert COM - COM0 == pytest.approx(pimp / M[0], rel=1e-13)
    textbook = (r[0] * pimp[1] - r[1] * pimp[0]) / I[0]
    assert ang[0] == pytest.approx(0.5 * textbook, rel=1e-13)
    assert ang[0] > 0  # pulled up at its right edge: counter-clockwise
    # the boundary vertex is placed on the moved body, not kept elastic
    c, s = np.cos(ang[0]), np.sin(ang[0])
    want = np.array([c * ri[2] - s * ri[3], s * ri[2] + c * ri[3]]) + COM
    assert xi[2:4] == pytest.approx(want, rel=1e-14)


def _contacts(vert, normal, point, compl):
    oc = O.OraContacts()
    keep = [np.asarray(vert, np.int32), np.ascontiguousarray(normal, np.float64),
            np.ascontiguousarray(point, np.float64), np.asarray(compl, np.float64)]
    oc.K = len(vert)
    oc.vertex, oc.normal, oc.point, oc.compliance = (P(keep[0], C.c_int32), P(keep[1]),
                                                     P(keep[2]), P(keep[3]))
    return oc, keep


def test_rigid_contact_pushes_along_the_normal_only_when_penetrating():
    p, mass, vb, M, COM, I, ri = _square_body(extra=0)
    h = 0.01
    for depth, moves in ((-0.02, False), (0.02, True)):
        ang = np.zeros(1)
        COMw = COM.copy()
        xi = p.copy()
        xi[1] -= depth  # vertex 0 at y = -depth
        # plane y = 0, normal +y, contact point the projection of vertex 0
        oc, keep = _contacts([0], [0.0, 1.0], [0.0, 0.0], [0.0])
        cl = np.zeros(1)
        O.lib().ora_contacts2d(1, P(M), P(COMw), P(ang), P(I), C.byref(oc), P(vb, C.c_int32),
                               P(ri), P(mass), P(cl), P(xi), C.c_double(h), C.c_double(1.0), 1)
        if not moves:
            assert cl[0] == 0 and np.array_equal(COMw, COM) and ang[0] == 0
        else:
            r = np.array([-0.5, -0.5])
            crn = r[0] * 1.0 - r[1] * 0.0
            w = 1.0 / M[0] + crn ** 2 / I[0]
            dl = depth / (2 * w)
            assert cl[0] == pytest.approx(dl, rel=1e-14)
            assert COMw - COM == pytest.approx([0.0, dl / M[0]], abs=1e-16, rel=1e-13)
            assert ang[0] == pytest.approx(0.5 * (r[0] * dl) / I[0], rel=1e-13)


@pytest.mark.parametrize("layers,iters", [((90, 50, 0), (3, 3, 6)), ((75, 0), (4, 8))])
def test_multilayer2d_oracle_vs_numpy(layers, iters):
    sc = scenes.cantilever2d(10, 4, cell=0.1)
    m = Mesh(sc)
    it = XpbdLayer2D()
    it.Gravity = -9.81
    it.setLayers(list(layers))
    it.iterations = list(iters)
    h = 0.01
    plane = PlaneContactFinder([0, 1], [0, -0.004])
    osim = O.OracleSim(m, copy.deepcopy(it), h, plane)
    ref = NumpyXpbd2DLayers(m, -9.81, h, iters, it.layersMap, contact=plane, restitution=1e-5,
                            contact_resistance=1.0, boundary_compliance=it.boundaryCompliance)
    rigid_contacts = 0
    bodies_seen = 0
    for k in range(25):
        prev = osim.p.copy()
        o = osim.step()
        ref.p = prev.reshape(-1, 2).copy()
        ref.v = osim_v_prev if k else np.zeros_like(ref.p)
        ref.step_layers(o["sets"])
        osim_v_prev = osim.v.reshape(-1, 2).copy()
        assert rel(osim.p, ref.p.reshape(-1)) < 1e-11, k
        assert rel(osim.v, ref.v.reshape(-1)) < 1e-9, k
        bodies_seen = max(bodies_seen, int(o["numRigids"].max()))
        c = ref.find_contacts()
        if c is not None and len(c["vertex"]):
            vbs = np.zeros(m.nV, int)
            col = int(it.layersMap[0]) - 1
            rig = o["sets"][:, col] > 0
            vbs[m.T0[rig].reshape(-1)] = 1
            rigid_contacts += int(vbs[c["vertex"]].sum())
    assert bodies_seen > 0
    assert rigid_contacts > 0
