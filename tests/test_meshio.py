"""Mesh readers on the reference's own mesh files (tests/golden/meshes:
3d/data/cube.*, arma_6.*, capsuleCoarse.mesh copied as data), and the
oracle on an irregular mesh: plain substeps against the numpy
restatement, layer maps against the breadth-first component definition."""
import os

import numpy as np
import pytest

import oracle as O
from helpers import rel
from multilayersolver2024_amd import meshio
from multilayersolver2024_amd.integrator import PlaneContactFinder, XpbdLayer3D
from multilayersolver2024_amd.mesh import Mesh
from numpy_xpbd import NumpyXpbd

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")


def _vol6(P, T):
    a, b, c, d = (P[T[:, k]] for k in range(4))
    return np.einsum("ij,ij->i", b - a, np.cross(c - a, d - a))


def test_cube_tetgen_known_answer():
    P, T, A = meshio.read_tetgen(os.path.join(MESHES, "cube"))
    assert P.shape == (8, 3) and T.shape == (6, 4) and A is None
    # node 4 is (1,1,1) and node 6 is (0,1,0) in the file (0-based numbering)
    assert np.array_equal(P[4], [1, 1, 1]) and np.array_equal(P[6], [0, 1, 0])
    assert np.array_equal(T[0], [0, 6, 4, 7])
    v = _vol6(P, T) / 6
    assert np.all(v > 0) and v.sum() == pytest.approx(1.0, rel=1e-15)


@pytest.mark.parametrize("name,nv,nt", [("arma_6", 259, 627)])
def test_tetgen_mesh_is_valid(name, nv, nt):
    P, T, A = meshio.read_tetgen(os.path.join(MESHES, name))
    assert P.shape == (nv, 3) and T.shape == (nt, 4)
    assert T.min() == 0 and T.max() == nv - 1
    assert np.all(_vol6(P, T) > 0)
    # every vertex is used, no duplicated element
    assert len(np.unique(T)) == nv
    assert len(np.unique(np.sort(T, axis=1), axis=0)) == nt


def test_medit_mesh_is_valid():
    P, T = meshio.read_medit(os.path.join(MESHES, "capsuleCoarse.mesh"))
    assert P.shape == (475, 3) and T.shape == (1753, 4)
    assert T.min() == 0 and T.max() == 474
    assert np.all(_vol6(P, T) > 0)
    # the capsule is about 1 x 1 x 3
    assert np.allclose(P.max(0) - P.min(0), [1, 1, 3], atol=0.02)


def _arma_scene(pin=True):
    P, T, _ = meshio.read_tetgen(os.path.join(MESHES, "arma_6"))
    pinned = np.nonzero(P[:, 1] > 1.0)[0] if pin else None
    return meshio.scene_from_mesh(P, T, E=2e5, nu=0.4, pinned=pinned, offset=(0, 0, 0.02),
                                  name="arma_6")


def test_arma_plain_oracle_vs_numpy():
    m = Mesh(_arma_scene())
    it = XpbdLayer3D()
    it.Gravity = -9.81
    it.iterations = [8]
    plane = PlaneContactFinder([0, 0, 1], [0, 0, 0])
    osim = O.OracleSim(m, it, 1 / 60, plane)
    ref = NumpyXpbd(m, -9.81, 1 / 60, 8, contact=plane, restitution=1e-2, contact_resistance=0.5)
    kmax = 0
    for k in range(10):
        osim.step()
        ref.step()
        kmax = max(kmax, osim.last_contacts)
        assert rel(osim.p, ref.p.reshape(-1)) < 1e-11, k
    assert kmax > 0


def test_arma_layer_maps_are_vertex_sharing_components():
    import ctypes as C
    from collections import deque
    m = Mesh(_arma_scene())
    rng = np.random.default_rng(5)
    order = rng.permutation(m.nT).astype(np.int32)
    layers = np.array([10.0, 50.0, 90.0])
    sets = np.zeros(m.nT * 3, np.int32)
    nr = np.zeros(3, np.int32)
    T = np.ascontiguousarray(m.T0, np.int32)
    O.lib().ora_union_find_by_vertice(O.P(order, C.c_int32), m.nT, O.P(layers), 3,
                                      O.P(np.ascontiguousarray(m.forcedElastic, np.uint8), C.c_uint8),
                                      O.P(T, C.c_int32), 4, m.nV, O.P(sets, C.c_int32),
                                      O.P(nr, C.c_int32))
    for l, pct in enumerate(layers):
        ins = np.zeros(m.nT, bool)
        ins[order[:int(np.floor(pct / 100 * m.nT))]] = True
        ins &= ~m.forcedElastic.astype(bool)
        by_v = [[] for _ in range(m.nV)]
        for e in np.nonzero(ins)[0]:
            for v in T[e]:
                by_v[v].append(e)
        comp = -np.ones(m.nT, int)
        roots = []
        for e in np.nonzero(ins)[0]:          # ascending: components by min element
            if comp[e] >= 0:
                continue
            comp[e] = len(roots)
            roots.append(e)
            q = deque([e])
            while q:
                a = q.popleft()
                for v in T[a]:
                    for b in by_v[v]:
                        if comp[b] < 0:
                            comp[b] = comp[e]
                            q.append(b)
        want = np.where(ins, comp + 1, 0)
        assert nr[l] == len(roots)
        assert np.array_equal(sets[l * m.nT:(l + 1) * m.nT], want)
