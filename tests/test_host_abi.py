"""Host-side parts of the C ABI (no GPU needed): the library loads and
exports every declared entry point, and its precompute / layer-map
functions agree with the oracle (integer outputs bit-exact)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
from multilayersolver2024_amd import _lib as L
from multilayersolver2024_amd import scenes
from multilayersolver2024_amd.mesh import Mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "mlxpbd.h")).read()
    declared = set(re.findall(r"\b(mlx_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(L.EXPORTED)
    lib = L.lib()
    for name in declared:
        assert hasattr(lib, name), name


def test_no_device_is_an_error_not_a_fallback():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU node is visible")
    assert L.lib().mlx_init(0) == L.MLX_ERR_NODEVICE


@pytest.mark.parametrize("scene", ["beam", "cantilever"])
def test_coloring_matches_oracle(scene):
    sc = scenes.beam3d(6, 3, 3) if scene == "beam" else scenes.cantilever2d()
    m = Mesh(sc)
    T = np.ascontiguousarray(sc["t"], dtype=np.int32)
    col = np.zeros(len(T), dtype=np.int32)
    O.lib().ora_naive_graph_coloring(O.P(T, C.c_int32), len(T), T.shape[1], len(sc["p"]),
                                     O.P(col, C.c_int32))
    assert np.array_equal(col, m.elementColor)
    assert m.numColors == col.max()


def test_precompute_matches_oracle():
    sc = scenes.beam3d(5, 3, 2, seed=4)
    m = Mesh(sc)
    nV, nT = len(sc["p"]), len(sc["t"])
    o = dict(DmInv=np.zeros(9 * nT), vol=np.zeros(nT), mass=np.zeros(3 * nV),
             alpha=np.zeros(36 * nT), beta=np.zeros(36 * nT))
    args = [np.ascontiguousarray(sc[k], dtype=np.float64) for k in ("rho", "mu", "lam", "alpha1")]
    O.lib().ora_precompute3d(nV, nT, O.P(m.p), O.P(np.ascontiguousarray(sc["t"], dtype=np.int32), C.c_int32),
                             *[O.P(a) for a in args], O.P(o["DmInv"]), O.P(o["vol"]), O.P(o["mass"]),
                             O.P(o["alpha"]), O.P(o["beta"]))
    np.testing.assert_allclose(m.DmInv, o["DmInv"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(m.elV, o["vol"], rtol=1e-13)
    np.testing.assert_allclose(m.mass, o["mass"], rtol=1e-13)
    scale = np.abs(o["alpha"]).max()
    np.testing.assert_allclose(m.alpha, o["alpha"], rtol=1e-10, atol=1e-14 * scale)
    np.testing.assert_allclose(m.beta, o["beta"], rtol=1e-15)


def test_union_find_and_boundary_match_oracle():
    rng = np.random.default_rng(30)
    sc = scenes.beam3d(6, 3, 3)
    T = np.ascontiguousarray(sc["t"], dtype=np.int32)
    nT, nV = len(T), len(sc["p"])
    t1 = L.fortran(T + 1, np.int32)
    for trial in range(6):
        order = rng.permutation(nT).astype(np.int32)
        forced = L.u8(rng.random(nT) < 0.05)
        layers = np.array([0.0, 12.5, 87.5, 98.4375, 100.0])
        Ls = len(layers)
        want = np.zeros(nT * Ls, dtype=np.int32)
        wn = np.zeros(Ls, dtype=np.int32)
        O.lib().ora_union_find_by_vertice(O.P(order, C.c_int32), nT, O.P(layers), Ls,
                                          O.P(forced, C.c_uint8), O.P(T, C.c_int32), 4, nV,
                                          O.P(want, C.c_int32), O.P(wn, C.c_int32))
        got = np.zeros(nT * Ls, dtype=np.int32)
        gn = np.zeros(Ls, dtype=np.int32)
        oi = L.i32(order + 1)
        L.check("uf", L.lib().mlx_union_find_by_vertice(L.ptr(oi, C.c_int32), nT, L.ptr(layers), Ls,
                                                        L.ptr(forced, C.c_uint8), L.ptr(t1, C.c_int32), 4,
                                                        nV, L.ptr(got, C.c_int32), L.ptr(gn, C.c_int32)))
        assert np.array_equal(got, want)
        assert np.array_equal(gn, wn)
        ob, orv, ovb = np.zeros(nV * Ls, np.uint8), np.zeros(nV * Ls, np.uint8), np.zeros(nV * Ls, np.int32)
        O.lib().ora_find_boundary_vertices(nV, nT, 4, O.P(T, C.c_int32), Ls, O.P(want, C.c_int32),
                                           O.P(ob, C.c_uint8), O.P(orv, C.c_uint8), O.P(ovb, C.c_int32))
        pb, prv, pvb = np.zeros(nV * Ls, np.uint8), np.zeros(nV * Ls, np.uint8), np.zeros(nV * Ls, np.int32)
        L.check("fbv", L.lib().mlx_find_boundary_vertices(L.ptr(t1, C.c_int32), nT, 4, nV, Ls,
                                                          L.ptr(got, C.c_int32), L.ptr(pb, C.c_uint8),
                                                          L.ptr(prv, C.c_uint8), L.ptr(pvb, C.c_int32)))
        assert np.array_equal(pb, ob) and np.array_equal(prv, orv) and np.array_equal(pvb, ovb)


def test_union_find_rejects_bad_permutation():
    T = L.fortran(np.array([[0, 1, 2, 3]]) + 1, np.int32)
    bad = L.i32([2])
    out = np.zeros(1, np.int32)
    n = np.zeros(1, np.int32)
    r = L.lib().mlx_union_find_by_vertice(L.ptr(bad, C.c_int32), 1, L.ptr(np.array([0.0])), 1,
                                          L.ptr(L.u8([0]), C.c_uint8), L.ptr(T, C.c_int32), 4, 4,
                                          L.ptr(out, C.c_int32), L.ptr(n, C.c_int32))
    assert r == L.MLX_ERR_ARG
    assert b"permutation" in L.lib().mlx_last_error()
