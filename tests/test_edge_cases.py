"""Degenerate and boundary inputs of the host-side entry points: an empty
mesh, a star of elements around one vertex, out-of-range indices (an
error, never a guess), layers at 0 % and 100 %, and every element forced
elastic.  Integer outputs are checked against the oracle or a closed form."""
import ctypes as C

import numpy as np

import oracle as O
from multilayersolver2024_amd import _lib as L
from multilayersolver2024_amd import scenes


def _color(T0, nV):
    """mlx_naive_graph_coloring on 0-based element rows (passed 1-based,
    column-major as Matlab holds them)."""
    T0 = np.asarray(T0, dtype=np.int64).reshape(-1, 4)
    t = L.fortran(T0 + 1, np.int32) if len(T0) else np.zeros(1, np.int32)
    col = np.zeros(max(len(T0), 1), np.int32)
    nc = np.zeros(1, np.int32)
    r = L.lib().mlx_naive_graph_coloring(L.ptr(t, C.c_int32), len(T0), 4, nV,
                                         L.ptr(col, C.c_int32), L.ptr(nc, C.c_int32))
    return r, col[:len(T0)], int(nc[0])


def test_coloring_empty_mesh_has_no_colors():
    r, col, nc = _color(np.zeros((0, 4)), 4)
    assert r == 0 and nc == 0 and len(col) == 0


def test_coloring_star_gives_every_element_its_own_color():
    n = 9
    T0 = [[0, 1 + 3 * e, 2 + 3 * e, 3 + 3 * e] for e in range(n)]  # all share vertex 0
    r, col, nc = _color(T0, 1 + 3 * n)
    assert r == 0 and nc == n and list(col) == list(range(1, n + 1))


def test_coloring_disjoint_elements_share_one_color():
    T0 = [[4 * e, 4 * e + 1, 4 * e + 2, 4 * e + 3] for e in range(6)]
    r, col, nc = _color(T0, 24)
    assert r == 0 and nc == 1 and np.all(col == 1)


def test_out_of_range_vertex_is_an_error():
    r, _, _ = _color([[0, 1, 2, 9]], 4)
    assert r == L.MLX_ERR_ARG
    assert b"outside" in L.lib().mlx_last_error()


def _uf(order0, layers, forced, T0, nV):
    nT, ns = len(T0), len(layers)
    sets_o = np.zeros(nT * ns, np.int32)
    nr_o = np.zeros(ns, np.int32)
    O.lib().ora_union_find_by_vertice(O.P(np.asarray(order0, np.int32), C.c_int32), nT,
                                      O.P(np.asarray(layers, np.float64)), ns,
                                      O.P(np.asarray(forced, np.uint8), C.c_uint8),
                                      O.P(np.ascontiguousarray(T0, np.int32), C.c_int32), 4, nV,
                                      O.P(sets_o, C.c_int32), O.P(nr_o, C.c_int32))
    return sets_o.reshape(ns, nT), nr_o


def test_union_find_layers_at_0_and_100_percent():
    sc = scenes.beam3d(4, 2, 2, seed=1)
    T0 = np.asarray(sc["t"], np.int64)
    nT, nV = len(T0), len(sc["p"])
    rng = np.random.default_rng(0)
    order0 = rng.permutation(nT)
    sets, nr = _uf(order0, [0.0, 100.0], np.zeros(nT, np.uint8), T0, nV)
    assert nr[0] == 0 and np.all(sets[0] == 0)
    # 100 %: the whole (connected) beam is one body
    assert nr[1] == 1 and np.all(sets[1] == 1)
    # every element forced elastic: no bodies at any percentage
    sets, nr = _uf(order0, [50.0, 100.0], np.ones(nT, np.uint8), T0, nV)
    assert np.all(nr == 0) and np.all(sets == 0)
