"""Layer maps on the device (mlx_layer_maps) against the oracle's
buildLayers chain: stable sort, union-find bodies, boundary vertices.
Bit-exact on keys with ties, zeros of both signs and NaNs, several sorted
layers and pinned elements, 3D beams and 2D cantilevers."""
import ctypes as C
import zlib

import numpy as np
import pytest

import oracle as O
from multilayersolver2024_amd import _lib as L
from multilayersolver2024_amd import scenes
from multilayersolver2024_amd.mesh import DeviceMesh, Mesh

pytestmark = pytest.mark.gpu


def _oracle_maps(m, keys, layers):
    nT, nV, ns = m.nT, m.nV, len(layers)
    order = np.zeros(nT, np.int32)
    O.lib().ora_stable_sort_ascend(O.P(keys), nT, O.P(order, C.c_int32))
    sets = np.zeros(nT * ns, np.int32)
    nr = np.zeros(ns, np.int32)
    T = np.ascontiguousarray(m.T0, np.int32)
    forced = np.ascontiguousarray(m.forcedElastic, np.uint8)
    lay = np.asarray(layers, np.float64)
    O.lib().ora_union_find_by_vertice(O.P(order, C.c_int32), nT, O.P(lay), ns,
                                      O.P(forced, C.c_uint8), O.P(T, C.c_int32), m.vpe, nV,
                                      O.P(sets, C.c_int32), O.P(nr, C.c_int32))
    isB = np.zeros(nV * ns, np.uint8)
    isR = np.zeros(nV * ns, np.uint8)
    vb = np.zeros(nV * ns, np.int32)
    O.lib().ora_find_boundary_vertices(nV, nT, m.vpe, O.P(T, C.c_int32), ns,
                                       O.P(sets, C.c_int32), O.P(isB, C.c_uint8),
                                       O.P(isR, C.c_uint8), O.P(vb, C.c_int32))
    return sets, nr, isB, vb


def _device_maps(dm, keys, layers):
    m = dm.mesh
    ns = len(layers)
    sets = np.zeros(m.nT * ns, np.int32)
    nr = np.zeros(ns, np.int32)
    isB = np.zeros(m.nV * ns, np.uint8)
    vb = np.zeros(m.nV * ns, np.int32)
    lay = np.asarray(layers, np.float64)
    L.check("mlx_layer_maps", L.lib().mlx_layer_maps(
        dm.handle, L.ptr(keys), L.ptr(lay), ns, L.ptr(sets, C.c_int32), L.ptr(nr, C.c_int32),
        L.ptr(isB, C.c_uint8), L.ptr(vb, C.c_int32)))
    return sets, nr, isB, vb


def _keys(rng, n, kind):
    if kind == "smooth":
        return rng.random(n) * 1e-3
    k = np.round(rng.random(n) * 7) * 0.25          # many ties
    k[rng.random(n) < 0.05] = np.nan
    k[rng.random(n) < 0.05] = -0.0
    k[rng.random(n) < 0.05] = 0.0
    k[rng.random(n) < 0.02] = np.inf
    return k


@pytest.mark.parametrize("case", ["beam", "beam_pinned", "cantilever"])
@pytest.mark.parametrize("kind", ["smooth", "ties"])
def test_device_layer_maps_bit_exact(case, kind, oracle_lib):
    """The device builds each layer's bodies from the previous layer's
    forest (the percentages ascend, so the inserted sets are nested): six
    layers, five increments."""
    rng = np.random.default_rng(zlib.crc32(f"{case}-{kind}".encode()))
    if case.startswith("beam"):
        sc = scenes.beam3d(12, 5, 4, seed=2)
        if case == "beam":
            sc["pinned"] = np.zeros(0, np.int32)
        m = Mesh(sc)
    else:
        m = Mesh(scenes.cantilever2d(30, 8))
    layers = [0.0, 12.5, 50.0, 87.5, 98.4375, 100.0]
    dm = DeviceMesh(m, L.MLX_FP64)
    try:
        for rep in range(3):
            keys = _keys(rng, m.nT, kind)
            got = _device_maps(dm, keys, layers)
            want = _oracle_maps(m, keys, layers)
            print(case, kind, rep, "bodies", list(want[1]), flush=True)
            for g, w, name in zip(got, want, ("sets", "numRigids", "isBoundary", "vertexRigidBody")):
                assert np.array_equal(g, w), name
            assert want[1].max() > 0
    finally:
        dm.close()


def test_device_and_host_maps_give_the_same_substeps(oracle_lib):
    """mlx_integrate with maps built on the device and on the host: equal
    bits, substep after substep, on the beam with rigid layers."""
    import copy
    from helpers import beam_case
    from multilayersolver2024_amd.integrator import Simulation
    m, integ, plane = beam_case(z0=0.01)
    a = Simulation(m, integ, 1 / 60, plane, precision=L.MLX_FP64)
    b = Simulation(m, copy.deepcopy(integ), 1 / 60, plane, precision=L.MLX_FP64)
    a.dmesh.set_option(L.MLX_OPT_DEVICE_MAPS, 0)
    b.dmesh.set_option(L.MLX_OPT_DEVICE_MAPS, 1)
    try:
        for k in range(6):
            ra, rb = a.step(), b.step()
            pa, va, la, ka = a.download()
            pb, vb, lb, kb = b.download()
            print("maps", k, ra.itersPerLayer, rb.itersPerLayer, ka, flush=True)
            assert ra.itersPerLayer == rb.itersPerLayer and ka == kb
            assert np.array_equal(pa, pb) and np.array_equal(la, lb) and np.array_equal(va, vb)
    finally:
        a.close()
        b.close()
