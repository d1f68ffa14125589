"""The 81,920-tet beam of the headline workload (64 x 16 x 16 hex cells x 5
tets, 18,785 vertices, x = 0 face pinned, per-element E and nu from the
seed) at full size on the GPU:

* FP64 against the oracle over substeps through touchdown, with the
  headline layers [98.4375, 87.5, 0], iterations [2, 4, 14] and the
  relative 1e-4 stop: equal iteration and contact counts, positions and
  lambdas to 1e-9.  The coarsest layer's body holds over 64 x 256 rigid
  vertices, so its restriction sums more than 64 chunks (lane-strided
  pass) -- the path the timed substeps take;
* FP32 from a state the oracle produced: one substep each, positions
  within 1e-4 relative of the FP64 oracle (the error is printed);
* the device layer maps (sort, union-find, boundary maps) bit-exact against
  the oracle at 81,920 and 655,360 elements on strain-rate keys of real
  substeps, repeated, with the nested headline layers and more."""
import copy
import ctypes as C
import time

import numpy as np
import pytest

import oracle as O
from helpers import rel
from multilayersolver2024_amd import _lib as L
from multilayersolver2024_amd import scenes
from multilayersolver2024_amd.integrator import PlaneContactFinder, Simulation, XpbdLayer3D
from multilayersolver2024_amd.mesh import DeviceMesh, Mesh
from test_gpu_maps import _device_maps, _oracle_maps

pytestmark = pytest.mark.gpu

H = 1 / 60
LAYERS = (98.4375, 87.5, 0)
ITERS = (2, 4, 14)


def _say(*a):
    print("[config1]", *a, flush=True)


def _integ():
    it = XpbdLayer3D()
    it.Gravity = -9.81
    it.setLayers(list(LAYERS))
    it.iterations = list(ITERS)
    it.computeResiduals = True
    it.StopAtImprovementPercent = 99.99
    it.hangingStop = 1000
    it.useRigidContacts = False
    return it


def test_c1_beam_fp64_matches_oracle_at_full_size(oracle_lib):
    """Substep 1 through mlx_integrate (device layer maps, contacts, the
    whole substep) against the oracle's integrate; substeps 2 and 3 --
    touchdown, about 1,100 ground contacts -- through the layerSolve seam
    from the oracle's own state and layer maps.  From substep 2 on, the
    coarse body moves rigidly, so its strain-rate keys are rounding noise of
    the previous positions: two implementations that agree to 1e-16 sort
    them differently and build other (equally valid) layer sets, which is
    why the later substeps are compared on the oracle's maps."""
    from multilayersolver2024_amd.integrator import layer_solve
    from test_gpu_parity import _maps_from_sets
    t0 = time.time()
    m = Mesh(scenes.beam3d(64, 16, 16, seed=0, z0=0.002))
    assert (m.nT, m.nV) == (81920, 18785)
    plane = PlaneContactFinder([0, 0, 1], [0, 0, 0])
    sim = Simulation(m, _integ(), H, plane, precision=L.MLX_FP64)
    osim = O.OracleSim(m, _integ(), H, plane)
    try:
        r = sim.step()
        o = osim.step()
        p, v, lam, K = sim.download()
        _say("fp64 integrate", r.itersPerLayer, o["iters"], K, osim.last_contacts,
             "p", rel(p, osim.p), "lam", rel(lam, osim.lambdas()), "t", round(time.time() - t0, 1))
        assert K == osim.last_contacts
        assert r.itersPerLayer == o["iters"]
        assert rel(p, osim.p) < 1e-9 and rel(v, osim.v) < 1e-7
        assert rel(lam, osim.lambdas()) < 1e-9
        np.testing.assert_allclose(r.residualArray, o["residuals"], rtol=1e-7, atol=1e-14)
    finally:
        sim.close()
    dm = DeviceMesh(m, L.MLX_FP64)
    contacts_seen = 0
    try:
        unp = np.ones(3 * m.nV, bool)
        unp[m.pinnedDOFs - 1] = False
        for k in (2, 3):
            p0, v0 = osim.p.copy(), osim.v.copy()
            f = np.zeros(3 * m.nV)
            f[2::3] = m.mass[2::3] * -9.81
            rv = v0 + np.where(unp, H * f / m.mass, 0.0)
            c = O.plane_contacts(3, p0, [0, 0, 1], [0, 0, 0], 1e-4, strict=False)
            o = osim.step()
            maps = _maps_from_sets(m, o["sets"], o["numRigids"])
            Kc = c["K"]
            contacts = dict(vertexID=c["vertex"] + 1, normals=c["normal"].reshape(Kc, 3),
                            points=c["point"].reshape(Kc, 3), compliance=c["compliance"])
            integ = _integ()
            integ.runUntilResSmallerThan = o["runUntilAtEntry"]
            xi, lam, _, out, res, _ = layer_solve(dm, integ, H, maps, p0, np.zeros((m.nT, 6)), p0,
                                                  residualv=rv, pinnedx=osim.pinnedx.copy(),
                                                  contacts=contacts)
            col = osim.integ.layersMap[0] - 1
            rigid_verts = len(np.unique(m.T0[o["sets"][:, col] == 1]))
            iters = list(out.itersPerLayer[:out.layersRun])
            _say("fp64 seam", k, iters, o["iters"], Kc, "bodies", list(o["numRigids"]),
                 "rigid verts of body 1", rigid_verts, "p", rel(xi, osim.p),
                 "lam", rel(lam, osim.lambdas()), "t", round(time.time() - t0, 1))
            assert iters == o["iters"]
            assert rel(xi, osim.p) < 1e-9
            assert rel(lam, osim.lambdas()) < 1e-9
            np.testing.assert_allclose(res, o["residuals"], rtol=1e-7, atol=1e-14)
            # the coarse body's restriction sums more than 64 chunks of 256
            assert rigid_verts > 64 * 256
            contacts_seen += Kc
        assert contacts_seen > 0
    finally:
        dm.close()


def test_c1_fp32_substep_error_from_an_oracle_state(oracle_lib):
    m = Mesh(scenes.beam3d(64, 16, 16, seed=0, z0=0.002))
    plane = PlaneContactFinder([0, 0, 1], [0, 0, 0])
    osim = O.OracleSim(m, _integ(), H, plane)
    for _ in range(3):
        osim.step()
    p0, v0 = osim.p.copy(), osim.v.copy()
    s32 = Simulation(m, _integ(), H, plane, precision=L.MLX_FP32, p0=p0, v0=v0)
    o2 = O.OracleSim(m, _integ(), H, plane, p0=p0, v0=v0)
    try:
        for k in range(2):
            r = s32.step()
            o = o2.step()
            p, v, lam, K = s32.download()
            e = rel(p, o2.p)
            disp = rel(p - p0, o2.p - p0)
            _say("fp32", k, r.itersPerLayer, o["iters"], K, o2.last_contacts, "p", e,
                 "displacement", disp, "lam", rel(lam, o2.lambdas()), "v", rel(v, o2.v))
            assert K == o2.last_contacts
            assert e < 1e-4
    finally:
        s32.close()


def _real_keys(m, steps=6):
    """Strain-rate keys (mexEdiffNorm3D of B*p at two consecutive substeps,
    through the oracle's B*p and EDiff) of a GPU FP32 run of the beam."""
    sim = Simulation(m, _integ(), H, PlaneContactFinder([0, 0, 1], [0, 0, 0]),
                     precision=L.MLX_FP32)
    states = []
    try:
        for _ in range(steps):
            sim.step()
            states.append(sim.download()[0])
    finally:
        sim.close()
    prev, cur = states[-2], states[-1]
    T = np.ascontiguousarray(m.T0, np.int32)
    Fa, Fb = np.zeros(9 * m.nT), np.zeros(9 * m.nT)
    DmInv = np.ascontiguousarray(m.DmInv)
    O.lib().ora_b_times_p3d(O.P(T, C.c_int32), m.nT, O.P(DmInv), O.P(np.ascontiguousarray(cur)), O.P(Fa))
    O.lib().ora_b_times_p3d(O.P(T, C.c_int32), m.nT, O.P(DmInv), O.P(np.ascontiguousarray(prev)), O.P(Fb))
    keys = np.zeros(m.nT)
    O.lib().ora_ediff_norm3d(O.P(Fa), O.P(Fb), m.nT, C.c_double(H), O.P(keys))
    return keys


@pytest.mark.parametrize("size", [(64, 16, 16), (128, 32, 32)])
def test_device_layer_maps_bit_exact_at_scale(size, oracle_lib):
    m = Mesh(scenes.beam3d(*size, seed=0, z0=0.002))
    keys = _real_keys(m)
    assert np.unique(keys).size > m.nT // 2
    dm = DeviceMesh(m, L.MLX_FP32)
    try:
        for layers in ([87.5, 98.4375], [0.0, 12.5, 50.0, 87.5, 98.4375, 100.0]):
            want = _oracle_maps(m, keys, layers)
            for rep in range(3):
                got = _device_maps(dm, keys, layers)
                _say("maps", m.nT, layers, rep, "bodies", list(want[1]))
                for g, w, name in zip(got, want, ("sets", "numRigids", "isBoundary",
                                                   "vertexRigidBody")):
                    assert np.array_equal(g, w), name
            assert want[1].max() > 0
    finally:
        dm.close()
