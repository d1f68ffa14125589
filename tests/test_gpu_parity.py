"""HIP path against the CPU oracle (FP64 to 1e-9 relative, FP32 to 1e-4)."""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import beam_case, cantilever_case, rel
from multilayersolver2024_amd import _lib as L
from multilayersolver2024_amd.integrator import Simulation

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_ediff_bit_exact(oracle_lib):
    import ctypes as C
    rng = np.random.default_rng(1)
    for nT in (1, 7, 1000):
        for dim, nf in ((3, 9), (2, 4)):
            _say("ediff", dim, nT)
            Fa = rng.normal(size=nf * nT)
            Fb = rng.normal(size=nf * nT)
            h = 1 / 60
            got = np.zeros(nT)
            fn = L.lib().mlx_ediff_norm3d if dim == 3 else L.lib().mlx_ediff_norm2d
            L.check("ediff", fn(L.ptr(Fa), L.ptr(Fb), nT, h, L.ptr(got)))
            want = np.zeros(nT)
            ofn = oracle_lib.ora_ediff_norm3d if dim == 3 else oracle_lib.ora_ediff_norm2d
            ofn(O.P(Fa), O.P(Fb), nT, C.c_double(h), O.P(want))
            assert np.array_equal(got, want)


@pytest.mark.parametrize("steps", [6])
def test_beam_fp64_matches_oracle(steps, oracle_lib):
    m, integ, plane = beam_case()
    _say("beam fp64", m.nV, m.nT)
    sim = Simulation(m, integ, 1 / 60, plane, precision=L.MLX_FP64)
    import copy
    ointeg = copy.deepcopy(integ)
    osim = O.OracleSim(m, ointeg, 1 / 60, plane)
    try:
        for k in range(steps):
            r = sim.step()
            o = osim.step()
            p, v, lam, K = sim.download()
            _say(k, r.itersPerLayer, o["iters"], K, osim.last_contacts, rel(p, osim.p))
            assert K == osim.last_contacts
            assert r.itersPerLayer == o["iters"]
            assert rel(p, osim.p) < 1e-9
            assert rel(v, osim.v) < 1e-7
            assert rel(lam, osim.lambdas()) < 1e-9
            np.testing.assert_allclose(r.residualArray, o["residuals"], rtol=1e-7, atol=1e-14)
    finally:
        sim.close()


@pytest.mark.parametrize("case", ["one_cube", "no_pins_no_contacts", "all_rigid_top"])
def test_small_and_degenerate_scenes_fp64_match_oracle(case, oracle_lib):
    """Boundary cases of the substep: a single cube (5 tets, one color
    each) that falls as one rigid body, lands and bounces; a free-falling
    beam with no pinned vertex and no contact; a pinned beam whose coarsest
    layer makes every insertable element rigid (100 %) before it hits the
    ground.

    The layer stacks are chosen so that the substep is well conditioned:
    the oracle moves by about 1e-15 (relative) after 8 substeps when its
    start positions are perturbed by 1e-15.  A stack that sorts a partial
    layer (50 %, 87.5 %) over a mesh at rest orders strain-rate keys that
    are pure rounding noise, so rounding-level differences between two
    correct implementations change the rigid sets (the oracle alone then
    moves by 1e-8 after two substeps and by 1e-1 after a cube lands); such
    stacks are covered from bit-identical start states by the beam cases.
    λ of an undeformed tet is rounding noise (|λ| ~ 1e-19 here, against
    1e-4 .. 1e-1 for a deformed one), so λ is compared with an absolute
    floor of 1e-15 besides the 1e-9 relative bound."""
    import copy
    from multilayersolver2024_amd import scenes
    from multilayersolver2024_amd.integrator import PlaneContactFinder, XpbdLayer3D
    from multilayersolver2024_amd.mesh import Mesh
    if case == "one_cube":
        sc = scenes.beam3d(1, 1, 1, seed=3, z0=0.004)
        sc["pinned"] = np.zeros(0, np.int32)
        layers, iters = (100.0, 0), (3, 14)
    elif case == "no_pins_no_contacts":
        sc = scenes.beam3d(6, 2, 2, seed=4, z0=5.0)
        sc["pinned"] = np.zeros(0, np.int32)
        layers, iters = (98.4375, 87.5, 0), (2, 4, 14)
    else:
        sc = scenes.beam3d(6, 3, 2, seed=5, z0=0.004)
        layers, iters = (100.0, 0), (3, 10)
    m = Mesh(sc)
    it = XpbdLayer3D()
    it.Gravity = -9.81
    it.setLayers(list(layers))
    it.iterations = list(iters)
    it.computeResiduals = True
    it.StopAtImprovementPercent = 99.99
    it.hangingStop = 200
    plane = PlaneContactFinder([0, 0, 1], [0, 0, 0])
    sim = Simulation(m, it, 1 / 60, plane, precision=L.MLX_FP64)
    osim = O.OracleSim(m, copy.deepcopy(it), 1 / 60, plane)
    contacts = 0
    try:
        for k in range(8):
            r = sim.step()
            o = osim.step()
            p, v, lam, K = sim.download()
            olam = osim.lambdas()
            dl = float(np.linalg.norm(lam - olam))
            _say(case, k, r.itersPerLayer, o["iters"], K, osim.last_contacts, rel(p, osim.p), dl)
            assert K == osim.last_contacts
            assert list(r.itersPerLayer) == list(o["iters"])
            assert rel(p, osim.p) < 1e-9
            assert dl <= 1e-9 * np.linalg.norm(olam) + 1e-15
            contacts += K
        if case == "no_pins_no_contacts":
            assert contacts == 0
        else:
            assert contacts > 0
    finally:
        sim.close()


def test_cantilever2d_fp64_200_steps(oracle_lib):
    m, integ = cantilever_case()
    sim = Simulation(m, integ, 0.01, None, precision=L.MLX_FP64)
    osim = O.OracleSim(m, integ, 0.01, None)
    try:
        for k in range(200):
            sim.step()
            osim.step()
        p, v, lam, _ = sim.download()
        _say("2d", rel(p, osim.p), rel(lam, osim.lambdas()))
        assert rel(p, osim.p) < 1e-9
        assert rel(lam, osim.lambdas()) < 1e-9
    finally:
        sim.close()


@pytest.mark.parametrize("layers,iters,pct", [((90, 50, 0), (3, 3, 6), 0.0),
                                              ((75, 0), (4, 8), 99.0)])
def test_cantilever2d_rigid_layers_fp64_match_oracle(layers, iters, pct, oracle_lib):
    """2D multi-layer substeps with rigid bodies (boundary coupling, rigid and
    elastic ground contacts, rotateAllVerts2D) against the oracle, with the
    residual monitor on (and the relative stop rule in the second case)."""
    import copy
    from multilayersolver2024_amd import scenes
    from multilayersolver2024_amd.integrator import PlaneContactFinder, XpbdLayer2D
    from multilayersolver2024_amd.mesh import Mesh
    m = Mesh(scenes.cantilever2d(20, 6, cell=0.05))
    integ = XpbdLayer2D()
    integ.Gravity = -9.81
    integ.setLayers(list(layers))
    integ.iterations = list(iters)
    integ.computeResiduals = True
    integ.StopAtImprovementPercent = pct
    plane = PlaneContactFinder([0, 1], [0, -0.004])
    sim = Simulation(m, integ, 0.01, plane, precision=L.MLX_FP64)
    osim = O.OracleSim(m, copy.deepcopy(integ), 0.01, plane)
    try:
        nb = 0
        for k in range(30):
            r = sim.step()
            o = osim.step()
            p, v, lam, K = sim.download()
            nb = max(nb, int(o["numRigids"].max()))
            _say("2d layers", k, r.itersPerLayer, o["iters"], K, rel(p, osim.p))
            assert r.itersPerLayer == o["iters"], k
            assert K == osim.last_contacts, k
            assert rel(p, osim.p) < 1e-9, k
            assert rel(v, osim.v) < 1e-7, k
            assert rel(lam, osim.lambdas()) < 1e-9, k
        assert nb > 0
    finally:
        sim.close()


def test_armadillo_multilayer_fp64_matches_oracle(oracle_lib):
    """The reference's own irregular TetGen mesh (3d/data/arma_6, 627 tets)
    with rigid layers, the relative stop rule and ground contacts, FP64,
    substep by substep through the mlx_layer_solve seam from the oracle's
    state and with the oracle's layer maps.  (Free-running comparisons are
    not meaningful on this mesh: its strain-rate keys have near-ties, so a
    last-bit difference flips the sorted order and with it the rigid sets;
    the oracle itself moves by 1e-9 after one substep when its start
    positions are perturbed by 1e-15.)"""
    import copy
    from multilayersolver2024_amd.integrator import PlaneContactFinder, XpbdLayer3D, layer_solve
    from multilayersolver2024_amd.mesh import DeviceMesh, Mesh
    from test_meshio import _arma_scene
    m = Mesh(_arma_scene())
    integ = XpbdLayer3D()
    integ.Gravity = -9.81
    integ.setLayers([90, 60, 0])
    integ.iterations = [3, 3, 10]
    integ.computeResiduals = True
    integ.StopAtImprovementPercent = 99.0
    plane = PlaneContactFinder([0, 0, 1], [0, 0, 0])
    h = 1 / 60
    osim = O.OracleSim(m, copy.deepcopy(integ), h, plane)
    dm = DeviceMesh(m, L.MLX_FP64)
    unp = np.ones(3 * m.nV, bool)
    unp[m.pinnedDOFs - 1] = False
    f = np.zeros(3 * m.nV)
    f[2::3] = m.mass[2::3] * integ.Gravity
    try:
        kmax, bodies = 0, 0
        for k in range(20):
            p0, v0 = osim.p.copy(), osim.v.copy()
            runUntil0 = osim.s.runUntil
            c = O.plane_contacts(3, p0, [0, 0, 1], [0, 0, 0], 1e-4, strict=False)
            o = osim.step()
            if k == 0:
                pinnedx = osim.pinnedx.copy()
            maps = _maps_from_sets(m, o["sets"], o["numRigids"])
            rv = np.where(unp, v0 + h * f / m.mass, v0)
            K = c["K"]
            contacts = None
            if K:
                contacts = dict(vertexID=c["vertex"] + 1, normals=c["normal"].reshape(K, 3),
                                points=c["point"].reshape(K, 3), compliance=c["compliance"])
            step_integ = copy.deepcopy(integ)
            step_integ.runUntilResSmallerThan = runUntil0
            xi, lam, rv2, out, res, cl = layer_solve(dm, step_integ, h, maps, p0, np.zeros((m.nT, 6)),
                                                     p0, residualv=rv, pinnedx=pinnedx,
                                                     contacts=contacts)
            kmax, bodies = max(kmax, K), max(bodies, int(o["numRigids"].max()))
            lo = osim.lambdas()
            _say("arma", k, list(out.itersPerLayer[:out.layersRun]), o["iters"], K, rel(xi, osim.p),
                 rel(lam, lo))
            assert list(out.itersPerLayer[:out.layersRun]) == o["iters"], k
            assert rel(xi, osim.p) < 1e-9, k
            assert rel(lam, lo) < 1e-8, k
        assert kmax > 0 and bodies > 0
    finally:
        dm.close()


def test_graph_replay_matches_eager(oracle_lib):
    """Iteration chunks replayed as a captured hipGraph give the same bits as
    direct launches, substep after substep (eager work and syncs between)."""
    m, integ, plane = beam_case()
    a = Simulation(m, integ, 1 / 60, plane, precision=L.MLX_FP64)
    import copy
    b = Simulation(m, copy.deepcopy(integ), 1 / 60, plane, precision=L.MLX_FP64)
    a.dmesh.set_option(L.MLX_OPT_GRAPHS, 0)
    b.dmesh.set_option(L.MLX_OPT_GRAPHS, 1)
    b.dmesh.set_option(L.MLX_OPT_GRAPH_ITERS, 2)
    for sim in (a, b):
        sim.dmesh.set_option(L.MLX_OPT_PERSISTENT, 0)
    try:
        for k in range(6):
            ra, rb = a.step(), b.step()
            pa, va, la, ka = a.download()
            pb, vb, lb, kb = b.download()
            _say("graph", k, ra.itersPerLayer, rb.itersPerLayer)
            assert ra.itersPerLayer == rb.itersPerLayer and ka == kb
            assert np.array_equal(pa, pb) and np.array_equal(la, lb) and np.array_equal(va, vb)
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("precision", [L.MLX_FP64, L.MLX_FP32])
def test_single_workgroup_sweeps_equal_per_color_launches(precision, oracle_lib):
    """Rigid layers with few elastic elements per color sweep all colors in
    one workgroup over compacted lists (MLX_OPT_SWEEP_WG); the projections
    and their order within the layer are the same as with one launch per
    color, so every state is bit-identical, with ground contacts, the
    relative stop rule and bodies in both coarse layers."""
    import copy
    m, integ, plane = beam_case(nx=24, ny=6, nz=6, z0=0.004, pct=99.99)
    a = Simulation(m, integ, 1 / 60, plane, precision=precision)
    b = Simulation(m, copy.deepcopy(integ), 1 / 60, plane, precision=precision)
    a.dmesh.set_option(L.MLX_OPT_SWEEP_WG, 0)
    b.dmesh.set_option(L.MLX_OPT_SWEEP_WG, 1)
    try:
        for k in range(8):
            ra, rb = a.step(), b.step()
            pa, va, la, ka = a.download()
            pb, vb, lb, kb = b.download()
            _say("sweep_wg", k, ra.itersPerLayer, rb.itersPerLayer, ka)
            assert ra.itersPerLayer == rb.itersPerLayer and ka == kb
            assert np.array_equal(pa, pb) and np.array_equal(la, lb) and np.array_equal(va, vb)
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("precision,wg,local,size", [
    (L.MLX_FP64, 0, 0, "small"), (L.MLX_FP64, 7, 0, "small"), (L.MLX_FP32, 0, 0, "small"),
    (L.MLX_FP64, 0, 1, "small"), (L.MLX_FP64, 7, 1, "small"), (L.MLX_FP32, 0, 1, "small"),
    (L.MLX_FP64, 0, 1, "large"), (L.MLX_FP64, 3, 1, "large"), (L.MLX_FP64, 0, 0, "large")])
def test_persistent_loop_matches_per_color_launches(precision, wg, local, size, oracle_lib):
    """The cooperative persistent loop (colors split over workgroups, grid
    barriers) runs the same element, contact and pin updates as per-color
    launches; only the residual is summed in another order.  In FP64 the
    iteration counts match and the states agree to 1e-12 (the two kernels
    may fuse multiply-adds differently); in FP32 the states agree to 1e-5.
    Runs the stop rule (relative 1e-4) with ground contacts.  local=1 runs
    the loop on one XCD's workgroups with L2-keeping barriers; the large
    beam (19,200 tets, colors of ~480 tets) spreads every color over
    several workgroups so each barrier carries positions between CUs."""
    import copy
    if size == "small":
        m, integ, plane = beam_case(nx=10, ny=4, nz=4, z0=0.004, pct=99.99)
    else:
        m, integ, plane = beam_case(nx=60, ny=8, nz=8, z0=0.004, pct=99.99)
    a = Simulation(m, integ, 1 / 60, plane, precision=precision)
    b = Simulation(m, copy.deepcopy(integ), 1 / 60, plane, precision=precision)
    a.dmesh.set_option(L.MLX_OPT_PERSISTENT, 0)
    b.dmesh.set_option(L.MLX_OPT_PERSISTENT, 1)
    b.dmesh.set_option(L.MLX_OPT_PERSIST_WG, wg)
    b.dmesh.set_option(L.MLX_OPT_PERSIST_LOCAL, local)
    try:
        kmax = 0
        for k in range(8 if size == "small" else 5):
            ra, rb = a.step(), b.step()
            pa, va, la, ka = a.download()
            pb, vb, lb, kb = b.download()
            kmax = max(kmax, ka)
            _say("persist", k, ra.itersPerLayer, rb.itersPerLayer, ka, rel(pb, pa), rel(lb, la))
            assert ka == kb
            if precision == L.MLX_FP64:
                assert ra.itersPerLayer == rb.itersPerLayer
                assert rel(pb, pa) < 1e-12 and rel(lb, la) < 1e-12 and rel(vb, va) < 1e-10
                np.testing.assert_allclose(rb.residualArray, ra.residualArray, rtol=1e-9,
                                           atol=1e-300)
            else:
                assert rel(pb, pa) < 1e-5
        assert kmax > 0
    finally:
        a.close()
        b.close()


def test_fp32_plain_within_1e4_of_fp64_oracle(oracle_lib):
    """Precision form on the plain method (continuous dynamics): FP32 bulk
    projection stays within 1e-4 relative of the FP64 oracle."""
    m, integ, plane = beam_case(nx=12, ny=4, nz=4, layers=(0,), iterations=(20,))
    import copy
    sim = Simulation(m, integ, 1 / 60, plane, precision=L.MLX_FP32)
    osim = O.OracleSim(m, copy.deepcopy(integ), 1 / 60, plane)
    try:
        for k in range(8):
            r = sim.step()
            o = osim.step()
            p, v, lam, K = sim.download()
            e = rel(p, osim.p)
            _say("fp32 plain", k, r.itersPerLayer, K, e, rel(lam, osim.lambdas()))
            assert e < 1e-4
    finally:
        sim.close()


def test_fp32_multilayer_error_is_bounded(oracle_lib):
    """With strain-rate layers the rigid sets are a discontinuous function of
    the positions: once FP32 rounding flips a near-tie in the strain-rate
    order, FP32 and FP64 pick different rigid elements and the trajectories
    part (measured: 1e-7 for three substeps, then growing to ~6e-3 by the
    eighth).  The FP32 path is then a valid trajectory of its own, so the
    distance is bounded loosely here; the precision of the FP32 solve on
    identical layer maps is checked tightly by the FP32 seam test."""
    m, integ, plane = beam_case(nx=12, ny=4, nz=4)
    import copy
    sim = Simulation(m, integ, 1 / 60, plane, precision=L.MLX_FP32)
    osim = O.OracleSim(m, copy.deepcopy(integ), 1 / 60, plane)
    try:
        for k in range(8):
            r = sim.step()
            o = osim.step()
            p, v, lam, K = sim.download()
            e = rel(p, osim.p)
            _say("fp32 layers", k, r.itersPerLayer, e, rel(lam, osim.lambdas()))
            assert e < (1e-6 if k < 3 else 2e-2)
    finally:
        sim.close()


def _maps_from_sets(m, sets, nr):
    """findBoundaryVertices through the C ABI for per-layer body ids."""
    import ctypes as C
    ns = sets.shape[1]
    flat = L.i32(sets.T.reshape(-1))
    isB = np.zeros(m.nV * ns, np.uint8)
    isR = np.zeros(m.nV * ns, np.uint8)
    vb = np.zeros(m.nV * ns, np.int32)
    L.check("fbv", L.lib().mlx_find_boundary_vertices(
        L.ptr(m.t, C.c_int32), m.nT, m.vpe, m.nV, ns, L.ptr(flat, C.c_int32),
        L.ptr(isB, C.c_uint8), L.ptr(isR, C.c_uint8), L.ptr(vb, C.c_int32)))
    return dict(isElasticElement=(sets == 0).astype(np.uint8),
                vertexRigidBody=vb.reshape(ns, m.nV).T, isBoundaryVertex=isB.reshape(ns, m.nV).T,
                isRigidVertex=isR.reshape(ns, m.nV).T, numRigids=L.i32(nr))


def _seam_case(**beam):
    """Frame 1 of the beam sunk 1 cm into the ground plane (45 contacts on
    rigid, boundary, pinned and elastic vertices for the default beam), with
    the layer sets the oracle built for that substep and the inputs
    integrate() hands to layerSolve at frame 1 (xi = oldp = p, residual
    velocity h f / m)."""
    import copy
    m, integ, plane = beam_case(z0=-0.01, **beam)
    h = 1 / 60
    osim = O.OracleSim(m, copy.deepcopy(integ), h, plane)
    c = O.plane_contacts(3, m.p, [0, 0, 1], [0, 0, 0], 1e-4, strict=False)
    o = osim.step()
    assert int(np.count_nonzero(o["numRigids"])) > 0
    maps = _maps_from_sets(m, o["sets"], o["numRigids"])
    f = np.zeros(3 * m.nV)
    f[2::3] = m.mass[2::3] * integ.Gravity
    unp = np.ones(3 * m.nV, bool)
    unp[m.pinnedDOFs - 1] = False
    rv = np.where(unp, h * f / m.mass, 0.0)
    K = c["K"]
    contacts = dict(vertexID=c["vertex"] + 1, normals=c["normal"].reshape(K, 3),
                    points=c["point"].reshape(K, 3), compliance=c["compliance"])
    return m, integ, h, osim, maps, rv, contacts


@pytest.mark.parametrize("precision,tol", [(L.MLX_FP64, 1e-9), (L.MLX_FP32, 1e-5)])
def test_layer_solve_seam_matches_oracle_substep(precision, tol, oracle_lib):
    """mlx_layer_solve fed the arrays integrate() builds at frame 1 returns
    the oracle's end-of-substep xi and lambda (FP64 to 1e-9; FP32 bulk
    projection to 1e-5 on the same layer maps)."""
    from multilayersolver2024_amd.integrator import layer_solve
    from multilayersolver2024_amd.mesh import DeviceMesh
    m, integ, h, osim, maps, rv, contacts = _seam_case()
    K = len(contacts["vertexID"])
    dm = DeviceMesh(m, precision)
    try:
        xi, lam, rv2, out, res, cl = layer_solve(dm, integ, h, maps, m.p, np.zeros((m.nT, 6)), m.p,
                                                 residualv=rv, pinnedx=m.p[m.pinnedDOFs - 1],
                                                 contacts=contacts)
        e, el = rel(xi, osim.p), rel(lam, osim.lambdas())
        _say("seam", precision, K, list(out.itersPerLayer[:out.layersRun]), e, el)
        assert K > 0
        assert e < tol
        assert el < (tol if precision == L.MLX_FP64 else 1e-3)
    finally:
        dm.close()


def test_neohookean2d_sweep_matches_oracle(oracle_lib):
    import ctypes as C
    from multilayersolver2024_amd import scenes
    from multilayersolver2024_amd.mesh import DeviceMesh, Mesh
    sc = scenes.cantilever2d(20, 6)
    m = Mesh(sc)
    om = O.OracleMesh(m)
    rng = np.random.default_rng(42)
    xi = m.p + 0.01 * rng.normal(size=m.p.shape)
    isEl = L.u8(rng.random(m.nT) < 0.8)
    mu, lamL = L.f64(sc["mu"]), L.f64(sc["lam"])
    xo = xi.copy()
    lo = np.zeros(2 * m.nT)
    dm = DeviceMesh(m, L.MLX_FP64)
    xg = xi.copy()
    lg = np.zeros(2 * m.nT)  # column-major nT x 2
    try:
        for sweep in range(5):
            O.lib().ora_neohookean_constraints2d(C.byref(om.s), O.P(isEl, C.c_uint8), C.c_double(0.01),
                                                 O.P(mu), O.P(lamL), O.P(lo), O.P(xo))
            L.check("nh", L.lib().mlx_neohookean_constraints2d(dm.handle, L.ptr(isEl, C.c_uint8), 0.01,
                                                               L.ptr(mu), L.ptr(lamL), L.ptr(lg), L.ptr(xg)))
        lg_rm = lg.reshape(2, m.nT).T.reshape(-1)
        _say("neohookean", rel(xg, xo), rel(lg_rm, lo))
        assert rel(xg, xo) < 1e-9 and rel(lg_rm, lo) < 1e-9
    finally:
        dm.close()


def test_pcg_matches_oracle_and_dense_solve(oracle_lib):
    import ctypes as C
    import scipy.sparse as sp
    rng = np.random.default_rng(51)
    for n in (3, 40, 3000):
        M = sp.random(n, n, density=min(1.0, 8.0 / n), random_state=52, format="csr")
        A = (M @ M.T + sp.eye(n) * 4).tocsc()
        b = rng.normal(size=n)
        pinv = 1 / A.diagonal()
        it = min(n, 60)
        xo = np.zeros(n)
        ro = np.zeros(it)
        O.lib().ora_solve_pcg(n, O.P(A.indptr.astype(np.int32), C.c_int32),
                              O.P(A.indices.astype(np.int32), C.c_int32), O.P(A.data.copy()),
                              O.P(b.copy()), O.P(pinv.copy()), it, O.P(xo), O.P(ro))
        xg = np.zeros(n)
        rg = np.zeros(it)
        cp, ri = L.i32(A.indptr), L.i32(A.indices)
        L.check("pcg", L.lib().mlx_pcg_solve(n, L.ptr(cp, C.c_int32), L.ptr(ri, C.c_int32),
                                             L.ptr(A.data.copy()), L.ptr(b.copy()), L.ptr(pinv.copy()),
                                             it, L.ptr(xg), L.ptr(rg)))
        _say("pcg", n, rel(xg, xo), rg[-1])
        assert rel(xg, xo) < 1e-9
        if n <= 40:
            assert rel(xg, np.linalg.solve(A.toarray(), b)) < 1e-9


@pytest.mark.gpu
def test_layer_solve_seam_bodies_over_many_chunks_match_oracle(oracle_lib):
    """A 10,240-tet beam with a 17 x 17-vertex cross-section and layers
    [90, 40, 0]: each coarse layer is one body of 940 and 2,096 rigid
    vertices with 598 and 544 boundary vertices, so the body restriction
    and the boundary accumulation run over several 256-vertex chunks per
    body (the small beams above fit one chunk).  FP64 to 1e-9 against the
    oracle's substep on the oracle's layer maps."""
    from multilayersolver2024_amd.integrator import layer_solve
    from multilayersolver2024_amd.mesh import DeviceMesh
    m, integ, h, osim, maps, rv, contacts = _seam_case(nx=8, ny=16, nz=16, layers=(90, 40, 0))
    vb, bd = maps["vertexRigidBody"], maps["isBoundaryVertex"]
    for col in (1, 2):
        assert int(maps["numRigids"][col]) == 1
        assert np.count_nonzero(vb[:, col] == 1) > 2 * 256
        assert np.count_nonzero((vb[:, col] == 1) & (bd[:, col] != 0)) > 2 * 256
    dm = DeviceMesh(m, L.MLX_FP64)
    try:
        xi, lam, rv2, out, res, cl = layer_solve(dm, integ, h, maps, m.p, np.zeros((m.nT, 6)), m.p,
                                                 residualv=rv, pinnedx=m.p[m.pinnedDOFs - 1],
                                                 contacts=contacts)
        e, el = rel(xi, osim.p), rel(lam, osim.lambdas())
        _say("seam chunks", m.nT, list(out.itersPerLayer[:out.layersRun]), e, el)
        assert list(out.itersPerLayer[:out.layersRun]) == [2, 4, 14]
        assert e < 1e-9
        assert el < 1e-9
    finally:
        dm.close()
