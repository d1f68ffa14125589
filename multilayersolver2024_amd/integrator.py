"""Host-side mirror of the reference's XPBD layer integrators.

``XpbdLayer3D`` / ``XpbdLayer2D`` carry the public properties of
xpbdLayer3D (3d/integrator/xpbdLayer3D.m:5-40) and xpbdLayer2D
(2d/integrator/xpbdLayer2D.m:7-44) under the same names and defaults;
``integrate()`` runs one substep on the GPU through ``mlx_integrate`` and
``layer_solve()`` is the drop-in for the static ``layerSolve`` seam
(``mlx_layer_solve``).  ``Simulation`` plays the role of simulate3D /
simulate for one mesh: contacts at the start of the step, the integrate
call and the position history.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .mesh import DeviceMesh


class _XpbdLayerBase:
    def setLayers(self, layersArray):
        """setLayers (xpbdLayer3D.m:595-599): [sortedLayers,~,ic] = unique(layers)."""
        self.layers = [float(x) for x in layersArray]
        self.sortedLayers = sorted(set(self.layers))
        self.layersMap = [self.sortedLayers.index(x) + 1 for x in self.layers]

    def rigid_contacts_enabled(self):
        """useRigidContacts (xpbdLayer3D.m:37) / useRigidConstacts, the 2D
        spelling (xpbdLayer2D.m:37)."""
        return bool(getattr(self, "useRigidContacts", getattr(self, "useRigidConstacts", True)))

    def settings(self):
        s = L.Settings()
        if len(self.layers) != len(self.iterations):
            raise ValueError("layers and iterations must have the same length")
        s.nLayers = len(self.layers)
        for i, (l, it) in enumerate(zip(self.layers, self.iterations)):
            s.layers[i] = l
            s.iterations[i] = int(it)
        s.gravity = self.Gravity
        s.airDrag = getattr(self, "airDrag", 0.0)
        s.boundaryCompliance = self.boundaryCompliance
        s.contactResistance = self.contactResistance
        s.restitution = self.restitution
        s.computeResiduals = int(bool(self.computeResiduals))
        s.hangingStop = int(self.hangingStop)
        s.giveUpEnabled = int(bool(self.giveUpEnabled))
        s.giveUpThreshold = self.giveUpThreshold
        s.lastLayerGiveUpEnabled = int(bool(self.LastLayerGiveUpEnabled))
        s.useRigidContacts = int(self.rigid_contacts_enabled())
        s.stopAtImprovementPercent = self.StopAtImprovementPercent
        s.runUntilResSmallerThan = self.runUntilResSmallerThan
        return s


class XpbdLayer3D(_XpbdLayerBase):
    """Properties of xpbdLayer3D (xpbdLayer3D.m:5-40) and Integrator.m:17-24."""

    def __init__(self):
        self.iterations = [10]
        self.layers = [0.0]
        self.sortedLayers = [0.0]
        self.layersMap = [1]
        self.boundaryCompliance = 1e-8
        self.computeResiduals = False
        self.giveUpEnabled = False
        self.LastLayerGiveUpEnabled = False
        self.giveUpThreshold = 0.0
        self.contactResistance = 0.5
        self.runUntilResSmallerThan = -1.0
        self.hangingStop = 1000
        self.airDrag = 0.0
        self.useRigidContacts = True
        self.StopAtImprovementPercent = 0.0
        self.Gravity = -9.8
        self.restitution = 1e-2  # contactVelocityUpdate(..., 3, 1e-2), xpbdLayer3D.m:740


class XpbdLayer2D(_XpbdLayerBase):
    """Properties of xpbdLayer2D (xpbdLayer2D.m:7-44)."""

    def __init__(self):
        self.iterations = [10]
        self.layers = [0.0]
        self.sortedLayers = [0.0]
        self.layersMap = [1]
        self.boundaryCompliance = 1e-8
        self.computeResiduals = False
        self.giveUpEnabled = False
        self.LastLayerGiveUpEnabled = False
        self.giveUpThreshold = 1e-8
        self.contactCompliance = 1e-4
        self.contactResistance = 1.0
        self.runUntilResSmallerThan = -1.0
        self.StopAtImprovementPercent = 0.0
        self.hangingStop = 1000
        self.useRigidConstacts = True
        self.Gravity = -9.8
        self.restitution = 1e-5  # contactVelocityUpdate default, xpbdLayer2D.m:578-580


class PlaneContactFinder:
    """PlaneContactFinder3D(normal, position) / PlaneContactFinder(normal,
    position) with ContactFinder.xpbdContactCompliance (ContactFinder.m:10)."""

    def __init__(self, normal, position, compliance=1e-4):
        self.normal = [float(x) for x in normal]
        self.position = [float(x) for x in position]
        self.compliance = float(compliance)

    def plane(self):
        pl = L.Plane()
        pl.enabled = 1
        for d in range(len(self.normal)):
            pl.normal[d] = self.normal[d]
            pl.position[d] = self.position[d]
        pl.compliance = self.compliance
        return pl

    def finder(self):
        f = L.ContactFinder()
        f.kind = L.MLX_CONTACT_PLANE
        f.compliance = self.compliance
        for d in range(len(self.normal)):
            f.normal[d] = self.normal[d]
            f.position[d] = self.position[d]
        return f


def eul2rotm_xyz(radians):
    """Rotation of the Euler angles (x, y, z) in the 'XYZ' sequence:
    Rx(a) Ry(b) Rz(c) (what RotationMatrixDegrees builds with eul2rotm)."""
    a, b, c = radians
    Rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    Rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
    return Rx @ Ry @ Rz


class PlatformContactDetector:
    """PlatformContactDetector(degrees, scale, position)
    (3d/contact/PlatformContactDetector.m:25-64): a box of size `scale`
    (width, depth, height) centred at `position`, rotated by the Euler angles
    `degrees` ('XYZ'); its axes are the rows of the rotation, normalised."""

    def __init__(self, degrees, scale, position, compliance=1e-4):
        R = eul2rotm_xyz(np.radians(np.asarray(degrees, dtype=np.float64)))
        self.axes = R / np.linalg.norm(R, axis=1, keepdims=True)
        self.halfExtent = np.asarray(scale, dtype=np.float64) / 2
        self.center = np.asarray(position, dtype=np.float64)
        self.compliance = float(compliance)

    def finder(self):
        f = L.ContactFinder()
        f.kind = L.MLX_CONTACT_BOX
        f.compliance = self.compliance
        for a in range(3):
            f.center[a] = self.center[a]
            f.halfExtent[a] = self.halfExtent[a]
        for k in range(9):
            f.axes[k] = self.axes.reshape(-1)[k]
        return f


class SdfContactFinder:
    """A collider given by its signed distance sampled on a regular grid:
    values[i, j, k] at origin + spacing * (i, j, k) (the samples stand in for
    gptoolbox's signed_distance query, 3d/contact/MeshSCD3D.m:139)."""

    def __init__(self, values, origin, spacing, compliance=1e-4):
        v = np.asarray(values, dtype=np.float64)
        self.dims = v.shape
        self.values = np.ascontiguousarray(v.ravel(order="F"))  # x fastest
        self.origin = np.asarray(origin, dtype=np.float64)
        self.spacing = float(spacing)
        self.compliance = float(compliance)

    def finder(self):
        f = L.ContactFinder()
        f.kind = L.MLX_CONTACT_SDF
        f.compliance = self.compliance
        f.sdf = self.values.ctypes.data_as(C.POINTER(C.c_double))
        for a in range(3):
            f.dims[a] = self.dims[a]
            f.origin[a] = self.origin[a]
        f.spacing = self.spacing
        return f


def finder_array(finders):
    arr = (L.ContactFinder * len(finders))()
    for i, f in enumerate(finders):
        arr[i] = f.finder()
    return arr


class StepResult:
    def __init__(self, out, residuals):
        self.residualArray = residuals
        self.runUntil = out.runUntilResSmallerThan
        self.itersPerLayer = list(out.itersPerLayer[: out.layersRun])
        self.layersRun = out.layersRun
        self.elasticProjections = int(out.elasticProjections)


class Simulation:
    """One mesh driven substep by substep on the GPU (the simulate3D /
    simulate frame loop restricted to the XPBD path)."""

    def __init__(self, mesh, integrator, h, contact=None, precision=L.MLX_FP64, device=0,
                 v0=None, residual_cap=4096, p0=None):
        self.mesh = mesh
        self.integrator = integrator
        self.h = float(h)
        self.contact = contact
        self.dmesh = DeviceMesh(mesh, precision, device)
        v0 = np.zeros_like(mesh.p) if v0 is None else L.f64(v0)
        p0 = mesh.p if p0 is None else L.f64(p0)
        if p0.shape != mesh.p.shape or v0.shape != mesh.p.shape:
            raise ValueError("p0 and v0 need dim * nV values")
        st = C.c_void_p()
        L.check("mlx_state_create",
                L.lib().mlx_state_create(self.dmesh.handle, L.ptr(p0), L.ptr(v0), C.byref(st)))
        self.state = st
        self.frame = 0
        self._settings = integrator.settings()
        # one plane: mlx_integrate; a list of finders (planes, platform
        # boxes, grid SDFs): mlx_integrate_finders
        self._finders = None
        self._plane = None
        if isinstance(contact, (list, tuple)):
            self._finders = finder_array(contact)
            self._finder_objs = list(contact)
        elif contact is not None:
            self._plane = contact.plane()
        self._resid = np.zeros(residual_cap)
        self._out = L.LayerSolveOut()
        self._out.residualArray = L.ptr(self._resid)
        self._out.residualCap = residual_cap

    def step(self):
        """One integrate() substep; returns a StepResult."""
        self.frame += 1
        if self._finders is not None:
            L.check("mlx_integrate_finders", L.lib().mlx_integrate_finders(
                self.state, C.byref(self._settings), self.h, self.frame, self._finders,
                len(self._finders), C.byref(self._out)))
        else:
            pl = C.byref(self._plane) if self._plane is not None else None
            L.check("mlx_integrate", L.lib().mlx_integrate(self.state, C.byref(self._settings),
                                                           self.h, self.frame, pl,
                                                           C.byref(self._out)))
        self.integrator.runUntilResSmallerThan = self._settings.runUntilResSmallerThan
        n = min(self._out.residualCount, len(self._resid))
        return StepResult(self._out, self._resid[:n].copy())

    def find_contacts(self, finders):
        """The finders' contacts at the current positions: (vertex 0-based,
        normals K x 3, points K x 3, compliance)."""
        arr = finder_array(finders)
        nV = self.mesh.nV
        K = np.zeros(1, np.int32)
        vert = np.zeros(nV, np.int32)
        nrm = np.zeros(3 * nV)
        pts = np.zeros(3 * nV)
        cc = np.zeros(nV)
        L.check("mlx_find_contacts", L.lib().mlx_find_contacts(
            self.state, arr, len(finders), L.ptr(K, C.c_int32), nV, L.ptr(vert, C.c_int32),
            L.ptr(nrm), L.ptr(pts), L.ptr(cc)))
        k = int(K[0])
        return vert[:k].copy(), nrm[:3 * k].reshape(k, 3), pts[:3 * k].reshape(k, 3), cc[:k].copy()

    def set_pinned_targets(self, pinnedx):
        """Prescribed positions of the pinned DOFs (pinnedDOFs order) for the
        next substeps (3D, frame >= 2): scripted pins such as a twisted end
        cap.  The reference freezes them at frame 1."""
        x = L.f64(pinnedx)
        if len(x) != len(self.mesh.pinnedDOFs):
            raise ValueError("set_pinned_targets: one value per pinned DOF")
        L.check("mlx_state_set_pinned_targets",
                L.lib().mlx_state_set_pinned_targets(self.state, L.ptr(x)))

    def snapshot(self):
        """(p, v, oldp, lambda nT x nc, frame, runUntil): everything the
        next substep reads from the state."""
        p, v, lam, _ = self.download()
        oldp = np.zeros_like(p)
        L.check("mlx_state_get_oldp", L.lib().mlx_state_get_oldp(self.state, L.ptr(oldp)))
        return dict(p=p, v=v, oldp=oldp, lam=lam, frame=self.frame,
                    runUntil=self._settings.runUntilResSmallerThan)

    def restore(self, snap):
        """Continue from a snapshot (of this or another Simulation of the same
        mesh): state, frame count, runUntil and, in 3D, the pinned targets
        (the pinned DOFs' values in the snapshot's positions)."""
        m = self.mesh
        lam = L.fortran(np.asarray(snap["lam"], dtype=np.float64), np.float64)
        L.check("mlx_state_set", L.lib().mlx_state_set(
            self.state, L.ptr(L.f64(snap["p"])), L.ptr(L.f64(snap["v"])),
            L.ptr(L.f64(snap["oldp"])), L.ptr(lam)))
        self.frame = int(snap["frame"])
        self._settings.runUntilResSmallerThan = float(snap["runUntil"])
        self.integrator.runUntilResSmallerThan = float(snap["runUntil"])
        if m.dim == 3 and len(m.pinnedDOFs) and self.frame >= 1:
            self.set_pinned_targets(L.f64(snap["p"])[np.asarray(m.pinnedDOFs) - 1])

    def download(self):
        m = self.mesh
        p = np.zeros(m.dim * m.nV)
        v = np.zeros(m.dim * m.nV)
        lam = np.zeros(m.nT * m.nc)
        k = np.zeros(1, dtype=np.int32)
        L.check("mlx_state_download", L.lib().mlx_state_download(
            self.state, L.ptr(p), L.ptr(v), L.ptr(lam), L.ptr(k, C.c_int32)))
        return p, v, lam.reshape(m.nc, m.nT).T.copy(), int(k[0])

    def close(self):
        if self.state:
            L.lib().mlx_state_destroy(self.state)
            self.state = None
        self.dmesh.close()


def layer_solve(dmesh, integrator, h, maps, xi, lambdai, oldp, residualv=None, pinnedx=None,
                contacts=None, residual_cap=4096):
    """Drop-in for xpbdLayer3D.layerSolve / xpbdLayer2D.layerSolve through
    mlx_layer_solve.  maps: dict with the nT x nSorted / nV x nSorted arrays
    isElasticElement, vertexRigidBody, isBoundaryVertex, isRigidVertex and
    numRigids (as findBoundaryVertices returns them).  contacts: dict with
    vertexID (1-based), normals (K x dim), points (K x dim), compliance (K).
    Returns (xi, lambdai, residualv, out, residuals, contactLambdas)."""
    m = dmesh.mesh
    xi = L.f64(xi).copy()
    lam = L.fortran(np.asarray(lambdai, dtype=np.float64), np.float64).copy()
    a = L.LayerSolveArgs()
    keep = []
    lm = L.i32(integrator.layersMap)
    it = L.i32(integrator.iterations)
    keep += [lm, it]
    a.nLayers = len(integrator.layers)
    a.layersMap, a.iterations = L.ptr(lm, C.c_int32), L.ptr(it, C.c_int32)
    ns = int(np.asarray(maps["numRigids"]).size)
    a.nSorted = ns
    el = L.fortran(np.asarray(maps["isElasticElement"]).reshape(m.nT, ns), np.uint8)
    vb = L.fortran(np.asarray(maps["vertexRigidBody"]).reshape(m.nV, ns), np.int32)
    bd = L.fortran(np.asarray(maps["isBoundaryVertex"]).reshape(m.nV, ns), np.uint8)
    rv_ = L.fortran(np.asarray(maps["isRigidVertex"]).reshape(m.nV, ns), np.uint8)
    nr = L.i32(maps["numRigids"])
    keep += [el, vb, bd, rv_, nr]
    a.isElasticElement = L.ptr(el, C.c_uint8)
    a.vertexRigidBody = L.ptr(vb, C.c_int32)
    a.isBoundaryVertex = L.ptr(bd, C.c_uint8)
    a.isRigidVertex = L.ptr(rv_, C.c_uint8)
    a.numRigids = L.ptr(nr, C.c_int32)
    a.h = float(h)
    a.boundaryCompliance = integrator.boundaryCompliance
    a.contactResistance = integrator.contactResistance
    a.computeResiduals = int(bool(integrator.computeResiduals))
    a.hangingStop = int(integrator.hangingStop)
    a.giveUpEnabled = int(bool(integrator.giveUpEnabled))
    a.giveUpThreshold = integrator.giveUpThreshold
    a.lastLayerGiveUpEnabled = int(bool(integrator.LastLayerGiveUpEnabled))
    a.useRigidContacts = int(integrator.rigid_contacts_enabled())
    a.stopAtImprovementPercent = integrator.StopAtImprovementPercent
    a.runUntilResSmallerThan = integrator.runUntilResSmallerThan
    a.residualArrayLast = 0.0
    oldp = L.f64(oldp)
    keep.append(oldp)
    a.xi, a.lambdai, a.oldp = L.ptr(xi), L.ptr(lam), L.ptr(oldp)
    rv = None
    if m.dim == 3:
        rv = L.f64(residualv).copy()
        px = L.f64(pinnedx if pinnedx is not None else np.zeros(0))
        keep.append(px)
        a.residualv, a.pinnedx = L.ptr(rv), L.ptr(px)
    K = 0
    if contacts is not None and len(contacts["vertexID"]) > 0:
        K = len(contacts["vertexID"])
        cv = L.i32(contacts["vertexID"])
        cn = L.fortran(np.asarray(contacts["normals"]).reshape(K, m.dim), np.float64)
        cp = L.fortran(np.asarray(contacts["points"]).reshape(K, m.dim), np.float64)
        cc = L.f64(contacts["compliance"])
        keep += [cv, cn, cp, cc]
        a.nContacts = K
        a.contactVertexID = L.ptr(cv, C.c_int32)
        a.contactNormals, a.pointColliders, a.contactCompliance = L.ptr(cn), L.ptr(cp), L.ptr(cc)
    out = L.LayerSolveOut()
    res = np.zeros(residual_cap)
    out.residualArray = L.ptr(res)
    out.residualCap = residual_cap
    cl = np.zeros(max(K, 1))
    out.contactLambdas = L.ptr(cl)
    L.check("mlx_layer_solve", L.lib().mlx_layer_solve(dmesh.handle, C.byref(a), C.byref(out)))
    lam_out = lam.reshape(m.nc, m.nT).T.copy()
    n = min(out.residualCount, residual_cap)
    return xi, lam_out, rv, out, res[:n].copy(), cl[:K].copy()
