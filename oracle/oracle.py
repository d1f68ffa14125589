"""ctypes binding of the CPU oracle (oracle/_build/libmlxoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.  The
product path (multilayersolver2024_amd) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmlxoracle.so")
MAX_LAYERS = 16

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_u8 = C.POINTER(C.c_uint8)


class OraMesh(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("nV", C.c_int), ("nT", C.c_int), ("vpe", C.c_int),
        ("T", _i32), ("DmInv", _d), ("alpha", _d), ("beta", _d), ("vol", _d), ("mass", _d),
        ("color", _i32), ("numColors", C.c_int), ("pinnedDOFs", _i32), ("nPinnedDOFs", C.c_int),
        ("forcedElastic", _u8), ("nhModel", _u8), ("nhMu", _d), ("nhLam", _d),
    ]


class OraSettings(C.Structure):
    _fields_ = [
        ("nLayers", C.c_int), ("layers", C.c_double * MAX_LAYERS),
        ("iterations", C.c_int * MAX_LAYERS), ("boundaryCompliance", C.c_double),
        ("contactResistance", C.c_double), ("gravity", C.c_double), ("airDrag", C.c_double),
        ("computeResiduals", C.c_int), ("runUntil", C.c_double), ("hangingStop", C.c_int),
        ("giveUpEnabled", C.c_int), ("giveUpThreshold", C.c_double),
        ("lastLayerGiveUpEnabled", C.c_int), ("stopAtImprovementPercent", C.c_double),
        ("useRigidContacts", C.c_int), ("restitution", C.c_double),
    ]


class OraContacts(C.Structure):
    _fields_ = [("K", C.c_int), ("vertex", _i32), ("normal", _d), ("point", _d),
                ("compliance", _d)]


class OraStepOut(C.Structure):
    _fields_ = [
        ("lambda_", _d), ("residualArray", _d), ("residualCap", C.c_int),
        ("residualCount", C.c_int), ("itersPerLayer", C.c_int * MAX_LAYERS),
        ("layersRun", C.c_int), ("elasticProjections", C.c_int64),
        ("setsPerLayer", _i32), ("numRigids", _i32),
        ("normresv", C.c_double * MAX_LAYERS), ("runUntilAtEntry", C.c_double),
        ("learnLam0", _d), ("learnKeyScale", _d), ("learnKept", C.c_int),
        ("learnRes", C.c_double * 2), ("givenSets", _i32),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.ora_residual3d.restype = C.c_double
        _lib.ora_residual2d.restype = C.c_double
        _lib.ora_learn_eq_residual3d.restype = C.c_double
    return _lib


def P(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


class OracleMesh:
    """ora_mesh for a product Mesh, with the oracle's OWN precompute: DmInv,
    volumes, masses, alpha/beta (ora_precompute3d/2d), colors
    (ora_naive_graph_coloring) and pinned DOFs / pinned elements are rebuilt
    here from the scene (rest positions, connectivity, materials, pinned
    vertices), never copied from the product.  The product's arrays are then
    checked against them bit for bit -- colors, pinned DOFs, pinned elements
    and the constants, after a 1e-12 relative check that names a gross
    mismatch (ProductPrecomputeMismatch otherwise) -- so every run through
    the oracle also checks the product's precompute and coloring."""

    def __init__(self, mesh, check=True):
        self.m = mesh
        dim, nT, nV = mesh.dim, mesh.nT, mesh.nV
        vpe, nc = dim + 1, (6 if dim == 3 else 3)
        sc = mesh.scene
        self.T = np.ascontiguousarray(np.asarray(sc["t"], dtype=np.int32).reshape(nT, vpe))
        P0 = np.ascontiguousarray(np.asarray(sc["p"], dtype=np.float64).reshape(-1))
        mat = [np.ascontiguousarray(np.broadcast_to(np.asarray(sc[k], dtype=np.float64), (nT,)))
               for k in ("rho", "mu", "lam", "alpha1")]
        self.DmInv = np.zeros(dim * dim * nT)
        self.vol = np.zeros(nT)
        self.mass = np.zeros(dim * nV)
        self.alpha = np.zeros(nc * nc * nT)
        self.beta = np.zeros(nc * nc * nT)
        fn = lib().ora_precompute3d if dim == 3 else lib().ora_precompute2d
        if fn(nV, nT, P(P0), P(self.T, C.c_int32), *(P(a) for a in mat), P(self.DmInv),
              P(self.vol), P(self.mass), P(self.alpha), P(self.beta)) != 0:
            raise RuntimeError("oracle precompute: singular element")
        self.color = np.zeros(nT, dtype=np.int32)
        if lib().ora_naive_graph_coloring(P(self.T, C.c_int32), nT, vpe, nV,
                                          P(self.color, C.c_int32)) != 0:
            raise RuntimeError("oracle coloring failed")
        self.numColors = int(self.color.max()) if nT else 0
        pv = np.asarray(mesh.pinnedInds, dtype=np.int64)
        if dim == 3:
            dofs = np.sort(np.concatenate([3 * pv, 3 * pv + 1, 3 * pv + 2]))
        else:
            dofs = np.concatenate([2 * pv, 2 * pv + 1])
        self.pinned = np.ascontiguousarray(dofs, dtype=np.int32)
        mark = np.zeros(nV, dtype=np.uint8)
        mark[pv] = 1
        self.forced = np.zeros(nT, dtype=np.uint8)
        lib().ora_pinned_elements(P(self.T, C.c_int32), nT, vpe, nV, P(mark, C.c_uint8),
                                  P(self.forced, C.c_uint8))
        if check:
            self.check_product(mesh)
        s = OraMesh()
        s.dim, s.nV, s.nT, s.vpe = dim, nV, nT, vpe
        s.T, s.DmInv, s.alpha, s.beta = P(self.T, C.c_int32), P(self.DmInv), P(self.alpha), P(self.beta)
        s.vol, s.mass = P(self.vol), P(self.mass)
        s.color, s.numColors = P(self.color, C.c_int32), self.numColors
        s.pinnedDOFs, s.nPinnedDOFs = P(self.pinned, C.c_int32), len(self.pinned)
        s.forcedElastic = P(self.forced, C.c_uint8)
        nh = getattr(mesh, "neohookean", None)
        if dim == 3 and nh is not None and nh.any():
            self.nh = np.ascontiguousarray(nh, dtype=np.uint8)
            self.nhMu = mat[1].copy()
            self.nhLam = mat[2].copy()
            s.nhModel = P(self.nh, C.c_uint8)
            s.nhMu, s.nhLam = P(self.nhMu), P(self.nhLam)
        self.s = s

    def check_product(self, mesh, tol=1e-12):
        """Raise ProductPrecomputeMismatch unless the product Mesh's
        precompute and coloring equal the oracle's."""
        def exact(name, a, b):
            if not np.array_equal(np.asarray(a), np.asarray(b)):
                raise ProductPrecomputeMismatch(f"{name} differs from the oracle's")

        def close(name, a, b):
            # entry by entry to tol relative, with a floor of 1e-15 of the
            # array's largest entry for entries that are zero up to rounding
            a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
            if a.shape != b.shape:
                raise ProductPrecomputeMismatch(f"{name} has shape {a.shape}, oracle {b.shape}")
            if not a.size:
                return
            scale = float(np.max(np.abs(b)))
            bound = tol * np.maximum(np.abs(a), np.abs(b)) + 1e-15 * scale
            if not np.all(np.abs(a - b) <= bound):
                raise ProductPrecomputeMismatch(f"{name} differs from the oracle's by more than {tol}")

        exact("elementColor", mesh.elementColor, self.color)
        exact("numColors", int(mesh.numColors), self.numColors)
        exact("pinnedDOFs", np.asarray(mesh.pinnedDOFs) - 1, self.pinned)
        exact("pinnedTets", np.asarray(mesh.forcedElastic, dtype=np.uint8), self.forced)
        exact("t", mesh.T0, self.T)
        # the constants follow the same elimination order on both sides, so
        # they agree bit for bit (the layer maps at rest depend on the last
        # bits of DmInv through the strain-rate keys)
        for name, a, b in (("DmInv", mesh.DmInv, self.DmInv), ("elV", mesh.elV, self.vol),
                           ("mass", mesh.mass, self.mass), ("alpha", mesh.alpha, self.alpha),
                           ("beta", mesh.beta, self.beta)):
            close(name, a, b)
            exact(name, a, b)


class ProductPrecomputeMismatch(AssertionError):
    pass


def settings_from(integ):
    s = OraSettings()
    s.nLayers = len(integ.layers)
    for i, (l, it) in enumerate(zip(integ.layers, integ.iterations)):
        s.layers[i] = l
        s.iterations[i] = int(it)
    s.boundaryCompliance = integ.boundaryCompliance
    s.contactResistance = integ.contactResistance
    s.gravity = integ.Gravity
    s.airDrag = getattr(integ, "airDrag", 0.0)
    s.computeResiduals = int(bool(integ.computeResiduals))
    s.runUntil = integ.runUntilResSmallerThan
    s.hangingStop = int(integ.hangingStop)
    s.giveUpEnabled = int(bool(integ.giveUpEnabled))
    s.giveUpThreshold = integ.giveUpThreshold
    s.lastLayerGiveUpEnabled = int(bool(integ.LastLayerGiveUpEnabled))
    s.stopAtImprovementPercent = integ.StopAtImprovementPercent
    s.useRigidContacts = int(integ.rigid_contacts_enabled())
    s.restitution = integ.restitution
    return s


def plane_contacts(dim, p, normal, position, compliance, strict):
    nV = len(p) // dim
    vert = np.zeros(nV, dtype=np.int32)
    cn = np.zeros(nV * dim)
    cp = np.zeros(nV * dim)
    cc = np.zeros(nV)
    n = np.asarray(normal, dtype=np.float64)[:dim]
    n = n / np.linalg.norm(n)
    pos = np.asarray(position, dtype=np.float64)[:dim]
    # D uses the normal as given (PlaneContactFinder3D.m:21); callers pass
    # unit normals, for which the two coincide.
    K = lib().ora_plane_contacts(dim, nV, P(np.ascontiguousarray(p)), P(n), P(pos),
                                 C.c_double(compliance), int(strict), P(vert, C.c_int32), P(cn),
                                 P(cp), P(cc))
    return dict(vertex=vert[:K].copy(), normal=cn[:K * dim].copy(), point=cp[:K * dim].copy(),
                compliance=cc[:K].copy(), K=K)


def box_contacts(p, center, axes, half, compliance):
    nV = len(p) // 3
    vert = np.zeros(nV, dtype=np.int32)
    cn, cp, cc = np.zeros(3 * nV), np.zeros(3 * nV), np.zeros(nV)
    K = lib().ora_box_contacts(nV, P(np.ascontiguousarray(p)), P(np.ascontiguousarray(center, dtype=np.float64)),
                               P(np.ascontiguousarray(axes, dtype=np.float64).reshape(-1)),
                               P(np.ascontiguousarray(half, dtype=np.float64)), C.c_double(compliance),
                               P(vert, C.c_int32), P(cn), P(cp), P(cc))
    return dict(vertex=vert[:K].copy(), normal=cn[:3 * K].copy(), point=cp[:3 * K].copy(),
                compliance=cc[:K].copy(), K=K)


def sdf_contacts(p, values_xfast, dims, origin, spacing, compliance):
    nV = len(p) // 3
    vert = np.zeros(nV, dtype=np.int32)
    cn, cp, cc = np.zeros(3 * nV), np.zeros(3 * nV), np.zeros(nV)
    d = np.ascontiguousarray(dims, dtype=np.int32)
    K = lib().ora_sdf_contacts(nV, P(np.ascontiguousarray(p)),
                               P(np.ascontiguousarray(values_xfast, dtype=np.float64)),
                               P(d, C.c_int32), P(np.ascontiguousarray(origin, dtype=np.float64)),
                               C.c_double(spacing), C.c_double(compliance), P(vert, C.c_int32),
                               P(cn), P(cp), P(cc))
    return dict(vertex=vert[:K].copy(), normal=cn[:3 * K].copy(), point=cp[:3 * K].copy(),
                compliance=cc[:K].copy(), K=K)


def finder_contacts(p, finders):
    """The contacts of a list of 3D finders (plane, platform box, grid SDF;
    the package's finder objects), concatenated in finder order."""
    parts = []
    for f in finders:
        kind = type(f).__name__
        if kind == "PlaneContactFinder":
            parts.append(plane_contacts(3, p, f.normal, f.position, f.compliance, strict=False))
        elif kind == "PlatformContactDetector":
            parts.append(box_contacts(p, f.center, f.axes, f.halfExtent, f.compliance))
        elif kind == "SdfContactFinder":
            parts.append(sdf_contacts(p, f.values, f.dims, f.origin, f.spacing, f.compliance))
        else:
            raise TypeError(f"unknown contact finder {kind}")
    out = dict(vertex=np.concatenate([c["vertex"] for c in parts]).astype(np.int32),
               normal=np.concatenate([c["normal"] for c in parts]),
               point=np.concatenate([c["point"] for c in parts]),
               compliance=np.concatenate([c["compliance"] for c in parts]))
    out["K"] = len(out["vertex"])
    return out


class OracleSim:
    """The oracle's integrate() loop for one mesh (simulate3D / simulate step
    order, contacts at the start of each step)."""

    def __init__(self, mesh, integ, h, contact=None, v0=None, residual_cap=4096, p0=None):
        self.om = OracleMesh(mesh)
        self.integ = integ
        self.h = h
        self.contact = contact
        self.dim = mesh.dim
        self.p = mesh.p.copy() if p0 is None else np.array(p0, dtype=np.float64)
        self.v = np.zeros_like(self.p) if v0 is None else np.array(v0, dtype=np.float64)
        self.oldp = self.p.copy()       # cache.oldp (3D) / prevPrevp source (2D)
        self.oldOldp = self.p.copy()
        self.pinnedx = np.zeros(len(mesh.pinnedDOFs))
        self.frame = 0
        self.s = settings_from(integ)
        self.resid = np.zeros(residual_cap)
        self.lam = np.zeros(mesh.nT * mesh.nc)
        self.last_contacts = 0

    def _contacts(self):
        if self.contact is None:
            return None, None
        if isinstance(self.contact, (list, tuple)):
            c = finder_contacts(self.p, self.contact)
        else:
            c = plane_contacts(self.dim, self.p, self.contact.normal, self.contact.position,
                               self.contact.compliance, strict=(self.dim == 2))
        self.last_contacts = c["K"]
        keep = c
        oc = OraContacts()
        oc.K = c["K"]
        nrm = np.zeros(c["K"] * 3)
        pts = np.zeros(c["K"] * 3)
        if self.dim == 3:
            nrm, pts = c["normal"], c["point"]
        else:
            nrm, pts = c["normal"], c["point"]
        keep["_n"], keep["_p"] = np.ascontiguousarray(nrm), np.ascontiguousarray(pts)
        oc.vertex = P(c["vertex"], C.c_int32)
        oc.normal, oc.point = P(keep["_n"]), P(keep["_p"])
        oc.compliance = P(c["compliance"])
        return oc, keep

    def step(self, lam0=None, key_scale=None, sets=None):
        """One substep; lam0 (nT x 6) and key_scale (nT) apply the learned
        warm start and key factors as the product's learned mode does; sets
        (nT x nSorted body ids, 0 elastic) replaces the strain-rate layers."""
        self.frame += 1
        oc, keep = self._contacts()
        out = OraStepOut()
        gs = None if sets is None else np.ascontiguousarray(np.asarray(sets, dtype=np.int32).T)
        out.givenSets = P(gs, C.c_int32)
        l0 = None if lam0 is None else np.ascontiguousarray(lam0, dtype=np.float64).reshape(-1)
        ks = None if key_scale is None else np.ascontiguousarray(key_scale, dtype=np.float64)
        out.learnLam0 = P(l0)
        out.learnKeyScale = P(ks)
        out.lambda_ = P(self.lam)
        out.residualArray = P(self.resid)
        out.residualCap = len(self.resid)
        ns = len(np.unique(np.asarray(self.integ.layers, dtype=np.float64)))
        sets = np.zeros(self.om.m.nT * ns, dtype=np.int32)
        nrig = np.zeros(ns, dtype=np.int32)
        out.setsPerLayer = P(sets, C.c_int32)
        out.numRigids = P(nrig, C.c_int32)
        L = lib()
        if self.dim == 3:
            r = L.ora_integrate3d(C.byref(self.om.s), C.byref(self.s), C.c_double(self.h),
                                  self.frame, P(self.p), P(self.v), P(self.oldp), P(self.oldOldp),
                                  P(self.pinnedx), C.byref(oc) if oc else None, C.byref(out))
        else:
            prevPrevp = self.oldp.copy()
            prevp = self.p.copy()
            r = L.ora_integrate2d(C.byref(self.om.s), C.byref(self.s), C.c_double(self.h),
                                  P(self.p), P(self.v), P(prevp), P(prevPrevp),
                                  C.byref(oc) if oc else None, C.byref(out))
            self.oldp = prevp
        if r != 0:
            raise RuntimeError(f"oracle integrate returned {r}")
        self.integ_runUntil = self.s.runUntil
        n = min(out.residualCount, len(self.resid))
        return dict(residuals=self.resid[:n].copy(), iters=list(out.itersPerLayer[:out.layersRun]),
                    projections=int(out.elasticProjections), runUntil=self.s.runUntil,
                    sets=sets.reshape(ns, self.om.m.nT).T.copy(), numRigids=nrig,
                    normresv=list(out.normresv[:out.layersRun]),
                    runUntilAtEntry=float(out.runUntilAtEntry), residualCount=int(out.residualCount),
                    learnKept=int(out.learnKept), learnRes=list(out.learnRes))

    def lambdas(self):
        m = self.om.m
        return self.lam.reshape(m.nT, m.nc).copy()
