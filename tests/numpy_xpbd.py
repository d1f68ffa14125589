"""A second, independent restatement of the plain (all-elastic) XPBD
substep in vectorised numpy, used only to cross-check the C oracle end to
end.  It is written from the mathematics of the method rather than from the
reference's expression trees: F = Ds DmInv, constraint gradient
vol * F P_i g_a, dense solves with numpy.linalg, colors processed as whole
pages like the reference's vectorised Matlab."""
import numpy as np

PAIRS3 = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
PAIRS2 = [(0, 0), (1, 1), (0, 1)]


def _sel(j, k, d):
    P = np.zeros((d, d))
    P[j, k] += 0.5
    P[k, j] += 0.5
    return P


def _grads(F, G, vol, pairs, d):
    """gradC[e, i, a, :] = vol_e * F_e P_i g_a."""
    out = np.zeros((F.shape[0], len(pairs), G.shape[1], d))
    for i, (j, k) in enumerate(pairs):
        Pm = _sel(j, k, d)
        out[:, i] = vol[:, None, None] * np.einsum("erc,cs,eas->ear", F, Pm, G)
    return out


class NumpyXpbd:
    def __init__(self, mesh, gravity, h, iterations, contact=None, restitution=None,
                 contact_resistance=None):
        self.m = mesh
        self.d = mesh.dim
        self.h = h
        self.g = gravity
        self.iters = iterations
        self.contact = contact
        self.res = contact_resistance
        self.e = restitution
        d, nT = self.d, mesh.nT
        self.T = mesh.T0
        self.DmInv = mesh.DmInv.reshape(nT, d, d).transpose(0, 2, 1)  # page e -> DmInv(e)
        nc = mesh.nc
        self.A = mesh.alpha.reshape(nT, nc, nc).transpose(0, 2, 1)
        self.B = mesh.beta.reshape(nT, nc, nc).transpose(0, 2, 1)
        if d == 3:
            self.G = np.concatenate([-self.DmInv.sum(axis=1, keepdims=True), self.DmInv], axis=1)
        else:
            self.G = np.concatenate([self.DmInv, -self.DmInv.sum(axis=1, keepdims=True)], axis=1)
        self.mass = mesh.mass.reshape(-1, d)
        self.p = mesh.p.reshape(-1, d).copy()
        self.v = np.zeros_like(self.p)
        self.pinned = np.asarray(mesh.pinnedDOFs) - 1
        self.pinnedx = self.p.reshape(-1)[self.pinned].copy()

    def F(self, x):
        X = x[self.T]
        if self.d == 3:
            Ds = (X[:, 1:] - X[:, :1]).transpose(0, 2, 1)
        else:
            Ds = (X[:, :2] - X[:, 2:]).transpose(0, 2, 1)
        return Ds @ self.DmInv

    def sweep(self, x, xo, lam, elastic=None):
        d, h, pairs = self.d, self.h, (PAIRS3 if self.d == 3 else PAIRS2)
        aT = self.A / h ** 2
        bT = self.B * h ** 2
        gam = aT @ bT / h
        color = self.m.elementColor
        for c in range(1, self.m.numColors + 1):
            es = np.nonzero((color == c) & (True if elastic is None else elastic))[0]
            if len(es) == 0:
                continue
            F = self.F(x)[es]
            E = 0.5 * (np.transpose(F, (0, 2, 1)) @ F - np.eye(d))
            vol = self.m.elV[es]
            Cv = vol[:, None] * np.stack([E[:, j, k] for j, k in pairs], axis=1)
            grad = _grads(F, self.G[es], vol, pairs, d).reshape(len(es), len(pairs), -1)
            minv = (1.0 / self.mass[self.T[es]]).reshape(len(es), -1)
            gM = grad * minv[:, None, :]
            dx = (x[self.T[es]] - xo[self.T[es]]).reshape(len(es), -1)
            ng = np.einsum("eij,ejk,ek->ei", gam[es], grad, dx)
            rhs = -Cv - np.einsum("eij,ej->ei", aT[es], lam[es]) - ng
            S = np.einsum("eik,ejk->eij", gM, grad)
            Dg = np.eye(len(pairs)) + gam[es]
            if d == 3:
                top = np.linalg.solve(Dg[:, :3, :3] @ S[:, :3, :3] + aT[es][:, :3, :3], rhs[:, :3, None])[..., 0]
                # shear rows: scalar solves with alpha~(4,4) (decomposed system)
                sh = (-Cv[:, 3:] - aT[es][:, 3, 3, None] * lam[es][:, 3:] - ng[:, 3:]) / \
                     (Dg[:, 3, 3, None] * np.stack([S[:, i, i] for i in (3, 4, 5)], 1) + aT[es][:, 3, 3, None])
                dl = np.concatenate([top, sh], axis=1)
            else:
                dl = np.linalg.solve(Dg @ S + aT[es], rhs[..., None])[..., 0]
            lam[es] += dl
            step = np.einsum("eik,ei->ek", gM, dl).reshape(len(es), -1, d)
            x[self.T[es]] = x[self.T[es]] + step
        return x, lam

    def contacts(self, x, cl, c):
        if c is None:
            return x
        h = self.h
        for i, v in enumerate(c["vertex"]):
            n = c["normal"][i]
            phi = min(n @ (x[v] - c["point"][i]), 0.0)
            C = self.res * phi
            aT = c["compliance"][i] / h ** 2
            minv = 1.0 / self.mass[v]
            if self.d == 3:
                dl = (-C - aT * cl[i]) / (2 * minv[0] + aT)
                x[v] = x[v] + minv * dl * n
            else:
                gm = n * 2 * minv
                dl = -(C + aT * cl[i]) / (gm @ n + aT)
                x[v] = x[v] + gm * dl
            cl[i] += dl
        return x

    def find_contacts(self):
        if self.contact is None:
            return None
        n = np.asarray(self.contact.normal[:self.d], dtype=np.float64)
        n = n / np.linalg.norm(n)
        D = n @ np.asarray(self.contact.position[:self.d])
        dist = self.p @ n - D
        idx = np.nonzero(dist <= 0 if self.d == 3 else dist < 0)[0]
        return dict(vertex=idx, normal=np.tile(n, (len(idx), 1)),
                    point=self.p[idx] - np.outer(dist[idx], n),
                    compliance=np.full(len(idx), self.contact.compliance))

    def step(self):
        d, h = self.d, self.h
        c = self.find_contacts()
        f = np.zeros_like(self.p)
        f[:, d - 1] = self.mass[:, d - 1] * self.g
        unp = np.ones(self.p.size, bool)
        unp[self.pinned] = False
        unp = unp.reshape(-1, d)
        x = self.p.copy()
        xo = self.p.copy()
        v = self.v.copy()
        if d == 3:
            rv = v + np.where(unp, h * f / self.mass, 0.0)
            x = x + h * rv  # all vertices elastic: release
        else:
            v = v + np.where(unp, h * f / self.mass, 0.0)
            x = x + np.where(unp, h * v, 0.0)
        lam = np.zeros((self.m.nT, self.m.nc))
        cl = np.zeros(len(c["vertex"])) if c is not None else None
        for it in range(self.iters):
            x, lam = self.sweep(x, xo, lam)
            x = self.contacts(x, cl, c)
            xf = x.reshape(-1)
            xf[self.pinned] = self.pinnedx if d == 3 else xo.reshape(-1)[self.pinned]
            x = xf.reshape(-1, d)
        self.v = (x - xo) / h
        if c is not None:
            e = self.e
            for i, vid in enumerate(c["vertex"]):
                vn = self.v[vid].copy()
                n = c["normal"][i]
                self.v[vid] = vn + n * (-vn + min(-e * (n @ vn), 0.0))
        self.p = x
        self.lam = lam
        return x


def _cross2(a, b):
    return a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]


class NumpyXpbd2DLayers(NumpyXpbd):
    """2D multi-layer substep with rigid layers, vectorised: per layer the
    bodies are the given element sets; each body is a rigid pose (COM, angle)
    with mass sum(m), inertia sum(m |r|^2) built from the current positions;
    per iteration the elastic elements are swept, the boundary vertices pull
    their bodies (mass split over the body's boundary vertices, half the
    angular impulse), contacts act on elastic vertices directly and on rigid
    vertices through their body, pins snap; at the end of the layer every
    rigid vertex takes its body's pose.  The layer sets come from outside
    (they are checked bit-exactly elsewhere); no residual monitor."""

    def __init__(self, mesh, gravity, h, iterations, layers_map, contact=None, restitution=1e-5,
                 contact_resistance=1.0, boundary_compliance=1e-8, use_rigid_contacts=True):
        super().__init__(mesh, gravity, h, 0, contact=contact, restitution=restitution,
                         contact_resistance=contact_resistance)
        self.layer_iters = list(iterations)
        self.layers_map = [int(k) - 1 for k in layers_map]
        self.alpha_b = boundary_compliance
        self.use_rigid_contacts = use_rigid_contacts

    @staticmethod
    def _rot(ang):
        c, s = np.cos(ang), np.sin(ang)
        return np.stack([np.stack([c, -s], -1), np.stack([s, c], -1)], -2)

    def _pose(self, verts, body, COM, ang, ri):
        return np.einsum("kij,kj->ki", self._rot(ang[body]), ri[verts]) + COM[body]

    def _apply(self, body, p, r, M, I, COM, ang, nb):
        for d in range(2):
            COM[:, d] += np.bincount(body, weights=p[:, d] / M[body], minlength=nb)
        ang += np.bincount(body, weights=0.5 / I[body] * _cross2(r, p), minlength=nb)

    def step_layers(self, sets):
        """One substep with per-layer body ids sets[:, col] (nT x nSorted)."""
        h, m = self.h, self.mass[:, 0]
        c = self.find_contacts()
        unp = np.ones(self.p.size, bool)
        unp[self.pinned] = False
        unp = unp.reshape(-1, 2)
        f = np.zeros_like(self.p)
        f[:, 1] = self.mass[:, 1] * self.g
        xo = self.p.copy()
        v = self.v + np.where(unp, h * f / self.mass, 0.0)
        x = self.p + np.where(unp, h * v, 0.0)
        lam = np.zeros((self.m.nT, 3))
        for layer, col in enumerate(self.layers_map):
            bid = sets[:, col]
            rigid_el = bid > 0
            nb = int(bid.max()) if rigid_el.any() else 0
            vb = np.zeros(self.m.nV, np.int64)
            vb[self.T[rigid_el].reshape(-1)] = np.repeat(bid[rigid_el], 3)
            on_elastic = np.zeros(self.m.nV, bool)
            on_elastic[self.T[~rigid_el].reshape(-1)] = True
            bverts = np.nonzero((vb > 0) & on_elastic)[0]
            rig = np.nonzero(vb > 0)[0]
            kb = vb[rig] - 1
            M = np.bincount(kb, weights=m[rig], minlength=nb)
            COM = np.stack([np.bincount(kb, weights=x[rig, d] * m[rig], minlength=nb)
                            for d in range(2)], 1) / np.maximum(M, 1e-300)[:, None]
            ri = np.zeros_like(x)
            ri[rig] = x[rig] - COM[kb]
            I = np.bincount(kb, weights=m[rig] * (ri[rig] ** 2).sum(1), minlength=nb)
            ang = np.zeros(nb)
            lamB = np.zeros(len(bverts))
            cl = np.zeros(len(c["vertex"])) if c is not None else None
            for it in range(self.layer_iters[layer]):
                x, lam = self.sweep(x, xo, lam, elastic=~rigid_el)
                if len(bverts):
                    k = vb[bverts] - 1
                    xr = self._pose(bverts, k, COM, ang, ri)
                    r = xr - COM[k]
                    dvec = xr - x[bverts]
                    C = np.sqrt((dvec ** 2).sum(1))
                    with np.errstate(invalid="ignore", divide="ignore"):
                        n = dvec / C[:, None]
                    n[C == 0] = 0
                    nbc = np.bincount(k, minlength=nb)[k]
                    w = (1.0 / M[k]) / nbc + _cross2(r, n) ** 2 / I[k]
                    aT = self.alpha_b / h ** 2
                    dl = (-C - aT * lamB) / (1.0 / m[bverts] + w + aT)
                    lamB += dl
                    self._apply(k, dl[:, None] * n, r, M, I, COM, ang, nb)
                    x[bverts] = self._pose(bverts, k, COM, ang, ri)
                if c is not None and len(c["vertex"]):
                    cv = c["vertex"]
                    el = vb[cv] == 0
                    for i in np.nonzero(el)[0]:
                        vv, nrm = cv[i], c["normal"][i]
                        phi = min(nrm @ (x[vv] - c["point"][i]), 0.0)
                        aT = c["compliance"][i] / h ** 2
                        gm = nrm * 2 / self.mass[vv]
                        dl = -(self.res * phi + aT * cl[i]) / (gm @ nrm + aT)
                        x[vv] = x[vv] + gm * dl
                        cl[i] += dl
                    ri_idx = np.nonzero(~el)[0]
                    if len(ri_idx) and self.use_rigid_contacts:
                        verts = cv[ri_idx]
                        k = vb[verts] - 1
                        xr = self._pose(verts, k, COM, ang, ri)
                        r = xr - COM[k]
                        nrm = c["normal"][ri_idx]
                        phi = np.minimum(((x[verts] - c["point"][ri_idx]) * nrm).sum(1), 0.0)
                        nbc = np.bincount(k, minlength=nb)[k]
                        w = (1.0 / M[k]) / nbc + _cross2(r, nrm) ** 2 / I[k]
                        aT = c["compliance"][ri_idx] / h ** 2
                        dl = (-self.res * phi - aT * cl[ri_idx]) / (2 * w + aT)
                        cl[ri_idx] += dl
                        self._apply(k, dl[:, None] * nrm, r, M, I, COM, ang, nb)
                        x[verts] = self._pose(verts, k, COM, ang, ri)
                xf = x.reshape(-1)
                xf[self.pinned] = xo.reshape(-1)[self.pinned]
                x = xf.reshape(-1, 2)
            if len(rig):
                x[rig] = self._pose(rig, kb, COM, ang, ri)
        self.v = (x - xo) / h
        if c is not None:
            e = self.e
            for i, vid in enumerate(c["vertex"]):
                vn = self.v[vid].copy()
                if np.linalg.norm(vn) <= 2 * self.g * h:
                    ee = 0.0
                else:
                    ee = e
                n = c["normal"][i]
                self.v[vid] = vn + n * (-vn + min(-ee * (n @ vn), 0.0))
        self.p = x
        self.lam = lam
        return x


def _hat(w):
    """Cross-product matrices of the rows of w (n x 3)."""
    z = np.zeros(len(w))
    return np.stack([np.stack([z, -w[:, 2], w[:, 1]], -1),
                     np.stack([w[:, 2], z, -w[:, 0]], -1),
                     np.stack([-w[:, 1], w[:, 0], z], -1)], -2)


def _rodrigues(omega, h):
    """Rotation by angle h |omega| about omega / |omega| (identity for
    omega = 0), by scipy's rotation-vector exponential: an independent route
    to the same matrix."""
    from scipy.spatial.transform import Rotation
    return Rotation.from_rotvec(h * omega).as_matrix()


class NumpyXpbd3DLayers(NumpyXpbd):
    """3D multi-layer substep with rigid layers, vectorised, written from the
    rigid-body mechanics rather than from the reference's code: per layer the
    bodies are the given element sets; a body's mass, centre of mass, linear
    and angular momentum and inertia tensor m(|r|^2 I - r r^T) are taken from
    the current positions and the residual velocity; elastic vertices are
    released (x += h v_res); bodies with rigid elements take one symplectic
    step (c += h v_c, R = exp(h omega)), the residual velocity of their
    vertices is rotated into the stepped frame less the stepped displacement
    over h; per iteration the elastic elements are swept, the boundary
    vertices pull their bodies (XPBD distance constraint with the body's
    generalised inverse mass, impulses summed per body, compliance used as
    given), contacts act on elastic vertices directly and on rigid vertices
    through their body (twice the body's generalised inverse mass and twice
    the compliance), pins snap; after the layer every rigid vertex takes its
    body's pose and the residual velocity is turned by the pose change made
    during the projection.  Fixed iteration counts (no residual monitor);
    the layer sets come from outside (checked bit-exactly elsewhere)."""

    def __init__(self, mesh, gravity, h, iterations, layers_map, contact=None, restitution=1e-2,
                 contact_resistance=0.5, boundary_compliance=1e-8, use_rigid_contacts=True):
        super().__init__(mesh, gravity, h, 0, contact=contact, restitution=restitution,
                         contact_resistance=contact_resistance)
        self.layer_iters = list(iterations)
        self.layers_map = [int(k) - 1 for k in layers_map]
        self.alpha_b = boundary_compliance
        self.use_rigid_contacts = use_rigid_contacts
        self.max_turn = 0.0  # largest |dR - I| of a projection's pose change so far

    @staticmethod
    def _pose(R, COM, ri, k):
        return np.einsum("kij,kj->ki", R[k], ri) + COM[k]

    def _apply(self, k, p, r, M, R, COM, I0, nb):
        """Sum the impulses p (at arms r) per body, move the centres of mass
        and turn the bodies by exp(h I^-1 sum(r x p) / h)."""
        h = self.h
        for d in range(3):
            COM[:, d] += np.bincount(k, weights=p[:, d] / M[k], minlength=nb)
        tq = np.cross(r, p) / h
        acc = np.stack([np.bincount(k, weights=tq[:, d], minlength=nb) for d in range(3)], 1)
        I = R @ I0 @ np.transpose(R, (0, 2, 1))
        om = np.linalg.solve(I, acc[..., None])[..., 0]
        for b in range(nb):
            R[b] = _rodrigues(om[b], h) @ R[b]

    @staticmethod
    def _w(M, I, r, n):
        """Generalised inverse mass 1/M + (r x n) . I^-1 (r x n)."""
        a = np.cross(r, n)
        return 1.0 / M + np.einsum("ki,ki->k", a, np.linalg.solve(I, a[..., None])[..., 0])

    def step_layers(self, sets):
        """One substep with per-layer body ids sets[:, col] (nT x nSorted)."""
        h, m = self.h, self.mass[:, 0]
        c = self.find_contacts()
        unp = np.ones(self.p.size, bool)
        unp[self.pinned] = False
        unp = unp.reshape(-1, 3)
        f = np.zeros_like(self.p)
        f[:, 2] = self.mass[:, 2] * self.g
        xo = self.p.copy()
        x = self.p.copy()
        rv = self.v + np.where(unp, h * f / self.mass, 0.0)
        lam = np.zeros((self.m.nT, 6))
        for layer, col in enumerate(self.layers_map):
            bid = sets[:, col]
            rigid_el = bid > 0
            nb = int(bid.max()) if rigid_el.any() else 0
            vb = np.zeros(self.m.nV, np.int64)
            vb[self.T[rigid_el].reshape(-1)] = np.repeat(bid[rigid_el], 4)
            on_elastic = np.zeros(self.m.nV, bool)
            on_elastic[self.T[~rigid_el].reshape(-1)] = True
            bverts = np.nonzero((vb > 0) & on_elastic)[0]
            rig = np.nonzero(vb > 0)[0]
            kb = vb[rig] - 1
            xprev = x.copy()
            # restriction from the positions and residual velocity at the layer's start
            M = np.bincount(kb, weights=m[rig], minlength=nb)
            COM = np.stack([np.bincount(kb, weights=m[rig] * x[rig, d], minlength=nb)
                            for d in range(3)], 1) / np.maximum(M, 1e-300)[:, None]
            Vc = np.stack([np.bincount(kb, weights=m[rig] * rv[rig, d], minlength=nb)
                           for d in range(3)], 1) / np.maximum(M, 1e-300)[:, None]
            ri = np.zeros_like(x)
            ri[rig] = x[rig] - COM[kb]
            rr = ri[rig]
            outer = m[rig, None, None] * ((rr ** 2).sum(1)[:, None, None] * np.eye(3)
                                          - rr[:, :, None] * rr[:, None, :])
            I0 = np.stack([np.stack([np.stack([np.bincount(kb, weights=outer[:, i, j], minlength=nb)
                                               for j in range(3)], -1) for i in range(3)], -2)])[0] \
                if nb else np.zeros((0, 3, 3))
            Lm = np.cross(rr, (rv[rig] - Vc[kb])) * m[rig, None]
            Lb = np.stack([np.bincount(kb, weights=Lm[:, d], minlength=nb) for d in range(3)], 1)
            omega = np.linalg.solve(I0, Lb[..., None])[..., 0] if nb else np.zeros((0, 3))
            R = np.tile(np.eye(3), (nb, 1, 1))
            # release the elastic vertices
            el = vb == 0
            x[el] = x[el] + h * rv[el]
            rv[el] = 0.0
            Rstep = R.copy()
            if rigid_el.any():
                COM = COM + h * Vc
                Rstep = np.stack([_rodrigues(omega[b], h) for b in range(nb)])
                R = Rstep.copy()
                x[rig] = self._pose(R, COM, ri[rig], kb)
                disp = (x[rig] - xprev[rig]) / h
                rv[rig] = np.einsum("kij,kj->ki", R[kb], rv[rig]) - disp
            lamB = np.zeros(len(bverts))
            cl = np.zeros(len(c["vertex"])) if c is not None else None
            for it in range(self.layer_iters[layer]):
                x, lam = self.sweep(x, xo, lam, elastic=~rigid_el)
                if len(bverts):
                    k = vb[bverts] - 1
                    I = R @ I0 @ np.transpose(R, (0, 2, 1))
                    xr = self._pose(R, COM, ri[bverts], k)
                    r = xr - COM[k]
                    dvec = xr - x[bverts]
                    C = np.sqrt((dvec ** 2).sum(1))
                    with np.errstate(invalid="ignore", divide="ignore"):
                        n = dvec / C[:, None]
                    n[~np.isfinite(n)] = 0
                    w = self._w(M[k], I[k], r, n)
                    aT = self.alpha_b
                    dl = (-C - aT * lamB) / (1.0 / m[bverts] + w + aT)
                    lamB += dl
                    self._apply(k, dl[:, None] * n, r, M, R, COM, I0, nb)
                    x[bverts] = self._pose(R, COM, ri[bverts], k)
                if c is not None and len(c["vertex"]):
                    cv = c["vertex"]
                    elc = vb[cv] == 0
                    xin = x.copy()
                    for i in np.nonzero(elc)[0]:
                        vv, nrm = cv[i], c["normal"][i]
                        phi = min(nrm @ (xin[vv] - c["point"][i]), 0.0)
                        aT = c["compliance"][i] / h ** 2
                        dl = (-self.res * phi - aT * cl[i]) / (2.0 / m[vv] + aT)
                        x[vv] = xin[vv] + dl * nrm / self.mass[vv]
                        cl[i] += dl
                    ri_idx = np.nonzero(~elc)[0]
                    if len(ri_idx) and self.use_rigid_contacts:
                        verts = cv[ri_idx]
                        k = vb[verts] - 1
                        I = R @ I0 @ np.transpose(R, (0, 2, 1))
                        r = self._pose(R, COM, ri[verts], k) - COM[k]
                        nrm = c["normal"][ri_idx]
                        phi = np.minimum(((x[verts] - c["point"][ri_idx]) * nrm).sum(1), 0.0)
                        aT = 2 * c["compliance"][ri_idx] / h ** 2
                        dl = (-self.res * phi - aT * cl[ri_idx]) / (2 * self._w(M[k], I[k], r, nrm) + aT)
                        cl[ri_idx] += dl
                        self._apply(k, dl[:, None] * nrm, r, M, R, COM, I0, nb)
                        x[verts] = self._pose(R, COM, ri[verts], k)
                xf = x.reshape(-1)
                xf[self.pinned] = self.pinnedx
                x = xf.reshape(-1, 3)
            if len(rig):
                x[rig] = self._pose(R, COM, ri[rig], kb)
            if rigid_el.any():
                dR = R @ np.transpose(Rstep, (0, 2, 1))
                if nb:
                    self.max_turn = max(self.max_turn, float(np.abs(dR - np.eye(3)).max()))
                rv[rig] = np.einsum("kij,kj->ki", dR[kb], rv[rig])
        self.v = (x - xo) / h
        if c is not None:
            e = self.e
            for i, vid in enumerate(c["vertex"]):
                vn = self.v[vid].copy()
                n = c["normal"][i]
                self.v[vid] = vn + n * (-vn + min(-e * (n @ vn), 0.0))
        self.p = x
        self.lam = lam
        self.rv = rv
        return x
