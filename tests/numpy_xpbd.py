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

    def sweep(self, x, xo, lam):
        d, h, pairs = self.d, self.h, (PAIRS3 if self.d == 3 else PAIRS2)
        aT = self.A / h ** 2
        bT = self.B * h ** 2
        gam = aT @ bT / h
        color = self.m.elementColor
        for c in range(1, self.m.numColors + 1):
            es = np.nonzero(color == c)[0]
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
