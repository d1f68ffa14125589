"""Scene builders for the benchmark and test workloads.

* ``cantilever2d``: a quad-grid cantilever of linear triangles (each quad
  split along its (i,j)-(i+1,j+1) diagonal), left column pinned.
* ``beam3d``: a box beam on a hex grid with 5 tetrahedra per cube in
  alternating parity (so that neighbouring cubes share face diagonals), one
  end (x = 0) pinned, z up.  The reference's own generator
  (3d/tools/meshReaders/generateBeam.m) uses 6 tets per cube; the 5-tet split
  is this package's.
* ``Pcg32``: O'Neill's PCG-XSH-RR 32-bit generator, used for every seeded
  material draw so that scenes are reproducible across languages.
"""
import math

import numpy as np


class Pcg32:
    """PCG-XSH-RR (O'Neill 2014): 64-bit LCG state, 32-bit output."""

    MULT = 6364136223846793005
    MASK = (1 << 64) - 1

    def __init__(self, seed=0, stream=54):
        self.inc = ((stream << 1) | 1) & self.MASK
        self.state = 0
        self.next_u32()
        self.state = (self.state + seed) & self.MASK
        self.next_u32()

    def next_u32(self):
        old = self.state
        self.state = (old * self.MULT + self.inc) & self.MASK
        xorshifted = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xorshifted >> rot) | (xorshifted << ((-rot) & 31))) & 0xFFFFFFFF

    def next_double(self):
        """Uniform in [0, 1) with 53 random bits."""
        a = self.next_u32() >> 5
        b = self.next_u32() >> 6
        return (a * 67108864.0 + b) / 9007199254740992.0

    def uniform(self, n):
        return np.array([self.next_double() for _ in range(n)], dtype=np.float64)


def to_lame(nu, E):
    """toLame.m:4-5."""
    mu = E / (2 * (1 + nu))
    lam = (E * nu) / ((1 + nu) * (1 - 2 * nu))
    return mu, lam


def cantilever2d(nx=40, ny=10, cell=0.1, E=1e5, nu=0.35, rho=1000.0, alpha1=0.02):
    """2-D cantilever: (nx+1) x (ny+1) vertices, vertex (i, j) at index
    i*(ny+1)+j, x = i*cell, y = j*cell; 2*nx*ny triangles; the x = 0 column
    is pinned.  Returns a dict with p (nV x 2), t (nT x 3, 0-based), the
    per-element material arrays and the pinned vertex ids."""
    vid = lambda i, j: i * (ny + 1) + j
    p = np.array([[i * cell, j * cell] for i in range(nx + 1) for j in range(ny + 1)],
                 dtype=np.float64)
    t = []
    for i in range(nx):
        for j in range(ny):
            a, b, c, d = vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)
            t.append([a, b, c])
            t.append([a, c, d])
    t = np.array(t, dtype=np.int32)
    nT = len(t)
    mu, lam = to_lame(nu, E)
    return dict(
        dim=2, p=p, t=t,
        rho=np.full(nT, rho), mu=np.full(nT, mu), lam=np.full(nT, lam),
        alpha1=np.full(nT, alpha1),
        pinned=np.array([vid(0, j) for j in range(ny + 1)], dtype=np.int32),
        name=f"cantilever2d_{nx}x{ny}",
    )


def _tet_vol6(P, tet):
    a, b, c, d = (P[k] for k in tet)
    return float(np.linalg.det(np.stack([b - a, c - a, d - a], axis=1)))


def beam3d(nx=64, ny=16, nz=16, cell=0.02, z0=0.05, seed=0, E=None, nu=None,
           E_range=(1e4, 1e7), nu_range=(0.30, 0.45), rho=1000.0, alpha1=0.02,
           pin_far_end=False, neohookean=False):
    """3-D box beam: (nx+1)(ny+1)(nz+1) vertices, vertex (i,j,k) at index
    (i*(ny+1)+j)*(nz+1)+k, position (i, j, k)*cell + (0, 0, z0); 5 tets per
    cube (alternating parity), each positively oriented (det(Dm) > 0).
    Per-element E is log-uniform in E_range and nu uniform in nu_range, drawn
    from Pcg32(seed) in element order as (E draw, nu draw) pairs, unless E and
    nu are given as scalars.  The x = 0 face is pinned (and the x = nx*cell
    face too with pin_far_end, for the twisting study).  neohookean=True
    projects every element with the NeoHookean pair instead of StVK-Voigt."""
    NY, NZ = ny + 1, nz + 1
    vid = lambda i, j, k: (i * NY + j) * NZ + k
    ii, jj, kk = np.meshgrid(np.arange(nx + 1), np.arange(NY), np.arange(NZ), indexing="ij")
    P = np.stack([ii.ravel() * cell, jj.ravel() * cell, kk.ravel() * cell + z0], axis=1)
    ci, cj, ck = (a.ravel() for a in np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz),
                                                 indexing="ij"))
    corner = {(a, b, d): vid(ci + a, cj + b, ck + d) for a in (0, 1) for b in (0, 1) for d in (0, 1)}
    even = [[(1, 0, 0), (0, 0, 0), (1, 1, 0), (1, 0, 1)],
            [(0, 1, 0), (0, 0, 0), (1, 1, 0), (0, 1, 1)],
            [(0, 0, 1), (0, 0, 0), (1, 0, 1), (0, 1, 1)],
            [(1, 1, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
            [(0, 0, 0), (1, 1, 0), (1, 0, 1), (0, 1, 1)]]
    odd = [[(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)],
           [(1, 1, 0), (1, 0, 0), (0, 1, 0), (1, 1, 1)],
           [(1, 0, 1), (1, 0, 0), (0, 0, 1), (1, 1, 1)],
           [(0, 1, 1), (0, 1, 0), (0, 0, 1), (1, 1, 1)],
           [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1)]]
    par = ((ci + cj + ck) % 2)[:, None]
    T = np.empty((len(ci), 5, 4), dtype=np.int64)
    for q in range(5):
        for r in range(4):
            T[:, q, r] = np.where(par[:, 0] == 0, corner[even[q][r]], corner[odd[q][r]])
    T = T.reshape(-1, 4).astype(np.int32)
    # orient positively: swap the last two vertices where det(Dm) < 0
    a, b, cc, d = (P[T[:, q]] for q in range(4))
    det = np.einsum("ij,ij->i", b - a, np.cross(cc - a, d - a))
    neg = det < 0
    T[neg, 2], T[neg, 3] = T[neg, 3].copy(), T[neg, 2].copy()
    nT = len(T)
    if E is None or nu is None:
        rng = Pcg32(seed)
        u = rng.uniform(2 * nT).reshape(nT, 2)
        lo, hi = math.log10(E_range[0]), math.log10(E_range[1])
        Es = 10.0 ** (lo + (hi - lo) * u[:, 0])
        nus = nu_range[0] + (nu_range[1] - nu_range[0]) * u[:, 1]
    else:
        Es = np.full(nT, float(E))
        nus = np.full(nT, float(nu))
    mu, lam = to_lame(nus, Es)
    ends = (0, nx) if pin_far_end else (0,)
    pinned = np.array([vid(i, j, k) for i in ends for j in range(NY) for k in range(NZ)],
                      dtype=np.int32)
    return dict(
        dim=3, p=P, t=T, rho=np.full(nT, rho), mu=mu, lam=lam,
        alpha1=np.full(nT, alpha1), pinned=pinned, E=Es, nu=nus,
        neohookean=bool(neohookean),
        name=f"beam3d_{nx}x{ny}x{nz}" + ("_nh" if neohookean else ""),
    )


def twist_pinned_targets(mesh, angle, axis_dim=0):
    """Pinned-DOF targets (pinnedDOFs order) of a beam whose far end cap
    (the pinned vertices at the largest coordinate along axis_dim) is turned
    by `angle` radians about the beam axis (through the centre of the
    cross-section), the other pinned vertices kept at their rest positions.
    The twisting animation of the precision study (end cap at 1 rad/s:
    angle = t).  3D meshes; uses mesh.p as the rest positions."""
    P = np.asarray(mesh.p, dtype=np.float64).reshape(-1, 3)
    pv = np.asarray(mesh.pinnedInds, dtype=np.int64)
    a = [d for d in range(3) if d != axis_dim]
    far = P[pv, axis_dim] >= P[:, axis_dim].max() - 1e-12
    cy = P[:, a[0]].min() + np.ptp(P[:, a[0]]) / 2
    cz = P[:, a[1]].min() + np.ptp(P[:, a[1]]) / 2
    Q = P.copy()
    c, s_ = np.cos(angle), np.sin(angle)
    idx = pv[far]
    y, z = P[idx, a[0]] - cy, P[idx, a[1]] - cz
    Q[idx, a[0]] = cy + c * y - s_ * z
    Q[idx, a[1]] = cz + s_ * y + c * z
    dofs = np.asarray(mesh.pinnedDOFs, dtype=np.int64) - 1  # 1-based
    return Q.reshape(-1)[dofs]
