"""Readers for the reference's tetrahedral mesh files.

The reference loads its 3D meshes through gptoolbox (`readNODE`/`readELE`
in polyLoader.m:21-22, `readMESH` in meshLoader.m:20), a third-party
library that is not vendored with it.  These readers follow the published
file formats instead:

* TetGen ``.node``: a header ``<#points> <dim> <#attributes> <#markers>``,
  then one line per point ``<index> <x> <y> <z> [attributes] [marker]``;
  ``.ele``: ``<#tets> <nodes per tet> <#region attributes>``, then
  ``<index> <n1> .. <n4> [attribute]``.  Point indices start at 0 or 1 (the
  first point's index decides), ``#`` starts a comment.
* Medit ``.mesh``: keyword sections ``Vertices`` (count, then
  ``x y z ref``) and ``Tetrahedra`` (count, then ``i j k l ref``, 1-based).

``scene_from_mesh`` turns positions and elements into the scene dict that
``Mesh`` takes, with one material (or one per region attribute), as
Mesh3D(V, T, A, materials) does (Mesh3D.m:123-170).
"""
import numpy as np


def _lines(path):
    with open(path, "r") as f:
        for raw in f:
            line = raw.split("#", 1)[0].strip()
            if line:
                yield line


def read_node(path):
    """Points of a TetGen .node file: (P (n x dim), first index)."""
    it = _lines(path)
    head = next(it).split()
    n, dim = int(head[0]), int(head[1])
    P = np.zeros((n, dim))
    first = None
    for i in range(n):
        parts = next(it).split()
        idx = int(parts[0])
        if first is None:
            first = idx
        k = idx - first
        if not 0 <= k < n:
            raise ValueError(f"{path}: point index {idx} out of range")
        P[k] = [float(x) for x in parts[1:1 + dim]]
    return P, (first if first is not None else 0)


def read_ele(path, first_index=0):
    """Elements of a TetGen/Triangle .ele file, 0-based against the .node
    numbering that starts at first_index: (T (m x nodes), attributes or None)."""
    it = _lines(path)
    head = next(it).split()
    m, k = int(head[0]), int(head[1])
    nattr = int(head[2]) if len(head) > 2 else 0
    T = np.zeros((m, k), dtype=np.int64)
    A = np.zeros(m) if nattr > 0 else None
    for i in range(m):
        parts = next(it).split()
        T[i] = [int(x) - first_index for x in parts[1:1 + k]]
        if nattr > 0:
            A[i] = float(parts[1 + k])
    return T.astype(np.int32), A


def read_tetgen(base):
    """``base.node`` + ``base.ele``: positions, 0-based elements, attributes."""
    P, first = read_node(base + ".node")
    T, A = read_ele(base + ".ele", first)
    if T.min() < 0 or T.max() >= len(P):
        raise ValueError(f"{base}: element index out of range")
    return P, T, A


def read_medit(path):
    """Vertices and tetrahedra of a Medit .mesh file (0-based)."""
    tok = []
    for line in _lines(path):
        tok.extend(line.split())
    P = T = None
    i = 0
    while i < len(tok):
        key = tok[i]
        if key == "Vertices":
            n = int(tok[i + 1])
            vals = np.array(tok[i + 2:i + 2 + 4 * n], dtype=np.float64).reshape(n, 4)
            P = vals[:, :3].copy()
            i += 2 + 4 * n
        elif key == "Tetrahedra":
            n = int(tok[i + 1])
            vals = np.array(tok[i + 2:i + 2 + 5 * n], dtype=np.int64).reshape(n, 5)
            T = (vals[:, :4] - 1).astype(np.int32)
            i += 2 + 5 * n
        elif key in ("Triangles", "Edges", "Corners", "Ridges", "RequiredVertices"):
            n = int(tok[i + 1])
            width = {"Triangles": 4, "Edges": 3}.get(key, 1)
            i += 2 + width * n
        else:
            i += 1
    if P is None or T is None:
        raise ValueError(f"{path}: no Vertices/Tetrahedra section")
    if T.min() < 0 or T.max() >= len(P):
        raise ValueError(f"{path}: element index out of range")
    return P, T


def scene_from_mesh(P, T, E=1e5, nu=0.35, rho=1000.0, alpha1=0.02, pinned=None, scale=1.0,
                    offset=(0.0, 0.0, 0.0), name="mesh"):
    """Scene dict for ``Mesh`` from positions (n x 3) and tets (m x 4,
    0-based): one isotropic material, optional pinned vertex ids, positions
    scaled then offset."""
    from .scenes import to_lame
    P = np.asarray(P, dtype=np.float64) * scale + np.asarray(offset, dtype=np.float64)
    T = np.asarray(T, dtype=np.int32)
    nT = len(T)
    mu, lam = to_lame(nu, E)
    return dict(dim=3, p=P, t=T, rho=np.full(nT, rho), mu=np.full(nT, mu),
                lam=np.full(nT, lam), alpha1=np.full(nT, alpha1),
                pinned=np.zeros(0, np.int32) if pinned is None else np.asarray(pinned, np.int32),
                name=name)
