"""Scene precompute: the arrays the reference's Mesh3D / Mesh constructors
build for the XPBD path (3d/@Mesh3D/Mesh3D.m:123-298, 2d/@Mesh/Mesh.m:108-227),
computed by the native library and kept in the reference's Matlab layouts
(column-major, 1-based indices) so they can be handed to the C ABI as a MEX
gateway would.
"""
import ctypes as C
import sys

import numpy as np

from . import _lib as L


class Mesh:
    """Precomputed scene.

    Attributes (Matlab names): ``t`` (nT x vpe, 1-based, column-major int32),
    ``DmInv`` (dim x dim x nT pages), ``elV`` (nT; elA in 2D), ``mass``
    (dim*nV), ``alpha``/``beta`` (perElementXPBDalphaMatrix / perElementXPBDbeta,
    nc x nc x nT pages), ``elementColor`` (1-based), ``numColors``,
    ``pinnedDOFs`` (1-based), ``pinnedTets`` mask (``forcedElastic``), ``p``
    (dim*nV initial positions), ``elMu``/``elLambda`` (Lame parameters) and
    ``neohookean`` (nT mask, 3D: elements projected with the NeoHookean
    pair instead of StVK-Voigt; scene key "neohookean", a mask or a bool).
    """

    def __init__(self, scene):
        lib = L.lib()
        self.scene = scene
        self.dim = int(scene["dim"])
        dim = self.dim
        P = np.asarray(scene["p"], dtype=np.float64)
        T0 = np.asarray(scene["t"], dtype=np.int32)
        self.nV, self.vpe = P.shape[0], dim + 1
        self.nT = T0.shape[0]
        self.nc = 6 if dim == 3 else 3
        self.T0 = np.ascontiguousarray(T0)
        self.t = L.fortran(T0 + 1, np.int32)
        self.p = L.f64(P.reshape(-1))
        nV, nT, nc = self.nV, self.nT, self.nc
        self.DmInv = np.zeros(dim * dim * nT)
        self.elV = np.zeros(nT)
        self.mass = np.zeros(dim * nV)
        self.alpha = np.zeros(nc * nc * nT)
        self.beta = np.zeros(nc * nc * nT)
        rho, mu, lam, a1 = (L.f64(scene[k]) for k in ("rho", "mu", "lam", "alpha1"))
        self.elMu, self.elLambda = mu.copy(), lam.copy()
        nh = scene.get("neohookean")
        self.neohookean = (L.u8(np.broadcast_to(np.asarray(nh, dtype=bool), (nT,)))
                           if nh is not None else np.zeros(nT, dtype=np.uint8))
        if dim != 3 and self.neohookean.any():
            raise ValueError("NeoHookean elements in the substep are 3D only")
        fn = lib.mlx_mesh3d_precompute if dim == 3 else lib.mlx_mesh2d_precompute
        L.check(fn.__name__, fn(L.ptr(self.p), nV, L.ptr(self.t, C.c_int32), nT, L.ptr(rho),
                                L.ptr(mu), L.ptr(lam), L.ptr(a1), L.ptr(self.DmInv),
                                L.ptr(self.elV), L.ptr(self.mass), L.ptr(self.alpha),
                                L.ptr(self.beta)))
        self.elementColor = np.zeros(nT, dtype=np.int32)
        nc_ = np.zeros(1, dtype=np.int32)
        L.check("mlx_naive_graph_coloring",
                lib.mlx_naive_graph_coloring(L.ptr(self.t, C.c_int32), nT, self.vpe, nV,
                                             L.ptr(self.elementColor, C.c_int32),
                                             L.ptr(nc_, C.c_int32)))
        self.numColors = int(nc_[0])
        self.set_pinned(scene.get("pinned", np.zeros(0, dtype=np.int32)))

    def set_pinned(self, pinned_vertices):
        """pin.m: pinnedDOFs, pinnedTets (every element touching a pinned
        vertex, elementAdjacencyMatrix.m:26-35)."""
        pv = np.asarray(pinned_vertices, dtype=np.int64)
        self.pinnedInds = pv
        if self.dim == 3:
            dofs = np.sort(np.concatenate([3 * pv + 1, 3 * pv + 2, 3 * pv + 3]))
        else:
            dofs = np.concatenate([2 * pv + 1, 2 * pv + 2])  # reshape([2v-1; 2v], 1, [])
        self.pinnedDOFs = L.i32(dofs)
        mark = np.zeros(self.nV, dtype=bool)
        mark[pv] = True
        self.forcedElastic = L.u8(mark[self.T0].any(axis=1))
        self.pinnedTets = np.nonzero(self.forcedElastic)[0] + 1

    def desc(self):
        """mlx_mesh_desc for mlx_mesh_create (keeps references alive)."""
        d = L.MeshDesc()
        d.dim, d.nV, d.nT = self.dim, self.nV, self.nT
        d.t = L.ptr(self.t, C.c_int32)
        d.DmInv, d.alpha, d.beta = L.ptr(self.DmInv), L.ptr(self.alpha), L.ptr(self.beta)
        d.volume, d.mass = L.ptr(self.elV), L.ptr(self.mass)
        d.elementColor = L.ptr(self.elementColor, C.c_int32)
        d.pinnedDOFs = L.ptr(self.pinnedDOFs, C.c_int32)
        d.nPinnedDOFs = len(self.pinnedDOFs)
        d.forcedElastic = L.ptr(self.forcedElastic, C.c_uint8)
        if self.neohookean.any():
            d.neohookean = L.ptr(self.neohookean, C.c_uint8)
            d.mu, d.lambda_ = L.ptr(self.elMu), L.ptr(self.elLambda)
        return d


class DeviceMesh:
    """A Mesh uploaded to the GPU (mlx_mesh_create); precision MLX_FP64 or
    MLX_FP32."""

    def __init__(self, mesh, precision=L.MLX_FP64, device=0):
        lib = L.lib()
        L.check("mlx_init", lib.mlx_init(device))
        self.mesh = mesh
        self.precision = precision
        self._desc = mesh.desc()
        h = C.c_void_p()
        L.check("mlx_mesh_create", lib.mlx_mesh_create(C.byref(self._desc), precision, C.byref(h)))
        self.handle = h

    def set_option(self, option, value):
        L.check("mlx_mesh_set_option", L.lib().mlx_mesh_set_option(self.handle, option, int(value)))

    def info(self):
        n = np.zeros(1, dtype=np.int32)
        L.check("mlx_mesh_info", L.lib().mlx_mesh_info(self.handle, L.ptr(n, C.c_int32), None))
        sizes = np.zeros(int(n[0]), dtype=np.int64)
        L.check("mlx_mesh_info", L.lib().mlx_mesh_info(self.handle, L.ptr(n, C.c_int32),
                                                        L.ptr(sizes, C.c_int64)))
        return int(n[0]), sizes

    def close(self):
        if self.handle:
            L.lib().mlx_mesh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        # A handle still open while the interpreter shuts down is left to
        # process exit: no HIP call is made from finalization, whose order
        # against the runtime's own teardown is not defined.
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass
