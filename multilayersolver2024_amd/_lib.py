"""ctypes binding of libmlxpbd.so (the C ABI declared in include/mlxpbd.h).

The library is built in-tree by ``multilayersolver2024_amd.build``.  Loading
fails loudly if it is missing: there is no CPU fallback for the GPU path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libmlxpbd.so")

MLX_OK = 0
MLX_ERR_ARG = -1
MLX_ERR_HIP = -2
MLX_ERR_NOMEM = -3
MLX_ERR_UNSUPPORTED = -4
MLX_ERR_NODEVICE = -5
MLX_ERR_STATE = -6
MLX_FP64 = 0
MLX_FP32 = 1
MLX_MAX_LAYERS = 16

# Every symbol include/mlxpbd.h declares.
EXPORTED = (
    "mlx_last_error", "mlx_version", "mlx_init", "mlx_naive_graph_coloring",
    "mlx_mesh3d_precompute", "mlx_mesh2d_precompute", "mlx_ediff_norm3d",
    "mlx_ediff_norm2d", "mlx_union_find_by_vertice", "mlx_find_boundary_vertices",
    "mlx_mesh_create", "mlx_mesh_destroy", "mlx_mesh_info",
    "mlx_stvk_voigt_constraints", "mlx_layer_solve", "mlx_state_create",
    "mlx_state_destroy", "mlx_state_download", "mlx_integrate",
    "mlx_profile_enable", "mlx_profile_read", "mlx_profile_phases", "mlx_profile_read_persistent", "mlx_layer_maps", "mlx_batch_set_split", "mlx_batch_set_lds", "mlx_batch_set_graph", "mlx_state_set_pinned_targets", "mlx_synchronize", "mlx_mesh_set_option",
    "mlx_batch_create", "mlx_batch_destroy", "mlx_batch_set_block", "mlx_batch_set_state",
    "mlx_batch_download", "mlx_batch_integrate", "mlx_neohookean_constraints2d",
    "mlx_pcg_solve", "mlx_pcg_solve_noalloc",
    "mlx_mlp_create", "mlx_mlp_destroy", "mlx_mlp_set", "mlx_mlp_get", "mlx_mlp_infer",
    "mlx_mlp_grad", "mlx_mlp_train", "mlx_mesh_set_learned", "mlx_learn_stats",
    "mlx_learn_outputs", "mlx_learn_features", "mlx_state_set", "mlx_state_get_oldp",
    "mlx_batch_set_rigid", "mlx_batch_autotune", "mlx_integrate_finders", "mlx_find_contacts",
    "mlx_batch_set_learned", "mlx_batch_learn_stats",
)
MLX_LEARN_WARMSTART = 1
MLX_LEARN_KEYS = 2
MLX_OPT_GRAPHS = 1
MLX_OPT_GRAPH_ITERS = 2
MLX_OPT_PERSISTENT = 3
MLX_OPT_PERSIST_WG = 4
MLX_OPT_DEVICE_MAPS = 5
MLX_OPT_PERSIST_LOCAL = 6
MLX_OPT_SWEEP_BLOCK = 7
MLX_OPT_SWEEP_WG = 8
MLX_OPT_FLOW = 9

_p = C.c_void_p
_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_u8 = C.POINTER(C.c_uint8)
_i64 = C.POINTER(C.c_int64)
_f = C.POINTER(C.c_float)


class MeshDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("nV", C.c_int64), ("nT", C.c_int64),
        ("t", _i32), ("DmInv", _d), ("alpha", _d), ("beta", _d),
        ("volume", _d), ("mass", _d), ("elementColor", _i32),
        ("pinnedDOFs", _i32), ("nPinnedDOFs", C.c_int64), ("forcedElastic", _u8),
        ("neohookean", _u8), ("mu", _d), ("lambda_", _d),
    ]


class LayerSolveArgs(C.Structure):
    _fields_ = [
        ("nLayers", C.c_int32), ("layersMap", _i32), ("iterations", _i32),
        ("nSorted", C.c_int32), ("isElasticElement", _u8), ("vertexRigidBody", _i32),
        ("isBoundaryVertex", _u8), ("isRigidVertex", _u8), ("numRigids", _i32),
        ("h", C.c_double), ("boundaryCompliance", C.c_double),
        ("contactResistance", C.c_double), ("computeResiduals", C.c_int32),
        ("hangingStop", C.c_int32), ("giveUpEnabled", C.c_int32),
        ("giveUpThreshold", C.c_double), ("lastLayerGiveUpEnabled", C.c_int32),
        ("useRigidContacts", C.c_int32), ("stopAtImprovementPercent", C.c_double),
        ("runUntilResSmallerThan", C.c_double), ("residualArrayLast", C.c_double),
        ("xi", _d), ("lambdai", _d), ("oldp", _d), ("pinnedx", _d), ("residualv", _d),
        ("nContacts", C.c_int32), ("contactVertexID", _i32), ("contactNormals", _d),
        ("pointColliders", _d), ("contactCompliance", _d),
    ]


class LayerSolveOut(C.Structure):
    _fields_ = [
        ("residualArray", _d), ("residualCap", C.c_int32), ("residualCount", C.c_int32),
        ("runUntilResSmallerThan", C.c_double),
        ("itersPerLayer", C.c_int32 * MLX_MAX_LAYERS), ("layersRun", C.c_int32),
        ("elasticProjections", C.c_int64), ("contactLambdas", _d),
    ]


class Settings(C.Structure):
    _fields_ = [
        ("nLayers", C.c_int32), ("layers", C.c_double * MLX_MAX_LAYERS),
        ("iterations", C.c_int32 * MLX_MAX_LAYERS), ("gravity", C.c_double),
        ("airDrag", C.c_double), ("boundaryCompliance", C.c_double),
        ("contactResistance", C.c_double), ("restitution", C.c_double),
        ("computeResiduals", C.c_int32), ("hangingStop", C.c_int32),
        ("giveUpEnabled", C.c_int32), ("giveUpThreshold", C.c_double),
        ("lastLayerGiveUpEnabled", C.c_int32), ("useRigidContacts", C.c_int32),
        ("stopAtImprovementPercent", C.c_double), ("runUntilResSmallerThan", C.c_double),
    ]


class Plane(C.Structure):
    _fields_ = [
        ("enabled", C.c_int32), ("normal", C.c_double * 3),
        ("position", C.c_double * 3), ("compliance", C.c_double),
    ]


MLX_CONTACT_PLANE = 0
MLX_CONTACT_BOX = 1
MLX_CONTACT_SDF = 2


class ContactFinder(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("compliance", C.c_double), ("normal", C.c_double * 3),
        ("position", C.c_double * 3), ("center", C.c_double * 3), ("axes", C.c_double * 9),
        ("halfExtent", C.c_double * 3), ("sdf", C.POINTER(C.c_double)),
        ("dims", C.c_int32 * 3), ("origin", C.c_double * 3), ("spacing", C.c_double),
    ]


class BatchOut(C.Structure):
    _fields_ = [
        ("iterations", _i32), ("lastResidual", _d), ("contacts", _i32),
        ("elasticProjections", C.c_int64), ("kernelMs", C.c_double),
    ]


class MlxError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


_lib = None


def lib():
    """Load libmlxpbd.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m multilayersolver2024_amd.build`"
        )
    L = C.CDLL(LIB_PATH)
    L.mlx_last_error.restype = C.c_char_p
    L.mlx_version.restype = C.c_int
    sig = {
        "mlx_init": [C.c_int32],
        "mlx_naive_graph_coloring": [_i32, C.c_int64, C.c_int32, C.c_int64, _i32, _i32],
        "mlx_mesh3d_precompute": [_d, C.c_int64, _i32, C.c_int64, _d, _d, _d, _d, _d, _d, _d, _d, _d],
        "mlx_mesh2d_precompute": [_d, C.c_int64, _i32, C.c_int64, _d, _d, _d, _d, _d, _d, _d, _d, _d],
        "mlx_ediff_norm3d": [_d, _d, C.c_int64, C.c_double, _d],
        "mlx_ediff_norm2d": [_d, _d, C.c_int64, C.c_double, _d],
        "mlx_union_find_by_vertice": [_i32, C.c_int64, _d, C.c_int32, _u8, _i32, C.c_int32,
                                      C.c_int64, _i32, _i32],
        "mlx_find_boundary_vertices": [_i32, C.c_int64, C.c_int32, C.c_int64, C.c_int32, _i32,
                                       _u8, _u8, _i32],
        "mlx_mesh_create": [C.POINTER(MeshDesc), C.c_int32, C.POINTER(_p)],
        "mlx_mesh_destroy": [_p],
        "mlx_mesh_info": [_p, _i32, _i64],
        "mlx_stvk_voigt_constraints": [_p, _u8, C.c_double, _d, _d, _d],
        "mlx_layer_solve": [_p, C.POINTER(LayerSolveArgs), C.POINTER(LayerSolveOut)],
        "mlx_state_create": [_p, _d, _d, C.POINTER(_p)],
        "mlx_state_destroy": [_p],
        "mlx_state_download": [_p, _d, _d, _d, _i32],
        "mlx_integrate": [_p, C.POINTER(Settings), C.c_double, C.c_int32, C.POINTER(Plane),
                          C.POINTER(LayerSolveOut)],
        "mlx_profile_enable": [_p, C.c_int32],
        "mlx_profile_read": [_p, _d, _i64, _i64],
        "mlx_profile_phases": [_p, _d, _d, _d, _i64],
        "mlx_profile_read_persistent": [_p, _d, _i64, _i64, _i64],
        "mlx_layer_maps": [_p, _d, _d, C.c_int32, _i32, _i32, _u8, _i32],
        "mlx_batch_set_split": [_p, C.c_int32],
        "mlx_batch_set_lds": [_p, C.c_int32],
        "mlx_batch_set_graph": [_p, C.c_int32],
        "mlx_state_set_pinned_targets": [_p, _d],
        "mlx_synchronize": [],
        "mlx_mesh_set_option": [_p, C.c_int32, C.c_int64],
        "mlx_neohookean_constraints2d": [_p, _u8, C.c_double, _d, _d, _d, _d],
        "mlx_pcg_solve": [C.c_int64, _i32, _i32, _d, _d, _d, C.c_int32, _d, _d],
        "mlx_pcg_solve_noalloc": [C.c_int64, C.c_int64, _p, _p, _p, _p, _p, C.c_int32, _p, _p,
                                  _p, _p],
        "mlx_batch_create": [C.POINTER(MeshDesc), C.c_int32, _i64, _i64, C.c_int32, C.POINTER(_p)],
        "mlx_batch_destroy": [_p],
        "mlx_batch_set_block": [_p, C.c_int32],
        "mlx_batch_set_state": [_p, _d, _d],
        "mlx_batch_download": [_p, _d, _d, _d],
        "mlx_batch_integrate": [_p, C.POINTER(Settings), C.c_double, C.c_int32, C.POINTER(Plane),
                                C.POINTER(BatchOut)],
        "mlx_mlp_create": [C.POINTER(_p)],
        "mlx_mlp_destroy": [_p],
        "mlx_mlp_set": [_p, _f, _f, _f, _f, _f, _f, _f, _f, _f, _f],
        "mlx_mlp_get": [_p, _f, _f, _f, _f, _f, _f],
        "mlx_mlp_infer": [_p, _f, C.c_int64, _f],
        "mlx_mlp_grad": [_p, _f, _f, _f, C.c_int64, _f, _d],
        "mlx_mlp_train": [_p, _f, _f, _f, C.c_int64, C.c_int32, C.c_int32, C.c_float,
                          C.c_uint64, _f],
        "mlx_mesh_set_learned": [_p, _p, C.c_int32, C.c_float],
        "mlx_learn_stats": [_p, _i64, _i64],
        "mlx_learn_outputs": [_p, _d, _d],
        "mlx_learn_features": [_p, C.c_double, _f],
        "mlx_state_set": [_p, _d, _d, _d, _d],
        "mlx_state_get_oldp": [_p, _d],
        "mlx_batch_set_rigid": [_p, _u8],
        "mlx_batch_set_learned": [_p, _p, C.c_float],
        "mlx_batch_learn_stats": [_p, _i64, _i64],
        "mlx_batch_autotune": [_p, C.POINTER(Settings), C.c_double, C.c_int32, C.POINTER(Plane),
                               _i32, _f, _i32],
        "mlx_integrate_finders": [_p, C.POINTER(Settings), C.c_double, C.c_int32,
                                  C.POINTER(ContactFinder), C.c_int32, C.POINTER(LayerSolveOut)],
        "mlx_find_contacts": [_p, C.POINTER(ContactFinder), C.c_int32, _i32, C.c_int32, _i32, _d,
                              _d, _d],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    _lib = L
    return L


def check(fn, code):
    if code != MLX_OK:
        msg = lib().mlx_last_error().decode(errors="replace")
        raise MlxError(fn, code, msg)


def ptr(a, ctype=C.c_double):
    """Pointer to a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"], "array must be contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)


def fortran(a, dtype):
    """Matlab (column-major) copy of a 2-D array, returned flattened."""
    return np.ascontiguousarray(np.asarray(a, dtype=dtype).T).reshape(-1)
