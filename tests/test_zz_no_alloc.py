"""mlx_pcg_solve_noalloc: the allocation-free PCG on caller-owned device
arrays (work vectors in the library's static device array) against
mlx_pcg_solve, which allocates its own buffers.  Two calls in a row, the
second on a smaller system, so the second reads a scratch array still
holding the first call's vectors.  The test's own device arrays and stream
come from the HIP runtime the library is linked against (libamdhip64.so.7,
called through ctypes)."""
import ctypes as C
import sys

import numpy as np
import pytest

from multilayersolver2024_amd import _lib as L

pytestmark = pytest.mark.gpu


_hip = None


def hip():
    global _hip
    if _hip is None:
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipStreamDestroy.argtypes = [C.c_void_p]
        h.hipStreamSynchronize.argtypes = [C.c_void_p]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamCreate", "hipStreamDestroy",
                  "hipStreamSynchronize"):
            getattr(h, f).restype = C.c_int
        _hip = h
    return _hip


H2D, D2H = 1, 2


class DevArrays:
    """Device copies of numpy arrays (hipMalloc'd by the test, freed on close)."""

    def __init__(self):
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert hip().hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p)
        if a.nbytes:
            assert hip().hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert hip().hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def close(self):
        for p in self.ptrs:
            hip().hipFree(p)
        self.ptrs = []


def _say(*a):
    print("[no-alloc]", *a, flush=True)
    sys.stdout.flush()


def _system(n, seed):
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=min(1.0, 8.0 / n), random_state=seed + 1, format="csr")
    A = (M @ M.T + sp.eye(n) * 4).tocsc()
    b = rng.normal(size=n)
    x0 = rng.normal(size=n) * 0.1
    return A, b, 1 / A.diagonal(), x0


def _host_solve(A, b, pinv, x0, it):
    n = A.shape[0]
    x = x0.copy()
    res = np.zeros(it)
    cp, ri = L.i32(A.indptr), L.i32(A.indices)
    L.check("mlx_pcg_solve", L.lib().mlx_pcg_solve(
        n, L.ptr(cp, C.c_int32), L.ptr(ri, C.c_int32), L.ptr(L.f64(A.data)), L.ptr(L.f64(b)),
        L.ptr(L.f64(pinv)), it, L.ptr(x), L.ptr(res)))
    return x, res


def _dev_solve(A, b, pinv, x0, it):
    n = A.shape[0]
    D = DevArrays()
    st = C.c_void_p()
    assert hip().hipStreamCreate(C.byref(st)) == 0
    try:
        cp, ri = D.put(L.i32(A.indptr)), D.put(L.i32(A.indices))
        va, bd, pv = D.put(L.f64(A.data)), D.put(L.f64(b)), D.put(L.f64(pinv))
        xd, rd = D.put(L.f64(x0)), D.put(np.zeros(it))
        sd = D.put(np.full(1, -1, dtype=np.int32))
        L.check("mlx_pcg_solve_noalloc", L.lib().mlx_pcg_solve_noalloc(
            n, int(A.nnz), cp, ri, va, bd, pv, it, xd, rd, sd, st))
        assert hip().hipStreamSynchronize(st) == 0
        return (D.get(xd, L.f64(x0)), D.get(rd, np.zeros(it)),
                int(D.get(sd, np.zeros(1, dtype=np.int32))[0]))
    finally:
        hip().hipStreamDestroy(st)
        D.close()


def test_noalloc_pcg_matches_allocating_entry_point():
    L.check("mlx_init", L.lib().mlx_init(0))
    for n, seed, it in ((5000, 61, 80), (700, 62, 40)):
        _say("case", n, seed, it)
        A, b, pinv, x0 = _system(n, seed)
        xh, rh = _host_solve(A, b, pinv, x0, it)
        xd, rd, status = _dev_solve(A, b, pinv, x0, it)
        err = float(np.max(np.abs(xd - xh)) / np.max(np.abs(xh)))
        rerr = float(np.max(np.abs(rd - rh)) / np.max(np.abs(rh)))
        _say("n", n, "status", status, "x rel", err, "resvec rel", rerr, "res last", rh[-1])
        assert status == 0
        # the same solve in the same order: FP64 to rounding
        assert err <= 1e-13 and rerr <= 1e-13
        assert rh[-1] < 1e-6 * np.linalg.norm(b)


def test_noalloc_pcg_rejects_sizes_beyond_its_scratch():
    L.check("mlx_init", L.lib().mlx_init(0))
    n = 65537
    # rejected on the host before any device array is touched
    rc = L.lib().mlx_pcg_solve_noalloc(n, 0, 1, None, None, 1, 1, 1, 1, None, None, None)
    assert rc == L.MLX_ERR_ARG
    assert b"65536" in L.lib().mlx_last_error()


def test_noalloc_pcg_flags_an_invalid_structure_on_the_device():
    """A row index out of range is found by the kernel's own check, which
    leaves x untouched and reports status 1."""
    L.check("mlx_init", L.lib().mlx_init(0))
    A, b, pinv, x0 = _system(64, 63)
    bad = L.i32(A.indices).copy()
    bad[5] = 64
    D = DevArrays()
    try:
        cp, ri = D.put(L.i32(A.indptr)), D.put(bad)
        va, bd, pv = D.put(L.f64(A.data)), D.put(L.f64(b)), D.put(L.f64(pinv))
        xd = D.put(L.f64(x0))
        sd = D.put(np.full(1, -1, dtype=np.int32))
        L.check("mlx_pcg_solve_noalloc", L.lib().mlx_pcg_solve_noalloc(
            64, int(A.nnz), cp, ri, va, bd, pv, 10, xd, None, sd, None))
        L.check("mlx_synchronize", L.lib().mlx_synchronize())
        assert int(D.get(sd, np.zeros(1, dtype=np.int32))[0]) == 1
        assert np.array_equal(D.get(xd, L.f64(x0)), x0)
    finally:
        D.close()
