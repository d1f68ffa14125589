"""Pins of the oracle's PCG (solvePCG.m restated): exact solution after n
iterations on small SPD systems (numpy.linalg.solve), a hand-computed first
iteration, the d'Ad == 0 guard."""
import ctypes as C

import numpy as np
import scipy.sparse as sp

import oracle as O

P = O.P


def _pcg(A, b, pinv, iters, x0=None):
    A = sp.csr_matrix(A)
    n = A.shape[0]
    x = np.zeros(n) if x0 is None else x0.copy()
    res = np.zeros(max(iters, 1))
    O.lib().ora_solve_pcg(n, P(A.indptr.astype(np.int32), C.c_int32), P(A.indices.astype(np.int32), C.c_int32),
                          P(A.data.astype(np.float64)), P(b.copy()), P(pinv.copy()), iters, P(x), P(res))
    return x, res[:iters]


def test_pcg_exact_after_n_iterations():
    rng = np.random.default_rng(50)
    for n in (2, 5, 12):
        M = rng.normal(size=(n, n))
        A = M @ M.T + n * np.eye(n)
        b = rng.normal(size=n)
        x, res = _pcg(A, b, 1 / np.diag(A), n)
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-12)
        assert res[-1] < 1e-9 * np.linalg.norm(b)


def test_pcg_first_iteration_by_hand():
    A = np.array([[4.0, 1.0], [1.0, 3.0]])
    b = np.array([1.0, 2.0])
    pinv = np.array([0.25, 1 / 3])
    r = b.copy()
    z = pinv * r
    d = z.copy()
    alpha = (r @ z) / (d @ A @ d)
    x, res = _pcg(A, b, pinv, 1)
    np.testing.assert_allclose(x, alpha * d, rtol=1e-15)
    assert res[0] == np.linalg.norm(A @ (alpha * d) - b) or abs(res[0] - np.linalg.norm(A @ (alpha * d) - b)) < 1e-15


def test_pcg_zero_curvature_guard():
    A = np.zeros((3, 3))
    A[0, 0] = 0.0
    b = np.zeros(3)
    x0 = np.array([1.0, 2.0, 3.0])
    # one iteration: d'Ad == 0 gives alpha = 0, so x stays x0 (later
    # iterations divide 0/0 in beta, as the reference does)
    x, _ = _pcg(sp.csr_matrix((np.zeros(1), (np.zeros(1, int), np.zeros(1, int))), shape=(3, 3)), b,
                np.ones(3), 1, x0=x0)
    assert np.array_equal(x, x0)
