"""Dense symmetric solvers of the second-derivative drivers (dense.hip),
host paths (no GPU): eigenvalues and eigenvectors of random symmetric
matrices against numpy's LAPACK (eigvalsh / eigh), the Cholesky solve
against numpy.linalg.solve, and the dposv_-style return code of a matrix
that is not positive definite."""
import ctypes

import numpy as np
import pytest


def _lib():
    import nabc_amd as na
    f = na.lib.nabc_dense_eigh
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, na.dp, na.dp, ctypes.c_int]
    g = na.lib.nabc_dense_posv
    g.restype = ctypes.c_int
    g.argtypes = [ctypes.c_int, na.dp, na.dp]
    return na, f, g


@pytest.mark.parametrize("n", [1, 2, 3, 7, 72, 300])
def test_eigh_matches_lapack(n):
    na, eigh, _ = _lib()
    rng = np.random.default_rng(n)
    B = rng.normal(size=(n, n))
    A = np.ascontiguousarray(B + B.T)
    A[0, 0] += 10.0
    w = np.zeros(n)
    a = A.copy()
    assert eigh(n, a.ctypes.data_as(na.dp), w.ctypes.data_as(na.dp), 0) == 0
    ref = np.linalg.eigvalsh(A)
    scale = max(1.0, np.abs(ref).max())
    assert np.max(np.abs(w - ref)) <= 1e-12 * scale * n
    a = A.copy()
    assert eigh(n, a.ctypes.data_as(na.dp), w.ctypes.data_as(na.dp), 1) == 0
    V = a.reshape(n, n)  # row j: eigenvector of w[j]
    assert np.max(np.abs(V @ A - w[:, None] * V)) <= 1e-10 * scale * n
    assert np.max(np.abs(V @ V.T - np.eye(n))) <= 1e-11 * n


def test_eigh_degenerate_and_diagonal():
    na, eigh, _ = _lib()
    A = np.diag([3.0, 1.0, 1.0, -2.0, 5.0])
    w = np.zeros(5)
    assert eigh(5, A.copy().ctypes.data_as(na.dp), w.ctypes.data_as(na.dp), 0) == 0
    assert np.allclose(w, [-2.0, 1.0, 1.0, 3.0, 5.0], atol=1e-14)


@pytest.mark.parametrize("n", [1, 5, 200])
def test_posv_matches_solve(n):
    na, _, posv = _lib()
    rng = np.random.default_rng(100 + n)
    B = rng.normal(size=(n, n))
    A = np.ascontiguousarray(B @ B.T + n * np.eye(n))
    b = rng.normal(size=n)
    x = b.copy()
    assert posv(n, A.ctypes.data_as(na.dp), x.ctypes.data_as(na.dp)) == 0
    assert np.max(np.abs(x - np.linalg.solve(A, b))) <= 1e-12 * np.max(np.abs(x))


def test_posv_reports_the_failing_minor():
    na, _, posv = _lib()
    A = np.array([[4.0, 2.0, 0.0], [2.0, 1.0, 0.0], [0.0, 0.0, 3.0]])  # second minor singular
    b = np.ones(3)
    assert posv(3, A.ctypes.data_as(na.dp), b.ctypes.data_as(na.dp)) == 2
