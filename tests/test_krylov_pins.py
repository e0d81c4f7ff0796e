"""Pins of the oracle's Krylov path (solver.f90): matmul_ell against the
dense matrix the DIA layout stands for (integer data, so the products are
exact and the comparison is bit for bit), and solver_BICG3 against a dense
direct solve of the solver_1st_member matrix, its stopping rule and its
iteration cap.  CPU only."""
import numpy as np
import pytest

from conftest import stretched_grid
from oracle import Oracle


def dense_from_dia(coef, jcoef, n):
    """The matrix matmul_ell applies: column 1 the diagonal; for offset
    jcoef(j), j = 2..4, columns 2j-2 (upper, row i reads i+col) and 2j-1
    (lower, row i reads i-col)."""
    A = np.zeros((n, n), dtype=coef.dtype)
    A[np.arange(n), np.arange(n)] = coef[:, 0]
    for j in range(2, len(jcoef) + 1):
        col = jcoef[j - 1]
        up, lo = coef[:, 2 * j - 3], coef[:, 2 * j - 2]
        i = np.arange(n - col)
        A[i, i + col] += up[:n - col]
        i = np.arange(col, n)
        A[i, i - col] += lo[col:]
    return A


def test_matmul_ell_is_the_dia_matrix():
    rng = np.random.default_rng(1)
    nx, ny, nz = 5, 4, 3
    n = nx * ny * nz
    jcoef = [0, 1, nx, nx * ny]
    coef = np.asfortranarray(rng.integers(-9, 10, (n, 7)).astype(np.float64))
    y = rng.integers(-9, 10, n).astype(np.float64)
    A = dense_from_dia(coef.astype(np.int64), jcoef, n)
    want = (A @ y.astype(np.int64)).astype(np.float64)
    assert np.array_equal(Oracle.matmul_ell(coef, jcoef, y), want)
    # fewer diagonals: m = 2 uses only the +-1 pair
    A2 = dense_from_dia(coef.astype(np.int64), jcoef[:2], n)
    assert np.array_equal(Oracle.matmul_ell(coef, jcoef[:2], y),
                          (A2 @ y.astype(np.int64)).astype(np.float64))


@pytest.mark.parametrize("shape", [(6, 5, 4), (9, 7, 5)])
def test_bicg3_solves_the_poisson_matrix(shape):
    g = stretched_grid(*shape)
    coef, jcoef = Oracle(g).poisson_coeffs()
    n = g.cells
    A = dense_from_dia(coef, jcoef, n)
    rng = np.random.default_rng(sum(shape))
    b = rng.standard_normal(n)
    want = np.linalg.solve(A, b)
    x, it, nrm = Oracle.solver_bicg3(coef, jcoef, b, np.zeros(n), 1e-13, 2000)
    assert 0 < it < 2000 and nrm <= 1e-13
    assert np.linalg.norm(x - want) <= 1e-10 * np.linalg.norm(want)
    # the recursive residual it stopped on is the true residual
    true = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert true < 1e-11


def test_bicg3_cap_and_initial_guess():
    g = stretched_grid(6, 5, 4)
    coef, jcoef = Oracle(g).poisson_coeffs()
    n = g.cells
    A = dense_from_dia(coef, jcoef, n)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(n)
    x, it, nrm = Oracle.solver_bicg3(coef, jcoef, b, np.zeros(n), 1e-14, 3)
    assert it == 3 and nrm > 1e-14
    # the exact solution as the initial guess: no iteration
    exact = np.linalg.solve(A, b)
    x0 = A @ exact
    x, it, nrm = Oracle.solver_bicg3(coef, jcoef, x0, exact.copy(), 1e-8, 100)
    assert it == 0 and np.array_equal(x, exact)
