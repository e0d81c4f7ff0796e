"""GPU parity of the Krylov path (solver.f90) against the CPU restatement:
matmul_ell bit for bit; solver_BICG3 to the same solution within the
stopping tolerance, with the same iteration count up to the rounding of its
inner products (the GPU sums them as deterministic trees, the restatement
serially), the true residual at the tolerance, the iteration cap, the
zero-iteration cases of the reference's loop test, run-to-run
reproducibility, and the refusal of malformed DIA layouts."""
import numpy as np
import pytest

from conftest import stretched_grid
from ntustcfdlab_amd.hip import NtustError, ProjectionSolver
from oracle import Oracle

pytestmark = pytest.mark.gpu


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.fixture(scope="module")
def system(hip_lib):
    g = stretched_grid(20, 16, 14)
    coef, jcoef = Oracle(g).poisson_coeffs()
    s = ProjectionSolver(g)
    yield g, coef, jcoef, s
    s.close()


def test_matmul_ell_bitwise(system):
    g, coef, jcoef, s = system
    rng = np.random.default_rng(5)
    y = rng.standard_normal(g.cells)
    assert same(s.matmul_ell(coef, jcoef, y), Oracle.matmul_ell(coef, jcoef, y))
    assert same(s.matmul_ell(coef, jcoef[:2], y), Oracle.matmul_ell(coef, jcoef[:2], y))


@pytest.mark.parametrize("eps", [1e-6, 1e-10])
def test_bicg3_matches_oracle(system, eps):
    g, coef, jcoef, s = system
    n = g.cells
    rng = np.random.default_rng(int(-np.log10(eps)))
    b = rng.standard_normal(n)
    x0 = 0.1 * rng.standard_normal(n)
    xr, itr, nr = Oracle.solver_bicg3(coef, jcoef, b, x0, eps, 5000)
    x = x0.copy()
    it, nrm = s.solver_bicg3(coef, jcoef, b, x, eps, 5000)
    assert nrm <= eps and nr <= eps
    # BiCGSTAB's residual history is irregular: the tree and serial sums'
    # rounding moves the iteration where it first drops below eps by a few
    # percent (176 against 186 at 1e-10 on this matrix)
    assert abs(it - itr) <= max(3, itr // 10), (it, itr)
    true = np.linalg.norm(b - Oracle.matmul_ell(coef, jcoef, x)) / np.linalg.norm(b)
    assert true <= 10 * eps
    assert np.linalg.norm(x - xr) <= 100 * eps * np.linalg.norm(xr)
    # deterministic sums: the same bits on a second run
    x2 = x0.copy()
    assert s.solver_bicg3(coef, jcoef, b, x2, eps, 5000) == (it, nrm)
    assert same(x2, x)


@pytest.mark.parametrize("k", [1, 2, 3, 8])
def test_bicg3_iterates_track_oracle(system, k):
    """Step by step: after k iterations the iterate equals the serial
    restatement's up to the rounding of the inner products (a wrong update
    formula differs at order one)."""
    g, coef, jcoef, s = system
    n = g.cells
    rng = np.random.default_rng(11)
    b = rng.standard_normal(n)
    xr, itr, nr = Oracle.solver_bicg3(coef, jcoef, b, np.zeros(n), 1e-30, k)
    x = np.zeros(n)
    it, nrm = s.solver_bicg3(coef, jcoef, b, x, 1e-30, k)
    assert it == itr == k
    assert np.linalg.norm(x - xr) <= 1e-11 * np.linalg.norm(xr)
    assert abs(nrm - nr) <= 1e-9 * nr


def test_bicg3_cap_and_loop_test(system):
    g, coef, jcoef, s = system
    n = g.cells
    rng = np.random.default_rng(9)
    b = rng.standard_normal(n)
    x = np.zeros(n)
    it, nrm = s.solver_bicg3(coef, jcoef, b, x, 1e-14, 5)
    assert it == 5 and nrm > 1e-14
    # already converged: no iteration, x untouched
    exact = np.zeros(n)
    s.solver_bicg3(coef, jcoef, b, exact, 1e-13, 5000)
    b2 = Oracle.matmul_ell(coef, jcoef, exact)
    x = exact.copy()
    it, nrm = s.solver_bicg3(coef, jcoef, b2, x, 1e-6, 100)
    assert it == 0 and same(x, exact)
    # b = 0 and x = 0: ||r||/||b|| is 0/0, the loop test is false
    x = np.zeros(n)
    it, nrm = s.solver_bicg3(coef, jcoef, np.zeros(n), x, 1e-6, 100)
    assert it == 0 and np.isnan(nrm) and not x.any()
    # b = 0 and x != 0: ||r||/||b|| is +inf, the loop runs to the cap, as the
    # restatement's (and the reference's) does
    x0 = rng.standard_normal(n)
    xr, itr, nr = Oracle.solver_bicg3(coef, jcoef, np.zeros(n), x0, 1e-6, 7)
    x = x0.copy()
    it, nrm = s.solver_bicg3(coef, jcoef, np.zeros(n), x, 1e-6, 7)
    assert it == itr == 7 and np.isinf(nrm) and np.isinf(nr)


def test_dia_layout_refused(system):
    g, coef, jcoef, s = system
    y = np.zeros(g.cells)
    with pytest.raises(NtustError, match="DIA"):
        s.matmul_ell(coef, list(jcoef) + [1], y)
    with pytest.raises(NtustError, match="outside"):
        s.matmul_ell(coef, [0, 1, g.cells, 5], y)
