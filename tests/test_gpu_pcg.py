"""GPU parity of SpMV and Jacobi-PCG (CGGP) against the oracle and the
reference's asserted behaviour (Tests/CGGP/*.cs, Tests/SymmetricLinearSolverInterfaceTests.cs)
and the workbook's CG iterates on the 8x8 system."""
import numpy as np
import pytest

import magla_amd as M
import oracle as O
from conftest import dense_coo, grid_matrix, random_matrix, tridiagonal, vector
from magla_amd import generators as G

pytestmark = pytest.mark.gpu


def build(n, r, c, v):
    return M.SymmetricCSRMatrix.from_coo(n, r, c, v), O.Csr.from_coo(n, r, c, v)


@pytest.mark.parametrize("args", [grid_matrix(6, 5, 4, 3), random_matrix(300, 0.02, 4), tridiagonal(1000, 4, -1)])
def test_spmv_matches_oracle(args):
    pm, oc = build(*args)
    x = vector(pm.size, 1)
    y = M.DeviceMatrix(pm).multiply(x)
    e = oc.spmv(x)
    np.testing.assert_allclose(y, e, rtol=1e-13, atol=1e-13 * np.abs(e).max())


@pytest.mark.parametrize("diag,b,x0,iters", [([2.0, 4.0, 5.0], [3.0, 1.0, 2.0], None, 1),
                                             ([2.0, 3.0], [5.0, -2.0], [5.0, -2.0], 0),
                                             ([4.0, 4.0, 4.0], [0.0, 0.0, 0.0], None, 0)])
def test_diagonal_exact_iteration_counts(diag, b, x0, iters):
    # ConjugateGradientGaussPreSolverTests.cs:54-77 (b given as the solution: A x = A*sol)
    n = len(diag)
    pm, oc = build(n, np.arange(n), np.arange(n), np.array(diag))
    sol = np.array(b)
    rhs = oc.spmv(sol)
    s = M.ConjugateGradientGaussPreSolver()
    x = s.solve(pm, rhs, initial_guess=None if x0 is None else np.array(x0))
    assert s.last_result.converged
    assert s.last_result.iterations == iters
    np.testing.assert_allclose(x, sol, atol=1e-9)


@pytest.mark.parametrize("jacobi", [True, False])
def test_tridiagonal_converges(jacobi):
    pm, oc = build(*tridiagonal(50, 4.0, -1.0))
    b = vector(50, 3)
    s = M.ConjugateGradientGaussPreSolver(relative_tolerance=1e-10, use_preconditioner=jacobi)
    x = s.solve(pm, b)
    r = s.last_result
    assert r.converged and r.relative_residual <= 1e-10
    np.testing.assert_allclose(oc.spmv(x), b, atol=1e-6)
    _, ores = O.pcg(oc, b, tol=1e-10, use_jacobi=jacobi)
    assert abs(r.iterations - ores["iterations"]) <= 1


def test_golden_cg_iterates(golden):
    A = np.array(golden["matrix"])
    pm, oc = build(8, *dense_coo(A))
    b = np.array(golden["rhs"])
    nz = golden["cg_nonzero_components"]
    for jacobi in (True, False):  # constant diagonal: Jacobi-PCG iterates equal CG iterates
        for k, key in ((3, "cg_x_after_3"), (4, "cg_x_after_4")):
            s = M.ConjugateGradientGaussPreSolver(relative_tolerance=1e-14, max_iterations=k,
                                                  use_preconditioner=jacobi)
            x = s.solve(pm, b)
            assert s.last_result.iterations == k and not s.last_result.converged
            np.testing.assert_allclose(x[nz], golden[key], rtol=1e-11)
            ox, _ = O.pcg(oc, b, tol=1e-14, max_iterations=k, use_jacobi=jacobi)
            np.testing.assert_allclose(x, ox, rtol=1e-12, atol=1e-15)


def test_c1_pcg_iterations_match_oracle():
    n, r, c, v, b = G.config_c1()
    pm, oc = build(n, r, c, v)
    s = M.ConjugateGradientGaussPreSolver(relative_tolerance=1e-8)
    x = s.solve(pm, b)
    ox, ores = O.pcg(oc, b, tol=1e-8)
    assert s.last_result.converged
    assert abs(s.last_result.iterations - ores["iterations"]) <= 1
    assert np.linalg.norm(b - oc.spmv(x)) / np.linalg.norm(b) <= 1.0001e-8


def test_not_positive_definite_throws():
    pm, _ = build(2, [0, 1, 0], [0, 1, 1], [1.0, 1.0, 3.0])
    s = M.ConjugateGradientGaussPreSolver(use_preconditioner=False)
    with pytest.raises(M.InvalidOperationError):
        s.solve(pm, np.array([1.0, -1.0]))


def test_negative_max_iterations_and_zero_diagonal():
    pm, _ = build(1, [0], [0], [1.0])
    with pytest.raises(M.ArgumentOutOfRangeError):
        M.ConjugateGradientGaussPreSolver(max_iterations=-1).solve(pm, np.array([1.0]))
    pm2, _ = build(2, [0, 0], [1, 0], [1.0, 2.0])  # A[1,1] absent
    with pytest.raises(M.InvalidOperationError):
        M.ConjugateGradientGaussPreSolver().solve(pm2, np.array([1.0, 1.0]))


def test_non_converged_returns_last_iterate():
    pm, oc = build(*grid_matrix(6, 6, 6, 1))
    b = vector(pm.size, 5)
    s = M.ConjugateGradientGaussPreSolver(relative_tolerance=1e-14, max_iterations=3)
    x = s.solve(pm, b)
    assert not s.last_result.converged and s.last_result.iterations == 3
    ox, ores = O.pcg(oc, b, tol=1e-14, max_iterations=3)
    np.testing.assert_allclose(x, ox, rtol=1e-12, atol=1e-14)
    assert s.last_result.residual_norm == pytest.approx(ores["residual_norm"], rel=1e-10)
