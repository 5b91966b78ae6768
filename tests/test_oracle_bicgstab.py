"""Pins the oracle's restatement of CSRMatrix, Ilu0Preconditioner and BiCgStabSolver
(CAESolvers/CSRMatrix.cs, BiCGSTAB/Ilu0Preconditioner.cs, BiCGSTAB/BiCgStabSolver.cs):
the workbook's ILU(0) factors (sheet BICGSTAB, held by Tests/BiCGSTAB/CgComparisonData.cs),
the reference's own test cases (Tests/BiCGSTAB/*Tests.cs), and what the mathematics fixes:
ILU(0) reproduces A on A's pattern, equals the exact LU when elimination makes no fill, and
BiCGSTAB's solution is A^-1 b."""
import numpy as np
import pytest

import oracle as O


def workbook(golden):
    return O.GCsr.from_dense(np.array(golden["matrix"]))


def lu_of(a: O.GCsr, factors):
    rp, ci, _ = a.arrays()
    n = a.shape[0]
    L, U = np.eye(n), np.zeros((n, n))
    for i in range(n):
        for p in range(rp[i], rp[i + 1]):
            if ci[p] < i:
                L[i, ci[p]] = factors[p]
            else:
                U[i, ci[p]] = factors[p]
    return L, U


def random_nonsymmetric(n, density, seed, dominance=1.5):
    g = np.random.default_rng(seed)
    A = (g.random((n, n)) < density) * g.uniform(-1, 1, (n, n))
    np.fill_diagonal(A, 0.0)
    A[np.arange(n), np.arange(n)] = dominance * np.abs(A).sum(axis=1) + 1.0
    return A


def test_ilu0_matches_workbook_factors(golden):
    a = workbook(golden)
    f = O.Ilu0(a).factors
    L, U = lu_of(a, f)
    for i, j, v in golden["ilu0_lower"]:
        assert L[i, j] == pytest.approx(v, rel=1e-14)
    for i, j, v in golden["ilu0_upper"]:
        assert U[i, j] == pytest.approx(v, rel=1e-14)
    # every factor entry is one of the workbook's (the rest of the pattern is zero there)
    listed = {(i, j) for i, j, _ in golden["ilu0_lower"] + golden["ilu0_upper"]}
    rp, ci, _ = a.arrays()
    for i in range(8):
        for p in range(rp[i], rp[i + 1]):
            if (i, ci[p]) not in listed:
                assert f[p] == 0.0


def test_ilu0_apply_reproduces_workbook_solution(golden):
    # Ilu0PreconditionerTests.Apply_ReproducesFactorsFromCgComparisonWorkbook
    a = workbook(golden)
    L, U = np.eye(8), np.zeros((8, 8))
    for i, j, v in golden["ilu0_lower"]:
        L[i, j] = v
    for i, j, v in golden["ilu0_upper"]:
        U[i, j] = v
    expected = np.array([1.0, -2.0, 3.0, -4.0, 5.0, -6.0, 7.0, -8.0])
    x = O.Ilu0(a).apply(L @ (U @ expected))
    np.testing.assert_allclose(x, expected, rtol=0, atol=1e-11)


def test_ilu0_zero_pivot_and_missing_diagonal():
    # Ilu0PreconditionerTests.Constructor_ThrowsWhenIlu0ProducesZeroPivot
    with pytest.raises(O.OracleError) as e:
        O.Ilu0(O.GCsr.from_dense(np.ones((2, 2))))
    assert e.value.code == 4 and e.value.row == 1
    with pytest.raises(O.OracleError) as e:
        O.Ilu0(O.GCsr.from_dense(np.array([[1.0, 2.0, 0.0], [3.0, 0.0, 1.0], [0.0, 1.0, 1.0]])))
    assert e.value.code == 4 and e.value.row == 1


@pytest.mark.parametrize("n,density,seed", [(30, 0.1, 1), (60, 0.05, 2), (25, 0.4, 3)])
def test_ilu0_reproduces_a_on_its_pattern(n, density, seed):
    A = random_nonsymmetric(n, density, seed)
    a = O.GCsr.from_dense(A)
    L, U = lu_of(a, O.Ilu0(a).factors)
    LU = L @ U
    mask = A != 0
    np.testing.assert_allclose(LU[mask], A[mask], rtol=1e-12, atol=1e-12 * np.abs(A).max())


def test_ilu0_is_exact_lu_without_fill():
    # tridiagonal (and upper/lower bidiagonal parts): Gaussian elimination makes no fill
    g = np.random.default_rng(5)
    n = 40
    A = np.diag(g.uniform(3, 4, n)) + np.diag(g.uniform(-1, 1, n - 1), 1) + np.diag(g.uniform(-1, 1, n - 1), -1)
    a = O.GCsr.from_dense(A)
    p = O.Ilu0(a)
    L, U = lu_of(a, p.factors)
    np.testing.assert_allclose(L @ U, A, rtol=0, atol=1e-14 * np.abs(A).max())
    b = g.uniform(-1, 1, n)
    np.testing.assert_allclose(p.apply(b), np.linalg.solve(A, b), rtol=1e-12)


def test_coo_semantics():
    # CSRMatrix(int, int, IEnumerable<MatrixEntry>): duplicates summed, cancellation dropped, no
    # symmetric normalisation (SparseMatrixCSRTests.CoordinateConstructor_SumsDuplicateEntries)
    a = O.GCsr.from_coo(3, 4, [0, 2, 0, 1, 1, 1], [1, 3, 1, 0, 2, 2], [1.0, 5.0, 2.0, -1.0, 1e-16, 1e-16])
    rp, ci, v = a.arrays()
    assert rp.tolist() == [0, 1, 2, 3]
    assert ci.tolist() == [1, 0, 3]
    assert v.tolist() == [3.0, -1.0, 5.0]
    with pytest.raises(O.OracleError):
        O.GCsr.from_coo(2, 2, [0], [2], [1.0])


def test_linecross_general():
    A = random_nonsymmetric(6, 0.5, 9)
    a = O.GCsr.from_dense(A)
    rhs = np.arange(6, dtype=np.float64)
    a.linecross(rhs, 2.5, 3)
    D = a.dense()
    exp_rhs = np.arange(6, dtype=np.float64) - A[:, 3] * 2.5
    exp_rhs[3] = 2.5
    np.testing.assert_allclose(rhs, exp_rhs, rtol=1e-15)
    assert D[3, 3] == 1.0 and not np.any(np.delete(D[3], 3)) and not np.any(np.delete(D[:, 3], 3))


# ---- BiCGSTAB: the reference's tests (Tests/BiCGSTAB/BiCgStabSolverTests.cs) ----

def three_by_three():
    return np.array([[4.0, 1.0, 0.0], [2.0, 3.0, 1.0], [0.0, -1.0, 2.0]])


def test_workbook_converges_within_three_iterations(golden):
    a = workbook(golden)
    b = np.array(golden["rhs"])
    x, res = O.bicgstab(a, b, x0=np.zeros(8), tol=1e-12, max_iterations=3, preconditioner=O.Ilu0(a))
    assert res["converged"] and res["iterations"] <= 3
    np.testing.assert_allclose(x, golden["exact_solution"], rtol=0, atol=1e-10)
    assert res["relative_residual"] == pytest.approx(res["residual_norm"] / np.linalg.norm(b), abs=1e-15)


def test_workbook_without_preconditioner(golden):
    a = workbook(golden)
    x, res = O.bicgstab(a, np.array(golden["rhs"]), tol=1e-12, max_iterations=100, use_preconditioner=False)
    assert res["converged"]
    np.testing.assert_allclose(x, golden["exact_solution"], rtol=0, atol=1e-10)


@pytest.mark.parametrize("use_pre", [True, False])
def test_nonsymmetric_three_by_three(use_pre):
    A = three_by_three()
    expected = np.array([1.0, -2.0, 3.0])
    x, res = O.bicgstab(O.GCsr.from_dense(A), A @ expected, tol=1e-12, max_iterations=20,
                        use_preconditioner=use_pre)
    np.testing.assert_allclose(x, expected, rtol=0, atol=1e-10)


def test_exact_initial_guess_needs_no_iteration():
    A = three_by_three()
    expected = np.array([1.0, -2.0, 3.0])
    x, res = O.bicgstab(O.GCsr.from_dense(A), A @ expected, x0=expected, tol=1e-12)
    assert res["iterations"] == 0 and res["converged"]
    assert np.array_equal(x, expected)


def test_negative_max_iterations_and_bad_tolerances():
    a = O.GCsr.from_dense(three_by_three())
    for kw in (dict(max_iterations=-1), dict(breakdown_tolerance=0.0), dict(tol=-1.0)):
        with pytest.raises(O.OracleError) as e:
            O.bicgstab(a, np.zeros(3), **kw)
        assert e.value.code == 3


def test_skew_symmetric_breaks_down_at_the_first_alpha():
    # (r0, A r0) = 0 for skew-symmetric A: "(r0, A M^-1 p)" breakdown at iteration 1
    with pytest.raises(O.OracleError) as e:
        O.bicgstab(O.GCsr.from_dense(np.array([[0.0, 1.0], [-1.0, 0.0]])), np.array([1.0, 2.0]),
                   use_preconditioner=False)
    assert e.value.fail_code == 3 and e.value.fail_it == 1


@pytest.mark.parametrize("use_pre", [True, False])
def test_random_nonsymmetric_matches_dense_solve(use_pre):
    A = random_nonsymmetric(80, 0.06, 11, dominance=1.2)
    b = np.random.default_rng(12).uniform(-1, 1, 80)
    x, res = O.bicgstab(O.GCsr.from_dense(A), b, tol=1e-12, max_iterations=500, use_preconditioner=use_pre)
    assert res["converged"]
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-11)
    assert np.linalg.norm(b - A @ x) <= 1.0001e-12 * np.linalg.norm(b)
