"""GPU parity of the general SpMV, ILU(0) and BiCGSTAB (libmagla.so) against the oracle's
restatement of CSRMatrix.Multiply, Ilu0Preconditioner and BiCgStabSolver, on the workbook
system, the reference's own test cases (Tests/BiCGSTAB/*Tests.cs) and nonsymmetric
convection-diffusion and random systems.

Bars: SpMV, ILU(0) factors and the ILU(0) solves are bit-identical to the oracle (one thread
per row in the reference's order, no fused multiply-add).  BiCGSTAB reduces its dot products
in a fixed tree order instead of the reference's serial sums, so its iterates agree to
rounding: iteration counts within 1, solutions within the stated tolerances."""
import numpy as np
import pytest

import magla_amd as M
import oracle as O

pytestmark = pytest.mark.gpu


def convection_diffusion(m, peclet, seed=0):
    """Upwind 5-point convection-diffusion on an m x m grid (natural order): nonsymmetric,
    an M-matrix, diagonally dominant."""
    g = np.random.default_rng(seed)
    n = m * m
    rows, cols, vals = [], [], []
    for y in range(m):
        for x in range(m):
            i = y * m + x
            wx, wy = peclet * (1 + 0.5 * g.random()), peclet * (0.5 + g.random())
            diag = 4.0 + wx + wy
            rows.append(i); cols.append(i); vals.append(diag)
            for dx, dy, w in ((-1, 0, -1.0 - wx), (1, 0, -1.0), (0, -1, -1.0 - wy), (0, 1, -1.0)):
                xx, yy = x + dx, y + dy
                if 0 <= xx < m and 0 <= yy < m:
                    rows.append(i); cols.append(yy * m + xx); vals.append(w)
    return n, np.array(rows), np.array(cols), np.array(vals)


def random_nonsymmetric(n, density, seed, dominance=1.2):
    g = np.random.default_rng(seed)
    A = (g.random((n, n)) < density) * g.uniform(-1, 1, (n, n))
    np.fill_diagonal(A, 0.0)
    A[np.arange(n), np.arange(n)] = dominance * np.abs(A).sum(axis=1) + 1.0
    return A


def pair_from_coo(n, r, c, v):
    return M.CSRMatrix.from_coo(n, n, r, c, v), O.GCsr.from_coo(n, n, r, c, v)


def pair_from_dense(A):
    return M.CSRMatrix.from_dense(A), O.GCsr.from_dense(A)


def cases(golden):
    yield "workbook", pair_from_dense(np.array(golden["matrix"]))
    yield "three", pair_from_dense(np.array([[4.0, 1.0, 0.0], [2.0, 3.0, 1.0], [0.0, -1.0, 2.0]]))
    yield "random", pair_from_dense(random_nonsymmetric(150, 0.04, 3))
    yield "convdiff", pair_from_coo(*convection_diffusion(40, 2.0, 1))


def test_spmv_ilu0_factors_and_solves_bit_exact(golden):
    g = np.random.default_rng(7)
    for name, (a, o) in cases(golden):
        n = a.row_count
        x = g.uniform(-1, 1, n)
        np.testing.assert_array_equal(a.multiply(x), o.spmv(x), err_msg=name)
        p, op = M.Ilu0Preconditioner(a), O.Ilu0(o)
        np.testing.assert_array_equal(p.factors(), op.factors, err_msg=name)
        np.testing.assert_array_equal(p.apply(x), op.apply(x), err_msg=name)


def test_ilu0_reproduces_workbook_factors(golden):
    a = M.CSRMatrix.from_dense(np.array(golden["matrix"]))
    f = M.Ilu0Preconditioner(a).factors()
    pos = {}
    for i in range(8):
        for p in range(a.row_pointers[i], a.row_pointers[i + 1]):
            pos[(i, int(a.column_indices[p]))] = p
    for i, j, v in golden["ilu0_lower"] + golden["ilu0_upper"]:
        assert f[pos[(i, j)]] == pytest.approx(v, rel=1e-14)


def test_ilu0_errors_name_the_first_bad_row():
    # Ilu0PreconditionerTests.Constructor_ThrowsWhenIlu0ProducesZeroPivot
    with pytest.raises(M.InvalidOperationError, match="pivot in row 1 "):
        M.Ilu0Preconditioner(M.CSRMatrix.from_dense(np.ones((2, 2))))
    A = random_nonsymmetric(50, 0.1, 4)
    A[17, 17] = 0.0  # structurally missing diagonal in row 17
    with pytest.raises(M.InvalidOperationError, match="Row 17 does not contain the diagonal"):
        M.Ilu0Preconditioner(M.CSRMatrix.from_dense(A))
    # a zero pivot produced by elimination, far down a level schedule
    n, r, c, v = convection_diffusion(12, 1.0, 2)
    a = M.CSRMatrix.from_coo(n, n, r, c, v)
    o = O.GCsr.from_coo(n, n, r, c, v)
    Ad = o.dense()
    row = 100
    # make u_row,row exactly zero: a_rr = sum over the eliminated part, from the oracle's factors
    f = O.Ilu0(o).factors
    rp, ci, vals = o.arrays()
    dpos = [p for p in range(rp[row], rp[row + 1]) if ci[p] == row][0]
    Ad[row, row] -= f[dpos]
    # the pivot left by rounding is far below 1e-8, every other pivot far above it
    with pytest.raises(O.OracleError) as e:
        O.Ilu0(O.GCsr.from_dense(Ad), pivot_tolerance=1e-8)
    assert e.value.row == row
    with pytest.raises(M.InvalidOperationError, match=f"pivot in row {row} "):
        M.Ilu0Preconditioner(M.CSRMatrix.from_dense(Ad), pivot_tolerance=1e-8)
    with pytest.raises(M.ArgumentError):
        M.Ilu0Preconditioner(M.CSRMatrix.from_dense(np.ones((2, 3))))


@pytest.mark.parametrize("use_pre", [True, False])
def test_bicgstab_iterations_match_oracle(golden, use_pre):
    g = np.random.default_rng(11)
    for name, (a, o) in cases(golden):
        n = a.row_count
        b = g.uniform(-1, 1, n)
        s = M.BiCgStabSolver(relative_tolerance=1e-10, max_iterations=1000, use_preconditioner=use_pre)
        x = s.solve(a, b)
        xo, ro = O.bicgstab(o, b, tol=1e-10, max_iterations=1000, use_preconditioner=use_pre)
        res = s.last_result
        assert res.converged and ro["converged"], name
        assert abs(res.iterations - ro["iterations"]) <= 1, (name, res.iterations, ro["iterations"])
        A = o.dense()
        assert np.linalg.norm(b - A @ x) <= 1.0001e-10 * np.linalg.norm(b), name
        assert res.relative_residual <= 1e-10
        np.testing.assert_allclose(x, xo, rtol=0, atol=1e-8 * np.abs(xo).max())


def test_reference_cases(golden):
    # BiCgStabSolverTests: workbook within three iterations with a prebuilt ILU(0)
    a = M.CSRMatrix.from_dense(np.array(golden["matrix"]))
    b = np.array(golden["rhs"])
    s = M.BiCgStabSolver(relative_tolerance=1e-12, max_iterations=3)
    x = s.solve_with_initial_guess(a, b, np.zeros(8), M.Ilu0Preconditioner(a))
    r = s.last_result
    assert r.converged and r.iterations <= 3
    np.testing.assert_allclose(x, golden["exact_solution"], rtol=0, atol=1e-10)
    assert r.relative_residual == pytest.approx(r.residual_norm / np.linalg.norm(b), abs=1e-15)
    # without a preconditioner
    s = M.BiCgStabSolver(relative_tolerance=1e-12, max_iterations=100)
    np.testing.assert_allclose(s.solve(a, b), golden["exact_solution"], rtol=0, atol=1e-10)
    assert s.last_result.converged
    # nonsymmetric 3x3, exact initial guess, negative MaxIterations
    A3 = np.array([[4.0, 1.0, 0.0], [2.0, 3.0, 1.0], [0.0, -1.0, 2.0]])
    a3 = M.CSRMatrix.from_dense(A3)
    expected = np.array([1.0, -2.0, 3.0])
    for use_pre in (True, False):
        s = M.BiCgStabSolver(relative_tolerance=1e-12, max_iterations=20, use_preconditioner=use_pre)
        np.testing.assert_allclose(s.solve(a3, A3 @ expected), expected, rtol=0, atol=1e-10)
    s = M.BiCgStabSolver(relative_tolerance=1e-12)
    x = s.solve_with_initial_guess(a3, A3 @ expected, expected)
    assert s.last_result.iterations == 0 and s.last_result.converged
    np.testing.assert_array_equal(x, expected)
    with pytest.raises(M.ArgumentOutOfRangeError, match="MaxIterations"):
        M.BiCgStabSolver(max_iterations=-1).solve(a3, np.zeros(3))
    with pytest.raises(M.ArgumentOutOfRangeError):
        M.BiCgStabSolver(breakdown_tolerance=0.0).solve(a3, np.ones(3))


def test_breakdown_and_argument_errors():
    skew = M.CSRMatrix.from_dense(np.array([[0.0, 1.0], [-1.0, 0.0]]))
    s = M.BiCgStabSolver(use_preconditioner=False)
    with pytest.raises(M.InvalidOperationError, match=r"\(r0, A M\^-1 p\) is zero or non-finite at iteration 1"):
        s.solve(skew, np.array([1.0, 2.0]))
    with pytest.raises(M.ArgumentError, match="square"):
        M.BiCgStabSolver().solve(M.CSRMatrix.from_dense(np.ones((2, 3))), np.ones(2))
    a = M.CSRMatrix.from_dense(np.array([[2.0, 1.0], [0.0, 3.0]]))
    other = M.CSRMatrix.from_dense(np.array([[2.0, 1.0], [0.0, 3.0]]))
    with pytest.raises(M.ArgumentError, match="another matrix"):
        M.BiCgStabSolver().solve_with_initial_guess(a, np.ones(2), None, M.Ilu0Preconditioner(other))


def test_not_converged_returns_last_iterate():
    n, r, c, v = convection_diffusion(30, 3.0, 5)
    a, o = pair_from_coo(n, r, c, v)
    b = np.random.default_rng(2).uniform(-1, 1, n)
    s = M.BiCgStabSolver(relative_tolerance=1e-14, max_iterations=3, use_preconditioner=False)
    x = s.solve(a, b)
    xo, ro = O.bicgstab(o, b, tol=1e-14, max_iterations=3, use_preconditioner=False)
    assert not s.last_result.converged and s.last_result.iterations == 3 == ro["iterations"]
    np.testing.assert_allclose(x, xo, rtol=1e-10, atol=1e-12)
    assert s.last_result.residual_norm == pytest.approx(ro["residual_norm"], rel=1e-9)
