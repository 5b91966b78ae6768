"""GPU parity of the UCHOL numeric phase and solves (libmagla.so on the device)
against the CPU oracle, the golden 8x8 system and the reference's own test
assertions (Tests/UCHOL/SymmetricUtduSolverTests.cs).

Tolerances: FP64 factor/solve results are compared at 1e-12 relative (the
device sums in a different order than the reference; the reference's own
tests use 1e-9 against dense elimination and 1e-12 relative residual).
"""
import numpy as np
import pytest

import magla_amd as M
import oracle as O
from conftest import dense_coo, grid_matrix, random_matrix, tridiagonal, vector
from magla_amd import generators as G

pytestmark = pytest.mark.gpu


def build(n, r, c, v):
    return M.SymmetricCSRMatrix.from_coo(n, r, c, v), O.Csr.from_coo(n, r, c, v)


def relres(oc, b, x):
    return np.linalg.norm(b - oc.spmv(x)) / max(np.linalg.norm(b), 1e-300)


def max_rel_diff(a, e):
    return np.max(np.abs(a - e)) / max(np.max(np.abs(e)), 1e-300)


def test_golden_8x8_natural_order_factor_and_solve(golden):
    A = np.array(golden["matrix"])
    pm, oc = build(8, *dense_coo(A))
    solver = M.SymmetricUtduSolver(M.UtduSolverOptions(ordering=M.NATURAL))
    f = solver.factorize(pm)
    # elimination order [0,1,3,2,4,...]: D[2], D[3] are equal, so D matches the workbook as listed
    np.testing.assert_allclose(f.diagonal, golden["ldlt_diagonal"], rtol=1e-13)
    x = f.solve(np.array(golden["rhs"]))
    np.testing.assert_allclose(x, golden["exact_solution"], rtol=0, atol=1e-13)
    assert f.is_positive_definite


@pytest.mark.parametrize("ordering", [M.NATURAL, M.AMD])
@pytest.mark.parametrize("n,density", [(40, 0.02), (55, 0.15), (45, 0.90)])
def test_random_sparse_matches_dense_elimination(n, density, ordering):
    # SymmetricUtduSolverTests.cs:280-301
    pm, oc = build(*random_matrix(n, density, seed=n * 31 + int(density * 1000)))
    b = vector(n, 7)
    x = M.SymmetricUtduSolver(M.UtduSolverOptions(ordering=ordering)).solve(pm, b)
    expected = np.linalg.solve(oc.dense(), b)
    assert max_rel_diff(x, expected) < 1e-9
    assert relres(oc, b, x) < 1e-12


def test_c1_laplacian_matches_oracle():
    n, r, c, v, b = G.config_c1()
    pm, oc = build(n, r, c, v)
    f = M.SymmetricUtduSolver().factorize(pm)
    of = O.Numeric(O.Symbolic(oc), oc)
    np.testing.assert_allclose(f.diagonal, of.diag, rtol=1e-12)
    x = f.solve(b)
    assert relres(oc, b, x) <= 1e-12
    assert max_rel_diff(x, of.solve(b)) < 1e-12
    assert f.smallest_pivot_magnitude == pytest.approx(of.smallest, rel=1e-12)
    assert f.largest_pivot_magnitude == pytest.approx(of.largest, rel=1e-12)


@pytest.mark.parametrize("scenario", [0, 1, 2, 3])
def test_various_structures(scenario):
    # SymmetricUtduSolverTests.cs:337-377
    if scenario == 0:
        args = tridiagonal(400, 4.0, -1.0)
    elif scenario == 1:
        r = np.concatenate([np.zeros(299, int), [0], np.arange(1, 300)])
        c = np.concatenate([np.arange(1, 300), [0], np.arange(1, 300)])
        v = np.concatenate([np.full(299, -1.0), [301.0], np.full(299, 2.0)])
        args = (300, r, c, v)
    elif scenario == 2:
        args = grid_matrix(15, 15, 2, 2)
    else:
        args = grid_matrix(8, 8, 8, 3)
    pm, oc = build(*args)
    b = vector(pm.size, scenario + 3)
    f = M.SymmetricUtduSolver().factorize(pm)
    x = f.solve(b)
    assert relres(oc, b, x) < 1e-10
    assert f.is_positive_definite and f.regularized_pivot_count == 0
    of = O.Numeric(O.Symbolic(oc), oc)
    np.testing.assert_allclose(f.diagonal, of.diag, rtol=1e-12)


@pytest.mark.parametrize("diag,expected", [([2.0, 4.0, 5.0], [3.0, 1.0, 2.0]), ([4.0], [2.0]),
                                           ([1e-6, 1e6], [1.0, -1.0])])
def test_diagonal_system_exact(diag, expected):
    n = len(diag)
    pm, oc = build(n, np.arange(n), np.arange(n), np.array(diag))
    b = oc.spmv(np.array(expected))
    x = M.SymmetricUtduSolver().solve(pm, b)
    for i in range(n):
        assert abs(x[i] - expected[i]) <= 1e-9 * abs(expected[i]) + 1e-12


def test_empty_system():
    pm = M.SymmetricCSRMatrix.from_coo(0, [], [], [])
    x = M.SymmetricUtduSolver().solve(pm, np.zeros(0))
    assert x.shape == (0,)


def test_zero_rhs_gives_exact_zeros():
    pm, _ = build(*grid_matrix(5, 5, 5, 2))
    x = M.SymmetricUtduSolver().solve(pm, np.zeros(pm.size))
    assert np.all(x == 0.0)


def test_singular_counts_regularization_or_throws():
    # [[1,1],[1,1]] + 2 + 3: last pivot of the block is exactly 0 -> one regularized, none negative
    r, c, v = [0, 1, 0, 2, 3], [0, 1, 1, 2, 3], [1.0, 1.0, 1.0, 2.0, 3.0]
    pm, oc = build(4, r, c, v)
    f = M.SymmetricUtduSolver(M.UtduSolverOptions(allow_diagonal_regularization=True)).factorize(pm)
    of = O.Numeric(O.Symbolic(oc), oc)
    assert (f.regularized_pivot_count, f.negative_pivot_count) == (of.regularized, of.negative) == (1, 0)
    assert not f.is_positive_definite
    with pytest.raises(M.InvalidOperationError):
        M.SymmetricUtduSolver(M.UtduSolverOptions(allow_diagonal_regularization=False)).factorize(pm)


def test_indefinite_reports_negative_pivot():
    r, c, v = [0, 1, 2, 0], [0, 1, 2, 1], [1.0, 1.0, 1.0, 3.0]
    pm, oc = build(3, r, c, v)
    f = M.SymmetricUtduSolver().factorize(pm)
    of = O.Numeric(O.Symbolic(oc), oc)
    assert (f.regularized_pivot_count, f.negative_pivot_count) == (of.regularized, of.negative) == (1, 1)
    assert not f.is_positive_definite


def test_reused_symbolic_after_values_change():
    n, r, c, v = grid_matrix(8, 8, 5, 2)
    pm, oc = build(n, r, c, v)
    solver = M.SymmetricUtduSolver()
    sym = solver.analyze(pm)
    b = vector(n, 31)
    first = solver.factorize(pm, sym).solve(b)
    assert relres(oc, b, first) < 1e-10
    diag_pos = pm.row_pointers[:-1]  # diagonal is the first entry of each row here
    pm.values[diag_pos] += 5.0
    oc2 = O.Csr.from_csr(n, pm.row_pointers, pm.column_indices, pm.values)
    second = solver.factorize(pm, sym).solve(b)
    assert relres(oc2, b, second) < 1e-10
    assert max_rel_diff(second, first) > 1e-6


def test_structure_mismatch_and_rhs_length_throw():
    solver = M.SymmetricUtduSolver()
    sym = solver.analyze(build(*tridiagonal(100, 4.0, -1.0))[0])
    other = build(*grid_matrix(5, 5, 4, 1))[0]
    with pytest.raises(M.ArgumentError):
        solver.factorize(other, sym)
    f = solver.factorize(build(*tridiagonal(8, 4.0, -1.0))[0])
    for length in (3, 12):
        with pytest.raises(M.ArgumentError):
            f.solve(np.zeros(length))


def test_multiple_rhs_on_one_factorization():
    n, r, c, v = grid_matrix(8, 8, 6, 2)
    pm, oc = build(n, r, c, v)
    f = M.SymmetricUtduSolver().factorize(pm)
    B = np.column_stack([vector(n, 100 + k) for k in range(5)])
    X = f.solve(B)
    for k in range(5):
        assert relres(oc, B[:, k], X[:, k]) < 1e-10
        np.testing.assert_allclose(X[:, k], f.solve(B[:, k]), rtol=0, atol=1e-14 * np.abs(X[:, k]).max())


@pytest.mark.parametrize("k", [9, 20, 37])
def test_refinement_with_many_rhs(k):
    # residual passes of 8 and 16 columns with a ragged last pass; every column refined on its own
    n, r, c, v = G.diffusion_3d_lognormal(12, seed=4)
    pm, oc = build(n, r, c, v)
    B = G.uniform_rhs(n, k, 21)
    B[:, 3] = 0.0  # a zero column stays exactly zero
    f = M.SymmetricUtduSolver(M.UtduSolverOptions(precision=M.FP32)).factorize(pm)
    rep = {}
    X = f.solve(B, tol=1e-10, max_refine=30, report=rep)
    assert rep["converged"] and rep["refinement_steps"] >= 1
    assert not X[:, 3].any()
    for j in range(k):
        if j != 3:
            assert relres(oc, B[:, j], X[:, j]) <= 1e-10
    X1 = f.solve(B[:, k - 1], tol=1e-10, max_refine=30)
    assert relres(oc, B[:, k - 1], X1) <= 1e-10


def test_diffusion_3d_matches_oracle():
    n, r, c, v = G.diffusion_3d_lognormal(20, seed=5)
    pm, oc = build(n, r, c, v)
    f = M.SymmetricUtduSolver().factorize(pm)
    of = O.Numeric(O.Symbolic(oc), oc)
    np.testing.assert_allclose(f.diagonal, of.diag, rtol=1e-11)
    b = G.uniform_rhs(n, 4, 9)
    X = f.solve(b)
    for k in range(4):
        assert relres(oc, b[:, k], X[:, k]) <= 1e-12


@pytest.mark.parametrize("case", ["c1", "grid", "diffusion"])
def test_fp32_factor_with_fp64_refinement(case):
    if case == "c1":
        n, r, c, v, _ = G.config_c1()
    elif case == "grid":
        n, r, c, v = grid_matrix(8, 8, 8, 3)
    else:
        n, r, c, v = G.diffusion_3d_lognormal(16, seed=3)
    pm, oc = build(n, r, c, v)
    b = G.uniform_rhs(n, 3, 11)
    f32 = M.SymmetricUtduSolver(M.UtduSolverOptions(precision=M.FP32)).factorize(pm)
    rep = {}
    X = f32.solve(b, tol=1e-10, max_refine=30, report=rep)
    assert rep["converged"] and rep["relative_residual"] <= 1e-10
    for k in range(3):
        assert relres(oc, b[:, k], X[:, k]) <= 1e-10
    # D of the FP32 factor agrees with the FP64 oracle to single precision
    of = O.Numeric(O.Symbolic(oc), oc)
    np.testing.assert_allclose(f32.diagonal, of.diag, rtol=2e-4)


@pytest.mark.parametrize("precision", [M.FP64, M.FP32])
def test_mfma_accumulator_layout(precision):
    d = M.test_mfma_layout(precision)
    i, j = np.meshgrid(np.arange(16), np.arange(16), indexing="ij")
    np.testing.assert_array_equal(d, i + 16 * j)


@pytest.mark.parametrize("ordering", [M.NATURAL, M.AMD])
@pytest.mark.parametrize("case", ["random", "grid", "dense"])
def test_factor_blocks_match_oracle(ordering, case):
    """Every factor column of every supernode against the oracle's packed blocks (covers fused
    chains, multi-panel and blocked Schur paths: the dense case is one 300-column front)."""
    if case == "random":
        n, r, c, v = random_matrix(55, 0.15, 7)
    elif case == "grid":
        n, r, c, v = grid_matrix(9, 9, 9, 2)
    else:
        n, r, c, v = random_matrix(300, 1.0, 8)
    pm, oc = build(n, r, c, v)
    solver = M.SymmetricUtduSolver(M.UtduSolverOptions(ordering=ordering))
    sym = solver.analyze(pm)
    f = solver.factorize(pm, sym)
    st = sym.structure()
    os_ = O.Symbolic(oc, O.options(ordering=ordering))
    of = O.Numeric(os_, oc)
    for s in range(os_.nsuper):
        fs = int(st["pat_ptr"][s + 1] - st["pat_ptr"][s])
        w = int(st["first"][s + 1] - st["first"][s])
        F = M.test_factor_block(f, s, fs, w)
        blk = of.factor[os_.block_off[s]:os_.block_off[s + 1]]
        off = 0
        for k in range(w):
            np.testing.assert_allclose(F[k:, k], blk[off:off + fs - k], rtol=1e-12, atol=1e-13)
            off += fs - k
