"""The benchmark workloads on the GPU, at reduced and at full size.

* 3D linear elasticity (trilinear hex elements, 3 DOF per node, x = 0 face clamped through
  SymmetricCSRMatrix.LineCross, SymmetricCSRMatrix.cs:125-167):
  - 10^3 elements (n = 3,993): every FP64 factor block and D against the oracle at 1e-12, and
    FP32 factor + FP64 refinement to 1e-10 against the oracle's FP64 solution.  Its widest
    front (711 columns) has outer Schur steps whose trailing part overlaps the next panels.
  - 40^3 elements (n = 206,763, 8 right-hand sides): FP32 + refinement to 1e-10 on every
    column, FP64 to 1e-12, and the two solutions within 1e-8 of each other.
* 1,024 uneven 2D systems (side 16..128, 1..32 right-hand sides): FP32 + refinement to 1e-10
  for every system and column, and the FP64 batch bitwise equal to single-system solves on a
  seeded sample of 16 systems.

Residuals are formed by the oracle's SpMV (SymmetricCSRMatrix.Multiply restated) in FP64.
"""
import numpy as np
import pytest

import magla_amd as M
import oracle as O
from magla_amd import generators as G

pytestmark = pytest.mark.gpu


def elasticity(m, k):
    n, r, c, v, clamped, b = G.config_c3(m=m, k=k)
    a = M.SymmetricCSRMatrix.from_coo(n, r, c, v)
    a.line_cross(np.zeros(n), 0.0, clamped)
    oc = O.Csr.from_csr(n, a.row_pointers, a.column_indices, a.values)
    return a, oc, np.asfortranarray(b)


def relres(oc, b, x):
    return np.linalg.norm(b - oc.spmv(x)) / np.linalg.norm(b)


def test_elasticity_small_fp64_factor_blocks_match_oracle():
    a, oc, b = elasticity(10, 3)
    solver = M.SymmetricUtduSolver()
    sym = solver.analyze(a)
    assert M.plan_stats(sym)["lookahead_tiles"] > 0  # the overlapped trailing updates run here
    f = solver.factorize(a, sym)
    os_ = O.Symbolic(oc)
    of = O.Numeric(os_, oc)
    np.testing.assert_allclose(f.diagonal, of.diag, rtol=1e-12)
    st = sym.structure()
    np.testing.assert_array_equal(st["first"], os_.first)
    for s in range(os_.nsuper):
        fs = int(st["pat_ptr"][s + 1] - st["pat_ptr"][s])
        w = int(st["first"][s + 1] - st["first"][s])
        F = M.test_factor_block(f, s, fs, w)
        blk = of.factor[os_.block_off[s]:os_.block_off[s + 1]]
        off = 0
        for k in range(w):
            np.testing.assert_allclose(F[k:, k], blk[off:off + fs - k], rtol=1e-12, atol=1e-12)
            off += fs - k
    X = f.solve(b)
    for k in range(b.shape[1]):
        assert relres(oc, b[:, k], X[:, k]) <= 1e-12
        xo = of.solve(b[:, k])
        assert np.linalg.norm(X[:, k] - xo) <= 1e-11 * np.linalg.norm(xo)


def test_elasticity_small_fp32_refined_matches_oracle_solution():
    a, oc, b = elasticity(10, 3)
    of = O.Numeric(O.Symbolic(oc), oc)
    f32 = M.SymmetricUtduSolver(M.UtduSolverOptions(precision=M.FP32)).factorize(a)
    np.testing.assert_allclose(f32.diagonal, of.diag, rtol=2e-3)
    rep = {}
    X = f32.solve(b, tol=1e-10, max_refine=30, report=rep)
    assert rep["converged"] and 1 <= rep["refinement_steps"] <= 6
    for k in range(b.shape[1]):
        assert relres(oc, b[:, k], X[:, k]) <= 1e-10
        xo = of.solve(b[:, k])
        assert np.linalg.norm(X[:, k] - xo) <= 1e-8 * np.linalg.norm(xo)


def test_elasticity_full_size_fp32_refined_and_fp64():
    a, oc, b = elasticity(40, 8)
    assert a.size == 206763 and b.shape == (206763, 8)
    sym = M.SymmetricUtduSolver().analyze(a)
    f64 = M.SymmetricUtduSolver(M.UtduSolverOptions(precision=M.FP64)).factorize(a, sym)
    X64 = f64.solve(b)
    f32 = M.SymmetricUtduSolver(M.UtduSolverOptions(precision=M.FP32)).factorize(a, sym)
    rep = {}
    X32 = f32.solve(b, tol=1e-10, max_refine=30, report=rep)
    assert rep["converged"] and rep["relative_residual"] <= 1e-10
    assert f64.is_positive_definite and f32.is_positive_definite
    for k in range(8):
        assert relres(oc, b[:, k], X64[:, k]) <= 1e-12
        assert relres(oc, b[:, k], X32[:, k]) <= 1e-10
        assert np.linalg.norm(X32[:, k] - X64[:, k]) <= 1e-8 * np.linalg.norm(X64[:, k])


def test_batched_full_size_every_system_and_column():
    sysl = G.config_c4()
    assert len(sysl) == 1024
    mats = [M.SymmetricCSRMatrix.from_coo(n, r, c, v) for n, r, c, v, _ in sysl]
    bs = [b for *_, b in sysl]
    b32 = M.UtduBatch(mats, M.UtduSolverOptions(precision=M.FP32))
    b32.factorize()
    xs = b32.solve(bs, tol=1e-10, max_refine=30)
    assert b32.last_report["converged"] and b32.last_report["relative_residual"] <= 1e-10
    for (n, r, c, v, b), x in zip(sysl, xs):
        oc = O.Csr.from_coo(n, r, c, v)
        R = b - np.column_stack([oc.spmv(x[:, k]) for k in range(b.shape[1])])
        assert np.all(np.linalg.norm(R, axis=0) <= 1.0001e-10 * np.linalg.norm(b, axis=0))
    b64 = M.UtduBatch(mats, M.UtduSolverOptions(precision=M.FP64))
    b64.factorize()
    x64 = b64.solve(bs, tol=0.0, max_refine=0)
    for t in np.random.default_rng(20260816).choice(len(sysl), 16, replace=False):
        x1 = M.SymmetricUtduSolver().factorize(mats[t]).solve(bs[t])
        np.testing.assert_array_equal(x64[t], x1)
