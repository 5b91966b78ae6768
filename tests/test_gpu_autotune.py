"""Schur tile shapes of the outer updates per front-size bucket: 64x64 and 128x128 MFMA output
tiles (and any mix of them over the buckets) give the same factorization to rounding, and the
autotuner's choice leaves a factorization that solves to the refinement target, for one system
and for a batch."""
import numpy as np
import pytest

import magla_amd as M
import oracle as O
from magla_amd import generators as G

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ALL64 = [64] * M.SCHUR_BUCKETS


def relres(oc, b, x):
    return np.linalg.norm(b - oc.spmv(x)) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def system():
    # 3D lognormal diffusion on 28^3: fronts in the buckets f < 128 ... < 2048
    n, r, c, v = G.diffusion_3d_lognormal(28, seed=5)
    return n, M.SymmetricCSRMatrix.from_coo(n, r, c, v), O.Csr.from_coo(n, r, c, v)


# D of the two shapes differs by rounding of the two summation orders only
@pytest.mark.parametrize("precision,tol", [(M.FP64, 1e-10), (M.FP32, 1e-4)])
@pytest.mark.parametrize("tiles", [[128] * M.SCHUR_BUCKETS, [128, 64, 128, 64, 128, 64], [64, 128, 128, 64, 64, 128]])
def test_tile_shapes_give_the_same_factorization(system, precision, tol, tiles):
    n, pm, oc = system
    solver = M.SymmetricUtduSolver(M.UtduSolverOptions(precision=precision))
    sym = solver.analyze(pm)
    f64 = solver.factorize(pm, sym)
    M.test_set_schur_tiles(f64, ALL64)
    f64 = solver.factorize(pm, sym)
    d_ref = f64.diagonal
    b = G.uniform_rhs(n, 2, 3)
    x_ref = f64.solve(b)
    M.test_set_schur_tiles(f64, tiles)
    f = solver.factorize(pm, sym)
    d = f.diagonal
    assert np.max(np.abs(d - d_ref) / np.abs(d_ref)) <= tol
    if precision == M.FP64:
        x = f.solve(b)
        assert np.linalg.norm(x - x_ref) <= 1e-8 * np.linalg.norm(x_ref)
    rep = {}
    x = f.solve(b, tol=1e-10, max_refine=30, report=rep)
    assert rep["converged"]
    for k in range(2):
        assert relres(oc, b[:, k], x[:, k]) <= 1e-10
    M.test_set_schur_tiles(f, ALL64)


def test_autotune_choice_solves(system):
    n, pm, oc = system
    solver = M.SymmetricUtduSolver(M.UtduSolverOptions(precision=M.FP32))
    sym = solver.analyze(pm)
    f = solver.factorize(pm, sym)
    vals = torch.from_numpy(np.ascontiguousarray(pm.values)).to("cuda:0")
    tiles, ms = f.autotune_dev(vals.data_ptr())
    torch.cuda.synchronize()
    assert set(tiles.tolist()) <= {64, 128}
    assert tiles[-1] == 128  # no front of 2048 rows here: the bucket keeps its untuned 128x128
    assert ms > 0
    b = G.uniform_rhs(n, 3, 5)
    rep = {}
    x = f.solve(b, tol=1e-10, max_refine=30, report=rep)
    assert rep["converged"]
    for k in range(3):
        assert relres(oc, b[:, k], x[:, k]) <= 1e-10
    M.test_set_schur_tiles(f, ALL64)


def test_batch_autotune_then_solve():
    sysl = G.config_c4(16, seed=13, side_lo=20, side_hi=60, max_rhs=3)
    mats = [M.SymmetricCSRMatrix.from_coo(n, r, c, v) for n, r, c, v, b in sysl]
    batch = M.UtduBatch(mats, M.UtduSolverOptions(precision=M.FP32))
    vals = torch.from_numpy(batch.values()).to("cuda:0")
    tiles, ms = batch.autotune_dev(vals.data_ptr())
    torch.cuda.synchronize()
    assert set(tiles.tolist()) <= {64, 128} and ms > 0
    xs = batch.solve([b for *_, b in sysl], tol=1e-10, max_refine=30)
    assert batch.last_report["converged"]
    for (n, r, c, v, b), x in zip(sysl, xs):
        oc = O.Csr.from_coo(n, r, c, v)
        for k in range(b.shape[1]):
            assert relres(oc, b[:, k], x[:, k]) <= 1e-10
