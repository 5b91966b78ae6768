"""Solves with the refinement test on the device (the host reads the norms once per solve)
and, for handles factored with mg_options.device_graph, replayed from a captured graph:
every replay gives the same bits as an eager solve of the same right-hand sides, with new
values in the same buffers between replays and host work and synchronizations between them
(UtduNumericFactorization.Solve, UCHOL/UtduNumericFactorization.cs:90-112, plus FP64
refinement to the requested tolerance)."""
import numpy as np
import pytest

import magla_amd as M
import oracle as O
from magla_amd import generators as G

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
DEV = torch.device("cuda:0")


def dev_cols(b):
    """(n, k) host array -> column-major device buffer (a (k, n) contiguous tensor)."""
    return torch.from_numpy(np.ascontiguousarray(np.asarray(b).T)).to(DEV)


@pytest.mark.parametrize("graph", [1, 2])  # factorization and solves captured, or the solves only
@pytest.mark.parametrize("precision", [M.FP32, M.FP64])
def test_captured_solve_replays_equal_eager_solves(precision, graph):
    n, r, c, v, clamped, _ = G.config_c3(m=8, k=4)
    a = M.SymmetricCSRMatrix.from_coo(n, r, c, v)
    a.line_cross(np.zeros(n), 0.0, clamped)
    oc = O.Csr.from_csr(n, a.row_pointers, a.column_indices, a.values)
    vals = torch.from_numpy(a.values).to(DEV)
    eager_s = M.SymmetricUtduSolver(M.UtduSolverOptions(precision=precision))
    graph_s = M.SymmetricUtduSolver(M.UtduSolverOptions(precision=precision, device_graph=graph))
    sym = eager_s.analyze(a)
    fe = eager_s.factorize_dev(sym, vals.data_ptr(), a.nonzero_count)
    fg = graph_s.factorize_dev(sym, vals.data_ptr(), a.nonzero_count)
    k = 5
    b_dev = torch.zeros((k, n), dtype=torch.float64, device=DEV)
    xe = torch.zeros_like(b_dev)
    xg = torch.zeros_like(b_dev)
    for rep in range(4):
        b = G.uniform_rhs(n, k, 700 + rep).reshape(n, k)
        if rep == 3:
            b[:, 2:] = 0.0  # trailing zero columns: their solutions are zero
        b_dev.copy_(dev_cols(b))
        re = fe.solve_dev(b_dev.data_ptr(), xe.data_ptr(), k, tol=1e-10, max_refine=30)
        rg = fg.solve_dev(b_dev.data_ptr(), xg.data_ptr(), k, tol=1e-10, max_refine=30)
        torch.cuda.synchronize()
        assert re == rg
        assert rg["converged"] and rg["relative_residual"] <= 1e-10
        np.testing.assert_array_equal(xg.cpu().numpy(), xe.cpu().numpy())
        x = xg.cpu().numpy().T
        for col in range(k):
            nb = np.linalg.norm(b[:, col])
            res = np.linalg.norm(oc.spmv(x[:, col]) - b[:, col])
            assert res <= 1e-10 * max(nb, 1e-300) if nb > 0 else np.all(x[:, col] == 0.0)
        xe.zero_()
        xg.fill_(1.0)  # eager device work between replays


def test_refinement_beyond_the_pre_launched_rounds():
    """A solve that needs more rounds than the handle's previous one continues round by round:
    max_refine=1 first (the handle learns one round), then a tolerance that needs more."""
    n, r, c, v = G.diffusion_3d_lognormal(12, seed=5, sigma=2.0)
    a = M.SymmetricCSRMatrix.from_coo(n, r, c, v)
    oc = O.Csr.from_coo(n, r, c, v)
    vals = torch.from_numpy(a.values).to(DEV)
    s = M.SymmetricUtduSolver(M.UtduSolverOptions(precision=M.FP32, device_graph=True))
    f = s.factorize_dev(s.analyze(a), vals.data_ptr(), a.nonzero_count)
    b = G.uniform_rhs(n, 2, 3).reshape(n, 2)
    b_dev = dev_cols(b)
    x = torch.zeros_like(b_dev)
    r1 = f.solve_dev(b_dev.data_ptr(), x.data_ptr(), 2, tol=1e-15, max_refine=1)
    assert r1["refinement_steps"] == 1 and not r1["converged"]
    r2 = f.solve_dev(b_dev.data_ptr(), x.data_ptr(), 2, tol=1e-13, max_refine=30)
    torch.cuda.synchronize()
    assert r2["converged"] and r2["refinement_steps"] >= 2
    xs = x.cpu().numpy().T
    for col in range(2):
        assert np.linalg.norm(oc.spmv(xs[:, col]) - b[:, col]) <= 1e-13 * np.linalg.norm(b[:, col])


def test_captured_batch_solve_replays_equal_eager():
    sysl = G.config_c4(24, side_lo=16, side_hi=40, max_rhs=6)
    mats = [M.SymmetricCSRMatrix.from_coo(n, r, c, v) for n, r, c, v, b in sysl]
    eager = M.UtduBatch(mats, M.UtduSolverOptions(precision=M.FP32))
    graph = M.UtduBatch(mats, M.UtduSolverOptions(precision=M.FP32))
    vals = torch.from_numpy(eager.values()).to(DEV)
    eager.factorize_dev(vals.data_ptr(), device_graph=False)
    graph.factorize_dev(vals.data_ptr(), device_graph=True)
    ntot = int(eager.row_begin[-1])
    k = 6
    B = np.zeros((ntot, k))
    for t, (n, r, c, v, b) in enumerate(sysl):
        b = np.asarray(b).reshape(n, -1)
        B[eager.row_begin[t]:eager.row_begin[t + 1], :b.shape[1]] = b
    b_dev = dev_cols(B)
    xe = torch.zeros_like(b_dev)
    xg = torch.zeros_like(b_dev)
    for rep in range(3):
        b_dev.mul_(-1.0 if rep % 2 else 1.0)
        re = eager.solve_dev(b_dev.data_ptr(), xe.data_ptr(), k)
        rg = graph.solve_dev(b_dev.data_ptr(), xg.data_ptr(), k)
        torch.cuda.synchronize()
        assert re == rg and rg["converged"]
        np.testing.assert_array_equal(xg.cpu().numpy(), xe.cpu().numpy())
        xg.fill_(3.0)
