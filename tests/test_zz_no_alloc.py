"""mg_solve_dev_noalloc: the solve with its workspace in the library's static device scratch
must give exactly what mg_solve_dev gives (same kernels, same order: bit-identical), on two
calls in a row, the second with a smaller input (the scratch keeps the first call's
contents), and must refuse a request larger than the scratch before touching the GPU."""
import numpy as np
import pytest

import magla_amd as M
from magla_amd import generators as G

pytestmark = pytest.mark.gpu


def _case(torch, n, r, c, v, k, precision, seed):
    a = M.SymmetricCSRMatrix.from_coo(n, r, c, v)
    f = M.SymmetricUtduSolver(M.UtduSolverOptions(precision=precision)).factorize(a)
    dev = torch.device("cuda:0")
    B = torch.from_numpy(np.ascontiguousarray(G.uniform_rhs(n, k, seed).reshape(n, -1).T)).to(dev)
    return f, B


def test_noalloc_solve_matches_pooled_solve_twice_smaller_second():
    torch = pytest.importorskip("torch")
    cases = [(G.diffusion_3d_lognormal(14, seed=21), 6, M.FP32, 5),
             (G.laplacian_2d(20, 0.05, True), 2, M.FP64, 6)]
    sizes = []
    for (n, r, c, v), k, prec, seed in cases:
        f, B = _case(torch, n, r, c, v, k, prec, seed)
        X_pool = torch.zeros_like(B)
        X_static = torch.full_like(B, np.nan)
        rep_pool = f.solve_dev(B.data_ptr(), X_pool.data_ptr(), k, tol=1e-10, max_refine=30)
        rep_static = f.solve_dev_noalloc(B.data_ptr(), X_static.data_ptr(), k, tol=1e-10, max_refine=30)
        torch.cuda.synchronize()
        assert rep_pool == rep_static
        assert rep_static["converged"] and rep_static["relative_residual"] <= 1e-10
        assert torch.equal(X_pool, X_static)
        sizes.append(n * k)
    assert sizes[1] < sizes[0]


def test_noalloc_refuses_more_than_the_scratch():
    torch = pytest.importorskip("torch")
    n, r, c, v = G.laplacian_2d(24, 0.0, False)
    k = 20000  # Y, Z and R alone take 3 x 8 x n x k = 276 MB of workspace for n = 576
    a = M.SymmetricCSRMatrix.from_coo(n, r, c, v)
    f = M.SymmetricUtduSolver().factorize(a)
    B = torch.ones((k, n), dtype=torch.float64, device="cuda:0")
    X = torch.zeros_like(B)
    with pytest.raises(M.ArgumentOutOfRangeError):
        f.solve_dev_noalloc(B.data_ptr(), X.data_ptr(), k, tol=1e-10, max_refine=30)
    torch.cuda.synchronize()
    assert not X.any()  # refused before any kernel ran
