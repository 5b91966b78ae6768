"""Batched independent systems (mg_batch_*): every system of the batch must come out
exactly as when it is factored and solved alone (same fronts, same kernels, same order:
bit-identical in FP64), FP32+IR must reach 1e-10 on every system and column, the pivot rule
must use each system's own scale, and a captured hipGraph must replay to the eager result."""
import numpy as np
import pytest

import magla_amd as M
import oracle as O
from magla_amd import generators as G

pytestmark = pytest.mark.gpu


def small_batch(count=12, seed=5):
    return G.config_c4(count, seed=seed, side_lo=6, side_hi=24, max_rhs=4)


def test_batch_equals_single_system_solves_bitwise():
    sysl = small_batch()
    mats = [M.SymmetricCSRMatrix.from_coo(n, r, c, v) for n, r, c, v, b in sysl]
    batch = M.UtduBatch(mats, M.UtduSolverOptions(precision=M.FP64))
    batch.factorize()
    xs = batch.solve([b for *_, b in sysl], tol=0.0, max_refine=0)
    for (n, r, c, v, b), m, x in zip(sysl, mats, xs):
        num = M.SymmetricUtduSolver().factorize(m)
        x1 = num.solve(b)
        np.testing.assert_array_equal(x, x1)
        oc = O.Csr.from_coo(n, r, c, v)
        for k in range(b.shape[1]):
            res = oc.spmv(x[:, k]) - b[:, k]
            assert np.linalg.norm(res) <= 1e-12 * np.linalg.norm(b[:, k])


def test_batch_fp32_refinement_reaches_1e10_everywhere():
    sysl = small_batch(20, seed=8)
    mats = [M.SymmetricCSRMatrix.from_coo(n, r, c, v) for n, r, c, v, b in sysl]
    batch = M.UtduBatch(mats, M.UtduSolverOptions(precision=M.FP32))
    batch.factorize()
    xs = batch.solve([b for *_, b in sysl], tol=1e-10, max_refine=30)
    assert batch.last_report["converged"]
    assert batch.last_report["relative_residual"] <= 1e-10
    for (n, r, c, v, b), x in zip(sysl, xs):
        oc = O.Csr.from_coo(n, r, c, v)
        for k in range(b.shape[1]):
            assert np.linalg.norm(oc.spmv(x[:, k]) - b[:, k]) <= 1e-10 * np.linalg.norm(b[:, k])


def test_converged_systems_leave_the_refinement():
    # well-conditioned systems (diagonal shift 1) converge after the first correction, the
    # shift-free 9-point Laplacians need more; the later corrections skip the converged ones,
    # whose solutions must then stay as they were and still meet the tolerance
    easy = [G.laplacian_2d(s, 1.0, False) for s in (12, 20, 9, 31)]
    hard = [G.laplacian_2d(s, 0.0, True) for s in (64, 48)]
    systems = [easy[0], hard[0], easy[1], easy[2], hard[1], easy[3]]
    g = np.random.default_rng(3)
    bs = [g.uniform(-1, 1, (n, 3)) for n, *_ in systems]
    mats = [M.SymmetricCSRMatrix.from_coo(n, r, c, v) for n, r, c, v in systems]
    batch = M.UtduBatch(mats, M.UtduSolverOptions(precision=M.FP32))
    batch.factorize()
    xs = batch.solve(bs, tol=1e-12, max_refine=30)
    assert batch.last_report["converged"] and batch.last_report["refinement_steps"] >= 2
    x1 = batch.solve(bs, tol=1e-12, max_refine=1)
    for (n, r, c, v), b, x, xa in zip(systems, bs, xs, x1):
        oc = O.Csr.from_coo(n, r, c, v)
        for k in range(3):
            assert np.linalg.norm(oc.spmv(x[:, k]) - b[:, k]) <= 1e-12 * np.linalg.norm(b[:, k])
    for i in (0, 2, 3, 5):  # converged after one correction: later corrections left them alone
        oc = O.Csr.from_coo(*systems[i])
        if all(np.linalg.norm(oc.spmv(x1[i][:, k]) - bs[i][:, k]) <= 1e-12 * np.linalg.norm(bs[i][:, k])
               for k in range(3)):
            np.testing.assert_array_equal(xs[i], x1[i])


def test_uneven_and_zero_right_hand_sides():
    # 1..5 right-hand sides per system (padded to 5 slots on the device) and one all-zero
    # column inside a system's count: the solves stop at each system's last nonzero column, and
    # in the refinement at its last unconverged one; every solution meets the tolerance and the
    # zero column's solution is zero
    sysl = small_batch(count=9, seed=11)
    mats = [M.SymmetricCSRMatrix.from_coo(n, r, c, v) for n, r, c, v, _ in sysl]
    batch = M.UtduBatch(mats, M.UtduSolverOptions(precision=M.FP32))
    batch.factorize()
    g = np.random.default_rng(4)
    counts = [1, 5, 3, 2, 4, 1, 5, 3, 2]
    bs = [g.uniform(-1, 1, (n, k)) for (n, *_), k in zip(sysl, counts)]
    bs[1][:, 2] = 0.0
    xs = batch.solve(bs, tol=1e-10, max_refine=30)
    assert batch.last_report["converged"] and batch.last_report["refinement_steps"] >= 1
    for (n, r, c, v, _), b, x in zip(sysl, bs, xs):
        oc = O.Csr.from_coo(n, r, c, v)
        for k in range(b.shape[1]):
            nb = np.linalg.norm(b[:, k])
            if nb == 0.0:
                assert not np.any(x[:, k])
            else:
                assert np.linalg.norm(oc.spmv(x[:, k]) - b[:, k]) <= 1.0001e-10 * nb


def test_pivot_scale_is_per_system():
    # system 0 is singular ([[1,1],[1,1]] (+) 2 (+) 3: one regularized pivot); system 1 is an SPD
    # matrix of the same pattern times 1e14.  With one scale for the whole batch (4e14) every
    # pivot of system 0 would fall under the tolerance 1e-13 * scale = 40; with per-system
    # scales only the singular pivot does (UtduNumericFactorization.cs:411-418,559-572).
    r = np.array([0, 0, 1, 2, 3]); c = np.array([0, 1, 1, 2, 3]); v = np.array([1.0, 1.0, 1.0, 2.0, 3.0])
    a0 = M.SymmetricCSRMatrix.from_coo(4, r, c, v)
    a1 = M.SymmetricCSRMatrix.from_coo(4, r, c, np.array([4.0, 1.0, 3.0, 2.0, 3.0]) * 1e14)
    batch = M.UtduBatch([a0, a1], M.UtduSolverOptions(precision=M.FP64))
    batch.factorize()
    info = batch.numeric_info()
    assert info["regularized_pivot_count"] == 1
    assert info["negative_pivot_count"] == 0


def test_graph_replay_matches_eager():
    torch = pytest.importorskip("torch")
    sysl = small_batch(10, seed=11)
    mats = [M.SymmetricCSRMatrix.from_coo(n, r, c, v) for n, r, c, v, b in sysl]
    k = max(b.shape[1] for *_, b in sysl)
    nt = sum(m.size for m in mats)
    B = np.zeros((k, nt))
    off = 0
    for (n, *_, b) in sysl:
        B[:b.shape[1], off:off + n] = b.T
        off += n
    dev = torch.device("cuda:0")
    Bd = torch.from_numpy(B).to(dev)
    base = np.concatenate([m.values for m in mats])
    eager = M.UtduBatch(mats, M.UtduSolverOptions(precision=M.FP64))
    graph = M.UtduBatch(mats, M.UtduSolverOptions(precision=M.FP64, device_graph=True))
    for rep, scale in enumerate([1.0, 2.0, 0.5, 3.0]):
        vals = torch.from_numpy(base * scale).to(dev)
        Xe = torch.zeros_like(Bd)
        Xg = torch.zeros_like(Bd)
        eager.factorize_dev(vals.data_ptr(), device_graph=False)
        eager.solve_dev(Bd.data_ptr(), Xe.data_ptr(), k, tol=0.0, max_refine=0)
        graph.factorize_dev(vals.data_ptr(), device_graph=True)
        graph.solve_dev(Bd.data_ptr(), Xg.data_ptr(), k, tol=0.0, max_refine=0)
        torch.cuda.synchronize()
        _ = (Bd * 2.0).sum().item()  # unrelated eager work between replays
        assert torch.equal(Xe, Xg), rep
