"""The drop-in boundary's contract beyond single calls (include/magla.h):

* one numeric factorization solved from several host threads at once, each on its own
  stream, gives the same bits as the same solves made one after the other
  (UCHOL/SymmetricUtduSolver.cs:39-43: the factorization is immutable and different
  right-hand sides may be solved with it in parallel; UtduNumericFactorization.cs:101,121:
  per-call scratch);
* a captured factorization graph is not replayed with other pivot options;
* a handle whose refactorization failed refuses to solve until a factorization succeeds
  (the reference never hands out a factorization after Factorize threw).
"""
import ctypes as C
import threading

import numpy as np
import pytest

import magla_amd as M
import oracle as O
from magla_amd import generators as G

pytestmark = pytest.mark.gpu


def hip():
    """The HIP runtime libmagla.so is linked against (ROCm's, by its soname), so that streams
    and buffers made here are valid in the library's calls."""
    M.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


H2D, D2H = 1, 2


def check(rc):
    assert rc == 0, f"HIP error {rc}"


def systems():
    n, r, c, v = G.diffusion_3d_lognormal(14, seed=12)
    a = M.SymmetricCSRMatrix.from_coo(n, r, c, v)
    # different right-hand-side counts: the workspace pool grows and is shared
    rhs = [G.uniform_rhs(n, k, 40 + t).reshape(n, -1) for t, k in enumerate([3, 1, 9, 2])]
    return a, O.Csr.from_coo(n, r, c, v), rhs


@pytest.mark.parametrize("precision", [M.FP32, M.FP64])
def test_concurrent_host_solves_match_serial_bitwise(precision):
    a, oc, rhs = systems()
    f = M.SymmetricUtduSolver(M.UtduSolverOptions(precision=precision)).factorize(a)
    serial = [f.solve(b, tol=1e-10, max_refine=30) for b in rhs]
    for b, x in zip(rhs, serial):
        for k in range(b.shape[1]):
            assert np.linalg.norm(oc.spmv(x[:, k]) - b[:, k]) <= 1e-10 * np.linalg.norm(b[:, k])
    out = [[None] * 3 for _ in rhs]
    errors = []
    start = threading.Barrier(len(rhs))

    def work(t):
        try:
            start.wait()
            for rep in range(3):
                out[t][rep] = f.solve(rhs[t], tol=1e-10, max_refine=30)
        except Exception as e:  # reported below
            errors.append(e)

    threads = [threading.Thread(target=work, args=(t,)) for t in range(len(rhs))]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    for t, x in enumerate(serial):
        for rep in range(3):
            np.testing.assert_array_equal(out[t][rep], x)


def test_concurrent_stream_solves_match_serial_bitwise():
    h = hip()
    a, oc, rhs = systems()
    solver = M.SymmetricUtduSolver(M.UtduSolverOptions(precision=M.FP32))
    f = solver.factorize(a)
    n = a.size
    serial = [f.solve(b, tol=1e-10, max_refine=30) for b in rhs]
    results = [None] * len(rhs)
    errors = []
    start = threading.Barrier(len(rhs))

    def work(t):
        st, db, dx = C.c_void_p(), C.c_void_p(), C.c_void_p()
        try:
            b = np.asfortranarray(rhs[t])
            k = b.shape[1]
            check(h.hipStreamCreate(C.byref(st)))
            check(h.hipMalloc(C.byref(db), b.nbytes))
            check(h.hipMalloc(C.byref(dx), b.nbytes))
            check(h.hipMemcpy(db, b.ctypes.data, b.nbytes, H2D))
            start.wait()
            for _ in range(3):
                rep = f.solve_dev(db, dx, k, tol=1e-10, max_refine=30, stream=st)
                assert rep["converged"]
            check(h.hipStreamSynchronize(st))
            x = np.zeros_like(b, order="F")
            check(h.hipMemcpy(x.ctypes.data, dx, b.nbytes, D2H))
            results[t] = x
        except Exception as e:  # reported below
            errors.append(e)
        finally:
            for p in (db, dx):
                if p.value:
                    h.hipFree(p)
            if st.value:
                h.hipStreamDestroy(st)

    threads = [threading.Thread(target=work, args=(t,)) for t in range(len(rhs))]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    for x, xs in zip(results, serial):
        np.testing.assert_array_equal(x, xs)
    assert n == a.size


def singular():
    # [[1,1],[1,1]] (+) 2 (+) 3: the last pivot of the block is exactly 0
    r, c, v = [0, 1, 0, 2, 3], [0, 1, 1, 2, 3], [1.0, 1.0, 1.0, 2.0, 3.0]
    return M.SymmetricCSRMatrix.from_coo(4, r, c, v)


def test_captured_graph_recaptured_when_pivot_options_change():
    torch = pytest.importorskip("torch")
    a = singular()
    allow = M.SymmetricUtduSolver(M.UtduSolverOptions(device_graph=True, allow_diagonal_regularization=True))
    refuse = M.SymmetricUtduSolver(M.UtduSolverOptions(device_graph=True, allow_diagonal_regularization=False))
    sym = allow.analyze(a)
    vals = torch.from_numpy(a.values).to("cuda:0")
    num = allow.factorize_dev(sym, vals.data_ptr(), a.nonzero_count)  # captured with regularization on
    assert num.regularized_pivot_count == 1
    num = allow.factorize_dev(sym, vals.data_ptr(), a.nonzero_count, numeric=num)  # replay
    assert num.regularized_pivot_count == 1
    with pytest.raises(M.InvalidOperationError):
        refuse.factorize_dev(sym, vals.data_ptr(), a.nonzero_count, numeric=num)


def test_failed_refactorization_refuses_to_solve():
    torch = pytest.importorskip("torch")
    spd = M.SymmetricCSRMatrix.from_coo(4, [0, 1, 0, 2, 3], [0, 1, 1, 2, 3], [2.0, 2.0, 1.0, 2.0, 3.0])
    bad = singular()
    solver = M.SymmetricUtduSolver(M.UtduSolverOptions(allow_diagonal_regularization=False))
    sym = solver.analyze(spd)
    dev = torch.device("cuda:0")
    good_vals = torch.from_numpy(spd.values).to(dev)
    bad_vals = torch.from_numpy(bad.values).to(dev)
    b = torch.ones(4, dtype=torch.float64, device=dev)
    x = torch.zeros_like(b)
    num = solver.factorize_dev(sym, good_vals.data_ptr(), spd.nonzero_count)
    num.solve_dev(b.data_ptr(), x.data_ptr(), 1)
    torch.cuda.synchronize()
    with pytest.raises(M.InvalidOperationError):
        solver.factorize_dev(sym, bad_vals.data_ptr(), bad.nonzero_count, numeric=num)
    with pytest.raises(M.ArgumentError):
        num.solve_dev(b.data_ptr(), x.data_ptr(), 1)
    num = solver.factorize_dev(sym, good_vals.data_ptr(), spd.nonzero_count, numeric=num)
    num.solve_dev(b.data_ptr(), x.data_ptr(), 1)
    torch.cuda.synchronize()
    xe = np.linalg.solve(O.Csr.from_csr(4, spd.row_pointers, spd.column_indices, spd.values).dense(), np.ones(4))
    np.testing.assert_allclose(x.cpu().numpy(), xe, rtol=1e-14)
