"""Training data and training for the learned preconditioner (LearnedPreconditioner).

Pairs (r, e = A^{-1} r) come from the project's own solvers on the GPU: residuals are
snapshots b - A x_t of Jacobi-PCG after t iterations from x0 = 0 for seeded uniform[-1,1]
right-hand sides b (t spread over 0 .. max_iterations, so the set covers the rough early
residuals and the smooth late ones that a preconditioner has to reduce), and the exact errors
are solved by the supernodal factorization (FP64).  Device buffers are torch tensors on the
current stream.
"""
from __future__ import annotations

import numpy as np

from . import (AMD, FP64, ConjugateGradientGaussPreSolver, DeviceMatrix, LearnedPreconditioner, SymmetricCSRMatrix,
               SymmetricUtduSolver, UtduSolverOptions)


def snapshot_iterations(count, max_iterations, seed):
    """Iteration counts t for the residual snapshots: a quarter spread uniformly over
    0..max_iterations, the rest log-uniformly (more of the early, fast-changing residuals)."""
    g = np.random.default_rng(seed)
    u = g.random(count)
    t_lin = np.floor(u * (max_iterations + 1))
    t_log = np.floor(np.expm1(u * np.log1p(max_iterations)))
    pick = g.random(count) < 0.25
    return np.where(pick, t_lin, t_log).astype(np.int64)


def make_pairs(a: SymmetricCSRMatrix, dm: DeviceMatrix, count, seed, max_iterations=256, chunk=64, stream=None):
    """Returns (R, E) as torch float64 tensors of shape (count, n) on cuda:0, rows = pairs."""
    import torch

    n = a.size
    dev = torch.device("cuda:0")
    sp = torch.cuda.current_stream().cuda_stream if stream is None else stream
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    R = torch.empty((count, n), dtype=torch.float64, device=dev)
    E = torch.empty((count, n), dtype=torch.float64, device=dev)
    its = snapshot_iterations(count, max_iterations, seed)
    b = torch.empty(n, dtype=torch.float64, device=dev)
    x = torch.empty(n, dtype=torch.float64, device=dev)
    ax = torch.empty(n, dtype=torch.float64, device=dev)
    for t in range(count):
        b.copy_((2.0 * torch.rand(n, generator=g, dtype=torch.float64) - 1.0).to(dev))
        if its[t] == 0:
            R[t].copy_(b)
            continue
        pcg = ConjugateGradientGaussPreSolver(relative_tolerance=0.0, max_iterations=int(its[t]))
        pcg.solve_dev(dm, b.data_ptr(), x.data_ptr(), stream=sp)
        dm.multiply_dev(x.data_ptr(), ax.data_ptr(), stream=sp)
        torch.sub(b, ax, out=R[t])
    solver = SymmetricUtduSolver(UtduSolverOptions(ordering=AMD, precision=FP64))
    sym = solver.analyze(a)
    vals = torch.from_numpy(a.values).to(dev)
    num = solver.factorize_dev(sym, vals.data_ptr(), a.nonzero_count, stream=sp)
    for c0 in range(0, count, chunk):
        k = min(chunk, count - c0)
        num.solve_dev(R[c0:c0 + k].data_ptr(), E[c0:c0 + k].data_ptr(), k, stream=sp)
    torch.cuda.synchronize()
    return R, E


def train(net: LearnedPreconditioner, dm: DeviceMatrix, R, E, epochs=20, batch=16, learning_rate=3e-3, kappa=1.0,
          seed=0, stream=None):
    """Adam on the GPU over the pairs (rows of R, E); returns the mean loss of each epoch."""
    count, n = R.shape
    return net.train_dev(dm, R.data_ptr(), E.data_ptr(), count, n, epochs=epochs, batch=batch,
                         learning_rate=learning_rate, kappa=kappa, seed=seed, stream=stream)
