"""GPU checks of the learned preconditioner: the bf16 MFMA operand layout, the FP32 and
BF16 network factors and the FP32 gradient (pair term and Kaporin term) against the oracle's
binary64 restatement (itself pinned by finite differences in test_oracle_mpnn.py), training on
pairs from the direct solver, the flexible PCG against Jacobi-PCG with the direct solve as the
reference for the error, and the Jacobi fallback."""
import numpy as np
import pytest

import magla_amd as M
import oracle as O
from magla_amd import generators as G

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def system(m, seed=11, length=0.3, variance=1.0):
    n, r, c, v = G.diffusion_3d_grf(m, seed, length=length, variance=variance)
    a = M.SymmetricCSRMatrix.from_coo(n, r, c, v)
    return a, O.Csr.from_coo(n, r, c, v), M.DeviceMatrix(a)


def random_params(seed, scale=0.2, readout=0.05):
    p = np.random.default_rng(seed).standard_normal(O.MP_P) * scale
    p[3 * O.MP_W + 6 * O.MP_W * O.MP_W:] *= readout / scale
    return p.astype(np.float32)


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")


def test_bf16_mfma_layout():
    g = np.random.default_rng(0)
    a = g.integers(-8, 9, (16, 16)).astype(np.float32)
    b = g.integers(-8, 9, (16, 16)).astype(np.float32)
    d = M.test_mfma_bf16(a, b)
    np.testing.assert_array_equal(d, a @ b)  # small integers: exact in bf16 and fp32


def test_param_count_matches_oracle():
    assert M.LearnedPreconditioner.param_count() == O.MP_P


@pytest.mark.parametrize("precision,tol", [(M.FP32, 2e-5), (M.BF16, 3e-2)])
def test_forward_matches_oracle(precision, tol):
    a, oc, dm = system(6)
    net = M.LearnedPreconditioner(seed=1)
    p = random_params(2)
    net.params = p
    r = np.random.default_rng(3).uniform(-1, 1, a.size)
    z_ref = O.mpnn_apply(oc, p.astype(np.float64), r)
    jac = r / a.diagonal()
    rd, zd = dev(r), torch.zeros(a.size, dtype=torch.float64, device="cuda:0")
    net.apply_dev(dm, rd.data_ptr(), zd.data_ptr(), precision=precision)
    torch.cuda.synchronize()
    z = zd.cpu().numpy()
    # error measured on the learned correction z - D^{-1} r
    err = np.linalg.norm((z - jac) - (z_ref - jac)) / np.linalg.norm(z_ref - jac)
    assert err <= tol, err


@pytest.mark.parametrize("kappa", [0.0, 0.5])
def test_loss_and_gradient_match_oracle(kappa):
    a, oc, dm = system(5)
    A = oc.dense()
    net = M.LearnedPreconditioner(seed=4)
    p = random_params(5)
    net.params = p
    g = np.random.default_rng(6)
    R = g.uniform(-1, 1, (3, a.size))
    E = np.linalg.solve(A, R.T).T
    loss_ref, grad_ref = O.mpnn_loss_grad(oc, p.astype(np.float64), R, E, kappa=kappa)
    Rd, Ed = dev(R), dev(E)
    loss, grad = net.loss_grad_dev(dm, Rd.data_ptr(), Ed.data_ptr(), 3, a.size, kappa=kappa)
    assert abs(loss - loss_ref) <= 1e-5 * abs(loss_ref), (loss, loss_ref)
    err = np.linalg.norm(grad - grad_ref) / np.linalg.norm(grad_ref)
    assert err <= 1e-4, err
    # every parameter block is exercised
    W = O.MP_W
    off_t = 3 * W + 6 * W * W
    for lo, hi in [(0, 3 * W), (3 * W, 3 * W + W * W), (3 * W + 5 * W * W, off_t), (off_t, off_t + W),
                   (off_t + W, off_t + 4 * W), (off_t + 4 * W, O.MP_P)]:
        assert np.abs(grad_ref[lo:hi]).max() > 0
        np.testing.assert_allclose(grad[lo:hi], grad_ref[lo:hi], rtol=2e-3, atol=2e-4 * np.abs(grad_ref).max())


@pytest.fixture(scope="module")
def trained():
    """A 16^3 GRF system, 192 pairs from Jacobi-PCG snapshots and the direct solver, trained."""
    from magla_amd import learned
    a, oc, dm = system(16, seed=21, length=0.1, variance=2.0)
    R, E = learned.make_pairs(a, dm, 192, seed=22, max_iterations=128)
    net = M.LearnedPreconditioner(seed=23)
    losses = learned.train(net, dm, R, E, epochs=30, batch=16, learning_rate=3e-3, kappa=1.0, seed=24)
    return a, dm, net, losses, R, E


def test_pairs_are_exact_errors(trained):
    a, dm, net, losses, R, E = trained
    e = E[:8].cpu().numpy()
    r = R[:8].cpu().numpy()
    for t in range(8):
        ae = dm.multiply(e[t])
        assert np.linalg.norm(ae - r[t]) <= 1e-10 * np.linalg.norm(r[t])


def test_training_reduces_loss(trained):
    a, dm, net, losses, R, E = trained
    assert np.all(np.isfinite(losses))
    # the loss includes the Kaporin term, whose floor (the best factor on this pattern) is well
    # above zero: a few percent is a real gain over the Jacobi-like start
    assert losses[-1] < 0.95 * losses[0], losses
    assert losses[-1] <= min(losses) + 0.02 * losses[0], losses


def test_learned_pcg_needs_fewer_iterations(trained):
    a, dm, net, losses, R, E = trained
    b = G.uniform_rhs(a.size, 1, 99)
    jac = M.ConjugateGradientGaussPreSolver(relative_tolerance=1e-8)
    xj = jac.solve(dm, b)
    lrn = M.LearnedConjugateGradientSolver(net, relative_tolerance=1e-8)
    xl = lrn.solve(dm, b)
    assert jac.last_result.converged and lrn.last_result.converged
    assert not lrn.fell_back
    assert lrn.last_result.iterations < jac.last_result.iterations, (lrn.last_result.iterations,
                                                                      jac.last_result.iterations)
    # the error of both is measured against the direct solve
    x = M.SymmetricUtduSolver().solve(a, b)
    for xi in (xj, xl):
        assert np.linalg.norm(dm.multiply(xi) - b) <= 1.0001e-8 * np.linalg.norm(b)
        assert np.linalg.norm(xi - x) <= 1e-5 * np.linalg.norm(x)


def test_fallback_to_jacobi_on_a_bad_network():
    a, oc, dm = system(10, seed=31, length=0.1, variance=2.0)
    net = M.LearnedPreconditioner(seed=32)
    p = random_params(33, scale=1.0, readout=3.0)  # a wildly scaled factor: G'G A is badly conditioned
    net.params = p
    b = G.uniform_rhs(a.size, 1, 34)
    lrn = M.LearnedConjugateGradientSolver(net, relative_tolerance=1e-8, stall_window=20)
    x = lrn.solve(dm, b)
    assert lrn.last_result.converged
    assert lrn.fell_back
    assert np.linalg.norm(dm.multiply(x) - b) <= 1.0001e-8 * np.linalg.norm(b)
