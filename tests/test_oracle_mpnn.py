"""The oracle's restatement of the learned preconditioner network, pinned by
mathematics rather than by retyping: its gradient against central finite
differences of its own loss, the loss against the dense A-norm formula, the
Jacobi limit (zero decoder) and scale equivariance of the correction."""
import numpy as np
import pytest

import oracle as O
from magla_amd import generators as G


def small_system(m=4, seed=7):
    n, r, c, v = G.diffusion_3d_grf(m, seed, length=0.3, variance=1.0)
    csr = O.Csr.from_coo(n, r, c, v)
    return csr, csr.dense()


def random_params(seed, scale=0.4):
    g = np.random.default_rng(seed)
    p = g.standard_normal(O.MP_P) * scale
    return p


def test_zero_decoder_is_jacobi():
    csr, A = small_system()
    p = random_params(1)
    p[-O.MP_W:] = 0.0
    r = np.random.default_rng(2).uniform(-1, 1, csr.n)
    z = O.mpnn_apply(csr, p, r)
    np.testing.assert_allclose(z, r / np.diag(A), rtol=0, atol=1e-15)


def test_scale_equivariance():
    csr, A = small_system()
    p = random_params(3)
    r = np.random.default_rng(4).uniform(-1, 1, csr.n)
    z1 = O.mpnn_apply(csr, p, r)
    z2 = O.mpnn_apply(csr, p, 3.5 * r)
    np.testing.assert_allclose(z2, 3.5 * z1, rtol=1e-12, atol=1e-14)


def test_loss_equals_dense_a_norm_error():
    csr, A = small_system()
    p = random_params(5)
    g = np.random.default_rng(6)
    R = g.uniform(-1, 1, (3, csr.n))
    E = np.linalg.solve(A, R.T).T
    loss, _ = O.mpnn_loss_grad(csr, p, R, E)
    ref = 0.0
    for r, e in zip(R, E):
        z = O.mpnn_apply(csr, p, r)
        d = z - e
        ref += (d @ A @ d) / (e @ A @ e)
    assert abs(loss - ref / 3) <= 1e-10 * abs(ref / 3)


def test_gradient_matches_finite_differences():
    csr, A = small_system()
    p = random_params(8)
    g = np.random.default_rng(9)
    R = g.uniform(-1, 1, (2, csr.n))
    E = np.linalg.solve(A, R.T).T
    _, grad = O.mpnn_loss_grad(csr, p, R, E)
    # every block of the layout: w_e, w_c, W1/W2 of each layer, w_d
    W = O.MP_W
    picks = [0, 5, W + 3, 2 * W + 7, 2 * W + W * W + 40, 2 * W + 2 * W * W + 100, 2 * W + 3 * W * W + 9,
             2 * W + 4 * W * W + 512, 2 * W + 5 * W * W + 1000, O.MP_P - 1, O.MP_P - 20]
    eps = 1e-6
    for k in picks:
        pp, pm = p.copy(), p.copy()
        pp[k] += eps
        pm[k] -= eps
        fd = (O.mpnn_loss_grad(csr, pp, R, E)[0] - O.mpnn_loss_grad(csr, pm, R, E)[0]) / (2 * eps)
        assert abs(fd - grad[k]) <= 1e-6 * max(1.0, abs(grad[k])), (k, fd, grad[k])
