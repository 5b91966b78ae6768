"""The oracle's restatement of the learned preconditioner (a network-predicted factor G with
z = G'G r), pinned by mathematics rather than by retyping: its gradient against central finite
differences of its own loss, the loss against the dense A-norm formula, the Jacobi limit
(zero readouts), linearity, symmetry and positive definiteness of r -> z."""
import numpy as np

import oracle as O
from magla_amd import generators as G

W = O.MP_W
OFF_T = 3 * W + 6 * W * W  # w_t, then (u1, v1, w1), (u2, v2, w2)


def small_system(m=4, seed=7):
    n, r, c, v = G.diffusion_3d_grf(m, seed, length=0.3, variance=1.0)
    csr = O.Csr.from_coo(n, r, c, v)
    return csr, csr.dense()


def random_params(seed, scale=0.4, readout=0.2):
    p = np.random.default_rng(seed).standard_normal(O.MP_P) * scale
    p[OFF_T:] *= readout / scale
    return p


def test_zero_readouts_are_jacobi():
    csr, A = small_system()
    p = random_params(1)
    p[OFF_T:] = 0.0
    r = np.random.default_rng(2).uniform(-1, 1, csr.n)
    z = O.mpnn_apply(csr, p, r)
    np.testing.assert_allclose(z, r / np.diag(A), rtol=1e-15, atol=0)


def test_linear_symmetric_positive_definite():
    csr, A = small_system()
    p = random_params(3)
    g = np.random.default_rng(4)
    u, v = g.uniform(-1, 1, (2, csr.n))
    zu, zv = O.mpnn_apply(csr, p, u), O.mpnn_apply(csr, p, v)
    np.testing.assert_allclose(O.mpnn_apply(csr, p, 2.5 * u - 0.5 * v), 2.5 * zu - 0.5 * zv, rtol=1e-12,
                               atol=1e-13 * np.abs(zu).max())
    assert abs(u @ zv - v @ zu) <= 1e-12 * abs(u @ zv)
    assert u @ zu > 0 and v @ zv > 0


def test_loss_equals_dense_a_norm_error():
    csr, A = small_system()
    p = random_params(5)
    g = np.random.default_rng(6)
    R = g.uniform(-1, 1, (3, csr.n))
    E = np.linalg.solve(A, R.T).T
    loss, _ = O.mpnn_loss_grad(csr, p, R, E)
    ref = 0.0
    for r, e in zip(R, E):
        d = O.mpnn_apply(csr, p, r) - e
        ref += (d @ A @ d) / (e @ A @ e)
    assert abs(loss - ref / 3) <= 1e-10 * abs(ref / 3)


def test_gradient_matches_finite_differences():
    csr, A = small_system()
    p = random_params(8, scale=0.2, readout=0.05)
    g = np.random.default_rng(9)
    R = g.uniform(-1, 1, (2, csr.n))
    E = np.linalg.solve(A, R.T).T
    _, grad = O.mpnn_loss_grad(csr, p, R, E)
    # every block of the layout: W_in, W1/W2 of each layer, w_t, w_p, w_q, w_r
    picks = [0, 4, 3 * W - 1, 3 * W + 7, 3 * W + W * W + 40, 3 * W + 2 * W * W + 100, 3 * W + 3 * W * W + 9,
             3 * W + 4 * W * W + 512, 3 * W + 5 * W * W + 1000, OFF_T + 3] + [OFF_T + s * W + 5 for s in range(1, 7)]
    eps = 1e-6
    for k in picks:
        pp, pm = p.copy(), p.copy()
        pp[k] += eps
        pm[k] -= eps
        fd = (O.mpnn_loss_grad(csr, pp, R, E)[0] - O.mpnn_loss_grad(csr, pm, R, E)[0]) / (2 * eps)
        assert abs(fd - grad[k]) <= 1e-6 * max(1.0, abs(grad[k])), (k, fd, grad[k])


def test_kaporin_term_equals_dense_formula():
    csr, A = small_system()
    p = random_params(10, scale=0.1, readout=0.01)  # moderate G: the dense log-determinant stays accurate
    Minv = np.array([O.mpnn_apply(csr, p, col) for col in np.eye(csr.n)]).T
    n = csr.n
    ref = np.log(np.trace(Minv @ A) / n) - np.linalg.slogdet(Minv)[1] / n
    R = np.zeros((0, n))
    loss, _ = O.mpnn_loss_grad(csr, p, R, R, kappa=1.0)
    assert abs(loss - ref) <= 1e-10 * abs(ref)
    assert ref >= -1e-12  # AM-GM: the Kaporin number is at least 1


def test_gradient_with_kaporin_matches_finite_differences():
    csr, A = small_system()
    p = random_params(11, scale=0.15, readout=0.05)
    g = np.random.default_rng(12)
    R = g.uniform(-1, 1, (2, csr.n))
    E = np.linalg.solve(A, R.T).T
    _, grad = O.mpnn_loss_grad(csr, p, R, E, kappa=0.7)
    eps = 1e-6
    for k in [1, 3 * W + 11, 3 * W + 5 * W * W + 3, OFF_T + 1, OFF_T + W + 2, OFF_T + 3 * W + 4, OFF_T + 5 * W + 6]:
        pp, pm = p.copy(), p.copy()
        pp[k] += eps
        pm[k] -= eps
        fd = (O.mpnn_loss_grad(csr, pp, R, E, 0.7)[0] - O.mpnn_loss_grad(csr, pm, R, E, 0.7)[0]) / (2 * eps)
        assert abs(fd - grad[k]) <= 1e-6 * max(1.0, abs(grad[k])), (k, fd, grad[k])
