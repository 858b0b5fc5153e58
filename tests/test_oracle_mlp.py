"""The learned preconditioner's oracle: analytic gradients against central
finite differences (FP64, no BF16 emulation), the BF16 rounding against
known values, the loss against its closed form, and Adam's first step."""
import numpy as np
import pytest

from oracle import mlp


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40], dtype=np.float32)
    r = mlp.bf16(x)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 is a tie -> even (1 + 2^-6).
    np.testing.assert_array_equal(r[:5], [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5])
    assert abs(r[5] - 3.0e38) / 3.0e38 < 2 ** -8
    assert np.isinf(mlp.bf16(np.array([np.inf], np.float32))[0])
    assert np.isnan(mlp.bf16(np.array([np.nan], np.float32))[0])


def test_nll_closed_form_and_minimiser():
    """For S = (I + U U^T)^{-1}, V = U minimises the loss (zero gradient), and
    the loss there equals tr(U^T S U) - logdet(I + U^T U)."""
    rng = np.random.default_rng(0)
    d, r = 9, 3
    U = rng.standard_normal((d, r))
    S = np.linalg.inv(np.eye(d) + U @ U.T)
    loss, dV = mlp.nll(U[None], S[None])
    assert np.abs(dV).max() < 1e-12
    assert loss[0] == pytest.approx(np.trace(U.T @ S @ U) - np.log(np.linalg.det(np.eye(r) + U.T @ U)))
    loss2, _ = mlp.nll((U * 1.1)[None], S[None])
    assert loss2[0] > loss[0]


def test_nll_gradient_finite_differences():
    rng = np.random.default_rng(1)
    B, d, r = 2, 6, 4
    V = rng.standard_normal((B, d, r))
    A = rng.standard_normal((B, d, d))
    S = A @ A.transpose(0, 2, 1) / d
    _, dV = mlp.nll(V, S)
    h = 1e-6
    for b in range(B):
        for i in range(d):
            for j in range(r):
                Vp, Vm = V.copy(), V.copy()
                Vp[b, i, j] += h
                Vm[b, i, j] -= h
                fd = (mlp.nll(Vp, S)[0][b] - mlp.nll(Vm, S)[0][b]) / (2 * h)
                assert dV[b, i, j] == pytest.approx(fd, rel=1e-6, abs=1e-7)


def test_network_gradient_finite_differences():
    """Gradient of the batch-mean loss through all three layers (FP64)."""
    rng = np.random.default_rng(2)
    F, H1, H2, d, r, B, s = 5, 7, 6, 4, 3, 3, 2.0
    n = H1 * F + H1 + H2 * H1 + H2 + d * r * H2 + d * r
    p = rng.standard_normal(n) * 0.5
    X = rng.uniform(0, 1, (B, F))
    A = rng.standard_normal((B, d, d))
    S = A @ A.transpose(0, 2, 1) / d + 0.1 * np.eye(d)
    _, g = mlp.loss_grad(p, X, S, F, H1, H2, d, r, s, bf=False)

    def f(q):
        return mlp.loss_grad(q, X, S, F, H1, H2, d, r, s, bf=False)[0].mean()

    h = 1e-6
    fd = np.empty(n)
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        fd[i] = (f(p + e) - f(p - e)) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-6)


def test_adam_first_step_is_signed_lr():
    """Step 1: m_hat = g, v_hat = g^2, so the update is -lr g / (|g| + eps)."""
    g = np.array([3.0, -0.5, 1e-3, 0.0])
    p, _, _ = mlp.adam(np.zeros(4), g, np.zeros(4), np.zeros(4), 1, lr=1e-2)
    np.testing.assert_allclose(p, -1e-2 * g / (np.abs(g) + 1e-8))


def test_nll_cov_minimiser_is_scale_free():
    """For S = tau (I + U U^T) the loss is stationary at V = U for every tau;
    scaling S shifts the loss by the constant d log(tau)."""
    rng = np.random.default_rng(3)
    d, r = 8, 3
    U = rng.standard_normal((d, r))
    base = np.eye(d) + U @ U.T
    l1, g1 = mlp.nll_cov(U[None], base[None])
    l2, g2 = mlp.nll_cov(U[None], (7.5 * base)[None])
    assert np.abs(g1).max() < 1e-10 and np.abs(g2).max() < 1e-10
    assert l2[0] - l1[0] == pytest.approx(d * np.log(7.5))
    l3, _ = mlp.nll_cov((0.8 * U)[None], base[None])
    assert l3[0] > l1[0]


def test_nll_cov_gradient_finite_differences():
    rng = np.random.default_rng(4)
    B, d, r = 2, 6, 3
    V = rng.standard_normal((B, d, r))
    A = rng.standard_normal((B, d, d))
    S = A @ A.transpose(0, 2, 1) / d + 0.2 * np.eye(d)
    _, dV = mlp.nll_cov(V, S)
    h = 1e-6
    for b in range(B):
        for i in range(d):
            for j in range(r):
                Vp, Vm = V.copy(), V.copy()
                Vp[b, i, j] += h
                Vm[b, i, j] -= h
                fd = (mlp.nll_cov(Vp, S)[0][b] - mlp.nll_cov(Vm, S)[0][b]) / (2 * h)
                assert dV[b, i, j] == pytest.approx(fd, rel=1e-6, abs=1e-7)
