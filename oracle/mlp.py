"""ORACLE — test infrastructure only: FP64 numpy restatement of the learned
preconditioner's network, loss, gradients and Adam step, with optional
emulation of the BF16 rounding the GPU applies to every matrix-product
operand. The analytic gradients here are pinned by finite differences in
tests/test_oracle_mlp.py.
"""
from __future__ import annotations

import numpy as np


def bf16(x):
    """Round to BF16 (round to nearest even, through FP32), returned as FP64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    special = (u & 0x7F800000) == 0x7F800000
    r = np.where(special, u & 0xFFFF0000, r)
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def split(p, F, H1, H2, O):
    p = np.asarray(p, dtype=np.float64)
    sizes = [H1 * F, H1, H2 * H1, H2, O * H2, O]
    shapes = [(H1, F), (H1,), (H2, H1), (H2,), (O, H2), (O,)]
    out, o = [], 0
    for n, sh in zip(sizes, shapes):
        out.append(p[o:o + n].reshape(sh))
        o += n
    return out


def forward(p, X, F, H1, H2, dim, rank, out_scale, bf=True):
    q = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = split(p, F, H1, H2, dim * rank)
    X = np.asarray(X, dtype=np.float64)
    h1 = np.maximum(q(X) @ q(W1).T + b1, 0.0)
    h2 = np.maximum(q(h1) @ q(W2).T + b2, 0.0)
    o = q(h2) @ q(W3).T + b3
    V = out_scale * o.reshape(-1, dim, rank)
    return V, dict(X=X, h1=h1, h2=h2, o=o)


def nll(V, S):
    """loss_b = tr(V^T S V) - logdet(I + V^T V); dL_b/dV = 2 S V - 2 V (I + V^T V)^{-1}."""
    V = np.asarray(V, dtype=np.float64)
    S = np.asarray(S, dtype=np.float64)
    B, d, r = V.shape
    loss = np.empty(B)
    dV = np.empty_like(V)
    for b in range(B):
        SV = S[b] @ V[b]
        M = np.eye(r) + V[b].T @ V[b]
        sign, ld = np.linalg.slogdet(M)
        loss[b] = np.sum(V[b] * SV) - ld
        dV[b] = 2.0 * SV - 2.0 * V[b] @ np.linalg.inv(M)
    return loss, dV


def nll_cov(V, S):
    """Scale-free covariance model C = tau (I + V V^T), tau profiled out:
    q = tr S - tr(M^{-1} V^T S V), loss_b = d log(q/d) + logdet M, written
    here through the explicit d x d inverse of I + V V^T (an independent route
    to the same value), with the gradient
    (d/q) (2 V M^{-1} G M^{-1} - 2 S V M^{-1}) + 2 V M^{-1}."""
    V = np.asarray(V, dtype=np.float64)
    S = np.asarray(S, dtype=np.float64)
    B, d, r = V.shape
    loss = np.empty(B)
    dV = np.empty_like(V)
    for b in range(B):
        C = np.eye(d) + V[b] @ V[b].T
        q = np.trace(np.linalg.solve(C, S[b]))
        M = np.eye(r) + V[b].T @ V[b]
        Mi = np.linalg.inv(M)
        G = V[b].T @ S[b] @ V[b]
        loss[b] = d * np.log(q / d) + np.linalg.slogdet(C)[1]
        dV[b] = (d / q) * (2 * V[b] @ Mi @ G @ Mi - 2 * S[b] @ V[b] @ Mi) + 2 * V[b] @ Mi
    return loss, dV


def loss_grad(p, X, S, F, H1, H2, dim, rank, out_scale, bf=True, model=0):
    """(per-condition loss, gradient of the batch-mean loss w.r.t. p);
    model 0 = precision (nll), 1 = scale-free covariance (nll_cov)."""
    q = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = split(p, F, H1, H2, dim * rank)
    V, c = forward(p, X, F, H1, H2, dim, rank, out_scale, bf)
    B = V.shape[0]
    loss, dV = (nll_cov if model == 1 else nll)(V, S)
    dO = (out_scale / B) * dV.reshape(B, -1)
    gW3 = q(dO).T @ q(c["h2"])
    gb3 = dO.sum(0)
    d2 = (q(dO) @ q(W3)) * (c["h2"] > 0)
    gW2 = q(d2).T @ q(c["h1"])
    gb2 = d2.sum(0)
    d1 = (q(d2) @ q(W2)) * (c["h1"] > 0)
    gW1 = q(d1).T @ q(c["X"])
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def adam(p, g, m1, m2, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1); returns (p, m1, m2)."""
    m1 = beta1 * m1 + (1 - beta1) * g
    m2 = beta2 * m2 + (1 - beta2) * g * g
    mhat = m1 / (1 - beta1 ** step)
    vhat = m2 / (1 - beta2 ** step)
    return p - lr * mhat / (np.sqrt(vhat) + eps), m1, m2
