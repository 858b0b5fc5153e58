"""Learned preconditioner on the GPU against the FP64 oracle (oracle/mlp.py):
BF16 MFMA products, forward, loss + gradient through all layers, the Adam
step, pilot covariances, and that training lowers the loss."""
import numpy as np
import pytest

from oracle import mlp as orc

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(70, 45, 37), (256, 512, 300), (1, 1, 1), (64, 64, 64)])
def test_gemm_bf16_matches_rounded_product(gpu, ta, tb, M, N, K):
    from naja_amd.learned import gemm_bf16

    rng = np.random.default_rng(M * 7 + N + K)
    A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    C = gemm_bf16(gpu.from_numpy(A).cuda(), gpu.from_numpy(B).cuda(), bool(ta), bool(tb), alpha=0.5).cpu().numpy()
    Aq = orc.bf16(A.T if ta else A)
    Bq = orc.bf16(B.T if tb else B)
    ref = 0.5 * Aq @ Bq
    bound = 0.5 * np.abs(Aq) @ np.abs(Bq)
    assert np.all(np.abs(C - ref) <= 2e-6 * bound + 1e-30), f"max err {np.abs(C - ref).max()}"


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(256, 512, 3200), (3200, 512, 256), (70, 45, 1000), (256, 512, 64), (5, 3, 129)])
def test_gemm_bf16_split_k_and_row_sums(gpu, ta, tb, M, N, K):
    """K split across workgroups (partial tiles summed in split order) keeps
    the same bound as the single pass; the row sums of op(A) are the FP32 sums
    of the unrounded operand."""
    from naja_amd.learned import gemm_bf16

    rng = np.random.default_rng(M + 3 * N + 11 * K)
    A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    out = gemm_bf16(gpu.from_numpy(A).cuda(), gpu.from_numpy(B).cuda(), bool(ta), bool(tb), alpha=0.5,
                    row_sums=bool(ta), split_k=True)
    C, rs = (out[0].cpu().numpy(), out[1].cpu().numpy()) if ta else (out.cpu().numpy(), None)
    Aq = orc.bf16(A.T if ta else A)
    Bq = orc.bf16(B.T if tb else B)
    ref = 0.5 * Aq.astype(np.float64) @ Bq.astype(np.float64)
    bound = 0.5 * np.abs(Aq).astype(np.float64) @ np.abs(Bq).astype(np.float64)
    assert np.all(np.abs(C - ref) <= 2e-6 * bound + 1e-30), f"max err {np.abs(C - ref).max()}"
    if ta:
        exact = A.astype(np.float64).sum(axis=0)
        assert np.all(np.abs(rs - exact) <= 1e-6 * np.abs(A).astype(np.float64).sum(axis=0) + 1e-30)
        # the split takes nothing from what one pass gives
        single = gemm_bf16(gpu.from_numpy(A).cuda(), gpu.from_numpy(B).cuda(), bool(ta), bool(tb), alpha=0.5,
                           row_sums=True)
        assert np.allclose(single[1].cpu().numpy(), rs, rtol=0, atol=1e-6 * np.abs(A).sum(axis=0).max())


def test_gemm_bf16_row_sums_need_transposed_a(gpu):
    from naja_amd.learned import gemm_bf16
    from naja_amd import NajaInvalidArgument

    with pytest.raises(NajaInvalidArgument):
        gemm_bf16(gpu.zeros((4, 3)).cuda(), gpu.zeros((3, 2)).cuda(), row_sums=True)


def test_pilot_covariance_matches_numpy(gpu):
    from naja_amd.learned import pilot_covariance

    rng = np.random.default_rng(3)
    nv, rows, d = 3, 777, 45
    X = (rng.standard_normal((nv, rows, d)) * rng.uniform(0.01, 2, d) + 50.0).astype(np.float32)
    S, mean = pilot_covariance(gpu.from_numpy(X).cuda(), scale=3.0)
    S, mean = S.cpu().numpy(), mean.cpu().numpy()
    for v in range(nv):
        np.testing.assert_allclose(mean[v], X[v].astype(np.float64).mean(0), rtol=1e-12)
        ref = 3.0 * np.cov(X[v].astype(np.float64).T)
        np.testing.assert_allclose(S[v], ref, rtol=2e-6, atol=2e-6 * np.abs(ref).max())


def make_problem(rng, B, F=300, d=100, r=32):
    X = np.where(rng.uniform(0, 1, (B, F)) < 0.1, rng.uniform(0.01, 1, (B, F)), 0.0).astype(np.float32)
    S = np.empty((B, d, d), np.float32)
    for b in range(B):
        U = rng.standard_normal((d, 5)) * rng.uniform(0.5, 5, 5)
        S[b] = np.linalg.inv(np.eye(d) + U @ U.T)
    return X, S


@pytest.fixture(scope="module")
def net(gpu):
    from naja_amd.learned import LearnedPreconditioner

    lp = LearnedPreconditioner(seed=11)
    yield lp
    lp.close()


def test_forward_matches_oracle(gpu, net):
    rng = np.random.default_rng(5)
    X, _ = make_problem(rng, 19)
    p = net.params().cpu().numpy()
    # Push the initial (near-zero) output layer to realistic magnitudes.
    p[-(net.dim * net.rank) * (net.hidden2 + 1):] *= 100.0
    net.set_params(gpu.from_numpy(p).cuda())
    V = net.forward(gpu.from_numpy(X).cuda()).cpu().numpy()
    Vo, _ = orc.forward(p, X, 300, 512, 512, 100, 32, net.out_scale, bf=True)
    err = rel_l2(V, Vo)
    print(f"forward rel L2 {err:.2e}", flush=True)
    assert err < 1e-4


@pytest.mark.parametrize("model", [0, 1])
def test_loss_and_gradients_match_oracle(gpu, net, model):
    from naja_amd.learned import LearnedPreconditioner

    rng = np.random.default_rng(6)
    X, S = make_problem(rng, 24)
    if model == 1:
        S = (S * rng.uniform(0.1, 10.0, (24, 1, 1))).astype(np.float32)  # arbitrary scale per condition
        net = LearnedPreconditioner(seed=12, model=1, out_scale=4.0)
    p = net.params().cpu().numpy()
    loss, g = net.loss_grad(gpu.from_numpy(X).cuda(), gpu.from_numpy(S).cuda())
    loss, g = loss.cpu().numpy(), g.cpu().numpy()
    lo, go = orc.loss_grad(p, X, S, 300, 512, 512, 100, 32, net.out_scale, bf=True, model=model)
    assert rel_l2(loss, lo) < 1e-4
    names = ["W1", "b1", "W2", "b2", "W3", "b3"]
    for name, a, b in zip(names, net.split(g), net.split(go)):
        err = rel_l2(a, b)
        print(f"model {model} grad {name} rel L2 {err:.2e}", flush=True)
        assert err < 2e-3, name
    if model == 1:
        net.close()


def test_adam_step_matches_oracle(gpu, net):
    rng = np.random.default_rng(7)
    X, S = make_problem(rng, 16)
    t = gpu
    p0 = net.params().cpu().numpy()
    net.set_params(t.from_numpy(p0).cuda())  # fresh Adam state, step 1 next
    _, g = net.loss_grad(t.from_numpy(X).cuda(), t.from_numpy(S).cuda())
    net.train_step(t.from_numpy(X).cuda(), t.from_numpy(S).cuda(), lr=1e-3)
    p1 = net.params().cpu().numpy()
    ref, _, _ = orc.adam(p0.astype(np.float64), g.cpu().numpy().astype(np.float64), 0.0, 0.0, 1, 1e-3)
    # FP32 update arithmetic: a few ulps of the parameter magnitude.
    np.testing.assert_allclose(p1, ref, rtol=1e-6, atol=3e-8)
    # The BF16 copies follow the masters: the forward pass uses the new weights.
    V = net.forward(t.from_numpy(X).cuda()).cpu().numpy()
    Vo, _ = orc.forward(p1, X, 300, 512, 512, 100, 32, net.out_scale, bf=True)
    assert rel_l2(V, Vo) < 1e-4


def test_training_lowers_loss(gpu):
    """A teacher problem: S_b = (I + U_b U_b^T)^{-1}, U_b built from the
    active features' directions; 1500 Adam steps must close at least half of
    the gap between the initial loss and the teacher's."""
    from naja_amd.learned import LearnedPreconditioner

    t = gpu
    rng = np.random.default_rng(8)
    B, F, d, r = 128, 300, 100, 32
    X = np.where(rng.uniform(0, 1, (B, F)) < 0.05, rng.uniform(0.2, 1, (B, F)), 0.0).astype(np.float32)
    dirs = rng.standard_normal((F, d))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    S = np.empty((B, d, d), np.float32)
    best = 0.0
    for b in range(B):
        act = np.nonzero(X[b])[0][:r]
        U = np.zeros((d, r))
        U[:, :len(act)] = (dirs[act] * (20.0 * X[b, act, None])).T
        S[b] = np.linalg.inv(np.eye(d) + U @ U.T)
        best += orc.nll(U[None], S[b][None].astype(np.float64))[0][0] / B
    lp = LearnedPreconditioner(seed=3)
    Xd, Sd = t.from_numpy(X).cuda(), t.from_numpy(S).cuda()
    first = float(lp.loss_grad(Xd, Sd)[0].mean())
    for it in range(1500):
        lp.train_step(Xd, Sd, lr=1e-3)
        if it % 300 == 299:
            print(f"step {it + 1}: loss {float(lp.loss_grad(Xd, Sd)[0].mean()):.3f}", flush=True)
    last = float(lp.loss_grad(Xd, Sd)[0].mean())
    lp.close()
    print(f"loss {first:.3f} -> {last:.3f} (teacher {best:.3f})", flush=True)
    assert last < first
    assert (first - last) > 0.5 * (first - best)


def test_invalid_arguments(gpu):
    from naja_amd import NajaError
    from naja_amd.learned import LearnedPreconditioner

    with pytest.raises(NajaError, match="rank"):
        LearnedPreconditioner(rank=16)
    with pytest.raises(NajaError, match="dim"):
        LearnedPreconditioner(dim=400)


@pytest.mark.parametrize("expand", [True, False])
@pytest.mark.parametrize("d,r,scale", [(100, 32, 0.3), (100, 32, 3.0), (50, 32, 30.0), (20, 4, 1.0), (160, 32, 1.0),
                                       (8, 8, 1e-3), (3, 1, 0.0)])
def test_lowrank_maps_device_matches_oracle(gpu, d, r, scale, expand):
    """P0 (I + V V^T)^{+-1/2} on the GPU (Newton-Schulz, FP64 inside) against
    P0 times the oracle's full d x d eigen-decomposition, including a
    rank-deficient factor and V = 0 (P = P0)."""
    import naja_amd
    from oracle import precond as orp

    rng = np.random.default_rng(d * 31 + r)
    nv = 4
    V = (rng.standard_normal((nv, d, r)) * scale).astype(np.float32)
    V[1, :, r // 2:] = 0.0
    P0 = rng.standard_normal((nv, d, d)).astype(np.float32)
    P = naja_amd.lowrank_maps(gpu.from_numpy(P0).cuda(), gpu.from_numpy(V).cuda(), expand=expand).cpu().numpy()
    for v in range(nv):
        ref = P0[v].astype(np.float64) @ orp.lowrank(V[v].astype(np.float64), expand)
        tol = 2e-6 * np.abs(P0[v]).astype(np.float64) @ np.abs(orp.lowrank(V[v].astype(np.float64), expand)) + 1e-30
        err = np.abs(P[v] - ref)
        assert np.all(err <= tol), f"variant {v}: max err {err.max():.3e}, worst ratio {(err / tol).max():.2f}"


def test_lowrank_maps_device_two_calls_and_errors(gpu):
    """The map agrees with the host FP64 routine on consecutive calls of
    different sizes; bad shapes are rejected before any launch."""
    import naja_amd
    from naja_amd import precond

    rng = np.random.default_rng(5)
    for nv, d in ((256, 100), (3, 40)):
        V = (rng.standard_normal((nv, d, 32)) * 0.5).astype(np.float32)
        P0 = np.broadcast_to(np.eye(d, dtype=np.float32), (nv, d, d)).copy()
        P = naja_amd.lowrank_maps(gpu.from_numpy(P0).cuda(), gpu.from_numpy(V).cuda()).cpu().numpy()
        ref = precond.lowrank(V.astype(np.float64), expand=True)
        np.testing.assert_allclose(P, ref, atol=2e-6 * np.abs(ref).max())
    with pytest.raises(naja_amd.NajaError):
        naja_amd.lowrank_maps(gpu.zeros((2, 5, 5)).cuda(), gpu.zeros((2, 5, 33)).cuda())
    with pytest.raises(naja_amd.NajaError):
        naja_amd.lowrank_maps(gpu.zeros((2, 200, 200)).cuda(), gpu.zeros((2, 200, 32)).cuda())
    with pytest.raises(naja_amd.NajaError):
        naja_amd.lowrank_maps(gpu.zeros((2, 6, 6)).cuda(), gpu.zeros((3, 6, 2)).cuda())
