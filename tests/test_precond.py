"""Host FP64 preconditioners (dikin_precondition.cpp:36-183 restated) against
numpy restatements, closed forms and invariants. Host-only library calls."""
import numpy as np
import pytest

from oracle import precond as orc

pc = pytest.importorskip("naja_amd.precond")


def conditioned_cube(d, nslab, w_lo, w_hi, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((nslab, d))
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    w = 10 ** rng.uniform(np.log10(w_lo), np.log10(w_hi), nslab)
    E = np.vstack([a, -a])
    f = np.concatenate([w, w])
    return np.vstack([np.eye(d), -np.eye(d)]), np.ones(2 * d), E, f


@pytest.mark.parametrize("n,kind", [(1, "rand"), (2, "rand"), (7, "rand"), (64, "rand"), (100, "rank3"),
                                    (30, "repeated"), (16, "diag"), (40, "identity"), (100, "barrier")])
def test_symmetric_eigen_matches_lapack(n, kind):
    rng = np.random.default_rng(n)
    if kind == "rand":
        a = rng.standard_normal((n, n))
        a = a + a.T
    elif kind == "rank3":
        u = rng.standard_normal((n, 3))
        a = u @ u.T
    elif kind == "repeated":
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        a = (Q * np.repeat([1.0, 2.0, 5.0], n // 3)) @ Q.T
    elif kind == "diag":
        a = np.diag(rng.standard_normal(n))
    elif kind == "barrier":
        # Rank-deficient Hessian with weights spanning 1e0..1e6 (a Dikin
        # tight set): most eigenvalues are zero up to rounding.
        N = rng.standard_normal((40, n))
        N /= np.linalg.norm(N, axis=1, keepdims=True)
        a = (N.T * 10 ** rng.uniform(0, 6, 40)) @ N
    else:
        a = np.eye(n)
    w, V = pc.symmetric_eigen(a)
    scale = max(1.0, np.abs(a).max())
    np.testing.assert_allclose(w, np.linalg.eigvalsh(a), atol=1e-11 * scale)
    assert np.all(np.diff(w) >= 0)
    np.testing.assert_allclose(V.T @ V, np.eye(n), atol=1e-12)
    np.testing.assert_allclose((V * w) @ V.T, a, atol=1e-11 * scale)


def test_dikin_closed_form_single_slab():
    """Box [-1,1]^d with |y_0| <= w: the tight set is the slab pair, H =
    (2/w^2) e0 e0^T, one corrected direction e0, target chord 2 (median of
    the axis chords), alpha = 1/w, so P_inv = I + (w - 1) e0 e0^T."""
    d, w = 12, 2e-3
    A = np.vstack([np.eye(d), -np.eye(d), np.eye(d)[:1], -np.eye(d)[:1]])
    b = np.concatenate([np.ones(2 * d), [w, w]])
    r = pc.dikin(A, b)
    expect = np.eye(d)
    expect[0, 0] = w
    assert r["n_corrected"] == 1
    np.testing.assert_allclose(r["P_inv"], expect, atol=1e-12)
    np.testing.assert_allclose(r["P"], np.linalg.inv(expect), rtol=1e-12)
    np.testing.assert_allclose(r["b_new"], b)
    np.testing.assert_allclose(r["A_new"], A @ expect)


def test_dikin_alpha_clipped():
    """alpha = clip(target/chord, 1, alpha_max): a slab thinner than
    2/alpha_max is stretched by alpha_max only."""
    d, w = 6, 1e-5
    A = np.vstack([np.eye(d), -np.eye(d), np.eye(d)[:1], -np.eye(d)[:1]])
    b = np.concatenate([np.ones(2 * d), [w, w]])
    r = pc.dikin(A, b, alpha_max=1000.0)
    assert r["P"][0, 0] == pytest.approx(1000.0)


def test_dikin_no_tight_set_is_identity():
    """Every facet at or behind x_c (no positive distance): identity, b_new = slack."""
    A = np.array([[1.0, 0.0], [0.0, 1.0]])
    b = np.array([-1.0, 0.0])
    r = pc.dikin(A, b)
    assert r["n_corrected"] == 0
    np.testing.assert_array_equal(r["P_inv"], np.eye(2))
    np.testing.assert_allclose(r["b_new"], b)


def test_dikin_empty_polytope_rejected():
    with pytest.raises(Exception, match="empty polytope"):
        pc.dikin(np.zeros((0, 3)), np.zeros(0))


@pytest.mark.parametrize("nslab,w_lo,w_hi,seed", [(3, 1e-3, 1e-1, 0), (20, 1e-3, 1e-1, 1), (80, 1e-2, 1.0, 2),
                                                  (150, 1e-3, 1e-1, 3)])
def test_dikin_matches_numpy_restatement(nslab, w_lo, w_hi, seed):
    d = 40
    A0, b0, E, f = conditioned_cube(d, nslab, w_lo, w_hi, seed)
    A, b = np.vstack([A0, E]), np.concatenate([b0, f])
    xc = np.random.default_rng(seed).uniform(-1e-4, 1e-4, d)
    r = pc.dikin(A, b, xc)
    o = orc.dikin(A, b, xc)
    assert r["n_corrected"] == o["n_corrected"]
    np.testing.assert_allclose(r["P_inv"], o["P_inv"], atol=1e-9)
    np.testing.assert_allclose(r["P"], o["P"], atol=1e-9 * np.abs(o["P"]).max())
    np.testing.assert_allclose(r["b_new"], o["b_new"], rtol=1e-13)
    # Invariants: P P_inv = I; y = x_c + P_inv z is feasible iff A_new z <= b_new.
    np.testing.assert_allclose(r["P"] @ r["P_inv"], np.eye(d), atol=1e-9)
    z = np.random.default_rng(9).standard_normal((200, d)) * 1e-3
    y = xc + z @ r["P_inv"].T
    np.testing.assert_allclose(z @ r["A_new"].T - r["b_new"], y @ A.T - b, atol=1e-12)


def test_dikin_batched_equals_single_calls():
    d = 30
    rng = np.random.default_rng(4)
    A0, b0 = np.vstack([np.eye(d), -np.eye(d)]), np.ones(2 * d)
    Es, fs = [], []
    for k in [0, 4, 30, 9]:
        a = rng.standard_normal((k, d))
        a /= np.maximum(np.linalg.norm(a, axis=1, keepdims=True), 1e-300)
        w = 10 ** rng.uniform(-3, -1, k)
        Es.append(np.vstack([a, -a]))
        fs.append(np.concatenate([w, w]))
    Pinv, nc = pc.dikin_batched(A0, b0, Es, fs, nthreads=3)
    for v in range(4):
        one = pc.dikin(np.vstack([A0, Es[v]]), np.concatenate([b0, fs[v]]))
        np.testing.assert_array_equal(Pinv[v], one["P_inv"])
        assert nc[v] == one["n_corrected"]


def test_ellipsoid_whitening():
    """P P^T = H^{-1}; on the cube H = 2 I, so P = I / sqrt(2)."""
    d = 10
    A0, b0 = np.vstack([np.eye(d), -np.eye(d)]), np.ones(2 * d)
    P = pc.ellipsoid_batched(A0, b0, [np.zeros((0, d))], [np.zeros(0)])[0]
    np.testing.assert_allclose(P, np.eye(d) / np.sqrt(2.0), atol=1e-15)
    _, _, E, f = conditioned_cube(d, 7, 1e-3, 1e-1, 5)
    P = pc.ellipsoid_batched(A0, b0, [E], [f])[0]
    A, b = np.vstack([A0, E]), np.concatenate([b0, f])
    H = (A.T / b ** 2) @ A
    np.testing.assert_allclose(P @ P.T @ H, np.eye(d), atol=1e-8)
    np.testing.assert_allclose(P, orc.ellipsoid(A, b), rtol=1e-9, atol=1e-12)
    assert np.allclose(np.tril(P, -1), 0.0)


def test_ellipsoid_rejects_exterior_centre():
    d = 3
    A0, b0 = np.vstack([np.eye(d), -np.eye(d)]), np.ones(2 * d)
    with pytest.raises(Exception, match="not interior"):
        pc.ellipsoid_batched(A0, b0, [np.zeros((0, d))], [np.zeros(0)], x_c=np.array([[1.0, 0.0, 0.0]]))


@pytest.mark.parametrize("expand", [False, True])
@pytest.mark.parametrize("d,r,scale", [(20, 4, 1.0), (50, 32, 30.0), (8, 8, 1e-3)])
def test_lowrank_sqrt(d, r, scale, expand):
    """P = (I + V V^T)^{-1/2}: P P (I + V V^T) = I; the expansion ^{+1/2}:
    P P = I + V V^T."""
    rng = np.random.default_rng(d + r)
    V = rng.standard_normal((3, d, r)) * scale
    V[1, :, r // 2:] = 0.0  # rank-deficient factor
    P = pc.lowrank(V, expand=expand)
    for v in range(3):
        M = np.eye(d) + V[v] @ V[v].T
        np.testing.assert_allclose(P[v], P[v].T, atol=1e-12 * max(1.0, np.abs(P[v]).max()))
        if expand:
            np.testing.assert_allclose(P[v] @ P[v], M, atol=1e-9 * np.abs(M).max())
        else:
            np.testing.assert_allclose(P[v] @ P[v] @ M, np.eye(d), atol=1e-9)
        np.testing.assert_allclose(P[v], orc.lowrank(V[v], expand), atol=1e-10 * max(1.0, np.abs(P[v]).max()))


def test_oracle_analytic_center_closed_forms():
    """The oracle's analytic centre: the middle of a box [l, u] (each
    coordinate's two barrier terms balance at (l + u) / 2), the centroid
    1/(n+1) of the simplex {x >= 0, sum x <= 1}, and a zero barrier gradient
    A^T (1/s) on a random polytope."""
    from oracle import precond as op

    rng = np.random.default_rng(0)
    d = 7
    lo, hi = -rng.uniform(0.1, 2, d), rng.uniform(0.1, 2, d)
    A = np.vstack([np.eye(d), -np.eye(d)])
    b = np.concatenate([hi, -lo])
    x, _ = op.analytic_center(A, b)
    np.testing.assert_allclose(x, (lo + hi) / 2, atol=1e-12)
    n = 5
    As = np.vstack([-np.eye(n), np.ones((1, n))])
    bs = np.concatenate([np.zeros(n), [1.0]])
    xs, _ = op.analytic_center(As, bs, x0=np.full(n, 0.1))
    np.testing.assert_allclose(xs, np.full(n, 1.0 / (n + 1)), atol=1e-12)
    Ar = np.vstack([rng.standard_normal((40, 6)), np.eye(6), -np.eye(6)])
    br = np.concatenate([np.ones(40), 3 * np.ones(12)])
    xr, _ = op.analytic_center(Ar, br)
    assert np.abs(Ar.T @ (1.0 / (br - Ar @ xr))).max() <= 1e-9
