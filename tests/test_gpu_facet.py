"""Facet-aligned rounding on the GPU (naja_precond_facet_device) against the
oracle's restatement of the same greedy rule, through the C ABI."""
import numpy as np
import pytest

import oracle
from oracle import precond as op

pytestmark = pytest.mark.gpu


def conditions(d, ks, seed, w_lo=1e-3, w_hi=1e-1):
    rng = np.random.default_rng(seed)
    EA, Eb = [], []
    for k in ks:
        a = rng.standard_normal((k // 2, d))
        a /= np.linalg.norm(a, axis=1, keepdims=True)
        w = 10.0 ** rng.uniform(np.log10(w_lo), np.log10(w_hi), k // 2)
        EA.append(np.vstack([a, -a]) if k else np.zeros((0, d)))
        Eb.append(np.concatenate([w, w]) if k else np.zeros(0))
    return EA, Eb


@pytest.mark.parametrize("lower", [False, True])
def test_facet_rounding_matches_oracle(gpu, lower):
    """Every variant's P equals the oracle's N^-1 (the same chosen rows) to
    FP32 rounding, for facet counts below, at and above d, and theta from
    'every independent facet' to 'none' (then P0 with permuted columns)."""
    import naja_amd

    t = gpu
    d = 40
    A0, b0 = np.vstack([np.eye(d), -np.eye(d)]), np.ones(2 * d)
    ks = [0, 10, 36, 80, 120, 200, 60, 90]
    thetas = np.array([0.1, 0.02, 0.1, 0.1, 0.02, 0.3, 1.5, 0.1], np.float32)
    EA, Eb = conditions(d, ks, seed=3)
    batch = naja_amd.VariantBatch(A0, b0, EA, Eb)
    P0 = naja_amd.ellipsoid_rounding(batch, None, None, None)
    P = naja_amd.facet_rounding(batch, None, None, None, P0, t.from_numpy(thetas), lower=lower)
    t.cuda.synchronize()
    P, P0 = P.cpu().numpy().astype(np.float64), P0.cpu().numpy().astype(np.float64)
    for v, k in enumerate(ks):
        Pref, sel, N = op.facet(A0, b0, EA[v], Eb[v], P0[v], thetas[v], lower=lower)
        nfac = int((sel < k).sum())
        err = np.abs(P[v] - Pref).max() / np.abs(Pref).max()
        ident = np.abs(N @ P[v] - np.eye(d)).max()
        print(f"[facet v{v}] k={k} theta={thetas[v]} facets={nfac} rel err {err:.2e} |NP - I| {ident:.2e}", flush=True)
        assert err <= 1e-5 and ident <= 1e-4
        if thetas[v] > 1 and not lower:
            assert nfac == 0
            np.testing.assert_allclose(P[v], P0[v][:, sel - k], rtol=0, atol=1e-5 * np.abs(P0[v]).max())
        if k // 2 >= 1.5 * d and thetas[v] <= 0.1:  # ample facets: the whole basis from facets
            assert nfac == d


@pytest.mark.parametrize("exchange_tol", [0.0, 1.02])
def test_facet_rounding_base_facets_matches_oracle(gpu, exchange_tol):
    """Completion from the base polytope's facets (orthonormalised in the free
    subspace), with and without the maximum-volume exchange, against the
    oracle's restatement: the same chosen rows (P equal to FP32 rounding),
    completion rows orthonormal and orthogonal to the chosen facets, and
    after the exchange no facet whose coefficient in the basis exceeds the
    tolerance. d = 40 with 0..150 slabs, and the conditioned-cube shape
    d = 100 with up to 300 extra rows (500 candidates)."""
    import naja_amd

    t = gpu
    for d, ks, seed in ((40, [0, 10, 36, 80, 120, 200, 300], 5), (100, [0, 60, 150, 198, 240, 300], 6)):
        A0, b0 = np.vstack([np.eye(d), -np.eye(d)]), np.ones(2 * d)
        EA, Eb = conditions(d, ks, seed=seed)
        batch = naja_amd.VariantBatch(A0, b0, EA, Eb)
        P = naja_amd.facet_rounding(batch, None, None, None, None, 0.02,
                                    completion=naja_amd.FACET_COMPLETE_BASE_FACETS, exchange_tol=exchange_tol)
        t.cuda.synchronize()
        P = P.cpu().numpy().astype(np.float64)
        for v, k in enumerate(ks):
            Pref, sel, N = op.facet(A0, b0, EA[v], Eb[v], None, 0.02, completion=op.COMPLETE_BASE_FACETS,
                                    exchange_tol=exchange_tol)
            fac = sel < k
            err = np.abs(P[v] - Pref).max() / np.abs(Pref).max()
            ident = np.abs(N @ P[v] - np.eye(d)).max()
            print(f"[facet-base d{d} v{v}] k={k} facets={int(fac.sum())} tol={exchange_tol} rel err {err:.2e} "
                  f"|NP - I| {ident:.2e}", flush=True)
            assert err <= 1e-5 and ident <= 1e-4
            F = N[~fac]
            if len(F):
                assert np.abs(F @ F.T - np.eye(len(F))).max() <= 1e-12
                if fac.any():
                    assert np.abs(F @ N[fac].T).max() <= 1e-12 * np.abs(N[fac]).max()
            if k == 0:  # the cube keeps its own axes
                assert np.allclose(np.abs(P[v]).max(axis=0), 1.0, atol=1e-6)
                assert np.allclose((np.abs(P[v]) > 0.5).sum(axis=0), 1)
            if exchange_tol and fac.all():
                U = EA[v] / Eb[v][:, None]
                C = U @ np.linalg.inv(N)
                assert np.abs(C).max() <= exchange_tol * (1 + 1e-3)


def test_facet_rounding_with_base_rows_and_centre(gpu):
    """include_base: the base polytope's own rows are candidates too (a
    polytope with no extra rows, as a model directory holds it), at an
    interior centre x_c != 0; the rounded chain stays inside."""
    import naja_amd

    t = gpu
    rng = np.random.default_rng(9)
    d, m = 24, 90
    A = np.vstack([rng.standard_normal((m, d)), np.eye(d), -np.eye(d)])
    b = np.concatenate([np.ones(m), 3.0 * np.ones(2 * d)])
    xc = (0.02 * rng.standard_normal(d)).astype(np.float32)
    assert (b - A @ xc > 0).all()
    batch = naja_amd.VariantBatch(A, b, [np.zeros((0, d))], [np.zeros(0)])
    P0 = naja_amd.ellipsoid_rounding(batch, None, None, None, x_c=t.from_numpy(xc[None]).cuda())
    P = naja_amd.facet_rounding(batch, None, None, None, P0, 0.1, x_c=t.from_numpy(xc[None]).cuda(),
                                include_base=True)
    t.cuda.synchronize()
    Pref, sel, N = op.facet(A, b, np.zeros((0, d)), np.zeros(0), P0.cpu().numpy()[0], 0.1, x_c=xc,
                            include_base=True)
    Pg = P.cpu().numpy()[0].astype(np.float64)
    assert np.abs(Pg - Pref).max() <= 1e-5 * np.abs(Pref).max()
    assert (sel < A.shape[0]).sum() == d  # 138 candidate facets: all d directions from facets
    Y = naja_amd.coordinate_hit_and_run_batched(
        batch, None, None, None, t.zeros((1, 3, d), device="cuda"), 400, 4, 3, seed=2, pair_prob=0.3,
        P=P, c=t.from_numpy(xc[None]).cuda())
    t.cuda.synchronize()
    y = Y.cpu().numpy()[0].astype(np.float64)
    assert (A @ y.T - b[:, None]).max() <= 1e-5


def test_facet_rounding_rejects_exterior_centre(gpu):
    import naja_amd

    t = gpu
    d = 6
    A0, b0 = np.vstack([np.eye(d), -np.eye(d)]), np.ones(2 * d)
    EA, Eb = conditions(d, [4], seed=1)
    batch = naja_amd.VariantBatch(A0, b0, EA, Eb)
    P0 = t.eye(d, device="cuda")[None].contiguous()
    xc = t.full((1, d), 5.0, device="cuda")
    with pytest.raises(naja_amd.NajaInvalidArgument):
        naja_amd.facet_rounding(batch, None, None, None, P0, 0.1, x_c=xc)
