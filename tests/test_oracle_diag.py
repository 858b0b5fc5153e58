"""Pins the CPU ESS / split-R-hat oracle (oracle/diag_oracle.c) against
numpy, closed forms and the reference's own kissfft (CPU only)."""
import numpy as np
import pytest

import oracle


def ar1(n, phi, rng, nseries=1):
    x = np.zeros((nseries, n))
    e = rng.standard_normal((nseries, n))
    x[:, 0] = e[:, 0] / np.sqrt(1 - phi * phi)
    for t in range(1, n):
        x[:, t] = phi * x[:, t - 1] + e[:, t]
    return x.astype(np.float32)


def np_autocov(x):
    """float64 biased autocovariance via numpy.correlate on the float-centred series."""
    x = np.asarray(x, dtype=np.float32)
    mean = np.float32(x.astype(np.float64).sum() / x.size)
    xc = (x - mean).astype(np.float64)
    n = x.size
    full = np.correlate(xc, xc, mode="full")[n - 1:]
    return full / n


@pytest.mark.parametrize("n", [4, 5, 17, 64, 1000, 2047])
def test_autocov_fft_matches_direct_and_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n).astype(np.float32) + 3.0
    a_fft = oracle.autocov(x)
    a_dir = oracle.autocov(x, direct=True)
    a_np = np_autocov(x)
    scale = a_np[0]
    assert np.abs(a_fft - a_np).max() <= 1e-6 * scale
    assert np.abs(a_dir - a_np).max() <= 1e-6 * scale


def test_autocov_matches_reference_kissfft():
    """oracle/_ref/libref_kiss.so runs mcmc.cpp:17-57 on the reference's kissfft."""
    if oracle.kiss() is None:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    rng = np.random.default_rng(5)
    for n in (12500, 2000, 333):
        x = ar1(n, 0.9, rng)[0]
        a_k = oracle.kiss_autocov(x)
        a_o = oracle.autocov(x)
        assert np.abs(a_k - a_o).max() <= 2e-5 * a_o[0]
        # the truncated ESS agrees to float noise
        assert abs(oracle.ess_from_acov(a_k) - oracle.ess_from_acov(a_o)) <= 1e-3 * oracle.ess_from_acov(a_o) + 0.5


def test_ess_ar1_closed_form():
    """AR(1) with coefficient phi: ESS -> n (1 - phi) / (1 + phi)."""
    rng = np.random.default_rng(1)
    n, phi = 20000, 0.8
    xs = ar1(n, phi, rng, nseries=8)
    ess = np.array([oracle.ess_one_chain(x) for x in xs])
    expect = n * (1 - phi) / (1 + phi)
    assert abs(ess.mean() / expect - 1) < 0.1


def test_ess_short_and_constant_chains():
    assert oracle.ess_one_chain(np.array([1, 2, 3], np.float32)) == 3.0
    assert oracle.ess_one_chain(np.full(100, 2.5, np.float32)) == 1.0


def test_ess_iid_near_n():
    rng = np.random.default_rng(9)
    x = rng.standard_normal(10000).astype(np.float32)
    e = oracle.ess_one_chain(x)
    assert 0.8 * 10000 < e <= 1.3 * 10000


def test_layouts_agree():
    rng = np.random.default_rng(3)
    nch, spc, dim = 4, 500, 3
    chains = ar1(spc, 0.5, rng, nseries=nch * dim).reshape(nch, dim, spc)
    inter = np.zeros((spc * nch, dim), np.float32)
    blocks = np.zeros((spc * nch, dim), np.float32)
    for c in range(nch):
        inter[c::nch] = chains[c].T
        blocks[c * spc:(c + 1) * spc] = chains[c].T
    e1, ec1 = oracle.ess(inter, nch, layout=0)
    e2, ec2 = oracle.ess(blocks, nch, layout=1)
    np.testing.assert_array_equal(ec1, ec2)
    r1 = oracle.split_rhat(inter, nch, layout=0)[0]
    r2 = oracle.split_rhat(blocks, nch, layout=1)[0]
    np.testing.assert_array_equal(r1, r2)
    # naja validate's grouping of an interleaved file differs (mcmc.cpp:103-110)
    e3, _ = oracle.ess(inter, nch, layout=1)
    assert not np.array_equal(e1, e3)


def test_rhat_closed_forms():
    nch, half = 3, 50
    spc = 2 * half
    # half chain j alternates mean_j -+ 1: sample var = half/(half-1), mean = mean_j
    means = np.array([0.0, 0.5, 1.0, 1.5, -1.0, 2.0])
    S = np.zeros((nch * spc, 1), np.float32)
    for c in range(nch):
        for h in range(2):
            j = 2 * c + h
            vals = means[j] + np.where(np.arange(half) % 2 == 0, -1.0, 1.0)
            S[c * spc + h * half:c * spc + (h + 1) * half, 0] = vals
    r, a11, a12 = oracle.split_rhat(S, nch, layout=1)
    W = half / (half - 1.0)
    B = half * np.var(means, ddof=1)
    expect = np.sqrt(((half - 1) / half * W + B / half) / W)
    assert abs(r[0] - expect) < 1e-6
    assert a11 == (expect > 1.1) and a12 == (expect > 1.2)
    # zero within-chain variance -> R = 1 (mcmc.cpp:177)
    r0, _, _ = oracle.split_rhat(np.ones((nch * spc, 2), np.float32), nch)
    assert (r0 == 1.0).all()


def test_rhat_iid_near_one():
    rng = np.random.default_rng(2)
    S = rng.standard_normal((8000, 5)).astype(np.float32)
    r, _, _ = oracle.split_rhat(S, 4)
    assert np.abs(r - 1).max() < 0.01


def test_percentile_matches_numpy_linear():
    rng = np.random.default_rng(4)
    v = rng.standard_normal(37).astype(np.float32)
    for q in (0.1, 0.5, 0.9):
        assert abs(oracle.percentile(v, q) - np.percentile(v.astype(np.float64), 100 * q)) < 1e-5


def test_backmap_oracle_matches_numpy():
    rng = np.random.default_rng(11)
    T = rng.standard_normal((9, 5))
    shift = rng.standard_normal(9)
    Y = rng.standard_normal((13, 5)).astype(np.float32)
    V = oracle.backmap(T, shift, Y)
    np.testing.assert_allclose(V, Y.astype(np.float64) @ T.T + shift, rtol=1e-12, atol=1e-12)
