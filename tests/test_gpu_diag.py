"""GPU ESS / split-R-hat vs the oracle's restatement of mcmc.cpp:17-194."""
import numpy as np
import pytest

import oracle
from tests.test_oracle_diag import ar1

pytestmark = pytest.mark.gpu


def gpu_diag(t, S, nch, layout=0):
    import naja_amd

    e, r, ec = naja_amd.diagnostics(t.from_numpy(np.ascontiguousarray(S, np.float32)), nch, layout)
    t.cuda.synchronize()
    return e.cpu().numpy(), r.cpu().numpy(), ec.cpu().numpy()


def rounding_sensitive(S, nch, eps=1e-5):
    """Per (dim, chain): True where the oracle's truncation could move under
    float rounding of the autocovariance, i.e. where a pair sum
    rho_{2k-1} + rho_{2k} up to and including the first negative one lies
    within eps of zero (mcmc.cpp:60-76). Elsewhere the ESS is a smooth
    function of the autocovariance and must agree to float rounding."""
    n_total, dim = S.shape
    spc = n_total // nch
    out = np.zeros((dim, nch), bool)
    if spc < 4:
        return out
    for d in range(dim):
        for c in range(nch):
            acov = oracle.autocov(S[c::nch, d]).astype(np.float64)
            if acov[0] < 1e-30:
                continue
            k = np.arange(1, spc // 2)
            ps = acov[2 * k - 1] / acov[0] + acov[2 * k] / acov[0]
            neg = np.flatnonzero(ps < 0)
            last = neg[0] + 1 if neg.size else ps.size
            out[d, c] = bool(np.abs(ps[:last]).min(initial=np.inf) < eps)
    return out


DIAG_CASES = [(12500, 4, 6, 0.9), (2000, 8, 5, 0.5), (333, 3, 4, 0.99), (2000, 4, 8, 0.6),
              (4, 2, 3, 0.1), (5, 1, 2, 0.0), (16384, 2, 2, 0.7), (200, 37, 300, 0.5), (77, 4, 100, 0.8),
              # truncation beyond the exact lags of the half-size pass: the full-size redo pass
              (8000, 2, 3, 0.999), (16000, 2, 2, 0.9995), (4000, 3, 2, 0.998), (1000, 4, 3, 0.9999),
              (2000, 3, 2, 1.0), (3001, 2, 3, 0.998),
              # beyond the LDS transform (N / 2 > 16384 points): the global-memory path
              (16385, 2, 2, 0.9), (20001, 3, 2, 0.95), (50000, 2, 3, 0.9), (40000, 2, 2, 1.0)]


@pytest.mark.parametrize("spc,nch,dim,phi", DIAG_CASES)
def test_diag_matches_oracle(gpu, spc, nch, dim, phi):
    rng = np.random.default_rng(spc + nch)
    if phi < 1.0:
        series = ar1(spc, phi, rng, nseries=nch * dim).reshape(nch, dim, spc)
    else:  # random walk: the truncation runs out to about n/2
        series = np.cumsum(rng.standard_normal((nch, dim, spc)), axis=2).astype(np.float32)
    series = series + rng.standard_normal((nch, dim, 1)).astype(np.float32)
    S = np.zeros((spc * nch, dim), np.float32)
    for c in range(nch):
        S[c::nch] = series[c].T
    e, r, ec = gpu_diag(gpu, S, nch)
    e_ref, ec_ref = oracle.ess(S, nch)
    if spc >= 4 and spc // 2 >= 2:
        r_ref, _, _ = oracle.split_rhat(S, nch)
        assert np.abs(r - r_ref).max() <= 1e-5 * np.abs(r_ref).max()
    rel = np.abs(ec - ec_ref) / ec_ref
    sens = rounding_sensitive(S, nch)
    tight = rel[~sens]
    print(f"ess spc={spc} phi={phi}: max rel {rel.max():.3e}, off the rounding-sensitive series "
          f"{tight.max(initial=0.0):.3e} ({int(sens.sum())} of {sens.size} sensitive)", flush=True)
    # The GPU transforms in float and sums the truncated pairs in FP64; the
    # oracle autocovariance is FP64 rounded to float and summed in float.
    assert tight.max(initial=0.0) <= 1e-5
    np.testing.assert_array_equal(e, ec.min(1))


@pytest.mark.parametrize("n,nch,dim", [(4000, 4, 3), (1400, 7, 150)])
def test_diag_blocks_layout(gpu, n, nch, dim):
    """Chain-major sample rows (layout 1), including several chain and dim
    tiles of the series transpose."""
    rng = np.random.default_rng(1)
    S = ar1(n, 0.8, rng, nseries=dim).T.copy()
    e0, r0, _ = gpu_diag(gpu, S, nch, layout=0)
    e1, r1, _ = gpu_diag(gpu, S, nch, layout=1)
    e1_ref, _ = oracle.ess(S, nch, layout=1)
    r1_ref, _, _ = oracle.split_rhat(S, nch, layout=1)
    assert np.abs(r1 - r1_ref).max() <= 1e-5
    assert (np.abs(e1 - e1_ref) / e1_ref).max() <= 1e-3


def test_diag_errors(gpu):
    import naja_amd

    t = gpu
    with pytest.raises(naja_amd.NajaInvalidArgument):
        naja_amd.diagnostics(t.zeros((10, 3)), 3)  # not divisible
    with pytest.raises(naja_amd.NajaInvalidArgument):
        naja_amd.diagnostics(t.zeros((6, 3)), 2)  # fewer than 4 per chain


def test_long_chains_batched_and_blocks_layout(gpu):
    """Chains longer than the LDS transform through naja_diag_batched and the
    chain-major layout: equal to per-variant calls and to the oracle."""
    import naja_amd

    t = gpu
    rng = np.random.default_rng(21)
    nv, nch, spc, dim = 2, 2, 30000, 3
    X = np.stack([ar1(spc * nch, 0.8, rng, nseries=dim).T for _ in range(nv)]).astype(np.float32)
    Xd = t.from_numpy(X).cuda()
    ess, rhat, emin, rmax = naja_amd.diagnostics_batched(Xd, nch, layout=1)
    for v in range(nv):
        e_ref, _ = oracle.ess(X[v], nch, layout=1)
        r_ref, _, _ = oracle.split_rhat(X[v], nch, layout=1)
        assert (np.abs(ess[v].cpu().numpy() - e_ref) / e_ref).max() <= 1e-5
        assert np.abs(rhat[v].cpu().numpy() - r_ref).max() <= 1e-5


def test_batched_diagnostics_equal_per_variant_calls(gpu):
    """naja_diag_batched over stacked variants equals naja_diag per variant,
    and its per-variant summaries are the min ESS / max R-hat."""
    import naja_amd

    t = gpu
    rng = np.random.default_rng(12)
    nv, nch, spc, dim = 5, 4, 300, 7
    X = np.cumsum(rng.standard_normal((nv, nch * spc, dim)), axis=1).astype(np.float32) * 0.1
    X += rng.standard_normal((nv, nch * spc, dim)).astype(np.float32)
    Xd = t.from_numpy(X).cuda()
    ess, rhat, emin, rmax = naja_amd.diagnostics_batched(Xd, nch)
    for v in range(nv):
        e1, r1, _ = naja_amd.diagnostics(Xd[v], nch)
        np.testing.assert_array_equal(ess[v].cpu().numpy(), e1.cpu().numpy())
        np.testing.assert_array_equal(rhat[v].cpu().numpy(), r1.cpu().numpy())
        assert float(emin[v]) == float(e1.min())
        assert float(rmax[v]) == float(r1.max())
