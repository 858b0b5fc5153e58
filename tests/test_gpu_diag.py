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


@pytest.mark.parametrize("spc,nch,dim,phi", [(12500, 4, 6, 0.9), (2000, 8, 5, 0.5), (333, 3, 4, 0.99), (2000, 4, 8, 0.6),
                                             (4, 2, 3, 0.1), (5, 1, 2, 0.0), (16384, 2, 2, 0.7),
                                             (200, 37, 300, 0.5), (77, 4, 100, 0.8)])
def test_diag_matches_oracle(gpu, spc, nch, dim, phi):
    rng = np.random.default_rng(spc + nch)
    series = ar1(spc, phi, rng, nseries=nch * dim).reshape(nch, dim, spc) + rng.standard_normal((nch, dim, 1))
    S = np.zeros((spc * nch, dim), np.float32)
    for c in range(nch):
        S[c::nch] = series[c].T
    e, r, ec = gpu_diag(gpu, S, nch)
    e_ref, ec_ref = oracle.ess(S, nch)
    if spc >= 4 and spc // 2 >= 2:
        r_ref, _, _ = oracle.split_rhat(S, nch)
        assert np.abs(r - r_ref).max() <= 1e-5 * np.abs(r_ref).max()
    rel = np.abs(ec - ec_ref) / ec_ref
    print(f"ess spc={spc}: max rel {rel.max():.3e}", flush=True)
    # float FFT rounding differs from the double oracle; the truncation point
    # may move by one pair only when a pair sum sits within rounding of zero
    assert np.median(rel) <= 1e-4
    assert (rel <= 1e-3).mean() >= 0.9
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
