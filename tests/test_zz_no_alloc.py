"""naja_diag_noalloc (the allocation-free diagnostics entry point, scratch in
a static per-GPU device buffer) against naja_diag, through the C ABI: the
same ESS, per-chain ESS and split-R-hat bit for bit, on two calls in a row
(the second smaller, and without a per-chain output, so it reads scratch the
first call left behind), and a clear error beyond the buffer's 64 MiB."""
import ctypes

import numpy as np
import pytest

from tests.test_oracle_diag import ar1

pytestmark = pytest.mark.gpu


def samples_of(spc, nch, dim, phi, seed):
    rng = np.random.default_rng(seed)
    series = ar1(spc, phi, rng, nseries=nch * dim).reshape(nch, dim, spc)
    S = np.zeros((spc * nch, dim), np.float32)
    for c in range(nch):
        S[c::nch] = series[c].T
    return S


def call(t, fn, S, nch, layout, with_chain):
    import naja_amd

    lib = naja_amd.load_library()
    n_total, dim = S.shape
    ess = t.full((dim,), -1.0, device="cuda")
    rhat = t.full((dim,), -1.0, device="cuda")
    ec = t.full((dim, nch), -1.0, device="cuda") if with_chain else None
    rc = getattr(lib, fn)(S.data_ptr(), dim, n_total, nch, layout, ess.data_ptr(), rhat.data_ptr(),
                          ec.data_ptr() if ec is not None else None, None)
    t.cuda.synchronize()
    return rc, ess.cpu().numpy(), rhat.cpu().numpy(), (ec.cpu().numpy() if ec is not None else None)


@pytest.mark.parametrize("layout", [0, 1])
def test_noalloc_diag_equals_naja_diag(gpu, layout):
    t = gpu
    cases = [(12500, 4, 100, 0.95, True), (333, 3, 7, 0.99, False)]  # second call smaller
    for spc, nch, dim, phi, with_chain in cases:
        print(f"[noalloc] layout={layout} spc={spc} chains={nch} dims={dim}", flush=True)
        S = t.from_numpy(samples_of(spc, nch, dim, phi, spc + layout)).cuda()
        rc0, e0, r0, c0 = call(t, "naja_diag", S, nch, layout, True)
        rc1, e1, r1, c1 = call(t, "naja_diag_noalloc", S, nch, layout, with_chain)
        assert rc0 == 0 and rc1 == 0
        np.testing.assert_array_equal(e1, e0)
        np.testing.assert_array_equal(r1, r0)
        if with_chain:
            np.testing.assert_array_equal(c1, c0)


def test_noalloc_diag_rejects_calls_beyond_its_buffer(gpu):
    import naja_amd

    t = gpu
    lib = naja_amd.load_library()
    dim, nch, spc = 2400, 4, 2000  # 2400 x 4 series of 2048 floats: 75 MiB > 64 MiB
    S = t.zeros((nch * spc, dim), device="cuda")
    ess = t.empty(dim, device="cuda")
    rhat = t.empty(dim, device="cuda")
    rc = lib.naja_diag_noalloc(S.data_ptr(), dim, nch * spc, nch, 0, ess.data_ptr(), rhat.data_ptr(), None, None)
    assert rc == naja_amd.NAJA_ERR_UNSUPPORTED
    assert b"static buffer" in lib.naja_last_error()
