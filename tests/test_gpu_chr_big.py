"""The large-polytope CHR kernel (16 wavefronts, 10-22 rows per lane, no
reciprocal column) against the FP32 oracle, bit for bit, up to the
large-scale shape (n = 10600 reactions: m = 21200 rows, d = 2400), and
against the register-resident kernels where both apply. [T; -T] polytopes
take the paired-row register kernels (8 wavefronts, one packed row per pair);
with their rows shuffled (no pairing) the same polytopes take the large
kernel, so both paths are checked at the large-scale shape."""
import numpy as np
import pytest

import oracle
from tests.polytopes import genome_like

pytestmark = pytest.mark.gpu


def run_gpu(t, A, b, nch, nspc, thin, tpb=-1, **kw):
    import naja_amd

    X0 = t.zeros((nch, A.shape[1]), dtype=t.float32, device="cuda")
    S = naja_amd.coordinate_hit_and_run(t.from_numpy(A.astype(np.float32)).cuda(), t.from_numpy(b).cuda(), X0,
                                        nspc, thin, nch, tpb_ss=tpb, **kw)
    t.cuda.synchronize()
    return S.cpu().numpy()


def shuffled(g, seed):
    perm = np.random.default_rng(seed).permutation(g["A"].shape[0])
    return dict(g, A=g["A"][perm], b=g["b"][perm])


@pytest.mark.parametrize("rows", ["paired", "shuffled"])
@pytest.mark.parametrize("n,d,nch,nspc,thin", [(10600, 2400, 4, 40, 5), (7000, 300, 3, 100, 7)])
def test_big_kernel_matches_oracle(gpu, n, d, nch, nspc, thin, rows):
    g = genome_like(n, d, seed=n)
    if rows == "shuffled":
        g = shuffled(g, n)
    A32 = g["A"].astype(np.float32)
    kw = dict(seed=3, pair_prob=0.3, ksparse_prob=0.03, ksparse_k=5, resync_interval=64)
    print(f"[big kernel] m={A32.shape[0]} d={d} chains={nch} steps={nspc * thin}", flush=True)
    S = run_gpu(gpu, A32, g["b"], nch, nspc, thin, **kw)
    ref = oracle.chr_run_f32(A32, g["b"], np.zeros((nch, d), np.float32), nspc, thin, **kw)["samples"]
    assert np.array_equal(S, ref), f"max |diff| {np.abs(S - ref).max()}"
    viol = (S.astype(np.float64) @ A32.T.astype(np.float64) - g["b"][None, :]).max()
    assert viol <= 1e-4


def test_big_kernel_equals_register_kernel(gpu):
    """m = 10000 rows fits both the 8-wavefront register kernel and (with a
    1024-thread request) the large kernel: identical samples."""
    g = shuffled(genome_like(5000, 200, seed=8), 8)
    kw = dict(seed=9, pair_prob=0.3, resync_interval=100)
    S_small = run_gpu(gpu, g["A"], g["b"], 4, 150, 6, tpb=-1, **kw)
    S_big = run_gpu(gpu, g["A"], g["b"], 4, 150, 6, tpb=1024, **kw)
    assert np.array_equal(S_small, S_big)
