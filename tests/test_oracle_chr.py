"""Pins the CPU CHR oracle (oracle/chr_oracle.c) against closed forms,
brute-force chords and the reference's own CHR tests (CPU only)."""
import numpy as np
import pytest

import oracle
from tests.polytopes import cube, random_polytope, simplex, thin_slab


def alpha_uniform(seed, variant, t, chain):
    k0 = seed & 0xFFFFFFFF
    k1 = ((seed >> 32) ^ variant) & 0xFFFFFFFF
    w = oracle.philox_block((t & 0xFFFFFFFF, t >> 32, chain >> 2, 0xFFFFFFFF), k0, k1)[chain & 3]
    return ((((int(w) >> 9) << 1) | 1) * 2.0 ** -24)


def brute_force_steps(A, b, y0, moves, seed, nsteps, chain=0):
    """Independent FP64 walk: chord endpoints from the half-space intersection
    of the line y + a u, then a = lo + U (hi - lo)."""
    y = np.array(y0, dtype=np.float64)
    out = []
    d = A.shape[1]
    for t in range(nsteps):
        typ, i, j = moves[t]
        u = np.zeros(d)
        if typ == 0:
            u[i] = 1.0
        else:
            u[i], u[j] = np.float32(0.70710678118654752440), -np.float32(0.70710678118654752440)
        au = A @ u
        s = b - A @ y
        pos, neg = au > 0, au < 0
        hi = np.min(s[pos] / au[pos])
        lo = np.max(s[neg] / au[neg])
        U = alpha_uniform(seed, 0, t, chain)
        y = y + (lo + U * (hi - lo)) * u
        out.append(y.copy())
    return np.array(out)


@pytest.mark.parametrize("pair_prob", [0.0, 0.5])
def test_oracle_matches_brute_force_chord(pair_prob):
    A, b, y0 = random_polytope(5, 12, seed=7)
    A32 = A.astype(np.float32)
    n = 300
    r = oracle.chr_run_f32(A32, b, y0[None, :], nspc=n, thin=1, seed=99, pair_prob=pair_prob, nlog=n)
    ref = brute_force_steps(A32.astype(np.float64), b, y0, r["move_log"], 99, n)
    err = np.abs(r["samples"].astype(np.float64) - ref).max()
    assert err < 1e-4, err


def test_first_sample_is_after_one_step_and_last_steps_unsaved():
    """chr.cu:283 saves when t % thin == 0: sample 0 is the state after step 0."""
    A, b, y0 = random_polytope(4, 10, seed=1)
    A32 = A.astype(np.float32)
    full = oracle.chr_run_f32(A32, b, y0[None, :], nspc=12, thin=1, seed=5)["samples"]
    thinned = oracle.chr_run_f32(A32, b, y0[None, :], nspc=4, thin=3, seed=5)["samples"]
    np.testing.assert_array_equal(thinned, full[[0, 3, 6, 9]])


def test_samples_chain_interleaved():
    """Row s*nchains + c holds sample s of chain c (chr.cu:284, chr.cu:460)."""
    A, b, y0 = cube(3)
    y0s = np.zeros((3, 3), dtype=np.float32)
    both = oracle.chr_run_f32(A.astype(np.float32), b, y0s, nspc=20, thin=2, seed=8)["samples"]
    for c in range(3):
        # chain c with G = 1 uses group c and alpha stream c
        one = oracle.chr_run_f32(A.astype(np.float32), b, y0s, nspc=20, thin=2, seed=8)["samples"][c::3]
        np.testing.assert_array_equal(both[c::3], one)
    assert not np.array_equal(both[0::3], both[1::3])


def _z_mean_var(x, rho):
    N = x.shape[0]
    infl = (1 + rho) / (1 - rho)
    z_mean = x.mean(0) / np.sqrt((1.0 / 3.0) / N * infl)
    z_var = (x.var(0) - 1.0 / 3.0) / np.sqrt((4.0 / 45.0) / N * infl)
    return z_mean, z_var


def test_cube_moments_within_mc_se():
    """test_gpu_sampling_cube_moments.cu:22-121 on the oracle: [-1,1]^8, 8 chains
    x 20000, thin 20. A coordinate move resamples x_i ~ U[-1,1] exactly, so the
    lag-1 sample correlation is rho = (1 - 1/d)^thin."""
    d, nch, nspc, thin = 8, 8, 20000, 20
    A, b, _ = cube(d)
    r = oracle.chr_run_f32(A.astype(np.float32), b, np.zeros((nch, d), np.float32), nspc, thin, seed=42)
    x = r["samples"].astype(np.float64)
    zm, zv = _z_mean_var(x, (1 - 1 / d) ** thin)
    assert np.all(np.abs(zm) < 3.5) and np.all(np.abs(zv) < 3.5), (zm, zv)
    assert abs(zm.mean()) * np.sqrt(d) < 3 and abs(zv.mean()) * np.sqrt(d) < 3
    corr = np.corrcoef(x.T)
    assert np.abs(corr - np.eye(d)).max() < 0.05  # the reference's own bound
    assert np.abs(x).max() <= 1.0 + 1e-5


def test_simplex_moments_within_mc_se():
    """test_gpu_sampling_simplex_moments.cu:25-158: {x >= 0, sum x <= 1} in R^6,
    8 x 30000, thin 20, start 1/(n+1)."""
    n, nch, nspc, thin = 6, 8, 30000, 20
    A, b, y0 = simplex(n)
    r = oracle.chr_run_f32(A.astype(np.float32), b, np.tile(y0, (nch, 1)).astype(np.float32), nspc, thin, seed=314)
    x = r["samples"].astype(np.float64)
    mean_t = 1.0 / (n + 1)
    var_t = n / ((n + 1) ** 2 * (n + 2))
    cov_t = -1.0 / ((n + 1) ** 2 * (n + 2))
    ess, _ = oracle.ess(r["samples"], nch)
    ess_tot = ess * nch
    se_mean = np.sqrt(var_t / ess_tot)
    assert np.all(np.abs(x.mean(0) - mean_t) < 3.5 * se_mean)
    assert np.all(np.abs(x.var(0) - var_t) / var_t < 0.20)
    cov = np.cov(x.T)
    off = cov[~np.eye(n, dtype=bool)]
    assert np.all(np.abs(off - cov_t) < 0.008)
    viol = (A @ x.T - b[:, None]).max()
    assert viol <= 1e-6


@pytest.mark.parametrize("case", ["cube10", "simplex8", "slab6", "rand15", "rand30"])
def test_feasibility(case):
    """test_gpu_sampling_feasibility.cu:22-193: every sample satisfies A y <= b + eps."""
    if case == "cube10":
        A, b, y0 = cube(10)
    elif case == "simplex8":
        A, b, y0 = simplex(8)
    elif case == "slab6":
        A, b, y0 = thin_slab(6, 1e-4)
    elif case == "rand15":
        A, b, y0 = random_polytope(15, 60, seed=999)
    else:
        A, b, y0 = random_polytope(30, 120, seed=2024)
    nch = 4
    r = oracle.chr_run_f32(A.astype(np.float32), b, np.tile(y0, (nch, 1)).astype(np.float32), 5000, 10,
                           seed=17, pair_prob=0.3)
    x = r["samples"].astype(np.float64)
    viol = (A.astype(np.float32).astype(np.float64) @ x.T - b[:, None]).max()
    assert viol <= 1e-5, viol


def test_thin_slab_pair_moves_mix():
    """test_gpu_thin_polytope.cu:48-141: coordinate moves are pinned by the slab,
    pair moves (e_i - e_j)/sqrt2 keep the free dimensions mixing."""
    d, eps = 10, 1e-4
    A, b, y0 = thin_slab(d, eps)
    A32 = A.astype(np.float32)
    for pp in (0.0, 0.3):
        r = oracle.chr_run_f32(A32, b, np.zeros((4, d), np.float32), 10000, 5, seed=77, pair_prob=pp)
        x = r["samples"].astype(np.float64)
        slab = x[:, :3].sum(1)
        assert np.abs(slab).max() <= eps + 1e-6
        assert (A32.astype(np.float64) @ x.T - b[:, None]).max() <= 1e-6
        if pp > 0:
            assert x[:, 3:].var(0).min() > 0.1


def test_pair_step_restatement_square():
    """tests/test_chr_pair_steps.cpp:103-163: |x|,|y| <= 1 with |x+y| <= 1e-4."""
    A = np.array([[1, 0], [-1, 0], [0, 1], [0, -1], [1, 1], [-1, -1]], dtype=np.float32)
    b = np.array([1, 1, 1, 1, 1e-4, 1e-4])
    coord = oracle.chr_run_f32(A, b, np.zeros((1, 2), np.float32), 200, 1, seed=123)
    pair = oracle.chr_run_f32(A, b, np.zeros((1, 2), np.float32), 200, 1, seed=123, pair_prob=1.0)
    xc, xp = coord["samples"], pair["samples"]
    assert np.abs(xc[:, 0] - xc[:, 1]).max() < 1e-2
    assert np.abs(xp[:, 0] - xp[:, 1]).max() > 1e-1
    for r in (coord, pair):
        assert r["s_final"].min() > -1e-6


def test_slack_tracks_b_minus_Ay_and_resync():
    A, b, y0 = random_polytope(12, 40, seed=3)
    A32 = A.astype(np.float32)
    for R in (0, 50):
        r = oracle.chr_run_f32(A32, b, np.zeros((2, 12), np.float32), 2000, 1, seed=4, pair_prob=0.3,
                               resync_interval=R)
        s_true = b[None, :] - r["y_final"].astype(np.float64) @ A32.astype(np.float64).T
        err = np.abs(r["s_final"] - s_true).max()
        assert err < 1e-4, (R, err)


def test_ksparse_moves_stay_feasible():
    A, b, y0 = random_polytope(10, 30, seed=8)
    A32 = A.astype(np.float32)
    r = oracle.chr_run_f32(A32, b, np.zeros((2, 10), np.float32), 3000, 2, seed=6, ksparse_prob=0.3,
                           ksparse_k=5, pair_prob=0.2, nlog=3000)
    assert (r["move_log"][:, 0] == 2).mean() > 0.2
    x = r["samples"].astype(np.float64)
    assert (A32.astype(np.float64) @ x.T - b[:, None]).max() < 1e-5


def test_fp32_drift_against_fp64():
    """FP32 state vs the all-FP64 walk on the same stream. CHR is chaotic: a
    1-ulp difference grows exponentially (about 1e-7 -> 1e-3 in 300 steps on
    this polytope), so the two agree only over a short horizon; long-run
    agreement is only possible bit for bit (the GPU kernel vs the FP32 oracle)."""
    A, b, y0 = random_polytope(10, 30, seed=21)
    n = 2000
    r32 = oracle.chr_run_f32(A.astype(np.float32), b, np.zeros((1, 10), np.float32), n, 1, seed=2,
                             pair_prob=0.3, resync_interval=1000)
    r64 = oracle.chr_run_f64(A.astype(np.float32).astype(np.float64), b, np.zeros((1, 10)), n, 1, seed=2,
                             pair_prob=0.3, resync_interval=1000)
    drift = np.abs(r32["samples"].astype(np.float64) - r64["samples"]).max(1)
    assert drift[:20].max() < 1e-5, drift[:20].max()
    # same chord structure: both stay inside the polytope
    for x in (r32["samples"].astype(np.float64), r64["samples"]):
        assert (A.astype(np.float32).astype(np.float64) @ x.T - b[:, None]).max() < 1e-5


def test_chains_per_group_shares_moves():
    A, b, y0 = cube(6)
    A32 = A.astype(np.float32)
    r = oracle.chr_run_f32(A32, b, np.zeros((4, 6), np.float32), 500, 1, seed=9, chains_per_group=4)
    x = r["samples"].reshape(500, 4, 6)
    # all four chains change the same coordinate every step
    dx = np.diff(np.concatenate([np.zeros((1, 4, 6), np.float32), x]), axis=0) != 0
    assert (dx[:, 0, :] == dx[:, 1, :]).all() and (dx[:, 0, :] == dx[:, 3, :]).all()
    assert not np.array_equal(x[:, 0], x[:, 1])
