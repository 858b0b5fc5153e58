"""End-to-end `naja sample run` / `naja validate` on the GPU against the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from tests.test_cli import NAJA, npy_header_rule, write_csv

pytestmark = pytest.mark.gpu


def make_model(tmp_path, name="GEM"):
    rng = np.random.default_rng(12)
    d, n = 6, 9
    T = rng.standard_normal((n, d))
    shift = rng.standard_normal(n)
    A = np.vstack([T, -T])
    m_hi = 10 ** rng.uniform(-1, 1, n)
    m_lo = 10 ** rng.uniform(-1, 1, n)
    b = np.concatenate([m_hi, m_lo])
    md = tmp_path / "models" / name
    (md / "rounding").mkdir(parents=True)
    (md / "gem").mkdir()
    pre = md / "rounding" / f"{name}_rounding_"
    write_csv(str(pre) + "A.csv", A)
    write_csv(str(pre) + "b.csv", b[:, None])
    write_csv(str(pre) + "start.csv", np.zeros((d, 1)))
    write_csv(str(pre) + "T.csv", T)
    write_csv(str(pre) + "shift.csv", shift[:, None])
    write_csv(md / "gem" / "l_bounds.csv", (shift - m_lo)[:, None])
    write_csv(md / "gem" / "u_bounds.csv", (shift + m_hi)[:, None])
    (md / "gem" / "reaction_ids.txt").write_text("".join(f"R{i}\n" for i in range(n)))
    return md, A, b, T, shift


def run_dir(out):
    runs = sorted(os.listdir(out))
    return out / runs[-1]


def test_sample_run_reduced_matches_oracle(gpu, tmp_path):
    md, A, b, T, shift = make_model(tmp_path)
    out = tmp_path / "out"
    r = subprocess.run([NAJA, "sample", "run", "--model-dir", str(md), "--out-root", str(out), "--gpu", "0",
                        "--n-chains", "3", "--n-samples", "400", "--thinning", "4", "--pair-prob", "0.3",
                        "--seed", "77", "--write-npy"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    rd = run_dir(out)
    raw = (rd / "samples.npy").read_bytes()
    hdr = npy_header_rule("<f4", "1200, 6")
    assert raw[:len(hdr)] == hdr
    S = np.frombuffer(raw[len(hdr):], dtype=np.float32).reshape(1200, 6)
    ref = oracle.chr_run_f32(A.astype(np.float32), b, np.zeros((3, 6), np.float32), 400, 4, seed=77,
                             pair_prob=0.3)
    assert np.abs(S.astype(np.float64) - ref["samples"]).max() <= 1e-5
    prof = json.loads((rd / "profile.json").read_text())
    assert prof["config"]["n_chains"] == 3 and prof["config"]["thinning"] == 4
    assert prof["performance"]["total_samples"] == 1200
    assert "THINNING=4" in (rd / "config_used.txt").read_text()


def test_sample_run_backmap_filter_and_validate(gpu, tmp_path):
    md, A, b, T, shift = make_model(tmp_path)
    out = tmp_path / "out"
    r = subprocess.run([NAJA, "sample", "run", "--model-dir", str(md), "--out-root", str(out), "--n-chains", "4",
                        "--n-samples", "500", "--seed", "3", "--write-npy", "--backmap", "--bounds-policy", "filter",
                        "--write-samples-valid", "--quiet"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rd = run_dir(out)
    V = np.load(rd / "samples.npy")
    assert V.shape == (2000, 9) and V.dtype == np.float32
    ref = oracle.chr_run_f32(A.astype(np.float32), b, np.zeros((4, 6), np.float32), 500, 1, seed=3)
    Vref = oracle.backmap(T.astype(np.float32).astype(np.float64), shift.astype(np.float32).astype(np.float64),
                          ref["samples"])
    assert np.abs(V - Vref).max() <= 1e-4
    mask = np.load(rd / "valid_mask.npy")
    assert mask.dtype == np.uint8 and mask.shape == (2000,)
    rep = json.loads((rd / "bounds_report.json").read_text())
    assert rep["n_total"] == 2000 and rep["n_valid"] == int(mask.sum())
    sv = np.load(rd / "samples_valid.npy")
    assert sv.shape == (9, int(mask.sum())) and sv.dtype == np.float64

    # naja validate on the reduced samples of the same run
    r2 = subprocess.run([NAJA, "sample", "run", "--model-dir", str(md), "--out-root", str(out), "--n-chains", "4",
                         "--n-samples", "500", "--seed", "3", "--write-npy", "--quiet"], capture_output=True,
                        text=True, timeout=120)
    assert r2.returncode == 0
    red = run_dir(out) / "samples.npy"
    v = subprocess.run([NAJA, "validate", "--samples", str(red), "--n-chains", "4", "--model-dir", str(md)],
                       capture_output=True, text=True, timeout=120)
    assert v.returncode == 0, v.stderr
    js = json.loads(v.stdout)
    S = np.load(red)
    e_ref, _ = oracle.ess(S, 4)
    r_ref, a11, a12 = oracle.split_rhat(S, 4)
    np.testing.assert_allclose(js["per_dim_ess"], np.round(e_ref, 1), rtol=2e-3, atol=0.6)
    assert js["rhat"]["above_1_1"] == a11 and js["chain_layout"] == "interleaved"
    assert js["bounds"]["violations"] == 0 and js["chords"]["positive_dims"] == 6
    vb = subprocess.run([NAJA, "validate", "--samples", str(red), "--n-chains", "4", "--chain-layout",
                         "naja-blocks"], capture_output=True, text=True, timeout=120)
    assert vb.returncode == 0 and json.loads(vb.stdout)["chain_layout"] == "naja-blocks"


@pytest.mark.parametrize("backmap", [False, True])
def test_sample_run_streamed_equals_whole(gpu, tmp_path, backmap):
    """`--stream-chunk` writes samples.npy chunk by chunk; with chunk steps a
    multiple of the resync window the file equals the one-shot run's."""
    md, A, b, T, shift = make_model(tmp_path)
    base = [NAJA, "sample", "run", "--model-dir", str(md), "--gpu", "0", "--n-chains", "3", "--n-samples", "400",
            "--thinning", "4", "--pair-prob", "0.3", "--seed", "77", "--write-npy", "--resync-interval", "200"]
    if backmap:
        base.append("--backmap")
    outs = []
    for extra in ([], ["--stream-chunk", "50"]):
        out = tmp_path / ("out" + "".join(extra))
        r = subprocess.run(base + ["--out-root", str(out)] + extra, capture_output=True, text=True, timeout=120)
        print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
        assert r.returncode == 0
        outs.append((run_dir(out) / "samples.npy").read_bytes())
    assert outs[0] == outs[1]


def make_hull_model(tmp_path):
    """Box [-1, 1]^4 with the equality y0 + y1 - y3 = 0.2 written as two
    opposite rows, a start on it, and a random transform to 7 fluxes."""
    rng = np.random.default_rng(5)
    d, n = 4, 7
    e = np.array([[1.0, 1.0, 0.0, -1.0]])
    A = np.vstack([np.eye(d), -np.eye(d), e, -e])
    b = np.concatenate([np.ones(2 * d), [0.2, -0.2]])
    x0 = np.array([0.1, 0.05, 0.3, -0.05])
    T = rng.standard_normal((n, d))
    shift = rng.standard_normal(n)
    md = tmp_path / "models" / "HULL"
    (md / "rounding").mkdir(parents=True)
    (md / "gem").mkdir()
    pre = md / "rounding" / "HULL_rounding_"
    write_csv(str(pre) + "A.csv", A)
    write_csv(str(pre) + "b.csv", b[:, None])
    write_csv(str(pre) + "start.csv", x0[:, None])
    write_csv(str(pre) + "T.csv", T)
    write_csv(str(pre) + "shift.csv", shift[:, None])
    return md, A, b, x0, T, shift


@pytest.mark.parametrize("backmap", [False, True])
def test_sample_run_affine_hull(gpu, tmp_path, backmap):
    """--affine-hull-tol (job_sampling.cu:196-235): sampling runs in the kernel
    of the tight rows and the output maps back through (B, x0), composed with
    (T, shift) under --backmap; equals the oracle chain on the reduced polytope."""
    import oracle.hull as ohull

    md, A, b, x0, T, shift = make_hull_model(tmp_path)
    out = tmp_path / "out"
    cmd = [NAJA, "sample", "run", "--model-dir", str(md), "--out-root", str(out), "--gpu", "0", "--n-chains", "4",
           "--n-samples", "600", "--seed", "21", "--write-npy", "--affine-hull-tol", "1e-9", "--verbose"]
    if backmap:
        cmd.append("--backmap")
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert "affine_hull_dim  3 (from 4)" in r.stdout
    rd = run_dir(out)
    V = np.load(rd / "samples.npy")
    prof = json.loads((rd / "profile.json").read_text())
    rows, B = ohull.affine_hull(A, b, x0, 1e-9)
    assert rows == [8, 9] and B.shape == (4, 3)
    assert prof["config"]["thinning"] == 1  # max(3 / 6, 1) on the reduced dimension
    # the reduced polytope in the CLI's own FP64 summation order (sequential dots)
    Ar = np.array([[sum((A[i, j] * B[j, l] for j in range(4)), 0.0) for l in range(3)] for i in range(10)])
    br = np.array([b[i] - sum((A[i, j] * x0[j] for j in range(4)), 0.0) for i in range(10)])
    ref = oracle.chr_run_f32(Ar.astype(np.float32), br, np.zeros((4, 3), np.float32), 600, 1, seed=21)
    Z = ref["samples"]
    if backmap:
        Tc, sc = T @ B, T @ x0 + shift
        assert V.shape == (2400, 7)
    else:
        Tc, sc = B, x0
        assert V.shape == (2400, 4)
        # every sample stays on the equality and inside the box
        assert np.abs(V @ np.array([1.0, 1.0, 0.0, -1.0]) - 0.2).max() <= 2e-6
        assert np.abs(V).max() <= 1.0 + 1e-6
        assert V[:, 0].std() > 0.2  # the chain moves along the face
    Vref = oracle.backmap(Tc.astype(np.float32).astype(np.float64), sc.astype(np.float32).astype(np.float64), Z)
    assert np.abs(V - Vref).max() <= 1e-5 * max(1.0, np.abs(Vref).max())


def test_validate_long_chains(gpu, tmp_path):
    """`naja validate` on chains longer than the LDS transform (2 chains x
    20000 samples, interleaved rows): per-dim ESS and R-hat counts as the
    oracle's (validate_cli.cpp:140-146 takes any length)."""
    from tests.test_oracle_diag import ar1

    rng = np.random.default_rng(4)
    nch, spc, dim = 2, 20000, 3
    series = ar1(spc, 0.97, rng, nseries=nch * dim).reshape(nch, dim, spc)
    S = np.zeros((spc * nch, dim), np.float32)
    for c in range(nch):
        S[c::nch] = series[c].T
    f = tmp_path / "long.npy"
    np.save(f, S)
    v = subprocess.run([NAJA, "validate", "--samples", str(f), "--n-chains", str(nch)], capture_output=True, text=True,
                       timeout=120)
    assert v.returncode == 0, v.stderr
    js = json.loads(v.stdout)
    e_ref, _ = oracle.ess(S, nch)
    _, a11, _ = oracle.split_rhat(S, nch)
    np.testing.assert_allclose(js["per_dim_ess"], np.round(e_ref, 1), rtol=2e-3, atol=0.6)
    assert js["rhat"]["above_1_1"] == a11
