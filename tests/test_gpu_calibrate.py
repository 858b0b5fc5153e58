"""`naja sample calibrate-rounding` (cmd_calibrate_rounding.cpp:24-195) on the
GPU: the Jacobi pair schedule equals the one the oracle's warmup chains give
(same pairs, same Philox streams, model-averaged moments, jacobi.cpp:8-21
angles), for conditioned models sharing a base (one batched launch per pass)
and for unrelated models (one launch per model)."""
import subprocess

import numpy as np
import pytest

import oracle
from tests.test_cli import NAJA, write_csv

pytestmark = pytest.mark.gpu


def write_model(root, name, A, b, x0, E=None, f=None):
    rd = root / name / "rounding"
    rd.mkdir(parents=True)
    write_csv(rd / f"{name}_rounding_A.csv", A)
    write_csv(rd / f"{name}_rounding_b.csv", b[:, None])
    write_csv(rd / f"{name}_rounding_start.csv", x0[:, None])
    if E is not None and len(f):
        write_csv(rd / f"{name}_rounding_extra_A.csv", E)
        write_csv(rd / f"{name}_rounding_extra_b.csv", f[:, None])


def oracle_schedule(models, pairs, warmup, passes, seed, pair_prob, shared):
    pi, pj = pairs
    c = np.ones(len(pi))
    s = np.zeros(len(pi))
    for p in range(passes):
        acc = np.zeros((3, len(pi)))
        for mi, (A, b, x0) in enumerate(models):
            sd = (seed ^ (p + 1)) if shared else (seed ^ mi ^ (p + 1))
            W = oracle.chr_run_f32(A.astype(np.float32), b, x0[None].astype(np.float32), warmup, 1, seed=sd,
                                   variant=mi if shared else 0, pair_prob=pair_prob, pair_mode=2,
                                   pair=(pi, pj, c.astype(np.float32), s.astype(np.float32)))["samples"]
            W = W.astype(np.float64)
            n = W.shape[0]
            for k in range(len(pi)):  # sequential sums, as the host code accumulates them
                di = W[:, pi[k]] - np.cumsum(W[:, pi[k]])[-1] / n
                dj = W[:, pj[k]] - np.cumsum(W[:, pj[k]])[-1] / n
                acc[:, k] += [np.cumsum(di * di)[-1] / n, np.cumsum(dj * dj)[-1] / n, np.cumsum(di * dj)[-1] / n]
        acc /= len(models)
        th = np.where(acc[2] == 0, 0.0, 0.5 * np.arctan2(2 * acc[2], acc[0] - acc[1]))
        c, s = np.cos(th), np.sin(th)
    return c, s


@pytest.mark.parametrize("shared", [True, False])
def test_calibrate_rounding_matches_oracle(gpu, tmp_path, shared):
    rng = np.random.default_rng(8)
    d, nm = 7, 3
    L = np.tril(rng.standard_normal((d, d))) + 2 * np.eye(d)  # correlated coordinates
    Ainv = np.linalg.inv(L)
    A = np.vstack([Ainv, -Ainv])
    root = tmp_path / "models"
    models = []
    for mi in range(nm):
        b = np.ones(2 * d) if shared else 10 ** rng.uniform(-0.3, 0.3, 2 * d)
        k = 2 + mi
        E = rng.standard_normal((k, d))
        f = 10 ** rng.uniform(-0.5, 0.5, k) + 0.5
        x0 = np.zeros(d)
        write_model(root, f"M{mi}", A, b, x0, E, f)
        models.append((np.vstack([A, E]), np.r_[b, f], x0))
    lst = tmp_path / "list.txt"
    lst.write_text("".join(f"M{mi}\n" for mi in range(nm)))
    out = tmp_path / "sched.csv"
    r = subprocess.run([NAJA, "sample", "calibrate-rounding", "--models-root", str(root), "--model-list", str(lst),
                        "--out", str(out), "--warmup", "1500", "--passes", "2", "--pair-prob", "0.4", "--seed", "5"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert f"batched          :: {'yes' if shared else 'no'}" in r.stdout
    rows = np.loadtxt(out, delimiter=",", ndmin=2)
    assert rows.shape == (d // 2, 4)
    pi, pj = rows[:, 0].astype(int), rows[:, 1].astype(int)
    assert len(set(pi) | set(pj)) == 2 * (d // 2)  # disjoint pairs
    c_ref, s_ref = oracle_schedule(models, (pi, pj), 1500, 2, 5, 0.4, shared)
    print("c", rows[:, 2], c_ref, "\ns", rows[:, 3], s_ref, flush=True)
    # the file holds 6 significant digits (schedule_io.cpp:85 default stream precision)
    np.testing.assert_allclose(rows[:, 2], c_ref, rtol=0, atol=2e-6)
    np.testing.assert_allclose(rows[:, 3], s_ref, rtol=0, atol=2e-6)
    # the schedule loads back into sample run's --pair-schedule parser
    assert np.allclose(rows[:, 2] ** 2 + rows[:, 3] ** 2, 1.0, atol=1e-5)
