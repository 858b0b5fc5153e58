"""`naja` host binary: flags, model-directory contract, dry run and file
formats (CPU; mirrors tests/test_sample_cli_dry_run.sh and the reference's
contract checks)."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NAJA = os.path.join(ROOT, "naja_amd", "bin", "naja")


def write_csv(path, arr):
    arr = np.atleast_2d(np.asarray(arr, dtype=np.float64))
    with open(path, "w") as f:
        for row in arr:
            f.write(",".join(repr(float(v)) for v in row) + "\n")


def make_model(tmp_path, name="TOY", d=3, extra=False, backmap=True):
    md = tmp_path / "models" / name
    rd = md / "rounding"
    rd.mkdir(parents=True)
    A = np.vstack([np.eye(d), -np.eye(d)])
    write_csv(rd / f"{name}_rounding_A.csv", A)
    write_csv(rd / f"{name}_rounding_b.csv", np.ones((2 * d, 1)))
    write_csv(rd / f"{name}_rounding_start.csv", np.zeros((d, 1)))
    if backmap:
        write_csv(rd / f"{name}_rounding_T.csv", np.eye(d))
        write_csv(rd / f"{name}_rounding_shift.csv", np.zeros((d, 1)))
    if extra:
        write_csv(rd / f"{name}_rounding_extra_A.csv", np.ones((1, d)))
        write_csv(rd / f"{name}_rounding_extra_b.csv", [[0.5]])
    return md


def run(args, **kw):
    return subprocess.run([NAJA] + args, capture_output=True, text=True, **kw)


def test_dry_run_phase_lines(tmp_path):
    """test_sample_cli_dry_run.sh:1-35: the dry run prints the contract phases and stops."""
    md = make_model(tmp_path)
    out = tmp_path / "out"
    r = run(["sample", "run", "--model-dir", str(md), "--out-root", str(out), "--gpu", "0", "--n-chains", "2",
             "--n-samples", "10", "--dry-run"])
    assert r.returncode == 0, r.stderr
    for line in ("> validate model contract", "> write generated config", "model :: TOY",
                 "> dry-run: not executing sampling"):
        assert line in r.stdout
    runs = os.listdir(out)
    assert len(runs) == 1 and runs[0].startswith("TOY_") and runs[0].endswith("_001")
    rd = out / runs[0]
    cfg = (rd / "config_generated.txt").read_text()
    assert "N_CHAINS=2" in cfg and "N_SAMPLES=10" in cfg and "BOUNDS_POLICY=ignore" in cfg
    man = json.loads((rd / "run_manifest.json").read_text())
    assert man["argv"][:3] == ["naja", "sample", "run"] and man["model"] == "TOY"
    # a second run allocates _002
    run(["sample", "run", "--model-dir", str(md), "--out-root", str(out), "--n-chains", "2", "--n-samples", "10",
         "--dry-run", "--quiet"])
    assert sorted(os.listdir(out))[1].endswith("_002")


def test_print_config(tmp_path):
    md = make_model(tmp_path)
    r = run(["sample", "run", "--model-dir", str(md), "--out-root", str(tmp_path / "o"), "--n-chains", "2",
             "--n-samples", "5", "--pair-prob", "0.25", "--thinning", "7", "--dry-run", "--print-config"])
    assert r.returncode == 0
    assert "PAIR_PROB=0.25" in r.stdout and "THINNING=7" in r.stdout and "---" in r.stdout


@pytest.mark.parametrize("args,msg", [
    (["--n-chains", "0"], "invalid --n-chains"),
    (["--n-samples", "-1"], "invalid --n-samples"),
    (["--pair-prob", "1.5"], "invalid --pair-prob"),
    (["--ksparse-k", "0"], "invalid --ksparse-k"),
    (["--start-policy", "lp"], "invalid --start-policy"),
    (["--extra-constraints", "maybe"], "invalid --extra-constraints"),
    (["--bounds-policy", "drop"], "invalid --bounds-policy"),
    (["--bogus"], "unknown flag: --bogus"),
    (["--thinning"], "missing value for --thinning"),
])
def test_invalid_flags_exit_2(tmp_path, args, msg):
    """cmd_run.cpp:104-133 + common.cpp:15-27: die_usage prints and exits 2."""
    md = make_model(tmp_path)
    base = ["sample", "run", "--model-dir", str(md), "--out-root", str(tmp_path / "o"), "--n-chains", "2",
            "--n-samples", "5"]
    r = run(base + args)
    assert r.returncode == 2
    assert msg in r.stderr and "usage:" in r.stderr


def test_contract_errors(tmp_path):
    md = make_model(tmp_path, backmap=False)
    base = ["sample", "run", "--model-dir", str(md), "--out-root", str(tmp_path / "o"), "--n-chains", "2",
            "--n-samples", "5", "--dry-run"]
    r = run(base + ["--backmap"])
    assert r.returncode == 1 and "missing rounding_T" in r.stderr
    (md / "rounding" / "TOY_rounding_extra_A.csv").write_text("1,1,1\n")
    r = run(base)
    assert r.returncode == 1 and "both-present or both-absent" in r.stderr
    (md / "rounding" / "TOY_rounding_extra_b.csv").write_text("")
    r = run(base)
    assert r.returncode == 1 and "empty extra_b" in r.stderr
    r = run(["sample", "run", "--model-dir", str(tmp_path / "nope"), "--out-root", str(tmp_path / "o"),
             "--n-chains", "2", "--n-samples", "5"])
    assert r.returncode == 1 and "not a directory" in r.stderr


def test_filter_requires_backmap(tmp_path):
    md = make_model(tmp_path)
    r = run(["sample", "run", "--model-dir", str(md), "--out-root", str(tmp_path / "o"), "--n-chains", "2",
             "--n-samples", "5", "--bounds-policy", "filter"])
    assert r.returncode == 1 and "requires --backmap" in r.stderr


def test_bad_pair_schedule_rejected_early(tmp_path):
    md = make_model(tmp_path)
    sched = tmp_path / "s.csv"
    sched.write_text("0,1,1.0\n")
    r = run(["sample", "run", "--model-dir", str(md), "--out-root", str(tmp_path / "o"), "--n-chains", "2",
             "--n-samples", "5", "--pair-schedule", str(sched), "--dry-run"])
    assert r.returncode == 1 and "expected 4 columns" in r.stderr


def npy_header_rule(descr, shape):
    """npy.h:27-40: dict, pad with spaces so 10 + len is a multiple of 64, then '\\n'."""
    h = "{'descr': '%s', 'fortran_order': False, 'shape': (%s), }" % (descr, shape)
    total = 10 + len(h)
    h += " " * ((64 - total % 64) % 64) + "\n"
    return b"\x93NUMPY\x01\x00" + len(h).to_bytes(2, "little") + h.encode()


def test_validate_reads_f4_and_f8(tmp_path):
    """validate_cli.cpp:79-136 loader: f4/f8, smaller axis = dim; the header
    written by the reference rule (npy.h) parses with numpy."""
    rng = np.random.default_rng(0)
    S = rng.standard_normal((400, 3)).astype(np.float32)
    path = tmp_path / "s.npy"
    path.write_bytes(npy_header_rule("<f4", "400, 3") + S.tobytes())
    np.testing.assert_array_equal(np.load(path), S)
    path8 = tmp_path / "s8.npy"
    path8.write_bytes(npy_header_rule("<f8", "3, 400") + S.T.astype(np.float64).tobytes())
    np.testing.assert_array_equal(np.load(path8), S.T)


def test_validate_usage_errors(tmp_path):
    r = run(["validate"])
    assert r.returncode == 2 and "missing --samples" in r.stderr
    r = run(["validate", "--samples", "x.npy", "--n-chains", "0"])
    assert r.returncode == 2
    r = run(["validate", "--samples", "x.npy", "--chain-layout", "zigzag"])
    assert r.returncode == 2
