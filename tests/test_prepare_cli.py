"""`naja prepare` (cmd_prepare.cpp:213-457) on the CPU: conditioner runs and
signatures, extra rows from tightened bounds (vs the oracle), empty-extra
normalisation, symlink/copy inheritance, and starts made feasible by the
cube-centre LP (feasible_start_files.cpp:41-90)."""
import os
import sys

import numpy as np
import pytest

import oracle.lp as olp
from tests.test_cli import NAJA, run, write_csv

COND = """
import os, sys, shutil
import numpy as np
a = sys.argv[1:]
base = a[a.index('--base-model-dir') + 1]; out = a[a.index('--out-model-dir') + 1]; row = a[a.index('--row-id') + 1]
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'calls.log'), 'a') as f:
    f.write(row + '\\n')
os.makedirs(out + '/gem', exist_ok=True)
shutil.copy(base + '/gem/reaction_ids.txt', out + '/gem/reaction_ids.txt')
lb = np.loadtxt(base + '/gem/l_bounds.csv', delimiter=',', ndmin=1)
ub = np.loadtxt(base + '/gem/u_bounds.csv', delimiter=',', ndmin=1)
spec = dict(x.split('=') for x in a if '=' in x)
if row == 'TIGHT':
    ub[0] = ub[0] - float(spec.get('frac', '0.5')) * (ub[0] - lb[0])
elif row == 'SHIFTED':
    lb[2] = lb[2] + 0.7 * (ub[2] - lb[2])
np.savetxt(out + '/gem/l_bounds.csv', lb[:, None], delimiter=',', fmt='%.17g')
np.savetxt(out + '/gem/u_bounds.csv', ub[:, None], delimiter=',', fmt='%.17g')
"""


def make_base(root, rng, d=4, n=6):
    md = root / "BASE"
    (md / "rounding").mkdir(parents=True)
    (md / "gem").mkdir()
    T = rng.standard_normal((n, d))
    shift = rng.standard_normal(n)
    m = 10 ** rng.uniform(-0.5, 0.5, n)
    pre = md / "rounding" / "BASE_rounding_"
    write_csv(str(pre) + "A.csv", np.vstack([T, -T]))
    write_csv(str(pre) + "b.csv", np.r_[m, m][:, None])
    write_csv(str(pre) + "start.csv", np.zeros((d, 1)))
    write_csv(str(pre) + "T.csv", T)
    write_csv(str(pre) + "shift.csv", shift[:, None])
    write_csv(md / "gem" / "l_bounds.csv", (shift - m)[:, None])
    write_csv(md / "gem" / "u_bounds.csv", (shift + m)[:, None])
    (md / "gem" / "reaction_ids.txt").write_text("".join(f"R{i}\n" for i in range(n)))
    return md, T, shift, m


def read_csv(p):
    return np.loadtxt(p, delimiter=",", ndmin=2)


@pytest.mark.parametrize("mode", ["symlink", "copy"])
def test_prepare_conditioned_models(tmp_path, mode):
    rng = np.random.default_rng(3)
    root = tmp_path / "models"
    base, T, shift, m = make_base(root, rng)
    for name in ("TIGHT", "SHIFTED", "SAME"):
        (root / name).mkdir()
    lst = tmp_path / "list.txt"
    lst.write_text("# conditions\nTIGHT\n\nSHIFTED\nSAME\n")
    cond = tmp_path / "cond.py"
    cond.write_text(COND)
    out_list = tmp_path / "prepared.txt"
    args = ["prepare", "--models-root", str(root), "--model-list", str(lst), "--base-model-dir", str(base),
            "--mode", mode, "--conditioner-cmd", f"{sys.executable} {cond}", "--conditioner-arg", "frac=0.5",
            "--out-model-list", str(out_list)]
    r = run(args, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "prepared        :: 3 / 3" in r.stdout
    assert out_list.read_text().split() == ["TIGHT", "SHIFTED", "SAME"]
    lb, ub = shift - m, shift + m
    for name in ("TIGHT", "SHIFTED", "SAME"):
        md = root / name
        pre = md / "rounding" / f"{name}_rounding_"
        assert (md / "rounding" / "INHERITED_FROM.txt").read_text().count(f"mode={mode}") == 1
        assert (md / "gem" / "conditioning.signature").read_text().startswith(f"{sys.executable} {cond}\nARG frac=0.5\nROW_ID {name}\n")
        lbn = np.loadtxt(md / "gem" / "l_bounds.csv", ndmin=1)
        ubn = np.loadtxt(md / "gem" / "u_bounds.csv", ndmin=1)
        EA, Eb = olp.extra_from_bounds(T, shift, lb, ub, lbn, ubn)
        if name == "SAME":
            assert not os.path.exists(str(pre) + "extra_A.csv") and not os.path.exists(str(pre) + "extra_b.csv")
            continue
        np.testing.assert_allclose(read_csv(str(pre) + "extra_A.csv"), EA, rtol=1e-11)
        np.testing.assert_allclose(read_csv(str(pre) + "extra_b.csv").ravel(), Eb, rtol=1e-11, atol=1e-12)
        A = np.vstack([read_csv(str(pre) + "A.csv"), read_csv(str(pre) + "extra_A.csv")])
        b = np.r_[read_csv(str(pre) + "b.csv").ravel(), read_csv(str(pre) + "extra_b.csv").ravel()]
        x = read_csv(str(pre) + "start.csv").ravel()
        assert (A @ x <= b + 1e-9).all()
        recentred = name == "SHIFTED"  # lb of flux 2 moved above the base start's flux
        assert os.path.islink(str(pre) + "start.csv") == (mode == "symlink" and not recentred)
        if recentred:
            _, r_ref = olp.cube_center(A, b)
            w = np.abs(A).sum(1)
            assert (b - A @ x).min() >= 0 and ((b - A @ x) / w).min() >= r_ref * (1 - 1e-7)
    # base files untouched by the recentring (the symlink was replaced, not written through)
    np.testing.assert_array_equal(read_csv(base / "rounding" / "BASE_rounding_start.csv").ravel(), np.zeros(4))
    # a second run finds matching signatures and does not call the conditioner again
    calls = (tmp_path / "calls.log").read_text().split()
    assert sorted(calls) == ["SAME", "SHIFTED", "TIGHT"]
    r = run(args, timeout=120)
    assert r.returncode == 0
    assert (tmp_path / "calls.log").read_text().split() == calls
    # changed arguments -> signature mismatch unless forced
    r = run(args[:-4] + ["--conditioner-arg", "frac=0.25", "--out-model-list", str(out_list)], timeout=120)
    assert r.returncode == 1 and "conditioning signature mismatch" in r.stderr
    r = run(args[:-4] + ["--conditioner-arg", "frac=0.25", "--force-conditioning"], timeout=120)
    assert r.returncode == 0
    EA, Eb = olp.extra_from_bounds(T, shift, lb, ub, np.loadtxt(root / "TIGHT" / "gem" / "l_bounds.csv"),
                                   np.loadtxt(root / "TIGHT" / "gem" / "u_bounds.csv"))
    np.testing.assert_allclose(read_csv(root / "TIGHT" / "rounding" / "TIGHT_rounding_extra_b.csv").ravel(), Eb,
                               rtol=1e-11)
    assert abs(Eb[0] - (ub[0] - 0.25 * 2 * m[0] - shift[0])) <= 1e-9


def test_prepare_errors(tmp_path):
    rng = np.random.default_rng(4)
    root = tmp_path / "models"
    base, T, shift, m = make_base(root, rng)
    (root / "X").mkdir()
    lst = tmp_path / "list.txt"
    lst.write_text("X\n")
    r = run(["prepare", "--models-root", str(root), "--model-list", str(lst), "--conditioner-cmd", "true"])
    assert r.returncode == 2 and "--conditioner-cmd requires --base-model-dir" in r.stderr
    r = run(["prepare", "--models-root", str(root), "--model-list", str(lst)])
    assert r.returncode == 1 and "gem/ missing and no conditioner provided" in r.stderr
    # a conditioner that leaves the reaction ids different from the base
    bad = tmp_path / "bad.py"
    bad.write_text("import sys, os\na = sys.argv\nout = a[a.index('--out-model-dir') + 1]\n"
                   "os.makedirs(out + '/gem', exist_ok=True)\nopen(out + '/gem/reaction_ids.txt', 'w').write('Z\\n')\n"
                   "open(out + '/gem/l_bounds.csv', 'w').write('0\\n')\nopen(out + '/gem/u_bounds.csv', 'w').write('1\\n')\n")
    r = run(["prepare", "--models-root", str(root), "--model-list", str(lst), "--base-model-dir", str(base),
             "--conditioner-cmd", f"{sys.executable} {bad}"])
    assert r.returncode == 1 and "reaction_ids mismatch vs base" in r.stderr
    r = run(["prepare", "--models-root", str(root), "--model-list", str(lst), "--out-bulk-config", "x"])
    assert r.returncode == 2 and "bulk job runner" in r.stderr
    r = run(["prepare", "--models-root", str(root), "--model-list", str(lst), "--base-model-dir", str(base),
             "--mode", "hardlink"])
    assert r.returncode == 2 and "invalid --mode" in r.stderr


def test_prepare_dry_run_touches_nothing(tmp_path):
    """--dry-run reports every listed model as prepared and writes nothing:
    no conditioner call, no rounding files, no model list."""
    rng = np.random.default_rng(5)
    root = tmp_path / "models"
    base, *_ = make_base(root, rng)
    (root / "TIGHT").mkdir()
    lst = tmp_path / "list.txt"
    lst.write_text("TIGHT\n")
    cond = tmp_path / "cond.py"
    cond.write_text(COND)
    out_list = tmp_path / "prepared.txt"
    r = run(["prepare", "--models-root", str(root), "--model-list", str(lst), "--base-model-dir", str(base),
             "--conditioner-cmd", f"{sys.executable} {cond}", "--out-model-list", str(out_list), "--dry-run"])
    assert r.returncode == 0 and "prepared        :: 1 / 1" in r.stdout and "dry_run         :: true" in r.stdout
    assert not (tmp_path / "calls.log").exists()
    assert not out_list.exists()
    assert sorted(os.listdir(root / "TIGHT")) == []
