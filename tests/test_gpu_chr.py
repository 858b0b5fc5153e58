"""GPU CHR kernel vs the FP32 oracle on the same Philox stream (through the C ABI)."""
import numpy as np
import pytest

import oracle
from tests.polytopes import cube, genome_like, random_polytope, simplex, thin_slab

pytestmark = pytest.mark.gpu


def run_gpu(torch, A, b, y0, nspc, thin, tpb=-1, **kw):
    import naja_amd

    out = naja_amd.coordinate_hit_and_run(torch.from_numpy(np.asarray(A, np.float32)),
                                          torch.from_numpy(np.asarray(b, np.float64)),
                                          torch.from_numpy(np.asarray(y0, np.float32)),
                                          nspc, thin, y0.shape[0], tpb_ss=tpb, return_final=True, **kw)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out]


def report(tag, gpu, ref):
    diff = np.abs(gpu.astype(np.float64) - ref.astype(np.float64))
    rel = diff / np.maximum(np.abs(ref.astype(np.float64)), 1e-3)
    exact = float((gpu == ref).mean())
    print(f"[{tag}] exact={exact:.6f} max_abs={diff.max():.3e} max_rel={rel.max():.3e}", flush=True)
    return rel.max(), exact


MOVE_CASES = {
    "coord": dict(),
    "pair1": dict(pair_prob=0.3),
    "ksparse": dict(pair_prob=0.2, ksparse_prob=0.1, ksparse_k=5),
}


@pytest.mark.parametrize("case", list(MOVE_CASES))
@pytest.mark.parametrize("tpb", [64, 256, 512, 1024])
def test_move_log_bit_exact(gpu, case, tpb):
    A, b, y0 = random_polytope(9, 40, seed=3)
    n = 4096
    kw = MOVE_CASES[case]
    _, _, mlog = run_gpu(gpu, A, b, np.zeros((2, 9)), n, 1, tpb=tpb, move_log_steps=n, **kw)
    ref = oracle.chr_run_f32(A.astype(np.float32), b, np.zeros((2, 9), np.float32), n, 1, nlog=n, **kw)
    np.testing.assert_array_equal(mlog, ref["move_log"])


def test_move_log_pair_mode2(gpu):
    import naja_amd  # noqa: F401

    d = 8
    A, b, y0 = random_polytope(d, 30, seed=4)
    pi = np.array([0, 2, 4, 6], np.int32)
    pj = np.array([1, 3, 5, 7], np.int32)
    th = np.array([0.3, 0.9, 1.7, 2.5])
    pc, ps = np.cos(th).astype(np.float32), np.sin(th).astype(np.float32)
    n = 3000
    t = gpu
    samples, fin, mlog = run_gpu(t, A, b, np.zeros((2, d)), n, 1, move_log_steps=n, pair_prob=0.5, pair_mode=2,
                                 pair_i=t.from_numpy(pi), pair_j=t.from_numpy(pj), pair_c=t.from_numpy(pc),
                                 pair_s=t.from_numpy(ps))
    ref = oracle.chr_run_f32(A.astype(np.float32), b, np.zeros((2, d), np.float32), n, 1, nlog=n, pair_prob=0.5,
                             pair_mode=2, pair=(pi, pj, pc, ps))
    np.testing.assert_array_equal(mlog, ref["move_log"])
    rel, _ = report("pair-mode-2", samples, ref["samples"])
    assert rel <= 1e-5


TRAJ_CASES = {
    # name: (polytope builder, tpb, extra kwargs)
    "rand-w1": (lambda: random_polytope(10, 30, seed=21), 64, dict(pair_prob=0.3)),
    "rand-w4": (lambda: random_polytope(24, 300, seed=22), 256, dict(pair_prob=0.3)),
    "rand-w16": (lambda: random_polytope(40, 2100, seed=23), 1024, dict(pair_prob=0.3)),
    "rand-auto-resync": (lambda: random_polytope(16, 200, seed=24), -1, dict(pair_prob=0.3, resync_interval=50)),
    "ksparse": (lambda: random_polytope(12, 60, seed=25), -1, dict(ksparse_prob=0.2, ksparse_k=6)),
    "slab": (lambda: thin_slab(10, 1e-4), -1, dict(pair_prob=0.3)),
    "simplex": (lambda: simplex(6), -1, dict()),
}


@pytest.mark.parametrize("case", list(TRAJ_CASES))
def test_trajectory_matches_oracle(gpu, case):
    """Indices are bit-exact by construction; states agree within 1e-5 relative
    over 1e4 steps (the kernel repeats the oracle's FP32 operations, so in
    practice they agree bit for bit)."""
    build, tpb, kw = TRAJ_CASES[case]
    A, b, y0 = build()
    d = A.shape[1]
    nch = 3
    Y0 = np.tile(y0, (nch, 1))
    n = 10000
    samples, fin = run_gpu(gpu, A, b, Y0, n, 1, tpb=tpb, **kw)
    ref = oracle.chr_run_f32(A.astype(np.float32), b, Y0.astype(np.float32), n, 1, **kw)
    rel, exact = report(case, samples, ref["samples"])
    assert rel <= 1e-5
    np.testing.assert_array_equal(fin, ref["y_final"]) if exact == 1.0 else None


@pytest.mark.parametrize("d,m", [(100, 300), (128, 200), (64, 100), (150, 200)])
@pytest.mark.parametrize("moves", ["coord", "pair-ksparse"])
def test_one_wavefront_state_paths_bit_exact(gpu, d, m, moves):
    """One-wavefront launches keep the chain states in VGPRs for d <= 128 and
    in LDS above; both repeat the oracle's operations bit for bit (samples,
    final states and moves), with every move kind."""
    A, b, _ = random_polytope(d, m, seed=d + m)
    kw = {} if moves == "coord" else dict(pair_prob=0.3, ksparse_prob=0.05, ksparse_k=5)
    nch, n, thin = 3, 700, 3
    Y0 = np.zeros((nch, d))
    samples, fin, mlog = run_gpu(gpu, A, b, Y0, n, thin, tpb=64, move_log_steps=n * thin, resync_interval=500, **kw)
    ref = oracle.chr_run_f32(A.astype(np.float32), b, Y0.astype(np.float32), n, thin, nlog=n * thin,
                             resync_interval=500, **kw)
    np.testing.assert_array_equal(mlog, ref["move_log"])
    np.testing.assert_array_equal(samples, ref["samples"])
    np.testing.assert_array_equal(fin, ref["y_final"])


def halves_polytope(d, n, seed):
    """{-b_lo <= T y <= b_hi}, T (n x d) N(0,1): A = [T; -T], every row with
    its negation (the layout of flux bounds and slab conditions)."""
    rng = np.random.default_rng(seed)
    T = rng.standard_normal((n, d))
    A = np.vstack([T, -T])
    b = np.concatenate([10.0 ** rng.uniform(-1, 0.5, n), 10.0 ** rng.uniform(-1, 0.5, n)])
    return A, b


@pytest.mark.parametrize("d,n", [(40, 150), (100, 100), (100, 250), (150, 200), (12, 300)])
@pytest.mark.parametrize("moves", ["coord", "pair-ksparse"])
def test_paired_rows_bit_exact(gpu, d, n, moves):
    """[T; -T] polytopes run in the paired-row layout (one packed row per
    pair, two slacks: the shape holds n rows, not 2n); samples, final states
    and moves equal the oracle's row-by-row run bit for bit, with the states
    in VGPRs (d <= 128) and in LDS."""
    import naja_amd

    t = gpu
    A, b = halves_polytope(d, n, seed=d * 7 + n)
    kw = {} if moves == "coord" else dict(pair_prob=0.3, ksparse_prob=0.05, ksparse_k=5)
    nch, nspc, thin = 3, 600, 3
    Y0 = np.zeros((nch, d))
    samples, fin, mlog = run_gpu(t, A, b, Y0, nspc, thin, move_log_steps=nspc * thin, resync_interval=400, **kw)
    ref = oracle.chr_run_f32(A.astype(np.float32), b, Y0.astype(np.float32), nspc, thin, nlog=nspc * thin,
                             resync_interval=400, **kw)
    np.testing.assert_array_equal(mlog, ref["move_log"])
    np.testing.assert_array_equal(samples, ref["samples"])
    np.testing.assert_array_equal(fin, ref["y_final"])
    plan = naja_amd.ChrPlan(t.from_numpy(A.astype(np.float32)).cuda(), t.from_numpy(b).cuda(), nspc, thin, nch,
                            **{k: v for k, v in kw.items() if k == "pair_prob"})
    assert plan.shape() == (1, 1, (n + 63) // 64)
    assert plan.layout() == (1, 1, 1)  # every row paired: scaled pairs
    plan.close()


@pytest.mark.parametrize("d,n", [(40, 150), (150, 200)])
def test_paired_rows_many_chains_bit_exact(gpu, d, n):
    """256 chains of a [T; -T] polytope: a launch of a workgroup per CU or
    more keeps the unscaled pairs (its padding runs skip memory), a launch of
    few chains takes scaled pairs; samples and final states equal the
    oracle's row-by-row run bit for bit, with every move kind, the states in
    VGPRs (d <= 128) and in LDS."""
    import naja_amd

    t = gpu
    A, b = halves_polytope(d, n, seed=d + 3 * n)
    kw = dict(pair_prob=0.3, ksparse_prob=0.05, ksparse_k=5)
    nch, nspc, thin = 256, 150, 2
    Y0 = np.zeros((nch, d))
    samples, fin = run_gpu(t, A, b, Y0, nspc, thin, resync_interval=100, **kw)
    ref = oracle.chr_run_f32(A.astype(np.float32), b, Y0.astype(np.float32), nspc, thin, resync_interval=100,
                             nthreads=8, **kw)
    np.testing.assert_array_equal(samples, ref["samples"])
    np.testing.assert_array_equal(fin, ref["y_final"])
    At, bt = t.from_numpy(A.astype(np.float32)).cuda(), t.from_numpy(b).cuda()
    for chains, lay in ((nch, (1, 1, 0)), (3, (1, 1, 1))):
        plan = naja_amd.ChrPlan(At, bt, nspc, thin, chains, pair_prob=0.3)
        assert plan.layout() == lay
        plan.close()


@pytest.mark.parametrize("tpb", [256, 512])
@pytest.mark.parametrize("moves", ["coord", "pair-ksparse"])
def test_paired_rows_multi_wave_scaled_bit_exact(gpu, tpb, moves):
    """The four- and eight-wavefront paired kernels on scaled pairs (every
    row of [T; -T] paired, a few chains): samples, final states and moves
    equal the oracle's row-by-row run bit for bit."""
    import naja_amd

    t = gpu
    d, n = 150, 700
    A, b = halves_polytope(d, n, seed=tpb + len(moves))
    kw = {} if moves == "coord" else dict(pair_prob=0.3, ksparse_prob=0.05, ksparse_k=5)
    nch, nspc, thin = 3, 300, 3
    Y0 = np.zeros((nch, d))
    samples, fin, mlog = run_gpu(t, A, b, Y0, nspc, thin, tpb=tpb, move_log_steps=nspc * thin, resync_interval=400,
                                 **kw)
    ref = oracle.chr_run_f32(A.astype(np.float32), b, Y0.astype(np.float32), nspc, thin, nlog=nspc * thin,
                             resync_interval=400, **kw)
    np.testing.assert_array_equal(mlog, ref["move_log"])
    np.testing.assert_array_equal(samples, ref["samples"])
    np.testing.assert_array_equal(fin, ref["y_final"])
    plan = naja_amd.ChrPlan(t.from_numpy(A.astype(np.float32)).cuda(), t.from_numpy(b).cuda(), nspc, thin, nch,
                            tpb_ss=tpb, **{k: v for k, v in kw.items() if k == "pair_prob"})
    assert plan.shape()[0] == tpb // 64 and plan.layout() == (1, 1, 1)
    plan.close()


def test_paired_batched_mixed_blocks_bit_exact(gpu):
    """A paired base ([I; -I]) with extra blocks of every kind: empty, [a; -a]
    (paired), an odd row count and an even one that is not a negation (rows
    packed alone, partner slack NaN). Every variant equals the oracle bit for
    bit."""
    import naja_amd

    t = gpu
    d = 16
    A, b, _ = cube(d)
    rng = np.random.default_rng(17)

    def slabs(k):
        a = rng.standard_normal((k, d))
        a /= np.linalg.norm(a, axis=1, keepdims=True)
        return np.vstack([a, -a]), np.concatenate([np.full(k, 0.3), np.full(k, 0.2)])

    blocks = [(np.zeros((0, d)), np.zeros(0)), slabs(10), (rng.standard_normal((7, d)), np.full(7, 0.8)),
              (rng.standard_normal((10, d)), np.full(10, 0.9)), slabs(20)]
    EA = [e for e, _ in blocks]
    Eb = [f for _, f in blocks]
    nch, nspc, thin = 3, 500, 2
    X0 = np.zeros((len(blocks), nch, d), np.float32)
    # An odd k-sparse count keeps every k-sparse direction nonzero (with an
    # even count the signed entries can cancel pairwise on the cube).
    kw = dict(pair_prob=0.3, ksparse_prob=0.05, ksparse_k=5, resync_interval=300)
    out = naja_amd.coordinate_hit_and_run_batched(
        t.from_numpy(A.astype(np.float32)), t.from_numpy(b), [t.from_numpy(e.astype(np.float32)) for e in EA],
        [t.from_numpy(f) for f in Eb], t.from_numpy(X0), nspc, thin, nch, seed=77, **kw)
    t.cuda.synchronize()
    out = out.cpu().numpy()
    for v in range(len(blocks)):
        Av = np.vstack([A, EA[v]]).astype(np.float32)
        bv = np.concatenate([b, Eb[v]])
        ref = oracle.chr_run_f32(Av, bv, X0[v], nspc, thin, seed=77, variant=v, **kw)
        np.testing.assert_array_equal(out[v], ref["samples"])


@pytest.mark.parametrize("layout", ["paired", "unpaired", "paired-d128", "paired-all", "paired-all-wide"])
def test_lds_column_chains_bit_exact(gpu, layout):
    """Coordinate-only batches with d <= 128 and a multiple of four chains run
    four chains of a variant per workgroup with the variant's A in LDS; every
    variant equals the oracle bit for bit, in both row layouts (the unpaired
    case holds up to 374 rows, near the LDS budget). With every block paired
    ("paired-all") the packs are in the scaled-pair form (each pair divided
    by a power of two above twice its width, slacks clamped by the update's
    own FMA clamp): the same bits again, for pair widths from 1e-3 to 1e3
    ("-wide")."""
    import naja_amd

    t = gpu
    rng = np.random.default_rng(5)
    if layout == "unpaired":
        d = 100
        A, b, _ = random_polytope(d, 150, seed=61)
    else:
        d = 60 if layout in ("paired", "paired-all", "paired-all-wide") else 128
        A, b = halves_polytope(d, 80 if d == 60 else 140, seed=62)

    def slabs(k, wide=False):
        a = rng.standard_normal((k, d))
        a /= np.linalg.norm(a, axis=1, keepdims=True)
        if wide:  # half-widths 1e-3 .. 1e3 around the start
            w = 10.0 ** rng.uniform(-3, 3, k)
            return np.vstack([a * 1.7, -a * 1.7]), np.concatenate([w, 0.5 * w])
        return np.vstack([a, -a]), np.concatenate([np.full(k, 0.4), np.full(k, 0.3)])

    if layout.startswith("paired-all"):
        wide = layout.endswith("wide")
        blocks = [(np.zeros((0, d)), np.zeros(0)), slabs(12, wide), slabs(3, wide), slabs(30, wide)]
    else:
        # unpaired: at most 350 + 24 rows (6 per lane, 150 KB of columns)
        blocks = [(np.zeros((0, d)), np.zeros(0)), slabs(12), (rng.standard_normal((5, d)), np.full(5, 0.7)),
                  slabs(8 if layout == "unpaired" else 30)]
    EA = [e for e, _ in blocks]
    Eb = [f for _, f in blocks]
    nch, nspc, thin = 8, 400, 3
    X0 = np.zeros((len(blocks), nch, d), np.float32)
    out = naja_amd.coordinate_hit_and_run_batched(
        t.from_numpy(A.astype(np.float32)), t.from_numpy(b), [t.from_numpy(e.astype(np.float32)) for e in EA],
        [t.from_numpy(f) for f in Eb], t.from_numpy(X0), nspc, thin, nch, seed=91, resync_interval=350)
    t.cuda.synchronize()
    out = out.cpu().numpy()
    for v in range(len(blocks)):
        Av = np.vstack([A, EA[v]]).astype(np.float32)
        bv = np.concatenate([b, Eb[v]])
        ref = oracle.chr_run_f32(Av, bv, X0[v], nspc, thin, seed=91, variant=v, resync_interval=350)
        np.testing.assert_array_equal(out[v], ref["samples"])
    # the layout the same batch runs in: paired or not, four chains per
    # workgroup, scaled pairs exactly when every block is paired
    plan = naja_amd.ChrPlan(naja_amd.VariantBatch(A, b, EA, Eb), None, nspc, thin, nch, seed=91,
                            resync_interval=350)
    assert plan.layout() == ((0, 4, 0) if layout == "unpaired" else (1, 4, int(layout.startswith("paired-all"))))
    plan.close()


def test_chains_per_group_matches_oracle(gpu):
    A, b, y0 = random_polytope(12, 70, seed=31)
    nch = 6  # one full group of 4 and a partial group of 2
    Y0 = np.zeros((nch, 12))
    samples, _ = run_gpu(gpu, A, b, Y0, 3000, 2, pair_prob=0.3, chains_per_group=4)
    ref = oracle.chr_run_f32(A.astype(np.float32), b, Y0.astype(np.float32), 3000, 2, pair_prob=0.3,
                             chains_per_group=4)
    rel, _ = report("G4", samples, ref["samples"])
    assert rel <= 1e-5


def test_batched_variants_match_oracle(gpu):
    import naja_amd

    t = gpu
    d = 10
    A, b, _ = cube(d)
    rng = np.random.default_rng(8)
    ks = [0, 3, 17, 40, 64, 1]
    EA, Eb = [], []
    for k in ks:
        rows = rng.standard_normal((k, d))
        rows /= np.maximum(np.linalg.norm(rows, axis=1, keepdims=True), 1e-12)
        EA.append(rows)
        Eb.append(10.0 ** rng.uniform(-2, 0, k))
    nch, nspc, thin = 4, 800, 3
    X0 = np.zeros((len(ks), nch, d), np.float32)
    out = naja_amd.coordinate_hit_and_run_batched(
        t.from_numpy(A.astype(np.float32)), t.from_numpy(b), [t.from_numpy(e.astype(np.float32)) for e in EA],
        [t.from_numpy(f) for f in Eb], t.from_numpy(X0), nspc, thin, nch, seed=1234, pair_prob=0.3,
        resync_interval=100)
    t.cuda.synchronize()
    out = out.cpu().numpy()
    for v, k in enumerate(ks):
        Av = np.vstack([A, EA[v]]).astype(np.float32)
        bv = np.concatenate([b, Eb[v]])
        ref = oracle.chr_run_f32(Av, bv, X0[v], nspc, thin, seed=1234, variant=v, pair_prob=0.3,
                                 resync_interval=100)
        rel, _ = report(f"variant{v}-k{k}", out[v], ref["samples"])
        assert rel <= 1e-5


def test_cube_d100_moments(gpu):
    """[-1,1]^100, 4 chains x 12500, thin 16, seed 0: mean 0 and variance 1/3
    per coordinate within 3 Monte Carlo standard errors (rho = 0.99^16)."""
    d, nch, nspc, thin = 100, 4, 12500, 16
    A, b, _ = cube(d)
    samples, _ = run_gpu(gpu, A, b, np.zeros((nch, d)), nspc, thin, seed=0)
    x = samples.astype(np.float64)
    N = x.shape[0]
    rho = (1 - 1 / d) ** thin
    infl = (1 + rho) / (1 - rho)
    zm = x.mean(0) / np.sqrt((1 / 3) / N * infl)
    zv = (x.var(0) - 1 / 3) / np.sqrt((4 / 45) / N * infl)
    print(f"cube d=100: max|z_mean|={np.abs(zm).max():.2f} max|z_var|={np.abs(zv).max():.2f} "
          f"frac(|z|<3)={(np.abs(zm) < 3).mean():.3f},{(np.abs(zv) < 3).mean():.3f}", flush=True)
    # 100 coordinates: expect ~0.3 outside 3 SE by chance; none beyond 4.5 SE
    assert (np.abs(zm) < 3).mean() >= 0.97 and (np.abs(zv) < 3).mean() >= 0.97
    assert np.abs(zm).max() < 4.5 and np.abs(zv).max() < 4.5
    assert np.abs(x).max() <= 1 + 1e-5


@pytest.mark.parametrize("rows", ["paired", "shuffled"])
def test_genome_scale_shape_matches_oracle(gpu, rows):
    """The genome-scale reduced polytope (n = 2583 reactions, d = 512,
    A = [T; -T] with m = 5166 rows, pair-prob 0.3, thinning 85, Philox seed 0)
    through the one-shot call, at the launch shape it picks for 4 chains: the
    paired-row layout (8 wavefronts, 6 packed rows per lane), and with the
    rows shuffled (no pairing) the unpaired one (11 rows per lane; odd R takes
    the one-row-run column loads). Samples, final states and chain 0's move
    log equal the oracle's bit for bit."""
    import naja_amd
    from naja_amd import workloads

    t = gpu
    g = workloads.genome_scale(n=2583, d=512, seed=0)
    if rows == "shuffled":
        perm = np.random.default_rng(1).permutation(g["A"].shape[0])
        g = dict(g, A=g["A"][perm], b=g["b"][perm])
    A32 = g["A"].astype(np.float32)
    m, d = A32.shape
    nch, nspc, thin, pp = 4, 200, 85, 0.3
    At, bt = t.from_numpy(A32).cuda(), t.from_numpy(g["b"]).cuda()
    X0 = t.zeros((nch, d), dtype=t.float32, device="cuda")
    plan = naja_amd.ChrPlan(At, bt, nspc, thin, nch, seed=0, pair_prob=pp)
    shape, lay = plan.shape(), plan.layout()
    plan.close()
    print(f"[genome scale] m={m} d={d} shape (W, G, R) = {shape}, layout {lay}", flush=True)
    assert shape == ((8, 1, 6) if rows == "paired" else (8, 1, 11))
    assert lay == ((1, 1, 1) if rows == "paired" else (0, 1, 0))
    nlog = nspc * thin
    S, fin, mlog = naja_amd.coordinate_hit_and_run(At, bt, X0, nspc, thin, nch, seed=0, pair_prob=pp,
                                                   return_final=True, move_log_steps=nlog)
    t.cuda.synchronize()
    ref = oracle.chr_run_f32(A32, g["b"], np.zeros((nch, d), np.float32), nspc, thin, seed=0, pair_prob=pp,
                             nlog=nlog, nthreads=nch)
    S = S.cpu().numpy()
    report("genome-scale", S, ref["samples"])
    np.testing.assert_array_equal(mlog.cpu().numpy(), ref["move_log"])
    np.testing.assert_array_equal(S, ref["samples"])
    np.testing.assert_array_equal(fin.cpu().numpy(), ref["y_final"])
    viol = (S.astype(np.float64) @ A32.T.astype(np.float64) - g["b"][None, :]).max()
    assert viol <= 1e-5


def test_genome_like_feasible(gpu):
    g = genome_like(300, 64, seed=5)
    A, b = g["A"], g["b"]
    nch = 4
    samples, _ = run_gpu(gpu, A, b, np.zeros((nch, 64)), 2000, 20, pair_prob=0.3)
    x = samples.astype(np.float64)
    viol = (A.astype(np.float32).astype(np.float64) @ x.T - b[:, None]) / np.maximum(1.0, np.abs(b))[:, None]
    assert viol.max() <= 1e-5, viol.max()


def test_invalid_arguments_raise(gpu):
    import naja_amd

    t = gpu
    A, b, _ = cube(4)
    with pytest.raises(naja_amd.NajaError):
        naja_amd.coordinate_hit_and_run(t.from_numpy(A.astype(np.float32)), t.from_numpy(b),
                                        t.zeros((2, 4)), 10, 0, 2)
    with pytest.raises(naja_amd.NajaInvalidArgument):
        naja_amd.coordinate_hit_and_run(t.from_numpy(A.astype(np.float32)), t.from_numpy(b[:-1]),
                                        t.zeros((2, 4)), 10, 1, 2)


def test_rounded_batched_matches_oracle(gpu):
    """Per-variant rounding transform y = c + P z: the packed A' = A P and
    b' = b - A c equal the oracle's element for element, so the z-chains are
    the oracle's; y comes back through the FP32 MFMA backmap."""
    import naja_amd

    t = gpu
    d = 8
    A, b, _ = cube(d)
    rng = np.random.default_rng(21)
    ks = [0, 6, 11]
    EA = [rng.standard_normal((k, d)) for k in ks]
    Eb = [np.ones(k) * 0.5 for k in ks]
    nv, nch, nspc, thin = len(ks), 3, 400, 2
    P = np.stack([np.eye(d) + 0.2 * rng.standard_normal((d, d)) for _ in ks]).astype(np.float32)
    c = (0.05 * rng.standard_normal((nv, d))).astype(np.float32)
    Z0 = np.zeros((nv, nch, d), np.float32)
    Y = naja_amd.coordinate_hit_and_run_batched(
        t.from_numpy(A.astype(np.float32)), t.from_numpy(b), [t.from_numpy(e.astype(np.float32)) for e in EA],
        [t.from_numpy(f) for f in Eb], t.from_numpy(Z0), nspc, thin, nch, seed=5, pair_prob=0.3,
        P=t.from_numpy(P), c=t.from_numpy(c))
    t.cuda.synchronize()
    Y = Y.cpu().numpy().astype(np.float64)
    for v in range(nv):
        Av = np.vstack([A, EA[v]]).astype(np.float32)
        bv = np.concatenate([b, Eb[v]])
        At, bt = oracle.transform(Av, bv, P[v], c[v])
        z = oracle.chr_run_f32(At, bt, Z0[v], nspc, thin, seed=5, variant=v, pair_prob=0.3)["samples"]
        y = c[v].astype(np.float64) + z.astype(np.float64) @ P[v].astype(np.float64).T
        scale = 1.0 + np.abs(z).astype(np.float64) @ np.abs(P[v].astype(np.float64)).T
        err = (np.abs(Y[v] - y) / scale).max()
        print(f"[rounded variant {v}] max scaled err {err:.3e}", flush=True)
        assert err <= 1e-5
        viol = (Av.astype(np.float64) @ Y[v].T - bv[:, None]).max()
        assert viol <= 1e-4


def test_rounded_scaled_pairs_match_oracle(gpu):
    """The rounding transform writes the scaled-pair form when every block is
    paired (coordinate moves, four chains per workgroup): with P = I and c = 0
    the transformed rows are the rows themselves and y = z exactly, so the
    samples equal the oracle's chain on the untransformed polytope bit for
    bit; a second rounding P = 2 I on the same plan equals the oracle on A P,
    b (the pairs' scales come from the right-hand sides, not from P)."""
    import naja_amd

    t = gpu
    d = 40
    A0, b0 = np.vstack([np.eye(d), -np.eye(d)]), np.ones(2 * d)
    rng = np.random.default_rng(31)
    EA, Eb = [], []
    for k in (0, 7, 30, 64):
        a = rng.standard_normal((k, d))
        a /= np.linalg.norm(a, axis=1, keepdims=True)
        w = 10.0 ** rng.uniform(-2, 0, k)
        EA.append(np.vstack([a, -a]) if k else np.zeros((0, d)))
        Eb.append(np.concatenate([w, w]) if k else np.zeros(0))
    batch = naja_amd.VariantBatch(A0, b0, EA, Eb)
    nv, nch, nspc, thin = len(EA), 4, 500, 3
    Z0 = t.zeros((nv, nch, d), dtype=t.float32, device="cuda")
    c = t.zeros((nv, d), dtype=t.float32, device="cuda")
    for scale in (1.0, 2.0):
        P = t.eye(d, dtype=t.float32, device="cuda").repeat(nv, 1, 1) * scale
        plan = naja_amd.ChrPlan(batch, None, nspc, thin, nch, P=P, c=c, seed=13, resync_interval=400)
        Y = plan.launch(Z0, plan.samples_like())
        t.cuda.synchronize()
        Y = Y.cpu().numpy()
        assert plan.layout() == (1, 4, 1)  # paired rows from LDS, scaled pairs
        plan.close()
        for v in range(nv):
            Av = np.vstack([A0, EA[v]]).astype(np.float32)
            bv = np.concatenate([b0, Eb[v]])
            z = oracle.chr_run_f32((Av * np.float32(scale)).astype(np.float32), bv, np.zeros((nch, d), np.float32),
                                   nspc, thin, seed=13, variant=v, resync_interval=400)["samples"]
            np.testing.assert_array_equal(Y[v], z * np.float32(scale))


@pytest.mark.parametrize("d", [12, 10])
def test_rounded_plan_equals_one_shot_call(gpu, d):
    """A rounding-transform plan (z in the plan's own buffer, then the
    backmap) and the one-shot rounded call (z sampled into the output and
    backmapped in place when d % 4 == 0) give identical samples, on repeated
    launches of the plan too."""
    import naja_amd

    t = gpu
    A, b, _ = cube(d)
    rng = np.random.default_rng(d)
    ks = [0, 5, 9, 3]
    EA = [rng.standard_normal((k, d)) for k in ks]
    Eb = [np.ones(k) * 0.4 for k in ks]
    nv, nch, nspc, thin = len(ks), 4, 700, 3
    P = np.stack([np.eye(d) + 0.1 * rng.standard_normal((d, d)) for _ in ks]).astype(np.float32)
    c = (0.02 * rng.standard_normal((nv, d))).astype(np.float32)
    Ad, bd = t.from_numpy(A.astype(np.float32)).cuda(), t.from_numpy(b).cuda()
    EAd = [t.from_numpy(e.astype(np.float32)).cuda() for e in EA]
    Ebd = [t.from_numpy(f).cuda() for f in Eb]
    Pd, cdv = t.from_numpy(P).cuda(), t.from_numpy(c).cuda()
    Z0 = t.zeros((nv, nch, d), device="cuda")
    one = naja_amd.coordinate_hit_and_run_batched(Ad, bd, EAd, Ebd, Z0, nspc, thin, nch, seed=3, pair_prob=0.2,
                                                  P=Pd, c=cdv)
    plan = naja_amd.ChrPlan(Ad, bd, nspc, thin, nch, extra_A=EAd, extra_b=Ebd, P=Pd, c=cdv, seed=3, pair_prob=0.2)
    for _ in range(2):
        out = plan.launch(Z0, plan.samples_like())
        t.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), one.cpu().numpy())
    plan.close()


def test_plan_set_rounding_equals_one_shot_calls(gpu):
    """A rounded plan given a new rounding (only A' = A P, b' = b - A c formed
    again) samples exactly what a one-shot call with that rounding samples,
    for several roundings in a row, in the paired layout and with LDS
    columns (coordinate moves, four chains) and with pair moves."""
    import naja_amd

    t = gpu
    d = 12
    A, b, _ = cube(d)
    rng = np.random.default_rng(77)
    ks = [0, 6, 10, 4]
    EA, Eb = [], []
    for k in ks:
        a = rng.standard_normal((k // 2, d))
        EA.append(np.vstack([a, -a]))
        Eb.append(np.full(k, 0.6))
    nv, nch, nspc, thin = len(ks), 4, 300, 3
    batch = naja_amd.VariantBatch(A, b, EA, Eb)
    Z0 = t.zeros((nv, nch, d), device="cuda")
    c = t.zeros((nv, d), device="cuda")
    for kw in (dict(), dict(pair_prob=0.3)):
        Ps = [t.from_numpy(np.stack([np.eye(d) + s * rng.standard_normal((d, d)) for _ in ks]).astype(np.float32))
              .cuda() for s in (0.1, 0.2, 0.05)]
        plan = naja_amd.ChrPlan(batch, None, nspc, thin, nch, P=Ps[0], c=c, seed=4, **kw)
        for P in Ps:
            plan.set_rounding(P, c)
            out = plan.launch(Z0, plan.samples_like())
            one = naja_amd.coordinate_hit_and_run_batched(batch, None, None, None, Z0, nspc, thin, nch, seed=4,
                                                          P=P, c=c, **kw)
            t.cuda.synchronize()
            np.testing.assert_array_equal(out.cpu().numpy(), one.cpu().numpy())
        plan.close()


def test_plan_autotune_relaunch_and_graph_replay(gpu):
    """A plan launches without allocating: relaunches and hipGraph replays
    (torch.cuda.CUDAGraph) reproduce the eager samples exactly, and the
    autotuned shape gives the same samples as the default one."""
    import naja_amd

    t = gpu
    A, b, y0 = random_polytope(20, 600, seed=41)
    nch, nspc, thin = 4, 300, 3
    At, bt = t.from_numpy(A.astype(np.float32)).cuda(), t.from_numpy(b).cuda()
    X0 = t.zeros((1, nch, 20), device="cuda")
    plan = naja_amd.ChrPlan(At, bt, nspc, thin, nch, seed=9, pair_prob=0.3)
    out0 = plan.launch(X0, plan.samples_like())
    t.cuda.synchronize()
    ref = oracle.chr_run_f32(A.astype(np.float32), b, np.zeros((nch, 20), np.float32), nspc, thin, seed=9,
                             pair_prob=0.3)["samples"]
    np.testing.assert_array_equal(out0.cpu().numpy()[0], ref)
    shape = plan.autotune(X0)
    print(f"autotuned shape (W, G, R) = {shape}", flush=True)
    assert shape == plan.shape()
    out1 = plan.launch(X0, plan.samples_like())
    t.cuda.synchronize()
    np.testing.assert_array_equal(out1.cpu().numpy()[0], ref)

    static_out = plan.samples_like()
    side = t.cuda.Stream()
    side.wait_stream(t.cuda.current_stream())
    with t.cuda.stream(side):
        plan.launch(X0, static_out)  # warm-up on the capture stream
    t.cuda.current_stream().wait_stream(side)
    t.cuda.synchronize()
    g = t.cuda.CUDAGraph()
    with t.cuda.graph(g):
        plan.launch(X0, static_out)
    for rep in range(3):
        static_out.fill_(-1.0)
        _ = (At * 2.0).sum()  # unrelated eager work between replays
        t.cuda.synchronize()
        g.replay()
        t.cuda.synchronize()
        np.testing.assert_array_equal(static_out.cpu().numpy()[0], ref)
    plan.close()


def test_plan_launches_on_other_streams(gpu):
    """A plan created on one stream and launched and autotuned on another:
    the launch stream waits for the packing and the creation stream for the
    launches before it repacks or frees, so every launch equals the oracle."""
    import naja_amd

    t = gpu
    A, b, _ = random_polytope(16, 900, seed=43)
    nch, nspc, thin = 4, 200, 3
    ref = oracle.chr_run_f32(A.astype(np.float32), b, np.zeros((nch, 16), np.float32), nspc, thin, seed=2,
                             pair_prob=0.3)["samples"]
    make, use = t.cuda.Stream(), t.cuda.Stream()
    with t.cuda.stream(make):
        At, bt = t.from_numpy(A.astype(np.float32)).cuda(), t.from_numpy(b).cuda()
        X0 = t.zeros((1, nch, 16), device="cuda")
        plan = naja_amd.ChrPlan(At, bt, nspc, thin, nch, seed=2, pair_prob=0.3, stream=make)
    use.wait_stream(make)
    with t.cuda.stream(use):
        out = plan.samples_like()
        outs = [plan.launch(X0, out, stream=use).clone()]
        plan.autotune(X0, stream=use)  # repacks on `make` between timed launches on `use`
        outs.append(plan.launch(X0, plan.samples_like(), stream=use))
    use.synchronize()
    for o in outs:
        np.testing.assert_array_equal(o.cpu().numpy()[0], ref)
    plan.close()
    make.synchronize()


def test_streamed_chunks_reproduce_one_run(gpu):
    """CoordinateHitAndRunStreamed semantics: chunks continue the Philox
    stream and the chain state; with chunk steps a multiple of the resync
    interval the concatenated chunks equal one long run bit for bit, the
    state is persisted, and chunk backmaps equal the backmap of the run."""
    import naja_amd

    t = gpu
    A, b, _ = random_polytope(12, 300, seed=17)
    nch, nspc, thin, chunk = 3, 250, 4, 60  # 240 steps per chunk, resync 120
    At, bt = t.from_numpy(A.astype(np.float32)).cuda(), t.from_numpy(b).cuda()
    kw = dict(seed=4, pair_prob=0.3, resync_interval=120)
    X0 = t.zeros((nch, 12), dtype=t.float32, device="cuda")
    full, yfin = naja_amd.coordinate_hit_and_run(At, bt, X0, nspc, thin, nch, return_final=True, **kw)
    t.cuda.synchronize()
    parts = {}
    X = X0.clone()
    naja_amd.coordinate_hit_and_run_streamed(At, bt, X, nspc, thin, nch, chunk,
                                             lambda rows, first: parts.__setitem__(first, rows), **kw)
    got = np.concatenate([parts[k] for k in sorted(parts)])
    assert sorted(parts) == [k * chunk * nch for k in range(5)]
    np.testing.assert_array_equal(got, full.cpu().numpy())
    np.testing.assert_array_equal(X.cpu().numpy(), yfin.cpu().numpy())
    rng = np.random.default_rng(2)
    T = t.from_numpy(rng.standard_normal((7, 12)).astype(np.float32)).cuda()
    sh = t.from_numpy(rng.standard_normal(7).astype(np.float32)).cuda()
    V = naja_amd.backmap(T, sh, full).cpu().numpy()
    parts.clear()
    X = X0.clone()
    naja_amd.coordinate_hit_and_run_streamed(At, bt, X, nspc, thin, nch, chunk,
                                             lambda rows, first: parts.__setitem__(first, rows), T=T, shift=sh, **kw)
    np.testing.assert_array_equal(np.concatenate([parts[k] for k in sorted(parts)]), V)


def test_kernel_timing_hook(gpu):
    """naja_chr_timing / naja_chr_kernel_times: one record per CHR launch while
    enabled (work items, positive ms), none when disabled, cleared on read."""
    import naja_amd

    t = gpu
    A, b, _ = cube(6)
    X0 = t.zeros((3, 6))
    naja_amd.chr_kernel_times()
    naja_amd.chr_kernel_timing(True)
    naja_amd.coordinate_hit_and_run(t.from_numpy(A.astype(np.float32)), t.from_numpy(b), X0, 50, 2, 3, seed=1)
    naja_amd.chr_kernel_timing(False)
    naja_amd.coordinate_hit_and_run(t.from_numpy(A.astype(np.float32)), t.from_numpy(b), X0, 50, 2, 3, seed=1)
    rec = naja_amd.chr_kernel_times()
    assert len(rec) == 1 and rec[0][0] == 3 and rec[0][1] > 0.0
    assert naja_amd.chr_kernel_times() == []
