"""Benchmark: min-ESS/sec at split-R-hat < 1.05, learned-rounded vs plain CHR.

Workload (one "step" = one pass of the hot path over one synthetic batch):
  256 held-out conditioned variants of the hypercube [-1, 1]^100 (A0 = [I; -I]),
  each adding k ~ U{0..300} extra rows: slabs |a . y| <= w around the start,
  a from a fixed pool of 300 unit directions, w log-uniform on [1e-3, 1e-1].
  4 chains per variant x 12500 samples x thinning 16, FP64 resync every 1000
  steps, all variants in one batched launch. The learned-rounded step: when
  validation accepted the network, Dikin ellipsoid P0 -> MLP inference ->
  learned low-rank update of P0 -> the facet-aligned rounding completed from
  it; else the base: the facet-aligned rounding completed from the cube's
  facets (orthonormalised) with the maximum-volume exchange -> main run (pack
  + rounding transform + CHR + backmap to y) -> per-variant ESS and
  split-R-hat -> the variants failing R-hat < 1.05 run again with the
  fallback (completed from the Dikin ellipsoid), all on the GPU. Inputs are
  resident in HBM when timing starts.
value = sum over variants with split-R-hat < 1.05 of the variant's min-ESS
  (minimum over dims of the minimum-over-chains ESS) / seconds per step.
Beside it, on the same variants, chains and steps and timed the same way:
plain CHR, the reference's Dikin rounding, the Dikin ellipsoid, the
facet-aligned rounding completed from the ellipsoid, and the facet-aligned
rounding without the network; the cube run (d = 100, 4 x 12500,
thinning 16); the batched conditions at their own half-widths (1e-2..1); the
genome-scale run (n = 2583, d = 512, 4 chains x 12500, thinning 85, pair-prob
0.3) and the large-scale run (n = 10600, d = 2400, 1024 chains x 2000,
thinning 400).

Prints one JSON line. `--gpus N` with N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published
# Counter summary of the main CHR launch (rocprofv3 FETCH_SIZE and WRITE_SIZE
# passes of this benchmark, scaled by what the same counters read for a copy
# kernel of known size); the command that made it is inside.
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "chr_main_launch_traffic.json")
METRIC = "min-ESS/sec at split-R-hat<1.05, learned-rounded vs plain CHR, 1 MI355X"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--variants", type=int, default=256)
    ap.add_argument("--nspc", type=int, default=12500)
    ap.add_argument("--pilots", type=int, default=2048)
    ap.add_argument("--epochs", type=int, default=150)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bounded CPU-baseline sample length")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-genome", action="store_true", help="skip the genome-scale plain run")
    ap.add_argument("--no-large", action="store_true", help="skip the large-scale runs (m = 21200, d = 2400)")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the comparison methods (for counter passes over the headline's launches)")
    ap.add_argument("--no-extra", action="store_true", help="skip the cube and batched-conditions lines")
    ap.add_argument("--only-large", action="store_true",
                    help="run only the large-scale line and print it (for profiling its kernel)")
    ap.add_argument("--large-nspc", type=int, default=2000, help="samples per chain of the large-scale run")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="after timing, copy a buffer of this size (a known-size kernel for counter scaling)")
    return ap.parse_args()


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described in DESIGN.md",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch

    if args.only_large:
        print(json.dumps({"large_scale_plain": large_scale(np, torch, args.large_nspc, rounded=False)}), flush=True)
        return

    import naja_amd
    from naja_amd import conditioning as cd
    from naja_amd import precond
    from naja_amd.learned import LearnedPreconditioner

    d, nch, thin, resync, seed = 100, 4, 16, 1000, 7
    A0, b0 = np.vstack([np.eye(d), -np.eye(d)]), np.ones(2 * d)
    pool = cd.make_pool(d)

    # Learned preconditioner: pilots + training (reported, not in the timed step).
    EA, Eb, F, _ = cd.sample_conditions(pool, args.pilots, 300, 1e-3, 1e-1, seed=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    S = cd.pilot_covariances(A0, b0, EA, Eb)
    torch.cuda.synchronize()
    t_pilot = time.perf_counter() - t0
    lp = LearnedPreconditioner(seed=0, model=LearnedPreconditioner.COVARIANCE, out_scale=1.0)
    t0 = time.perf_counter()
    hist = cd.train(lp, F, S, epochs=args.epochs)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del S
    log(f"pilots {t_pilot:.2f} s, training {t_train:.2f} s (NLL {hist[0]:.2f} -> {hist[-1]:.2f})")
    # Model-level acceptance on fresh training-distribution conditions.
    t0 = time.perf_counter()
    val = cd.validate_learned(lp, A0, b0, pool)
    t_val = time.perf_counter() - t0
    use_net = val["accept"]
    log(f"validation ({val['conditions']} conditions, {t_val:.2f} s): learned completion sum min-ESS "
        f"{val['learned_min_ess_sum']} ({val['learned_n_pass']} pass) vs base {val['base_min_ess_sum']} "
        f"({val['base_n_pass']} pass): learned map {'accepted' if use_net else 'rejected'}")

    EAh, Ebh, Fh, nsh = cd.sample_conditions(pool, args.variants, 300, 1e-3, 1e-1, seed=99)
    nv, nspc = len(EAh), args.nspc
    # The held-out polytopes are resident in HBM before timing starts.
    from naja_amd import VariantBatch

    batch = VariantBatch(A0, b0, EAh, Ebh)
    Fh_dev = torch.as_tensor(Fh, dtype=torch.float32, device="cuda")  # condition features, resident like the rows

    def fallback(sub, idx):
        return cd.facet_maps(sub, cd.base_rounding(sub, None, None, None))

    # Rejected by validation: the step is the base pipeline, one replay of a
    # persistent hipGraph (rounding, transform, CHR + backmap, diagnostics),
    # with event nodes around its CHR launch; no network work.
    run = None if use_net else cd.base_run(batch, nch, nspc, thin, seed, resync, timed=True)

    def learned_step():
        if use_net:
            P, _ = cd.learned_rounding(lp, Fh_dev, batch)
            return cd.sample_accepted(batch, P, fallback, nch, nspc, thin, seed, resync)
        return cd.sample_accepted(batch, None, fallback, nch, nspc, thin, seed, resync, first=cd.graph_first(run))

    # The other methods on the same variants, chains and steps, timed the same
    # way (warmup, then the mean of `steps` synchronised repetitions).
    others = cd.compare(A0, b0, EAh, Ebh, Fh, lp, nchains=nch, nspc=nspc, thin=thin, seed=seed, resync=resync,
                        methods=() if args.no_compare else ("plain", "dikin", "ellipsoid", "facet_ellipsoid",
                                                            "facet"),
                        warmup=args.warmup, steps=args.steps, log=log)
    plain = others.get("plain", {"metric": float("nan"), "metric_all": float("nan")})

    # Headline: the learned-rounded step, with HIP events around every CHR
    # step-kernel launch (recorded by the library on the launch stream) for
    # the main launch's kernel time.
    for _ in range(args.warmup):
        learned_step()
    torch.cuda.synchronize()
    naja_amd.chr_kernel_times()
    naja_amd.chr_kernel_timing(True)
    graph_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        Y, e, r, n_again = learned_step()
        if run is not None:
            graph_ms.append(run.sampling_ms())
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    naja_amd.chr_kernel_timing(False)
    launches = naja_amd.chr_kernel_times()
    chr_ms = graph_ms if run is not None else [ms for items, ms in launches if items == nv * nch]
    chr_kernel_ms = float(np.mean(chr_ms)) if chr_ms else float("nan")
    ok = r < 1.05
    value = float(e[ok].sum()) / sec

    # Zero bound violations: max over variants and samples of [A0; E_v] y - b_v.
    worst = -np.inf
    for v in range(nv):
        A = torch.as_tensor(np.vstack([A0, EAh[v]]), dtype=torch.float64, device="cuda")
        b = torch.as_tensor(np.concatenate([b0, Ebh[v]]), dtype=torch.float64, device="cuda")
        worst = max(worst, float((Y[v].double() @ A.T - b).max()))
    del Y
    log(f"learned-rounded: {sec * 1e3:.1f} ms/step (CHR main launch {chr_kernel_ms:.1f} ms), {int(ok.sum())}/{nv} "
        f"variants pass R-hat<1.05, sum min-ESS {e.sum():.1f}, {n_again} run again with the fallback, "
        f"max(Ay-b) {worst:.2e}")

    # Roofline of the dominant kernel (the batched CHR step). Algorithmic bytes
    # per chain-step: what must cross from memory to the CU, the column
    # A[:, i] of each move (4 m_v bytes, m_v = 200 + k_v rows, coordinate
    # moves) and the amortised sample store (4 d / t). The slack (8 m_v bytes
    # read and written per step in the reference) stays in registers here.
    m_v = 2 * d + np.array([e_.shape[0] for e_ in EAh])
    steps_per_chain = nspc * thin
    total_bytes = float(np.sum(nch * steps_per_chain * (4.0 * m_v + 4.0 * d / thin)))
    achieved = total_bytes / (chr_kernel_ms * 1e-3) / 1e9
    traffic, issue = measured_traffic(nv, nch, nspc, thin)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "issue_active_frac": issue,
                "kernel": "chr_step_kernel (batched, rounded variants)", "kernel_ms": round(chr_kernel_ms, 3),
                "launches_timed": len(chr_ms), "bytes_per_chain_step_mean": round(total_bytes / (nv * nch * steps_per_chain), 1),
                "chain_steps": int(nv * nch * steps_per_chain),
                "note": "achieved: the algorithm's column + sample bytes / the CHR kernel's own time (HIP events "
                        "around the launch: event nodes of the step's graph); traffic: FETCH_SIZE + WRITE_SIZE per "
                        "launch (the paired-row layout "
                        "packs each [row; -row] pair once and the variant's A sits in LDS, so about 4 % of the "
                        "algorithmic bytes cross HBM) and issue_active_frac: SQ_ACTIVE_INST_ANY / SQ_WAVE_CYCLES, "
                        "both from profiles/chr_main_launch_traffic.json (one wavefront per SIMD: instruction "
                        "issue and the step's latency bind, not bandwidth)"}

    cube = None if args.no_extra else cube_plain(np, torch, args)
    batched = None if args.no_extra else batched_conditions(np, torch, cd, pool, lp, A0, b0, args, use_net)
    genome = None if args.no_genome else genome_scale(np, torch, rounded=False)
    genome_r = None if args.no_genome else genome_scale(np, torch, rounded=True)
    large = None if args.no_large else large_scale(np, torch, args.large_nspc, rounded=False)
    large_r = None if args.no_large else large_scale(np, torch, args.large_nspc, rounded=True)

    if args.copy_probe_mb > 0:
        nel = args.copy_probe_mb * (1 << 20) // 4
        src = torch.ones(nel, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()

    cpu = None
    if not args.no_cpu:
        P = (cd.learned_rounding(lp, Fh_dev, batch)[0] if use_net
             else cd.facet_maps(batch))
        cpu = cpu_baseline(np, A0, b0, EAh, Ebh, P.cpu().numpy(), nch, nspc, thin, seed, resync,
                           float(e[ok].sum()), args.cpu_seconds)
    lp.close()

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "min-ESS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{nv} held-out conditioned variants of the cube [-1,1]^100, 0-300 extra slab rows "
                        f"(half-widths 1e-3..1e-1), {nch} chains x {nspc} x thin {thin}, one batched launch; "
                        f"learned-rounded (facet-aligned rounding completed from the learned low-rank update of "
                        f"the Dikin ellipsoid when validation accepts the network, else the base: completed from "
                        f"the cube's own facets, orthonormalised, with the maximum-volume exchange, the whole step "
                        f"one replay of a persistent hipGraph; R-hat acceptance with fallback)",
            "variants": nv, "chains": nch, "nspc": nspc, "thin": thin, "resync": resync,
            "n_pass_rhat_1_05": int(ok.sum()), "min_ess_sum": round(float(e.sum()), 1),
            "min_ess_sum_passing": round(float(e[ok].sum()), 1), "rerun_with_fallback": n_again,
            "max_bound_violation": worst, "bound_violations": bool(worst > 1e-5),
            "ratio_vs_plain": round(value / plain["metric"], 3) if plain["metric"] > 0 else None,
            "ratio_vs_plain_all_variants": (round((float(e.sum()) / sec) / plain["metric_all"], 3)
                                            if plain["metric_all"] > 0 else None),
            "methods": {name: {"value": round(o["metric"], 3), "ms_per_step": round(o["seconds"] * 1e3, 3),
                               "n_pass": o["n_pass"], "min_ess_sum": round(o["ess_sum"], 1),
                               "min_ess_sum_passing": round(o["ess_sum_pass"], 1)}
                        for name, o in others.items()},
            "methods_note": "every method timed like the headline (same warmup and steps); plain: no rounding; "
                            "dikin: the reference's dikin_precondition (host FP64); ellipsoid: Dikin ellipsoid; "
                            "facet_ellipsoid: facet-aligned rounding completed from the Dikin ellipsoid; facet: "
                            "facet-aligned rounding completed from the cube's facets with the maximum-volume "
                            "exchange (the method without the network)",
            "pilot_seconds": round(t_pilot, 3), "training_seconds": round(t_train, 3),
            "learned_validation": dict(val, seconds=round(t_val, 3)),
            "training_nll_first_last": [round(hist[0], 3), round(hist[-1], 3)],
            "cube_plain": cube,
            "batched_conditions": batched,
            "genome_scale_plain": genome,
            "genome_scale_rounded": genome_r,
            "large_scale_plain": large,
            "large_scale_rounded": large_r,
            "parallelism": "single GPU; variants x chains as workgroups of one batched launch",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cube_plain(np, torch, args):
    """Plain CHR on the hypercube [-1, 1]^100 (A = [I; -I], b = 1, y0 = 0),
    4 chains x 12500 x thinning 16, Philox seed 0, GPU ESS / split-R-hat; a
    coordinate move resamples x_i exactly, so the per-chain ESS has the closed
    form n (1 - rho) / (1 + rho), rho = (1 - 1/d)^t."""
    import naja_amd
    from naja_amd import conditioning as cd

    d, nch, nspc, thin = 100, 4, 12500, 16
    A = torch.from_numpy(np.vstack([np.eye(d), -np.eye(d)]).astype(np.float32)).cuda()
    b = torch.ones(2 * d, dtype=torch.float64, device="cuda")
    Y0 = torch.zeros((nch, d), dtype=torch.float32, device="cuda")

    def step():
        S = naja_amd.coordinate_hit_and_run(A, b, Y0, nspc, thin, nch, seed=0)
        ess, rhat, ec = naja_amd.diagnostics(S, nch)
        return S, ess, rhat, ec

    sec, (S, ess, rhat, ec) = cd.timed_steps(step, args.warmup, args.steps)
    rho = (1.0 - 1.0 / d) ** thin
    closed = nspc * (1.0 - rho) / (1.0 + rho)
    x = S.double()
    min_ess, rmax = float(ess.min()), float(rhat.max())
    log(f"cube d=100: {sec * 1e3:.2f} ms per step, min-ESS {min_ess:.1f} (per-chain mean {float(ec.mean()):.1f}, "
        f"closed form {closed:.1f}), max R-hat {rmax:.4f}")
    return {"value_min_ess_per_s": round(min_ess / sec, 2) if rmax < 1.05 else 0.0, "ms_per_step": round(sec * 1e3, 3),
            "min_ess": round(min_ess, 2), "ess_chain_mean": round(float(ec.mean()), 2),
            "ess_chain_closed_form": round(closed, 2), "rhat_max": round(rmax, 5),
            "mean_abs_max": round(float(x.mean(0).abs().max()), 4), "var_mean": round(float(x.var(0).mean()), 5),
            "var_closed_form": round(1.0 / 3.0, 5)}


def batched_conditions(np, torch, cd, pool, lp, A0, b0, args, use_net):
    """256 conditioned variants of the cube at their own slab half-widths,
    log-uniform on [1e-2, 1], 4 chains x 12500 x thinning 16 per variant, all
    in one launch with FP64 resync every 1000 steps: plain CHR, the
    facet-aligned rounding and the learned-rounded pipeline (its network was
    trained on half-widths 1e-3..1e-1), each timed like the headline."""
    EA, Eb, F, _ = cd.sample_conditions(pool, 256, 300, 1e-2, 1.0, seed=3)
    res = cd.compare(A0, b0, EA, Eb, F, lp, nchains=4, nspc=12500, thin=16, seed=11, resync=1000,
                     methods=("plain", "facet", "learned"), warmup=args.warmup, steps=max(1, args.steps // 4),
                     log=lambda m: log("batched conditions (1e-2..1) " + m), use_network=use_net)
    return {name: {"value": round(o["metric"], 3), "ms_per_step": round(o["seconds"] * 1e3, 3),
                   "n_pass": o["n_pass"], "min_ess_sum": round(o["ess_sum"], 1)} for name, o in res.items()}


def measured_traffic(nv, nch, nspc, thin):
    """(HBM bytes per main CHR launch, fraction of wave cycles issuing) from
    the committed counter summary, when it was measured on this workload
    (else (None, None))."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None, None
    if [t.get("variants"), t.get("chains"), t.get("nspc"), t.get("thin")] != [nv, nch, nspc, thin]:
        return None, None
    return float(t["fetch_bytes"]) + float(t["write_bytes"]), t.get("issue_active_frac")


def scale_run(np, torch, n, d, nch, nspc, thin, pp, rounded, label):
    """CHR on a genome-like reduced polytope (workloads.genome_scale: T with
    8-40 N(0,1) nonzeros per row densified by a random rotation, margins
    log-uniform on [1e-3, 1e2], A = [T; -T]), Philox seed 0, pair-prob pp,
    then the FP32 MFMA backmap to the n reactions and the GPU ESS /
    split-R-hat over the d reduced coordinates. rounded: the Dikin ellipsoid
    at the analytic centre (the start y = 0 lies 1e-3 from its closest
    facet), chains started at the centre; timed from the polytope on the
    device to samples + diagnostics on the device, the centre and rounding
    included. HIP events time each stage."""
    import naja_amd
    from naja_amd import VariantBatch, workloads

    g = workloads.genome_scale(n=n, d=d, seed=0)
    m = g["A"].shape[0]
    A = torch.from_numpy(g["A"].astype(np.float32)).cuda()
    b = torch.from_numpy(g["b"]).cuda()
    T = torch.from_numpy(g["T"].astype(np.float32)).cuda()
    shift = torch.from_numpy(g["shift"].astype(np.float32)).cuda()
    batch = VariantBatch(g["A"], g["b"], [np.zeros((0, d))], [np.zeros(0)]) if rounded else None
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
    log(f"{label}: m={m} d={d}, {nch} chains x {nspc} x thin {thin} ({nspc * thin} steps per chain)")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record(stream)
    steps = 0
    if rounded:
        xc, steps = naja_amd.analytic_center(A, b)
        P = naja_amd.ellipsoid_rounding_large(A, b, x_c=xc)
        ev[1].record(stream)
        S = naja_amd.coordinate_hit_and_run_batched(batch, None, None, None, torch.zeros((1, nch, d), device="cuda"),
                                                    nspc, thin, nch, seed=0, pair_prob=pp, P=P[None].contiguous(),
                                                    c=xc[None].contiguous())[0]
    else:
        ev[1].record(stream)
        S = naja_amd.coordinate_hit_and_run(A, b, torch.zeros((nch, d), dtype=torch.float32, device="cuda"), nspc,
                                            thin, nch, seed=0, pair_prob=pp)
    ev[2].record(stream)
    V = naja_amd.backmap(T, shift, S)
    ev[3].record(stream)
    ess, rhat, _ = naja_amd.diagnostics(S, nch)
    ev[4].record(stream)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    del V
    pre_ms, chr_ms, bm_ms, dg_ms = (ev[i].elapsed_time(ev[i + 1]) for i in range(4))
    chain_steps = nch * nspc * thin
    # Bound check on every sample (FP64, in row blocks): max over rows of A y - b.
    viol = max(float((S[i:i + 65536].double() @ A.double().T - b[None, :]).max()) for i in range(0, S.shape[0], 65536))
    min_ess, rmax = float(ess.min()), float(rhat.max())
    # Algorithmic bytes per chain-step that leave the chip: the columns of
    # the move (one per coordinate move, two per pair move) and the amortised
    # sample store; the slack stays in registers.
    col_bytes = 4.0 * m * (1.0 + pp) + 4.0 * d / thin
    log(f"{label}: preconditioner {pre_ms:.1f} ms ({steps} Newton steps), sampling {chr_ms / 1e3:.2f} s, backmap "
        f"{bm_ms:.1f} ms, diag {dg_ms:.1f} ms; min-ESS {min_ess:.2f}, median {float(ess.median()):.1f}, max R-hat "
        f"{rmax:.3f}, dims with R-hat >= 1.05: {int((rhat >= 1.05).sum())}, max(Ay-b) {viol:.2e}")
    return {"value_min_ess_per_s": round(min_ess / sec, 4), "seconds": round(sec, 3),
            "preconditioner_ms": round(pre_ms, 2), "newton_steps": steps, "sampling_s": round(chr_ms / 1e3, 3),
            "backmap_ms": round(bm_ms, 2), "diag_ms": round(dg_ms, 2), "min_ess": round(min_ess, 3),
            "median_ess": round(float(ess.median()), 2), "rhat_max": round(rmax, 4),
            "dims_rhat_ge_1_05": int((rhat >= 1.05).sum()), "dims": d, "max_bound_violation": viol,
            "chains": nch, "nspc": nspc, "thin": thin, "chain_steps_per_s": round(chain_steps / (chr_ms * 1e-3), 1),
            "column_GBs": round(col_bytes * chain_steps / (chr_ms * 1e-3) / 1e9, 1),
            "backmap_TFLOPs": round(2.0 * d * n * nch * nspc / (bm_ms * 1e-3) / 1e12, 2)}


def genome_scale(np, torch, rounded):
    """n = 2583, d = 512, m = 5166, 4 chains x 12500 x thinning 85, pair-prob 0.3."""
    return scale_run(np, torch, 2583, 512, 4, 12500, 85, 0.3, rounded,
                     "genome-scale " + ("rounded" if rounded else "plain"))


def large_scale(np, torch, nspc, rounded):
    """n = 10600, d = 2400, m = 21200 (203.5 MB FP32), 1024 chains x nspc x thinning 400, pair-prob 0.3."""
    return scale_run(np, torch, 10600, 2400, 1024, nspc, 400, 0.3, rounded,
                     "large-scale " + ("rounded" if rounded else "plain"))


def cpu_baseline(np, A0, b0, EA, Eb, P, nch, nspc, thin, seed, resync, ess_pass_sum, budget_s):
    """The FP32 oracle (same Philox stream and arithmetic as the GPU kernel)
    on the host cores, over a bounded sample of the learned-rounded main run:
    whole variants (every chain, every step) chosen at random, as many as fill
    about budget_s seconds on 16 threads (4 variants at a time x 4 chains);
    extrapolated linearly in chain-steps to all variants."""
    import concurrent.futures as cf

    import oracle

    y0 = np.zeros((nch, A0.shape[1]), np.float32)

    def job(v):
        A = np.vstack([A0, EA[v]]).astype(np.float32)
        b = np.concatenate([b0, Eb[v]])
        At, bt = oracle.transform(A, b, P[v].astype(np.float32), np.zeros(A.shape[1], np.float32))
        return At, bt, int(v)

    rng = np.random.default_rng(0)
    order = rng.permutation(len(EA))
    first = job(order[0])
    probe = oracle.chr_run_f32(first[0], first[1], y0, 50, thin, seed=seed, variant=first[2],
                               resync_interval=resync, nthreads=nch)
    per_variant = probe["seconds"] / 50 * nspc  # all chains of one variant, whole run
    n_pick = int(min(len(EA), max(4, round(budget_s * 4 / max(per_variant, 1e-9)))))
    jobs = [first] + [job(v) for v in order[1:n_pick]]
    ns = nspc

    def one(job):
        At, bt, v = job
        return oracle.chr_run_f32(At, bt, y0, ns, thin, seed=seed, variant=v, resync_interval=resync,
                                  nthreads=nch)["seconds"]

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        list(ex.map(one, jobs))
    wall = time.perf_counter() - t0
    chain_steps = len(jobs) * nch * ns * thin
    rate = chain_steps / wall
    total_chain_steps = len(EA) * nch * nspc * thin
    t_total = total_chain_steps / rate
    return {"value": round(ess_pass_sum / t_total, 4), "unit": "min-ESS/s", "cores": 16, "kind": "port",
            "sample": f"oracle CHR (FP32 slack, FP64 resync, same Philox stream) on the learned-rounded "
                      f"polytopes of {len(jobs)} of the {len(EA)} variants, whole runs ({nch} chains x {ns} "
                      f"samples x thin {thin}; {wall:.1f} s wall, 4 variants at a time x 4 chain threads, "
                      f"{rate:.0f} chain-steps/s), extrapolated to all {len(EA)} variants: {t_total:.1f} s; "
                      f"diagnostics and the rounding solves not included",
            "chain_steps_per_s": round(rate, 1), "seconds_total": round(t_total, 1)}


if __name__ == "__main__":
    main()
