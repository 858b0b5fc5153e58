"""Benchmark: min-ESS/sec of plain CHR on the genome-scale reduced polytope.

Workload (one "step" = one pass of the hot path over one synthetic batch):
  genome-scale reduced polytope, n = 2583 reactions, d = 512, A = [T; -T]
  (m = 5166) FP32, 4 chains x 12500 samples, thinning 85, pair_prob 0.3,
  Philox seed 0; then the backmap V = T y + shift to n = 2583 (FP32 MFMA) and
  the GPU ESS / split-R-hat of the reduced samples. Inputs are resident in HBM
  before the timed region.
value = min over dims of per-dim ESS (minimum over the 4 chains, chains taken
from the interleaved rows as they are written) / seconds per step.

Prints one JSON line (rank 0). `--gpus N` with N > 1 exits: this build runs on
one GPU.
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bounded CPU-baseline sample length")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nspc", type=int, default=12500)
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="after timing, copy a buffer of this size (a known-size kernel for rocprof counter scaling)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described in DESIGN.md",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch

    import naja_amd
    from naja_amd import workloads

    n, d, nch, nspc, thin, pair_prob, seed = 2583, 512, 4, args.nspc, 85, 0.3, 0
    g = workloads.genome_scale(n=n, d=d, seed=0)
    m = g["A"].shape[0]
    dev = torch.device("cuda:0")
    A = torch.from_numpy(g["A"].astype(np.float32)).to(dev)
    b = torch.from_numpy(g["b"]).to(dev)
    Y0 = torch.zeros((nch, d), dtype=torch.float32, device=dev)
    T = torch.from_numpy(g["T"].astype(np.float32)).to(dev)
    shift = torch.from_numpy(g["shift"].astype(np.float32)).to(dev)
    stream = torch.cuda.current_stream()

    def step(evs=None):
        if evs is not None:
            evs[0].record(stream)
        S = naja_amd.coordinate_hit_and_run(A, b, Y0, nspc, thin, nch, seed=seed, pair_prob=pair_prob)
        if evs is not None:
            evs[1].record(stream)
        V = naja_amd.backmap(T, shift, S)
        if evs is not None:
            evs[2].record(stream)
        ess, rhat, ess_chain = naja_amd.diagnostics(S, nch)
        if evs is not None:
            evs[3].record(stream)
        return S, V, ess, rhat, ess_chain

    print(f"bench: genome-scale n={n} d={d} m={m}, {nch} chains x {nspc} x thin {thin}; warmup {args.warmup}",
          file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        out = step(events[k])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    S, V, ess, rhat, ess_chain = out
    sec = (t1 - t0) / args.steps
    chr_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in events]))
    bm_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in events]))
    dg_ms = float(np.mean([e[2].elapsed_time(e[3]) for e in events]))
    ess_np = ess.cpu().numpy()
    rhat_np = rhat.cpu().numpy()
    min_ess = float(ess_np.min())
    value = min_ess / sec
    print(f"bench: {sec*1e3:.1f} ms/step (chr {chr_ms:.1f}, backmap {bm_ms:.2f}, diag {dg_ms:.2f}); "
          f"min-ESS {min_ess:.1f}, max R-hat {rhat_np.max():.3f}", file=sys.stderr, flush=True)

    # Roofline of the dominant kernel (CHR step): algorithmic bytes per
    # chain-step B_cs = 4 m (2 + kbar/G) + 4 d / t, kbar = 1 + pair_prob, G = 1.
    chain_steps = nch * nspc * thin
    b_cs = 4.0 * m * (2.0 + (1.0 + pair_prob) / 1.0) + 4.0 * d / thin
    achieved = b_cs * chain_steps / (chr_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "kernel": "chr_step_kernel", "kernel_ms": round(chr_ms, 3),
                "bytes_per_chain_step": round(b_cs, 1), "chain_steps": chain_steps,
                "us_per_step": round(chr_ms * 1e3 / (nspc * thin), 4)}

    if args.copy_probe_mb > 0:
        nel = args.copy_probe_mb * (1 << 20) // 4
        src = torch.ones(nel, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(g, nch, thin, pair_prob, seed, S.cpu().numpy(), min_ess, args.cpu_seconds, nspc)

    line = {
        "metric": "min-ESS/sec at split-R-hat<1.05, learned-rounded vs plain CHR, 1 MI355X",
        "value": round(value, 4),
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
        "config": {"workload": "genome-scale reduced polytope, plain CHR (n=2583, d=512, m=5166, "
                               "4 chains x 12500, thin 85, pair_prob 0.3, backmap, GPU ESS/R-hat)",
                   "nspc": nspc, "min_ess": round(min_ess, 3), "rhat_max": round(float(rhat_np.max()), 4),
                   "rhat_gate_1_05": bool(rhat_np.max() < 1.05),
                   "chr_ms": round(chr_ms, 3), "backmap_ms": round(bm_ms, 3), "diag_ms": round(dg_ms, 3),
                   "parallelism": "single GPU, one workgroup per chain"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(g, nch, thin, pair_prob, seed, samples_gpu, min_ess, budget_s, nspc):
    """The FP32 oracle (same stream, bit-identical chains) on the host cores,
    on a bounded prefix of the run, extrapolated linearly in steps; plus the
    oracle diagnostics on the full sample set and the FP64 backmap on a row
    sample, both scaled."""
    import numpy as np

    import oracle

    A32 = g["A"].astype(np.float32)
    y0 = np.zeros((nch, g["A"].shape[1]), np.float32)
    probe = oracle.chr_run_f32(A32, g["b"], y0, 20, thin, seed=seed, pair_prob=pair_prob, nthreads=nch)
    per_sample = probe["seconds"] / 20
    ns = max(20, min(nspc, int(budget_s / max(per_sample, 1e-9))))
    r = oracle.chr_run_f32(A32, g["b"], y0, ns, thin, seed=seed, pair_prob=pair_prob, nthreads=nch)
    rate = nch * ns * thin / r["seconds"]  # chain-steps / s
    t_sample = nch * nspc * thin / rate
    t0 = time.perf_counter()
    oracle.ess(samples_gpu, nch)
    oracle.split_rhat(samples_gpu, nch)
    t_diag = time.perf_counter() - t0
    rows = 200
    t0 = time.perf_counter()
    oracle.backmap(g["T"], g["shift"], samples_gpu[:rows])
    t_bm = (time.perf_counter() - t0) * samples_gpu.shape[0] / rows
    total = t_sample + t_diag + t_bm
    return {"value": round(min_ess / total, 6), "unit": "min-ESS/s", "cores": nch, "kind": "port",
            "sample": f"oracle CHR {nch} chains x {ns} samples x thin {thin} ({r['seconds']:.1f} s, "
                      f"{rate:.0f} chain-steps/s, extrapolated to {nspc} samples: {t_sample:.1f} s) + oracle "
                      f"ESS/R-hat on the full samples ({t_diag:.1f} s) + FP64 backmap of {rows} rows scaled "
                      f"({t_bm:.1f} s)",
            "chain_steps_per_s": round(rate, 1), "seconds_total": round(total, 2)}


if __name__ == "__main__":
    main()
