"""Conditioned-variant workflow: pilot chains, learned preconditioner training,
the ESS/R-hat acceptance gate, and the plain / Dikin / learned comparison.

A condition adds slabs |a_j . y| <= w_j to a base polytope A0 y <= b0, with
the directions a_j taken from a fixed pool (as naja's extra rows are fixed
+-T[i, :] rows whose bounds vary per condition, cmd_prepare.cpp:111-156). Its
feature vector holds w_ref / w_j at the pool index of each active slab and 0
elsewhere.

Base rounding (and fallback): the Dikin ellipsoid at the start point, y = P0 z
with P0 P0^T = H^{-1}, H the log-barrier Hessian (FP64, host). Its
orientation comes exactly from the constraint rows, which thin slabs need
(whitening a slab of half-width w takes its normal to within ~w).

Learned map: the MLP turns features into a factor V (dim x 32) of the
covariance model C = tau (I + V V^T) in the base frame, trained on pilot
covariances measured in that frame; the variant is sampled with
y = P0 (I + V V^T)^{1/2} z. Any fixed linear map leaves the uniform target on
the polytope exact, so the map only changes mixing. Per variant, a short probe
run of the learned map and the fallback decides which one the main run uses:
the learned map is accepted when its probe min-ESS is at least the
fallback's and its probe split-R-hat is no worse (or below 1.05).

The reference's Dikin rounding (dikin_precondition.cpp:36-183) is reported
beside them as the drop-in preconditioner.

Everything that touches samples runs on the GPU through libnaja_amd.so: the
batched CHR launches, the MLP, the covariances and the diagnostics. The FP64
eigen/Cholesky work of the roundings runs on the host.
"""
from __future__ import annotations

import time

import numpy as np

from . import (NAJA_ERR_UNSUPPORTED, NajaError, VariantBatch, coordinate_hit_and_run_batched, diagnostics_batched,
               ellipsoid_rounding, lowrank_maps)
from . import precond
from .learned import LearnedPreconditioner, pilot_covariance

W_REF = 1e-3  # feature scale: the thinnest half-width of the pilot distribution


def make_pool(d: int, n: int = 300, seed: int = 2024):
    """n unit directions in R^d (the candidate slab normals)."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, d))
    return a / np.linalg.norm(a, axis=1, keepdims=True)


def sample_conditions(pool, n: int, kmax: int = 300, w_lo: float = 1e-3, w_hi: float = 1e-1, seed: int = 0):
    """n conditions, each with k ~ U{0..kmax} extra rows (k/2 slabs, k even),
    slab directions drawn from the pool without repetition, half-widths
    log-uniform on [w_lo, w_hi]. Returns (extra_A list, extra_b list,
    features (n, pool size) float32, slab counts)."""
    rng = np.random.default_rng(seed)
    npool, d = pool.shape
    EA, Eb, nslab = [], [], []
    F = np.zeros((n, npool), np.float32)
    for v in range(n):
        k = int(rng.integers(0, kmax + 1)) // 2 * 2
        ns = min(k // 2, npool)
        idx = rng.choice(npool, size=ns, replace=False)
        w = 10.0 ** rng.uniform(np.log10(w_lo), np.log10(w_hi), ns)
        a = pool[idx]
        EA.append(np.vstack([a, -a]) if ns else np.zeros((0, d)))
        Eb.append(np.concatenate([w, w]) if ns else np.zeros(0))
        F[v, idx] = (W_REF / w).astype(np.float32)
        nslab.append(ns)
    return EA, Eb, F, np.array(nslab)


def _sync():
    import torch

    torch.cuda.synchronize()


def _shape(A0, EA):
    if isinstance(A0, VariantBatch):
        return A0.n_variants, A0.d
    return len(EA), A0.shape[1]


def sample_rounded(A0, b0, EA, Eb, P, nchains, nspc, thin, seed, resync=1000):
    """Batched CHR of every variant in z with y = P_v z (P None: plain).
    A0 may be a VariantBatch (b0, EA, Eb None). Returns y samples
    (n_variants, nchains * nspc, d) on the GPU."""
    import torch

    nv, d = _shape(A0, EA)
    X0 = torch.zeros((nv, nchains, d), dtype=torch.float32, device="cuda")
    kw = {}
    if P is not None:
        Pt = P if torch.is_tensor(P) else torch.as_tensor(np.asarray(P, np.float32))
        kw = dict(P=Pt.to(device="cuda", dtype=torch.float32).contiguous(),
                  c=torch.zeros((nv, d), dtype=torch.float32, device="cuda"))
    if not isinstance(A0, VariantBatch):
        A0 = VariantBatch(A0, b0, EA, Eb)
    return coordinate_hit_and_run_batched(A0, None, None, None, X0, nspc, thin, nchains, seed=seed,
                                          resync_interval=resync, **kw)


def variant_diagnostics(Y, nchains):
    """Per variant: (min over dims of the min-over-chains ESS, max split-R-hat),
    all variants in one set of launches."""
    _, _, ess_min, rhat_max = diagnostics_batched(Y, nchains)
    return ess_min.cpu().numpy(), rhat_max.cpu().numpy()


def base_rounding(A0, b0, EA, Eb):
    """The Dikin-ellipsoid rounding of every variant (Cholesky of the
    log-barrier Hessian at the start point, FP64): the learned map's base and
    its fallback. On the GPU (P as an FP32 device tensor); the host FP64
    routine when d or the rows exceed the kernel's LDS budget."""
    import torch

    try:
        return ellipsoid_rounding(A0, b0, EA, Eb)
    except NajaError as e:
        if e.code != NAJA_ERR_UNSUPPORTED or isinstance(A0, VariantBatch):
            raise
        return torch.as_tensor(precond.ellipsoid_batched(A0, b0, EA, Eb), dtype=torch.float32).cuda()


def pilot_covariances(A0, b0, EA, Eb, P0=None, nchains=32, steps=2000, thin=4, seed=1, chunk=512):
    """Pilot chains of every condition in its base frame (y = P0 z), `steps`
    steps per chain; returns the covariance of the pilot samples in that frame,
    S_z = P0^{-1} S_y P0^{-T}, as (n, d, d) float32 on the GPU (S_y from the
    GPU covariance kernel, the change of frame in FP64 on the host)."""
    import torch

    nv, d = len(EA), A0.shape[1]
    if P0 is None:
        P0 = base_rounding(A0, b0, EA, Eb)
    A0 = VariantBatch(A0, b0, EA, Eb)
    S = torch.empty((nv, d, d), dtype=torch.float32, device="cuda")
    for lo in range(0, nv, chunk):
        hi = min(nv, lo + chunk)
        Y = sample_rounded(A0.slice(lo, hi), None, None, None, P0[lo:hi], nchains, steps // thin, thin, seed + lo)
        Sy, _ = pilot_covariance(Y)
        del Y
        P0c = P0[lo:hi].double().cpu().numpy() if torch.is_tensor(P0) else np.asarray(P0[lo:hi], np.float64)
        Pinv = np.linalg.inv(P0c)
        Sz = Pinv @ Sy.double().cpu().numpy() @ Pinv.transpose(0, 2, 1)
        S[lo:hi] = torch.as_tensor(Sz, dtype=torch.float32).cuda()
    return S


def train(lp: LearnedPreconditioner, F, S, epochs=150, batch=256, lr=1e-3, seed=0, log=None):
    """Adam over shuffled mini-batches; returns the per-epoch mean loss."""
    import torch

    n = F.shape[0]
    Fd = torch.as_tensor(F, dtype=torch.float32).cuda()
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):
        order = torch.as_tensor(rng.permutation(n), device="cuda")
        tot = torch.zeros((), dtype=torch.float64, device="cuda")
        for lo in range(0, n - batch + 1, batch):
            idx = order[lo:lo + batch]
            loss = lp.train_step(Fd[idx].contiguous(), S[idx].contiguous(), lr=lr)
            tot += loss.double().sum()
        hist.append(float(tot) / (n - n % batch))
        if log and (ep % 25 == 0 or ep == epochs - 1):
            log(f"epoch {ep}: mean NLL {hist[-1]:.3f}")
    return hist


def learned_maps(lp: LearnedPreconditioner, F, P0, expand=True):
    """P_v = P0_v (I + V_v V_v^T)^{1/2} (expand) or P0_v (I + V_v V_v^T)^{-1/2}:
    the learned low-rank update of the base (P0 a device tensor). MLP
    inference and the FP64 matrix function plus product (naja_amd.lowrank_maps)
    both on the device."""
    import torch

    V = lp.forward(torch.as_tensor(F, dtype=torch.float32).cuda())
    return lowrank_maps(P0, V, expand=expand)


def gate(ess_l, rhat_l, ess_b, rhat_b):
    """Accept the learned map where its probe min-ESS is at least the
    fallback's and its probe split-R-hat is no worse (or below 1.05)."""
    return (ess_l >= ess_b) & ((rhat_l <= rhat_b) | (rhat_l < 1.05))


def compare(A0, b0, EA, Eb, F, lp, nchains=4, nspc=12500, thin=16, seed=7, probe_nspc=500, resync=1000,
            methods=("plain", "dikin", "ellipsoid", "learned", "learned_ungated"), expand=True, log=None):
    """Sampling methods on the same variants, chains and steps:
    plain CHR; the reference's Dikin rounding (dikin_precondition); the
    Dikin-ellipsoid rounding (the learned map's base and fallback); the gated
    learned map; the learned map on every variant. Each is timed from the
    polytopes on the device to samples + diagnostics on the device, including
    its preconditioner (FP64 solves; MLP inference; probe runs + gate).
    Returns per-method results."""
    out = {}

    def finish(name, t0, Y, extra=None):
        e, r = variant_diagnostics(Y, nchains)
        _sync()
        sec = time.perf_counter() - t0
        ok = r < 1.05
        res = dict(seconds=sec, ess_min=e, rhat_max=r, n_pass=int(ok.sum()),
                   ess_sum=float(e.sum()), ess_sum_pass=float(e[ok].sum()),
                   metric=float(e[ok].sum()) / sec, metric_all=float(e.sum()) / sec)
        if extra:
            res.update(extra)
        out[name] = res
        if log:
            log(f"{name}: {sec:.2f} s, variants passing R-hat<1.05: {res['n_pass']}/{len(e)}, "
                f"sum min-ESS {res['ess_sum']:.1f} (passing {res['ess_sum_pass']:.1f}), "
                f"median min-ESS {np.median(e):.1f}")

    def timed(name, make_P):
        _sync()
        t0 = time.perf_counter()
        P, extra = make_P()
        Y = sample_rounded(batch, None, None, None, P, nchains, nspc, thin, seed, resync)
        finish(name, t0, Y, extra)
        del Y

    nv = len(EA)
    # The polytopes go to the device once, before any method is timed.
    batch = VariantBatch(A0, b0, EA, Eb)
    batch2 = batch.repeat(2)
    if "plain" in methods:
        timed("plain", lambda: (None, None))
    if "dikin" in methods:
        timed("dikin", lambda: (precond.dikin_batched(A0, b0, EA, Eb)[0], None))
    if "ellipsoid" in methods:
        timed("ellipsoid", lambda: (base_rounding(batch, None, None, None), None))

    def gated():
        import torch

        P0 = base_rounding(batch, None, None, None)
        Pl = learned_maps(lp, F, P0, expand)
        # One probe launch holds both maps of every variant.
        Yp = sample_rounded(batch2, None, None, None, torch.cat([Pl, P0]), nchains, probe_nspc, thin, seed + 1,
                            resync)
        ep, rp = variant_diagnostics(Yp, nchains)
        del Yp
        acc = gate(ep[:nv], rp[:nv], ep[nv:], rp[nv:])
        sel = torch.as_tensor(acc, device=P0.device)[:, None, None]
        return torch.where(sel, Pl, P0), dict(accepted=int(acc.sum()), accepted_mask=acc)

    if "learned" in methods:
        timed("learned", gated)
    if "learned_ungated" in methods:
        timed("learned_ungated", lambda: (learned_maps(lp, F, base_rounding(batch, None, None, None), expand), None))
    return out
