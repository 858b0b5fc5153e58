"""Conditioned-variant workflow: roundings, pilot chains, learned
preconditioner training, the R-hat acceptance, and the comparison of methods.

A condition adds slabs |a_j . y| <= w_j to a base polytope A0 y <= b0, with
the directions a_j taken from a fixed pool (as naja's extra rows are fixed
+-T[i, :] rows whose bounds vary per condition, cmd_prepare.cpp:111-156). Its
feature vector holds w_ref / w_j at the pool index of each active slab and 0
elsewhere.

Roundings (y = P z; any fixed linear map leaves the uniform target exact and
changes only how fast coordinate hit-and-run mixes):

* Dikin ellipsoid P0 at the start point (P0 P0^T = H^-1, H the log-barrier
  Hessian): whitens the body, with the orientation of a Cholesky factor.
* Facet-aligned rounding (the pipeline's base): coordinate axes along the
  edges of a parallelotope of the variant's slab facets, chosen greedily
  (naja_amd.facet_rounding) and, when the slabs fill the basis, exchanged
  towards the smallest such parallelotope; the directions the slabs leave
  free are completed by the base polytope's own facets (the cube's axes),
  orthonormalised in the free subspace. On an axis-aligned box a coordinate
  move is an exact draw, which a whitened but rotated box does not give.
  Completing from the dual rows of the Dikin ellipsoid instead (the earlier
  base, kept as the R-hat fallback) mixes about a fifth slower.
* Learned update: the MLP turns features into a factor V (dim x 32) of the
  covariance model C = tau (I + V V^T) in the Dikin-ellipsoid frame, trained
  by the Gaussian NLL on covariances of pilot chains the sampler runs (in the
  facet-aligned frame, where they mix); its map P0 (I + V V^T)^{1/2} is the
  rounding whose dual rows complete the learned variant of the facet-aligned
  rounding.

Acceptance, at two levels. After training, validate_learned runs both
completions on fresh conditions of the training distribution, and the learned
map is used only if it loses neither summed min-ESS nor R-hat passes there
(else the pipeline runs its base and the network costs nothing). After each
main run, a variant whose split-R-hat is 1.05 or more is run again with the
fallback (the facet-aligned rounding completed from the Dikin ellipsoid),
and the run with the better R-hat and ESS is kept.

The reference's Dikin rounding (dikin_precondition.cpp:36-183) is reported
beside them as the drop-in preconditioner.

Everything that touches samples runs on the GPU through libnaja_amd.so: the
batched CHR launches, the roundings, the MLP, the covariances and the
diagnostics.
"""
from __future__ import annotations

import time

import numpy as np

from . import (FACET_COMPLETE_BASE_FACETS, NAJA_ERR_UNSUPPORTED, ChrPlan, ConditionedRun, NajaError, VariantBatch,
               coordinate_hit_and_run_batched, diagnostics_batched, ellipsoid_rounding, facet_rounding, lowrank_maps)
from . import precond
from .learned import LearnedPreconditioner, pilot_covariance

W_REF = 1e-3  # feature scale: the thinnest half-width of the pilot distribution
# A facet candidate is chosen while it keeps at least this fraction of its own
# norm after projection on the chosen ones (every independent slab first).
THETA = 0.02
# Maximum-volume exchange: a slab replaces a basis row while its coefficient in
# the basis exceeds this (each exchange shrinks the parallelotope by that factor).
EXCHANGE_TOL = 1.02


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
    A0 may be a VariantBatch (b0, EA, Eb None): rounded runs then keep a plan
    on the batch per run shape and only set the new rounding each call (no
    repacking, allocation or synchronisation). Returns y samples
    (n_variants, nchains * nspc, d) on the GPU."""
    import torch

    nv, d = _shape(A0, EA)
    if P is not None and isinstance(A0, VariantBatch):
        return _batch_plan(A0, nchains, nspc, thin, seed, resync).sample(P)
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


class _RoundedPlan:
    """A rounded sampling plan of one batch and run shape (chains from z = 0,
    c = 0), reused with a new rounding at every call."""

    def __init__(self, batch, nchains, nspc, thin, seed, resync):
        import torch

        nv, d = batch.n_variants, batch.d
        self.X0 = torch.zeros((nv, nchains, d), dtype=torch.float32, device="cuda")
        self.c = torch.zeros((nv, d), dtype=torch.float32, device="cuda")
        self.plan = None
        self.args = (batch, nspc, thin, nchains, seed, resync)

    def sample(self, P):
        import torch

        Pt = P if torch.is_tensor(P) else torch.as_tensor(np.asarray(P, np.float32))
        Pt = Pt.to(device="cuda", dtype=torch.float32).contiguous()
        if self.plan is None:
            batch, nspc, thin, nchains, seed, resync = self.args
            self.plan = ChrPlan(batch, None, nspc, thin, nchains, P=Pt, c=self.c, seed=seed,
                                resync_interval=resync)
        else:
            self.plan.set_rounding(Pt, self.c)
        return self.plan.launch(self.X0, self.plan.samples_like())


def _batch_plan(batch, nchains, nspc, thin, seed, resync):
    plans = batch.__dict__.setdefault("_rounded_plans", {})
    key = (nchains, nspc, thin, seed, resync)
    if key not in plans:
        plans[key] = _RoundedPlan(batch, nchains, nspc, thin, seed, resync)
    return plans[key]


def base_run(batch, nchains, nspc, thin, seed, resync=1000, timed=False):
    """The pipeline's base step over a VariantBatch (facet-aligned rounding
    completed from the base polytope's facets with the exchange, CHR from
    z = 0, backmap, diagnostics) as one persistent hipGraph, kept on the
    batch per run shape (ConditionedRun)."""
    runs = batch.__dict__.setdefault("_cond_runs", {})
    key = (nchains, nspc, thin, seed, resync, bool(timed))
    if key not in runs:
        runs[key] = ConditionedRun(batch, nspc, thin, nchains, seed=seed, resync_interval=resync, theta=THETA,
                                   exchange_tol=EXCHANGE_TOL, timed=timed)
    return runs[key]


def graph_first(run):
    """A first-run callable for sample_accepted: one replay of the run's
    graph; (samples, per-variant min ESS, per-variant max R-hat)."""
    def first():
        Y, emin, rmax = run.step()
        return Y, emin.cpu().numpy(), rmax.cpu().numpy()
    return first


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


def facet_maps(batch, P_base=None, theta=THETA, lower=False):
    """Facet-aligned rounding of every variant of a VariantBatch. P_base None
    (the pipeline's base): the free directions completed by the base
    polytope's facets, orthonormalised in the free subspace, with the
    maximum-volume exchange when the slabs fill the basis. Else completed from
    the dual rows of P_base (n_variants, d, d) (lower: of the lower Cholesky
    factor of P_base P_base^T)."""
    if P_base is None:
        return facet_rounding(batch, None, None, None, None, theta, completion=FACET_COMPLETE_BASE_FACETS,
                              exchange_tol=EXCHANGE_TOL)
    return facet_rounding(batch, None, None, None, P_base, theta, lower=lower)


def pilot_covariances(A0, b0, EA, Eb, P0=None, nchains=32, steps=20000, thin=40, seed=1, chunk=256):
    """Pilot chains of every condition, `steps` steps per chain, run in its
    facet-aligned frame (where they mix); returns the covariance of the pilot
    samples in the Dikin-ellipsoid frame P0, S_z = P0^{-1} S_y P0^{-T}, as
    (n, d, d) float32 on the GPU (S_y from the GPU covariance kernel, the
    change of frame in FP64)."""
    import torch

    nv = len(EA)
    batch = VariantBatch(A0, b0, EA, Eb)
    if P0 is None:
        P0 = base_rounding(batch, None, None, None)
    d = batch.d
    S = torch.empty((nv, d, d), dtype=torch.float32, device="cuda")
    for lo in range(0, nv, chunk):
        hi = min(nv, lo + chunk)
        sub = batch.slice(lo, hi)
        P0c = P0[lo:hi]
        Y = sample_rounded(sub, None, None, None, facet_maps(sub), nchains, steps // thin, thin, seed + lo)
        Sy, _ = pilot_covariance(Y)
        del Y
        Pinv = torch.linalg.inv(P0c.double())
        S[lo:hi] = (Pinv @ Sy.double() @ Pinv.transpose(1, 2)).float()
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


def validate_learned(lp: LearnedPreconditioner, A0, b0, pool, n=256, w_lo=1e-3, w_hi=1e-1, seed=12345,
                     nchains=4, nspc=12500, thin=16, resync=1000):
    """Model-level acceptance of the trained network, on n fresh conditions of
    the training distribution (not the evaluation variants): full-length runs
    of the facet-aligned rounding completed from the learned map and of the
    pipeline's base (completed from the base polytope's facets). The learned
    map is accepted when it loses neither summed min-ESS nor variants passing
    split-R-hat < 1.05."""
    import torch

    EA, Eb, F, _ = sample_conditions(pool, n, 300, w_lo, w_hi, seed=seed)
    batch = VariantBatch(A0, b0, EA, Eb)
    P0 = base_rounding(batch, None, None, None)
    PL, _ = learned_rounding(lp, torch.as_tensor(F, dtype=torch.float32).cuda(), batch, P0)
    out = {}
    for name, P in (("learned", PL), ("base", facet_maps(batch))):
        Y = sample_rounded(batch, None, None, None, P, nchains, nspc, thin, seed, resync)
        e, r = variant_diagnostics(Y, nchains)
        del Y
        out[name] = (float(e.sum()), int((r < 1.05).sum()))
    accept = out["learned"][0] >= out["base"][0] and out["learned"][1] >= out["base"][1]
    return dict(accept=bool(accept), learned_min_ess_sum=round(out["learned"][0], 1),
                base_min_ess_sum=round(out["base"][0], 1), learned_n_pass=out["learned"][1],
                base_n_pass=out["base"][1], conditions=n)


def learned_rounding(lp: LearnedPreconditioner, F, batch, P0=None):
    """The learned-rounded maps of every variant: the facet-aligned rounding
    completed from the learned map P0 (I + V V^T)^{1/2}; also returns P0."""
    if P0 is None:
        P0 = base_rounding(batch, None, None, None)
    return facet_maps(batch, learned_maps(lp, F, P0)), P0


def sample_accepted(batch, P, P_fallback, nchains, nspc, thin, seed, resync=1000, first=None):
    """Main run with P (or `first`, a callable returning the main run's
    samples, per-variant min ESS and max R-hat, e.g. graph_first); variants
    whose split-R-hat is >= 1.05 run again with P_fallback (callable giving
    the fallback maps of a sub-batch), keeping the better run (passing R-hat
    first, then the larger min-ESS).
    Returns (Y, ess_min, rhat_max, number of variants run again)."""
    import torch

    if first is not None:
        Y, e, r = first()
    else:
        Y = sample_rounded(batch, None, None, None, P, nchains, nspc, thin, seed, resync)
        e, r = variant_diagnostics(Y, nchains)
    bad = np.flatnonzero(r >= 1.05)
    if bad.size:
        sub = batch.take(bad)
        Y2 = sample_rounded(sub, None, None, None, P_fallback(sub, bad), nchains, nspc, thin, seed + 1, resync)
        e2, r2 = variant_diagnostics(Y2, nchains)
        for j, v in enumerate(bad):
            if (r2[j] < 1.05, e2[j]) > (r[v] < 1.05, e[v]):
                Y[v] = Y2[j]
                e[v], r[v] = e2[j], r2[j]
        del Y2
    return Y, e, r, int(bad.size)


def timed_steps(step, warmup, steps):
    """Mean wall seconds of `steps` calls of step() after `warmup` untimed
    ones, bracketed by device synchronisations; returns (seconds, last result)."""
    out = None
    for _ in range(warmup):
        out = step()
    _sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = step()
    _sync()
    return (time.perf_counter() - t0) / max(steps, 1), out


def compare(A0, b0, EA, Eb, F, lp, nchains=4, nspc=12500, thin=16, seed=7, resync=1000,
            methods=("plain", "dikin", "ellipsoid", "facet_ellipsoid", "facet", "learned"), warmup=1, steps=1, log=None,
            use_network=True):
    """Sampling methods on the same variants, chains and steps, each timed
    the same way (timed_steps) from the polytopes on the device to samples +
    diagnostics on the device, including its preconditioner:
    plain CHR; the reference's Dikin rounding (dikin_precondition, host
    FP64); the Dikin ellipsoid; the facet-aligned rounding completed from the
    Dikin ellipsoid (facet_ellipsoid); the facet-aligned rounding completed
    from the base polytope's facets with the exchange (facet: the method
    without the network); the learned-rounded pipeline (facet-aligned
    rounding completed from the learned map, or the base when use_network is
    False, R-hat acceptance). Returns per-method results."""
    import torch

    batch = VariantBatch(A0, b0, EA, Eb)  # the polytopes go to the device once, before any timing
    Fd = torch.as_tensor(F, dtype=torch.float32).cuda()

    def facet_fallback(sub, idx):
        return facet_maps(sub, base_rounding(sub, None, None, None))

    def plain():
        Y = sample_rounded(batch, None, None, None, None, nchains, nspc, thin, seed, resync)
        return variant_diagnostics(Y, nchains) + (0,)

    def rounded(make_P):
        def step():
            Y = sample_rounded(batch, None, None, None, make_P(), nchains, nspc, thin, seed, resync)
            return variant_diagnostics(Y, nchains) + (0,)
        return step

    def learned():
        if use_network:
            P, _ = learned_rounding(lp, Fd, batch)
            _, e, r, n_again = sample_accepted(batch, P, facet_fallback, nchains, nspc, thin, seed, resync)
        else:  # the network was rejected by validate_learned: the step runs its base, as one graph
            run = base_run(batch, nchains, nspc, thin, seed, resync)
            _, e, r, n_again = sample_accepted(batch, None, facet_fallback, nchains, nspc, thin, seed, resync,
                                               first=graph_first(run))
        return e, r, n_again

    steps_of = {
        "plain": plain,
        "dikin": rounded(lambda: precond.dikin_batched(A0, b0, EA, Eb)[0]),
        "ellipsoid": rounded(lambda: base_rounding(batch, None, None, None)),
        "facet_ellipsoid": rounded(lambda: facet_maps(batch, base_rounding(batch, None, None, None))),
        "facet": rounded(lambda: facet_maps(batch)),
        "learned": learned,
    }
    out = {}
    for name in methods:
        sec, (e, r, n_again) = timed_steps(steps_of[name], warmup, steps)
        ok = r < 1.05
        out[name] = dict(seconds=sec, ess_min=e, rhat_max=r, n_pass=int(ok.sum()), ess_sum=float(e.sum()),
                         ess_sum_pass=float(e[ok].sum()), metric=float(e[ok].sum()) / sec,
                         metric_all=float(e.sum()) / sec, rerun=n_again)
        if log:
            log(f"{name}: {sec * 1e3:.1f} ms per step, variants passing R-hat<1.05: {out[name]['n_pass']}/{len(e)}, "
                f"sum min-ESS {out[name]['ess_sum']:.1f} (passing {out[name]['ess_sum_pass']:.1f}), "
                f"median min-ESS {np.median(e):.1f}" + (f", {n_again} run again" if n_again else ""))
    return out
