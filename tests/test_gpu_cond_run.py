"""The conditioned run as one persistent hipGraph (naja_cond_run_*, through the
C ABI): every replay equals the eager pipeline it captures — the facet-aligned
rounding, the rounded batched CHR with its backmap and the batched
diagnostics, each checked against the oracle by its own tests — bit for bit,
with eager work and synchronisations between the replays; the timing event
nodes report the CHR launch; a start outside a variant raises."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def conditions(d, ks, seed, w_lo=1e-3, w_hi=1e-1):
    rng = np.random.default_rng(seed)
    EA, Eb = [], []
    for k in ks:
        a = rng.standard_normal((k // 2, d))
        a /= np.linalg.norm(a, axis=1, keepdims=True)
        w = 10.0 ** rng.uniform(np.log10(w_lo), np.log10(w_hi), k // 2)
        EA.append(np.vstack([a, -a]) if k else np.zeros((0, d)))
        Eb.append(np.concatenate([w, w]) if k else np.zeros(0))
    return EA, Eb


def eager_step(naja_amd, t, batch, nspc, thin, nch, seed, resync):
    P = naja_amd.facet_rounding(batch, None, None, None, None, 0.02,
                                completion=naja_amd.FACET_COMPLETE_BASE_FACETS, exchange_tol=1.02)
    nv, d = batch.n_variants, batch.d
    X0 = t.zeros((nv, nch, d), dtype=t.float32, device="cuda")
    c = t.zeros((nv, d), dtype=t.float32, device="cuda")
    plan = naja_amd.ChrPlan(batch, None, nspc, thin, nch, P=P, c=c, seed=seed, resync_interval=resync)
    Y = plan.launch(X0, plan.samples_like())
    ess, rhat, emin, rmax = naja_amd.diagnostics_batched(Y, nch)
    t.cuda.synchronize()
    out = [x.cpu().numpy() for x in (P, Y, ess, rhat, emin, rmax)]
    plan.close()
    return out


@pytest.mark.parametrize("d,ks,nch", [(40, [0, 6, 30, 52, 80, 120, 160, 200, 2, 90, 140, 60], 4),
                                      (100, [0, 40, 120, 200, 260, 300, 150, 88], 4),
                                      # d % 4 != 0: no series copy from the backmap, the diagnostics transpose
                                      (30, [0, 10, 44, 70, 120], 4),
                                      # 3 chains: the backmap's series copy by its general path
                                      (40, [0, 20, 64, 100, 150], 3)])
def test_graph_replays_equal_the_eager_pipeline(gpu, d, ks, nch):
    import naja_amd

    t = gpu
    A0, b0 = np.vstack([np.eye(d), -np.eye(d)]), np.ones(2 * d)
    EA, Eb = conditions(d, ks, seed=d)
    batch = naja_amd.VariantBatch(A0, b0, EA, Eb)
    nspc, thin, seed, resync = 700, 4, 19, 300
    ref = eager_step(naja_amd, t, batch, nspc, thin, nch, seed, resync)
    names = ("P", "samples", "ess", "rhat", "ess_min", "rhat_max")
    print(f"[cond-run d{d} c{nch}] {len(ks)} variants, {nch} x {nspc} x thin {thin}: creating", flush=True)
    run = naja_amd.ConditionedRun(batch, nspc, thin, nch, seed=seed, resync_interval=resync, timed=True)
    side = t.randn((512, 512), device="cuda")
    for rep in range(4):
        if rep > 0:
            # eager work on torch's stream and a synchronisation between replays
            side = (side @ side).clamp_(-1, 1)
            t.cuda.synchronize()
            Y, emin, rmax = run.step()
        got = [run.P, run.samples, run.ess, run.rhat, run.ess_min, run.rhat_max]
        got = [x.cpu().numpy() for x in got]
        for name, a, b in zip(names, got, ref):
            assert np.array_equal(a, b), f"replay {rep}: {name} differs from the eager pipeline"
        if rep == 0:  # the step creation ran eagerly, before the capture: no timing yet
            print(f"[cond-run d{d}] created: the eager step equals the separate eager pipeline", flush=True)
            continue
        ms = run.sampling_ms()
        print(f"[cond-run d{d}] replay {rep}: identical to the eager step; CHR launch {ms:.3f} ms", flush=True)
        assert ms > 0.0
    # every sample inside its variant
    Yh = ref[1].astype(np.float64)
    for v in range(len(ks)):
        A = np.vstack([A0, EA[v]])
        b = np.concatenate([b0, Eb[v]])
        assert (Yh[v] @ A.T - b).max() <= 1e-5
    run.close()


def test_bounds_rewritten_in_place_are_sampled(gpu):
    """New half-widths for the same slabs, written into the batch's device
    buffer between steps: the next replay equals the eager pipeline built
    from scratch on the new conditions (the graph packs the rows again)."""
    import naja_amd

    t = gpu
    d = 40
    ks = [0, 16, 40, 90, 150]
    A0, b0 = np.vstack([np.eye(d), -np.eye(d)]), np.ones(2 * d)
    EA, Eb = conditions(d, ks, seed=5)
    batch = naja_amd.VariantBatch(A0, b0, EA, Eb)
    nspc, thin, nch, seed, resync = 500, 4, 4, 23, 256
    run = naja_amd.ConditionedRun(batch, nspc, thin, nch, seed=seed, resync_interval=resync)
    rng = np.random.default_rng(8)
    Eb2 = []
    for e in Eb:
        w = 10.0 ** rng.uniform(-2, 0, len(e) // 2)
        Eb2.append(np.concatenate([w, w]) if len(e) else np.zeros(0))
    batch.be.copy_(t.from_numpy(np.concatenate(Eb2)))
    t.cuda.synchronize()
    run.step()
    got = [x.cpu().numpy() for x in (run.P, run.samples, run.ess, run.rhat, run.ess_min, run.rhat_max)]
    ref = eager_step(naja_amd, t, naja_amd.VariantBatch(A0, b0, EA, Eb2), nspc, thin, nch, seed, resync)
    for name, a, b in zip(("P", "samples", "ess", "rhat", "ess_min", "rhat_max"), got, ref):
        assert np.array_equal(a, b), f"{name} after the in-place bounds differs from the eager pipeline"
    print("[cond-run] rewritten bounds: the replay equals the eager pipeline on the new conditions", flush=True)
    run.close()


def test_cond_run_rejects_a_start_outside_a_variant(gpu):
    import naja_amd

    d = 8
    A0, b0 = np.vstack([np.eye(d), -np.eye(d)]), np.ones(2 * d)
    a = np.zeros(d)
    a[0] = 1.0
    EA = [np.zeros((0, d)), np.vstack([a, -a])]
    Eb = [np.zeros(0), np.array([-0.1, 0.5])]  # 0.1 <= -y_0 ... the origin violates y_0 <= -0.1
    batch = naja_amd.VariantBatch(A0, b0, EA, Eb)
    with pytest.raises(naja_amd.NajaInvalidArgument):
        naja_amd.ConditionedRun(batch, 64, 2, 4, seed=1)
