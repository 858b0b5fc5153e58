"""Conditioned variants produced from bound conditions (build_extra_from_bounds,
cmd_prepare.cpp:111-156, plus cube-centre starts, feasible_start_files.cpp:41-90)
sampled in one batched launch: every variant equals the oracle chain on its
augmented polytope from its start, and no sample leaves its conditioned bounds."""
import numpy as np
import pytest

import oracle
from tests.test_variants import gem_like

pytestmark = pytest.mark.gpu


def test_bound_conditions_in_one_batched_launch(gpu):
    import naja_amd
    from naja_amd import variants as V

    t = gpu
    rng = np.random.default_rng(21)
    d, n, nv = 12, 30, 10
    T, shift, m_lo, m_hi = gem_like(rng, d, n)
    A0, b0 = np.vstack([T, -T]), np.r_[m_hi, m_lo]
    lb, ub = shift - m_lo, shift + m_hi
    LB, UB = np.tile(lb, (nv, 1)), np.tile(ub, (nv, 1))
    for v in range(nv):
        if v % 2:  # lower bound beyond the base start's flux -> cube-centre start
            u = rng.standard_normal(d)
            Au = A0 @ u
            y1 = 0.5 * (b0[Au > 0] / Au[Au > 0]).min() * u
            j0 = int(np.argmax(T @ y1))
            LB[v, j0] = shift[j0] + 0.5 * (T[j0] @ y1)
        else:
            j = rng.choice(n, 5, replace=False)
            UB[v, j] = shift[j] + rng.uniform(0.1, 0.9, 5) * m_hi[j]
    cv = V.conditioned_variants(A0, b0, np.zeros(d), T, shift, lb, ub, LB, UB)
    assert cv["recentred"].sum() == nv // 2
    nch, nspc, thin = 4, 400, 2
    X0 = np.repeat(cv["starts"][:, None, :], nch, axis=1).astype(np.float32)
    batch = naja_amd.VariantBatch(t.from_numpy(A0.astype(np.float32)), t.from_numpy(b0), cv["extra_A"],
                                  cv["extra_b"])
    print(f"variants {nv}, extra rows {batch.ks}", flush=True)
    out = naja_amd.coordinate_hit_and_run_batched(batch, None, None, None, t.from_numpy(X0), nspc, thin, nch,
                                                  seed=77, pair_prob=0.3)
    t.cuda.synchronize()
    out = out.cpu().numpy().astype(np.float64)
    for v in range(nv):
        Av = np.vstack([A0, cv["extra_A"][v]])
        bv = np.r_[b0, cv["extra_b"][v]]
        ref = oracle.chr_run_f32(Av.astype(np.float32), bv, X0[v], nspc, thin, seed=77, variant=v, pair_prob=0.3)
        err = np.abs(out[v] - ref["samples"]).max() / max(1.0, np.abs(ref["samples"]).max())
        print(f"variant {v}: k={len(bv) - len(b0)} max rel {err:.2e}", flush=True)
        assert err <= 1e-5
        flux = out[v] @ T.T + shift
        tol = 1e-5 * (1 + np.abs(flux).max())
        assert (flux <= UB[v] + tol).all() and (flux >= LB[v] - tol).all()
