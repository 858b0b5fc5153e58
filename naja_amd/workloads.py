"""Synthetic polytopes of the benchmark configurations (numpy, seeded PCG64).

There is no network for real genome-scale models, so the benchmark builds
the reduced polytopes from their published recipes:

* hypercube [-1, 1]^d: A = [I; -I], b = 1, start 0;
* genome-scale reduced polytope: T (n x d) with 8-40 N(0,1) nonzeros per row,
  densified by a random orthogonal rotation; shift ~ N(0,1); bounds
  lb = T y0 + shift - m_lo, ub = T y0 + shift + m_hi with y0 = 0 and margins
  log-uniform on [1e-3, 1e2]; A = [T; -T], b = [m_hi; m_lo] (the reduced form
  of lb <= T y + shift <= ub, as naja's rounding files describe it);
* conditioned variants of the hypercube: each adds k ~ U{0..kmax} rows drawn
  as +-a slab pairs of random unit vectors a with half-width w log-uniform on
  [w_lo, w_hi] around the start (a single half-space per row would not be a
  slab; each pair counts two rows).
"""
from __future__ import annotations

import numpy as np


def cube(d: int):
    A = np.vstack([np.eye(d), -np.eye(d)])
    b = np.ones(2 * d)
    return dict(A=A, b=b, y0=np.zeros(d))


def genome_scale(n: int = 2583, d: int = 512, seed: int = 0, nnz=(8, 40)):
    rng = np.random.default_rng(seed)
    T0 = np.zeros((n, d))
    counts = rng.integers(nnz[0], nnz[1] + 1, size=n)
    for r in range(n):
        cols = rng.choice(d, size=int(counts[r]), replace=False)
        T0[r, cols] = rng.standard_normal(int(counts[r]))
    Q, R = np.linalg.qr(rng.standard_normal((d, d)))
    Q = Q * np.sign(np.diag(R))
    T = T0 @ Q
    shift = rng.standard_normal(n)
    m_lo = 10.0 ** rng.uniform(-3, 2, n)
    m_hi = 10.0 ** rng.uniform(-3, 2, n)
    A = np.vstack([T, -T])
    b = np.concatenate([m_hi, m_lo])
    return dict(A=A, b=b, y0=np.zeros(d), T=T, shift=shift, lb=shift - m_lo, ub=shift + m_hi)


def conditioned_variants(d: int, n_variants: int, kmax: int, w_lo: float, w_hi: float, seed: int,
                         pool: np.ndarray | None = None):
    """Extra rows per variant: k/2 slabs |a . y| <= w (k even, 0 <= k <= kmax).

    With `pool` (P x d unit rows) the slab directions come from that shared
    pool, so each variant is described by P half-widths (0 = inactive): the
    feature vector of the learned preconditioner.
    """
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_variants):
        k = int(rng.integers(0, kmax + 1)) // 2 * 2
        nslab = k // 2
        if pool is None:
            a = rng.standard_normal((nslab, d))
            a /= np.linalg.norm(a, axis=1, keepdims=True)
            idx = None
        else:
            idx = rng.choice(pool.shape[0], size=nslab, replace=False)
            a = pool[idx]
        w = 10.0 ** rng.uniform(np.log10(w_lo), np.log10(w_hi), nslab)
        E = np.vstack([a, -a]) if nslab else np.zeros((0, d))
        f = np.concatenate([w, w]) if nslab else np.zeros(0)
        out.append(dict(E=E, f=f, slab_idx=idx, w=w))
    return out
