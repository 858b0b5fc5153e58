"""ORACLE — test infrastructure only: FP64 numpy restatements of the rounding
preconditioners, used by tests/ as the checker of naja_amd.precond.

Eigenproblems and factorisations go through numpy's LAPACK routines
(numpy.linalg.eigh / cholesky), independent of the library's own solver.
"""
from __future__ import annotations

import numpy as np


def chord_along(A, b, x_c, v):
    """dikin_precondition.cpp:13-32."""
    slack = b - A @ x_c
    Av = A @ v
    tmin, tmax = -np.inf, np.inf
    for i in range(A.shape[0]):
        if abs(Av[i]) < 1e-15:
            continue
        t = slack[i] / Av[i]
        if Av[i] > 0:
            tmax = min(tmax, t)
        else:
            tmin = max(tmin, t)
    return 0.0 if tmax <= tmin else tmax - tmin


def dikin(A, b, x_c=None, k_directions=64, tight_factor=100.0, target_chord=0.0, alpha_max=1000.0):
    """dikin_precondition (dikin_precondition.cpp:36-183)."""
    A = np.asarray(A, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    m, d = A.shape
    if m == 0 or d == 0:
        raise ValueError("dikin_precondition: empty polytope")
    x_c = np.zeros(d) if x_c is None else np.asarray(x_c, dtype=np.float64)
    slack = b - A @ x_c
    norms = np.linalg.norm(A, axis=1)
    deltas = np.full(m, np.inf)
    N = np.zeros_like(A)
    ok = norms >= 1e-15
    deltas[ok] = slack[ok] / norms[ok]
    N[ok] = A[ok] / norms[ok, None]
    pos = (deltas > 0) & np.isfinite(deltas)
    min_delta = deltas[pos].min() if pos.any() else np.inf
    tight = np.where((deltas > 0) & (deltas <= tight_factor * min_delta))[0]
    I = np.eye(d)
    if tight.size == 0:
        return dict(P=I, P_inv=I.copy(), A_new=A.copy(), b_new=slack, n_corrected=0)
    H = (N[tight].T * (1.0 / deltas[tight] ** 2)) @ N[tight]
    ev, V = np.linalg.eigh(H)
    k = min(k_directions, d)
    med = ev[d // 2]
    kk = 0
    for j in range(d - 1, -1, -1):
        if kk < k and ev[j] > 10.0 * max(med, 1e-10):
            kk += 1
        else:
            break
    kk = max(kk, 1)
    Vk = V[:, ::-1][:, :kk]
    if target_chord <= 0:
        cs = sorted(c for c in (chord_along(A, b, x_c, I[j]) for j in range(min(d, 50))) if c > 0)
        target_chord = cs[len(cs) // 2] if cs else 1.0
    al = np.empty(kk)
    for j in range(kk):
        c = chord_along(A, b, x_c, Vk[:, j])
        al[j] = min(alpha_max, 10.0) if c < 1e-15 else min(max(target_chord / c, 1.0), alpha_max)
    P = I + (Vk * (al - 1.0)) @ Vk.T
    Pinv = I + (Vk * (1.0 / al - 1.0)) @ Vk.T
    return dict(P=P, P_inv=Pinv, A_new=A @ Pinv, b_new=slack, n_corrected=kk)


def ellipsoid(A, b, x_c=None):
    """P with P P^T = H^{-1}, H = A^T diag(1/s^2) A (upper-triangular L^{-T})."""
    A = np.asarray(A, dtype=np.float64)
    d = A.shape[1]
    x_c = np.zeros(d) if x_c is None else np.asarray(x_c, dtype=np.float64)
    s = np.asarray(b, dtype=np.float64) - A @ x_c
    H = (A.T * (1.0 / s ** 2)) @ A
    L = np.linalg.cholesky(H)
    return np.linalg.inv(L).T


def lowrank(V, expand=False):
    """(I + V V^T)^{-1/2} (or ^{+1/2}) through the full d x d eigenproblem."""
    V = np.asarray(V, dtype=np.float64)
    M = np.eye(V.shape[0]) + V @ V.T
    ev, Q = np.linalg.eigh(M)
    return (Q * (np.sqrt(ev) if expand else 1.0 / np.sqrt(ev))) @ Q.T


COMPLETE_ROUNDING, COMPLETE_LOWER, COMPLETE_BASE_FACETS = 0, 1, 2


def facet(A0, b0, E, f, P0, theta, x_c=None, include_base=False, lower=False, completion=None, exchange_tol=0.0):
    """Facet-aligned rounding of [A0; E] y <= [b0; f] (the rule of
    naja_precond_facet_device): candidate dual rows u_r = a_r / s_r at x_c of
    the extra rows (and the base rows with include_base), chosen greedily by
    modified Gram-Schmidt residual norm (FP32 residuals) while the best keeps
    at least theta of its own norm, then the completion candidates: the rows
    of P0^-1 (FP64; COMPLETE_LOWER or lower: of chol(P0 P0^T)^-1), or with
    COMPLETE_BASE_FACETS the base rows' dual rows, whose chosen residuals are
    made orthonormal (FP64 classical Gram-Schmidt, twice, in the order of
    choice) and kept as rows of N. exchange_tol > 0: when the facets fill the
    basis, C = U N^-1 in FP32 (summed in column order) and, while the largest
    |C_ik| (first in row-major order) exceeds exchange_tol, basis row k
    becomes candidate i, C -= (C[:, k] / C_ik) (C_i - e_k) in FP32. Returns
    (P = N^-1 in FP64, the chosen candidate indices, N). Ties go to the
    smaller index."""
    if completion is None:
        completion = COMPLETE_LOWER if lower else COMPLETE_ROUNDING
    A0 = np.asarray(A0, dtype=np.float64)
    d = A0.shape[1]
    m0 = A0.shape[0]
    E = np.asarray(E, dtype=np.float64).reshape(-1, d)
    A = np.vstack([A0, E])
    b = np.concatenate([np.asarray(b0, np.float64), np.asarray(f, np.float64).reshape(-1)])
    x_c = np.zeros(d) if x_c is None else np.asarray(x_c, dtype=np.float64)
    rows = np.arange(A.shape[0]) if include_base else np.arange(m0, A.shape[0])
    s = b[rows] - A[rows] @ x_c
    if not (s > 0).all():
        raise ValueError("facet rounding: the centre is not interior")
    U = A[rows] / s[:, None]
    if completion == COMPLETE_BASE_FACETS:
        if include_base:
            raise ValueError("facet rounding: base rows cannot be both facets and the completion")
        s0 = b[:m0] - A0 @ x_c
        if not (s0 > 0).all():
            raise ValueError("facet rounding: the centre is not interior")
        comp = A0 / s0[:, None]
    else:
        P0 = np.asarray(P0, dtype=np.float64)
        if completion == COMPLETE_LOWER:  # the lower Cholesky factor of the same ellipsoid (LAPACK)
            P0 = np.linalg.cholesky(P0 @ P0.T)
        comp = np.linalg.inv(P0)
    cand = np.vstack([U, comp])
    nf = U.shape[0]
    R = cand.astype(np.float32)
    n0 = (R * R).sum(1)
    taken = np.zeros(len(cand), bool)
    th2 = np.float32(theta) * np.float32(theta)
    tiny = np.float32(1e-12)
    sel = []
    n2 = n0.copy()
    for _ in range(d):
        elig_f = (~taken[:nf]) & (n2[:nf] >= th2 * n0[:nf]) & (n2[:nf] > tiny * n0[:nf])
        if elig_f.any():
            p = int(np.argmax(np.where(elig_f, n2[:nf], -1.0)))
        else:
            elig_b = ~taken[nf:]
            if completion == COMPLETE_BASE_FACETS:
                elig_b &= n2[nf:] > tiny * n0[nf:]
            if not elig_b.any():
                raise RuntimeError("facet rounding: no independent completion")
            p = nf + int(np.argmax(np.where(elig_b, n2[nf:], -1.0)))
        taken[p] = True
        sel.append(p)
        q = R[p] / np.sqrt(n2[p])
        c = R @ q
        R = R - np.outer(c, q).astype(np.float32)
        n2 = (R * R).sum(1)
    sel = np.array(sel)
    if exchange_tol > 0 and (sel < nf).all():
        Ni32 = np.linalg.inv(U[sel]).astype(np.float32)
        U32 = U.astype(np.float32)
        C = np.zeros((nf, d), np.float32)
        for j in range(d):
            C = C + U32[:, j:j + 1] * Ni32[j][None, :]
        for _ in range(4 * d):
            flat = int(np.argmax(np.abs(C)))
            i, k = divmod(flat, d)
            if not abs(C[i, k]) > np.float32(exchange_tol):
                break
            fcol = C[:, k] / C[i, k]
            row = C[i].copy()
            row[k] = row[k] - np.float32(1.0)
            C = C - np.outer(fcol, row).astype(np.float32)
            sel[k] = i
    N = cand[sel].copy()
    if completion == COMPLETE_BASE_FACETS and not (sel < nf).all():
        Q = np.zeros((d, d))
        for t in range(d):
            w = cand[sel[t]].copy()
            for _ in range(2):
                w = w - (Q[:t] @ w) @ Q[:t]
            Q[t] = w / np.linalg.norm(w)
            if sel[t] >= nf:
                N[t] = Q[t]
    return np.linalg.inv(N), sel, N


def facet_ordered(A, b, theta, x_c=None):
    """Facet-aligned rounding of one polytope by the rule of
    naja_precond_facet_large: dual rows u_r = a_r / s_r in order of tightness
    ||u_r|| (stable), each accepted while its residual against the accepted
    ones (FP64 modified Gram-Schmidt, twice) keeps at least theta of its norm,
    until d are accepted; returns (P = N^-1, accepted row indices)."""
    A = np.asarray(A, dtype=np.float64)
    m, d = A.shape
    x_c = np.zeros(d) if x_c is None else np.asarray(x_c, dtype=np.float64)
    s = np.asarray(b, dtype=np.float64) - A @ x_c
    if not (s > 0).all():
        raise ValueError("facet rounding: the centre is not interior")
    U = A / s[:, None]
    nrm = np.linalg.norm(U, axis=1)
    order = np.argsort(-nrm, kind="stable")
    Q = np.zeros((d, d))
    k = 0
    acc = []
    for r in order:
        if k == d:
            break
        u = U[r].copy()
        for _ in range(2):
            u -= Q[:, :k] @ (Q[:, :k].T @ u)
        nu = np.linalg.norm(u)
        if nu > 0 and nu >= theta * nrm[r]:
            Q[:, k] = u / nu
            k += 1
            acc.append(r)
    if k < d:
        raise ValueError("facet rounding: fewer than d independent facets")
    N = U[acc]
    return np.linalg.inv(N), np.array(acc)


def analytic_center(A, b, x0=None, max_iter=200, tol=1e-12):
    """Minimiser of -sum log(b - A x) by damped Newton (numpy FP64): step
    1 / (1 + lambda) while the Newton decrement lambda >= 1/4, full steps
    after; returns (x, steps)."""
    A = np.asarray(A, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    x = np.zeros(A.shape[1]) if x0 is None else np.asarray(x0, dtype=np.float64).copy()
    for it in range(max_iter):
        s = b - A @ x
        if not (s > 0).all():
            raise ValueError("analytic centre: not interior")
        g = A.T @ (1.0 / s)
        H = (A.T / s ** 2) @ A
        dx = -np.linalg.solve(H, g)
        lam2 = max(-g @ dx, 0.0)
        x = x + (1.0 if lam2 < 0.0625 else 1.0 / (1.0 + np.sqrt(lam2))) * dx
        if lam2 < tol:
            return x, it + 1
    raise ValueError("analytic centre: no convergence")
