"""ORACLE — test infrastructure only: pure-Python restatement of the
affine-hull reduction of run_sampling_job (job_sampling.cu:196-235), used by
tests/ as the checker of naja_amd.hull.

The reference calls Eigen::FullPivLU (threshold 1e-10) and its kernel();
Eigen is not vendored (version unpinned), so the restatement follows the
published full-pivoting LU: pivot = largest |entry| of the trailing block,
ties to the first in column-major order, stop at an exactly zero block; rank
counts pivots with |u_kk| > threshold * max |pivot|; the kernel basis puts 1
at each free column (in permuted order) and back-solves the pivot columns.
Its independent pins (in tests/test_hull.py) are scipy's SVD null space
(span and dimension), closed forms on small systems, and A_eq B = 0.
"""
from __future__ import annotations

import numpy as np


def full_piv_lu(Aeq):
    """(lu, q, nonzero, maxpivot): row-major LU with L strictly below, U on/above."""
    lu = np.array(Aeq, dtype=np.float64, copy=True)
    r, n = lu.shape
    q = list(range(n))
    size = min(r, n)
    nonzero, maxpivot = size, 0.0
    for k in range(size):
        best, bi, bj = -1.0, k, k
        for j in range(k, n):  # column-major scan, strict '>' keeps the first maximum
            for i in range(k, r):
                v = abs(lu[i, j])
                if v > best:
                    best, bi, bj = v, i, j
        if best == 0.0:
            nonzero = k
            break
        maxpivot = max(maxpivot, best)
        lu[[k, bi]] = lu[[bi, k]]
        lu[:, [k, bj]] = lu[:, [bj, k]]
        q[k], q[bj] = q[bj], q[k]
        for i in range(k + 1, r):
            lu[i, k] = lu[i, k] / lu[k, k]
        for i in range(k + 1, r):
            for j in range(k + 1, n):
                lu[i, j] = lu[i, j] - lu[i, k] * lu[k, j]
    return lu, q, nonzero, maxpivot


def kernel_basis(Aeq, threshold=1e-10):
    Aeq = np.atleast_2d(np.asarray(Aeq, dtype=np.float64))
    r, n = Aeq.shape
    if r == 0:
        return np.eye(n)
    lu, q, nonzero, maxpivot = full_piv_lu(Aeq)
    piv = [k for k in range(nonzero) if abs(lu[k, k]) > maxpivot * threshold]
    free = [j for j in range(n) if j not in piv]
    B = np.zeros((n, len(free)))
    for f, col in enumerate(free):
        z = [-lu[p, col] if col >= p else 0.0 for p in piv]
        for t in range(len(piv) - 1, -1, -1):
            z[t] = z[t] / lu[piv[t], piv[t]]
            for u in range(t):
                z[u] = z[u] - lu[piv[u], piv[t]] * z[t]
        for t, p in enumerate(piv):
            B[q[p], f] = z[t]
        B[q[col], f] = 1.0
    return B


def tight_rows(A, b, x0, tol):
    """job_sampling.cu:197-204: rows with b_i - a_i.x0 <= tol (sequential dot)."""
    A, b, x0 = (np.asarray(v, dtype=np.float64) for v in (A, b, x0))
    rows = []
    for i in range(A.shape[0]):
        ax = 0.0
        for j in range(A.shape[1]):
            ax += A[i, j] * x0[j]
        if b[i] - ax <= tol:
            rows.append(i)
    return rows


def affine_hull(A, b, x0, tol, threshold=1e-10):
    A = np.asarray(A, dtype=np.float64)
    rows = tight_rows(A, b, x0, tol)
    B = kernel_basis(A[rows] if rows else np.zeros((0, A.shape[1])), threshold)
    return rows, B
