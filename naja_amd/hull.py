"""Affine-hull reduction (job_sampling.cu:196-235) over libnaja_amd.so.

At the start point x0 the rows of A x <= b with slack <= tol are treated as
equalities. When their kernel has dimension 0 < k < n the sampler runs on
z in R^k with A' = A B, b' = b - A x0, z0 = 0, and x = x0 + B z; a transform
(T, shift) to the model's variables composes to (T B, T x0 + shift)
(job_sampling.cu:298-302). Without a transform the samples are mapped back to
x by (B, x0) (job_sampling.cu:343-348).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _check, load_library

# job_sampling.cu:211 (lu.setThreshold(1e-10))
PIVOT_THRESHOLD = 1e-10


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def kernel_basis(Aeq, threshold=PIVOT_THRESHOLD):
    """Kernel basis (n x k) of Aeq (r x n) from a full-pivot LU."""
    Aeq = _f64(np.atleast_2d(Aeq))
    r, n = Aeq.shape
    buf = np.empty(n * n)
    k = ctypes.c_int(0)
    _check(load_library().naja_kernel_basis(_ptr(Aeq) if r else None, r, n, float(threshold), _ptr(buf),
                                            ctypes.byref(k)))
    return buf[:n * k.value].reshape(n, k.value).copy()


def affine_hull(A, b, x0, tol, threshold=PIVOT_THRESHOLD):
    """Reduce {x : A x <= b} to its affine hull through x0.

    Returns a dict with `enabled` (0 < k < n), `basis` (n x k), `shift` (x0),
    `n_tight`, and the reduced `A`, `b`, `x0` (the inputs when not enabled).
    """
    A, b, x0 = _f64(A), _f64(b).ravel(), _f64(x0).ravel()
    m, n = A.shape
    if b.shape != (m,) or x0.shape != (n,):
        raise ValueError("affine_hull: b must have m entries and x0 n entries")
    buf = np.empty(n * n)
    k, nt = ctypes.c_int(0), ctypes.c_int(0)
    _check(load_library().naja_affine_hull(_ptr(A), _ptr(b), _ptr(x0), m, n, float(tol), float(threshold),
                                           _ptr(buf), ctypes.byref(k), ctypes.byref(nt)))
    B = buf[:n * k.value].reshape(n, k.value).copy()
    enabled = nt.value > 0 and 0 < k.value < n
    if not enabled:
        return dict(enabled=False, basis=B, shift=x0, n_tight=nt.value, A=A, b=b, x0=x0)
    return dict(enabled=True, basis=B, shift=x0, n_tight=nt.value, A=A @ B, b=b - A @ x0, x0=np.zeros(k.value))


def compose_transform(T, shift, hull):
    """(T B, T x0 + shift) for a transform T (n_orig x n) of the unreduced variables."""
    T, shift = _f64(T), _f64(shift).ravel()
    if not hull["enabled"]:
        return T, shift
    return T @ hull["basis"], T @ hull["shift"] + shift
