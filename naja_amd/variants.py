"""Producers of conditioned variants for the batched sampler (host, FP64).

* `extra_from_bounds`: build_extra_from_bounds (cmd_prepare.cpp:111-156) --
  tightened flux bounds of a conditioned model as extra rows over the base
  model's rounded variables;
* `cube_center`: the start-point LP (feasible_start_lp.cpp:33-87), solved by
  the library's own interior-point method instead of Gurobi;
* `conditioned_variants`: many bound conditions of one base model at once ->
  the ragged extra rows, a start per variant that is feasible for its
  augmented polytope (feasible_start_files.cpp:41-90: keep the base start when
  it is feasible to 1e-9, else the cube centre), ready for `VariantBatch`.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _check, load_library

# cmd_prepare.cpp:129 (double eps = 1e-9)
BOUNDS_EPS = 1e-9
# feasible_start_files.cpp:78 / feasibility.h:12-35 tolerance
START_EPS = 1e-9


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def cube_center(A, b, x_hint=None, nthreads=0):
    """(x, r): centre and half-width of the largest axis-aligned cube in A x <= b."""
    A, b = _f64(A), _f64(b).ravel()
    m, n = A.shape
    if b.shape != (m,):
        raise ValueError("axis_aligned_cube_center: A.rows != b.size")
    xh = None if x_hint is None else _f64(x_hint).ravel()
    x = np.empty(n)
    r = ctypes.c_double(0.0)
    _check(load_library().naja_cube_center(_ptr(A), _ptr(b), m, n, _ptr(xh), _ptr(x), ctypes.byref(r),
                                           int(nthreads)))
    return x, float(r.value)


def extra_from_bounds(T, shift, lb_base, ub_base, lb_new, ub_new, eps=BOUNDS_EPS):
    """(extra_A k x d, extra_b k) for one condition of the base model."""
    T, shift = _f64(T), _f64(shift).ravel()
    n, d = T.shape
    vecs = [_f64(v).ravel() for v in (lb_base, ub_base, lb_new, ub_new)]
    for v in vecs:
        if v.shape != (n,):
            raise ValueError(f"bounds length mismatch: got {v.shape[0]} expected {n}")
    EA, Eb = np.empty((2 * n, d)), np.empty(2 * n)
    k = ctypes.c_int(0)
    _check(load_library().naja_extra_from_bounds(_ptr(T), _ptr(shift), n, d, *(_ptr(v) for v in vecs), float(eps),
                                                 _ptr(EA), _ptr(Eb), ctypes.byref(k)))
    return EA[:k.value].copy(), Eb[:k.value].copy()


def feasible(A, b, x, eps=START_EPS):
    return bool(np.all(_f64(A) @ _f64(x) <= _f64(b) + eps))


def conditioned_variants(A0, b0, x0, T, shift, lb_base, ub_base, LB, UB, eps=BOUNDS_EPS, nthreads=0):
    """Extra rows and feasible starts for conditions LB[v], UB[v] (nv x n).

    Returns dict(extra_A=[k_v x d], extra_b=[k_v], starts (nv x d),
    recentred (bool per variant: the cube-centre LP supplied the start)).
    """
    A0, b0, x0 = _f64(A0), _f64(b0).ravel(), _f64(x0).ravel()
    LB, UB = np.atleast_2d(_f64(LB)), np.atleast_2d(_f64(UB))
    if LB.shape != UB.shape:
        raise ValueError("l_bounds size != u_bounds size")
    EA, Eb, starts, rec = [], [], [], []
    for v in range(LB.shape[0]):
        ea, eb = extra_from_bounds(T, shift, lb_base, ub_base, LB[v], UB[v], eps)
        A = np.vstack([A0, ea]) if len(eb) else A0
        b = np.concatenate([b0, eb]) if len(eb) else b0
        if feasible(A, b, x0):
            starts.append(x0.copy())
            rec.append(False)
        else:
            x, _ = cube_center(A, b, x_hint=x0, nthreads=nthreads)
            if not feasible(A, b, x):
                raise RuntimeError(f"variant {v}: cube-centre start is not feasible")
            starts.append(x)
            rec.append(True)
        EA.append(ea)
        Eb.append(eb)
    return dict(extra_A=EA, extra_b=Eb, starts=np.array(starts), recentred=np.array(rec))
