"""Host FP64 rounding preconditioners (ctypes over libnaja_amd.so).

* `dikin` / `dikin_batched`: dikin_precondition (dikin_precondition.cpp:36-183)
  with the library's own symmetric eigensolver;
* `ellipsoid_batched`: whitening by the full Dikin ellipsoid at the centre
  (H = A^T diag(1/s^2) A = L L^T, P = L^{-T});
* `lowrank`: P = (I + V V^T)^{-1/2}, or the expansion (I + V V^T)^{1/2}, from
  the learned preconditioner's factors.

Every P returned maps the sampler's variable to the model's: y = x_c + P z,
which is the form `coordinate_hit_and_run_batched(..., P=, c=)` takes.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _check, load_library


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def symmetric_eigen(a):
    """(eigenvalues ascending, eigenvectors as columns) of a symmetric matrix."""
    a = _f64(a)
    n = a.shape[0]
    w = np.empty(n)
    V = np.empty((n, n))
    _check(load_library().naja_symmetric_eigen(_ptr(a), n, _ptr(w), _ptr(V)))
    return w, V


def dikin(A, b, x_c=None, k_directions=64, tight_factor=100.0, target_chord=0.0, alpha_max=1000.0):
    """DikinResult of dikin_precondition (dikin_precondition.h:7-13) plus P_inv."""
    A, b = _f64(A), _f64(b)
    m, d = A.shape
    xc = None if x_c is None else _f64(x_c)
    P, Pinv, An, bn = np.empty((d, d)), np.empty((d, d)), np.empty((m, d)), np.empty(m)
    k = ctypes.c_int(0)
    _check(load_library().naja_precond_dikin(_ptr(A), _ptr(b), _ptr(xc), m, d, int(k_directions),
                                             float(tight_factor), float(target_chord), float(alpha_max),
                                             _ptr(P), _ptr(Pinv), _ptr(An), _ptr(bn), ctypes.byref(k)))
    return dict(P=P, P_inv=Pinv, A_new=An, b_new=bn, x0_new=np.zeros(d), n_corrected=int(k.value))


def _ragged(A0, b0, extra_A, extra_b):
    A0, b0 = _f64(A0), _f64(b0)
    ks = np.array([np.asarray(e).shape[0] for e in extra_A], dtype=np.int32)
    d = A0.shape[1]
    Ae = _f64(np.vstack([np.asarray(e).reshape(-1, d) for e in extra_A])) if ks.sum() else None
    be = _f64(np.concatenate([np.asarray(f).reshape(-1) for f in extra_b])) if ks.sum() else None
    return A0, b0, ks, Ae, be


def dikin_batched(A0, b0, extra_A, extra_b, x_c=None, k_directions=64, tight_factor=100.0, target_chord=0.0,
                  alpha_max=1000.0, nthreads=0):
    """P_inv (n_variants, d, d) and n_corrected of every variant [A0; E_v]."""
    A0, b0, ks, Ae, be = _ragged(A0, b0, extra_A, extra_b)
    nv, (m0, d) = len(ks), A0.shape
    xc = None if x_c is None else _f64(x_c)
    Pinv = np.empty((nv, d, d))
    nc = np.empty(nv, dtype=np.int32)
    _check(load_library().naja_precond_dikin_batched(
        _ptr(A0), _ptr(b0), m0, d, nv, _ptr(ks), _ptr(Ae), _ptr(be), _ptr(xc), int(k_directions),
        float(tight_factor), float(target_chord), float(alpha_max), _ptr(Pinv), _ptr(nc), int(nthreads)))
    return Pinv, nc


def ellipsoid_batched(A0, b0, extra_A, extra_b, x_c=None, nthreads=0):
    """P (n_variants, d, d), upper triangular, P P^T = H^{-1}."""
    A0, b0, ks, Ae, be = _ragged(A0, b0, extra_A, extra_b)
    nv, (m0, d) = len(ks), A0.shape
    xc = None if x_c is None else _f64(x_c)
    P = np.empty((nv, d, d))
    _check(load_library().naja_precond_ellipsoid_batched(_ptr(A0), _ptr(b0), m0, d, nv, _ptr(ks), _ptr(Ae),
                                                         _ptr(be), _ptr(xc), _ptr(P), int(nthreads)))
    return P


def lowrank(V, expand=False, nthreads=0):
    """P_v = (I + V_v V_v^T)^{-1/2} (or ^{+1/2} with expand) for V (n_variants, d, rank)."""
    V = _f64(V)
    nv, d, r = V.shape
    P = np.empty((nv, d, d))
    _check(load_library().naja_precond_lowrank(_ptr(V), nv, d, r, int(bool(expand)), _ptr(P), int(nthreads)))
    return P
