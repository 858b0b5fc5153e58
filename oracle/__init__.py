"""ORACLE — test infrastructure only.

ctypes bindings of the CPU restatement in this directory (liboracle.so built
from chr_oracle.c, diag_oracle.c, backmap_oracle.c by oracle/Makefile) and of
the reference-built cross-check oracle/_ref/libref_kiss.so. Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg use this module, and
only as the checker; the product library never imports or links it.

Each function names the reference code it restates in its C source.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_KISS_PATH = os.path.join(_HERE, "_ref", "libref_kiss.so")

_lib = None
_kiss = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, u32, u64, f64 = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_uint32,
                                       ctypes.c_uint64, ctypes.c_double)
        L.orc_chr_run_f32.restype = f64
        L.orc_chr_run_f32.argtypes = [vp, vp, i32, i32, vp, i32, i32, i32, u64, u32, f64, i32,
                                      vp, vp, vp, vp, i32, f64, i32, i32, i32, vp, vp, vp, vp, i64, i32]
        L.orc_chr_run_f64.restype = f64
        L.orc_chr_run_f64.argtypes = [vp, vp, i32, i32, vp, i32, i32, i32, u64, u32, f64, i32,
                                      vp, vp, vp, vp, i32, f64, i32, i32, i32, vp, i32]
        L.orc_philox_block.argtypes = [vp, u32, u32, vp]
        L.orc_move_stream.argtypes = [i32, u64, u32, u32, f64, i32, vp, vp, vp, vp, i32, f64, i32, i64, vp]
        L.orc_autocov.argtypes = [vp, i32, vp, i32]
        L.orc_center.argtypes = [vp, i32, vp]
        L.orc_ess_from_acov.restype = ctypes.c_float
        L.orc_ess_from_acov.argtypes = [vp, i32]
        L.orc_ess_one_chain.restype = ctypes.c_float
        L.orc_ess_one_chain.argtypes = [vp, i32, i32]
        L.orc_ess.restype = i32
        L.orc_ess.argtypes = [vp, i32, i32, i32, i32, vp, vp, i32]
        L.orc_split_rhat.restype = i32
        L.orc_split_rhat.argtypes = [vp, i32, i32, i32, i32, vp, vp]
        L.orc_percentile.restype = ctypes.c_float
        L.orc_percentile.argtypes = [vp, i32, ctypes.c_float]
        L.orc_transform.argtypes = [vp, vp, i32, i32, vp, vp, vp, vp]
        L.orc_backmap.restype = i32
        L.orc_backmap.argtypes = [vp, vp, vp, vp, i32, i32, i64]
        _lib = L
    return _lib


def kiss():
    """The reference's kissfft autocovariance (None if oracle/_ref was not built)."""
    global _kiss
    if _kiss is None and os.path.exists(REF_KISS_PATH):
        K = ctypes.CDLL(REF_KISS_PATH)
        K.ref_kiss_autocov.restype = ctypes.c_int
        K.ref_kiss_autocov.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        _kiss = K
    return _kiss


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def philox_block(ctr, k0, k1):
    c = np.asarray(ctr, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox_block(_p(c), int(k0) & 0xFFFFFFFF, int(k1) & 0xFFFFFFFF, _p(out))
    return out


def _pairs(pair):
    if pair is None:
        return None, None, None, None, 0
    pi, pj, pc, ps = pair
    pi = np.ascontiguousarray(pi, dtype=np.int32)
    pj = np.ascontiguousarray(pj, dtype=np.int32)
    pc = np.ascontiguousarray(pc, dtype=np.float32)
    ps = np.ascontiguousarray(ps, dtype=np.float32)
    return pi, pj, pc, ps, int(pi.size)


def move_stream(d, n, seed=0, variant=0, group=0, pair_prob=0.0, pair_mode=1, pair=None,
                ksparse_prob=0.0, ksparse_k=8):
    pi, pj, pc, ps, npairs = _pairs(pair)
    out = np.zeros((n, 3), dtype=np.int32)
    lib().orc_move_stream(d, seed, variant, group, pair_prob, pair_mode, _p(pi), _p(pj), _p(pc), _p(ps),
                          npairs, ksparse_prob, ksparse_k, n, _p(out))
    return out


def chr_run_f32(A, b, y0, nspc, thin, seed=0, variant=0, pair_prob=0.0, pair_mode=1, pair=None,
                ksparse_prob=0.0, ksparse_k=8, resync_interval=0, chains_per_group=1, nlog=0,
                nthreads=None):
    """Scalar FP32 CHR (chr.cu:281-482 rule). A (m, d); y0 (nchains, d).

    Returns dict(samples (nchains*nspc, d), y_final (nchains, d),
    s_final (nchains, m), move_log (nlog, 3) or None, seconds).
    """
    A = np.asarray(A, dtype=np.float32)
    m, d = A.shape
    A_cm = np.ascontiguousarray(A.T)  # column-major m x d
    b = np.ascontiguousarray(b, dtype=np.float64)
    y0 = np.ascontiguousarray(y0, dtype=np.float32)
    nchains = y0.shape[0]
    samples = np.zeros((nchains * nspc, d), dtype=np.float32)
    y_final = np.zeros((nchains, d), dtype=np.float32)
    s_final = np.zeros((nchains, m), dtype=np.float32)
    mlog = np.zeros((nlog, 3), dtype=np.int32) if nlog > 0 else None
    pi, pj, pc, ps, npairs = _pairs(pair)
    nthreads = nthreads or min(nchains, os.cpu_count() or 1)
    t = lib().orc_chr_run_f32(_p(A_cm), _p(b), m, d, _p(y0), nchains, nspc, thin, seed, variant, pair_prob,
                              pair_mode, _p(pi), _p(pj), _p(pc), _p(ps), npairs, ksparse_prob, ksparse_k,
                              resync_interval, chains_per_group, _p(samples), _p(y_final), _p(s_final),
                              _p(mlog), nlog, nthreads)
    if t < 0:
        raise ValueError(f"oracle rejected arguments ({t})")
    return dict(samples=samples, y_final=y_final, s_final=s_final, move_log=mlog, seconds=t)


def chr_run_f64(A, b, y0, nspc, thin, seed=0, variant=0, pair_prob=0.0, pair_mode=1, pair=None,
                ksparse_prob=0.0, ksparse_k=8, resync_interval=0, chains_per_group=1, nthreads=None):
    """All-FP64 CHR on the same random stream (drift reference)."""
    A = np.asarray(A, dtype=np.float64)
    m, d = A.shape
    A_cm = np.ascontiguousarray(A.T)
    b = np.ascontiguousarray(b, dtype=np.float64)
    y0 = np.ascontiguousarray(y0, dtype=np.float64)
    nchains = y0.shape[0]
    samples = np.zeros((nchains * nspc, d), dtype=np.float64)
    pi, pj, pc, ps, npairs = _pairs(pair)
    nthreads = nthreads or min(nchains, os.cpu_count() or 1)
    t = lib().orc_chr_run_f64(_p(A_cm), _p(b), m, d, _p(y0), nchains, nspc, thin, seed, variant, pair_prob,
                              pair_mode, _p(pi), _p(pj), _p(pc), _p(ps), npairs, ksparse_prob, ksparse_k,
                              resync_interval, chains_per_group, _p(samples), nthreads)
    if t < 0:
        raise ValueError(f"oracle rejected arguments ({t})")
    return dict(samples=samples, seconds=t)


def autocov(x, direct=False):
    """Biased autocovariance of the centred float series (mcmc.cpp:17-57)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    xc = np.zeros_like(x)
    lib().orc_center(_p(x), x.size, _p(xc))
    out = np.zeros_like(x)
    lib().orc_autocov(_p(xc), x.size, _p(out), 1 if direct else 0)
    return out


def ess_from_acov(acov):
    acov = np.ascontiguousarray(acov, dtype=np.float32)
    return float(lib().orc_ess_from_acov(_p(acov), acov.size))


def ess_one_chain(x, direct=False):
    x = np.ascontiguousarray(x, dtype=np.float32)
    return float(lib().orc_ess_one_chain(_p(x), x.size, 1 if direct else 0))


def ess(samples, nchains, layout=0, direct=False):
    """compute_ess (mcmc.cpp:91-124): (per-dim min ESS, per-chain ESS (dim, nchains))."""
    S = np.ascontiguousarray(samples, dtype=np.float32)
    n_total, dim = S.shape
    e = np.zeros(dim, dtype=np.float32)
    ec = np.zeros((dim, nchains), dtype=np.float32)
    rc = lib().orc_ess(_p(S), dim, n_total, nchains, layout, _p(e), _p(ec), 1 if direct else 0)
    if rc:
        raise ValueError("compute_ess: fewer than 4 samples per chain")
    return e, ec


def split_rhat(samples, nchains, layout=0):
    """compute_split_rhat (mcmc.cpp:126-194): (rhat[dim], above_1_1, above_1_2)."""
    S = np.ascontiguousarray(samples, dtype=np.float32)
    n_total, dim = S.shape
    r = np.zeros(dim, dtype=np.float32)
    above = np.zeros(2, dtype=np.int32)
    rc = lib().orc_split_rhat(_p(S), dim, n_total, nchains, layout, _p(r), _p(above))
    if rc:
        raise ValueError("compute_split_rhat: chain too short to split")
    return r, int(above[0]), int(above[1])


def percentile(v, q):
    v = np.ascontiguousarray(v, dtype=np.float32)
    return float(lib().orc_percentile(_p(v), v.size, q))


def backmap(T, shift, Y):
    """FP64 V = Y T^T + shift (backmap.cu:21-48)."""
    T = np.ascontiguousarray(T, dtype=np.float64)
    shift = np.ascontiguousarray(shift, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float32)
    n, d = T.shape
    V = np.zeros((Y.shape[0], n), dtype=np.float64)
    if lib().orc_backmap(_p(T), _p(shift), _p(Y), _p(V), n, d, Y.shape[0]):
        raise ValueError("backmap_into: shape mismatch")
    return V


def kiss_autocov(x):
    """Autocovariance through the reference's own kissfft (float, size 2n)."""
    K = kiss()
    if K is None:
        return None
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros_like(x)
    if K.ref_kiss_autocov(_p(x), x.size, _p(out)):
        raise RuntimeError("kissfft allocation failed")
    return out


def transform(A, b, P, c):
    """A' = A P and b' = b - A c exactly as the GPU packs a rounded variant."""
    A = np.asarray(A, dtype=np.float32)
    m, d = A.shape
    A_cm = np.ascontiguousarray(A.T)
    b = np.ascontiguousarray(b, dtype=np.float64)
    P = np.ascontiguousarray(P, dtype=np.float32)
    c = np.ascontiguousarray(c, dtype=np.float32)
    Aout = np.zeros((d, m), dtype=np.float32)
    bout = np.zeros(m, dtype=np.float64)
    lib().orc_transform(_p(A_cm), _p(b), m, d, _p(P), _p(c), _p(Aout), _p(bout))
    return np.ascontiguousarray(Aout.T), bout
