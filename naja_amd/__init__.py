"""naja_amd: MI355X (gfx950) coordinate hit-and-run sampler.

Python host mirror of the reference's GPU sampler API, `naja::gpu` in
src/gpu/gpusamplers.h:30-86, over the C ABI of ``lib/libnaja_amd.so``
(include/naja_amd.h). Device memory and streams come from PyTorch tensors;
every computation runs in the HIP library. There is no CPU fallback: if the
shared library is missing or fails to load, importing the compute entry points
raises.

Tensor conventions (natural torch shapes; the library sees the reference's
column-major buffers):
  A        (m, d)  polytope rows, any float dtype; stored FP32 column-major
  b        (m,)    right-hand side, kept FP64 for the FP64 slack computation
  X0       (nchains, d) start points (the reference's X_d, d x nchains)
  samples  (nchains * nspc, d) float32, row s*nchains + c = sample s of chain c
           (exactly samples.npy, npy.h:96-141)
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libnaja_amd.so")

NAJA_OK = 0
NAJA_ERR_INVALID_ARGUMENT = -1
NAJA_ERR_RUNTIME = -2
NAJA_ERR_HIP = -3
NAJA_ERR_UNSUPPORTED = -4

LAYOUT_INTERLEAVED = 0
LAYOUT_BLOCKS = 1

EXPORTED_SYMBOLS = (
    "naja_last_error",
    "naja_version",
    "naja_chr_sample",
    "naja_chr_sample_batched",
    "naja_chr_sample_batched_rounded",
    "naja_chr_plan_create",
    "naja_chr_plan_set_rounding",
    "naja_chr_plan_autotune",
    "naja_chr_plan_shape",
    "naja_chr_plan_layout",
    "naja_chr_plan_launch",
    "naja_chr_plan_destroy",
    "naja_chr_sample_streamed",
    "naja_backmap",
    "naja_diag",
    "naja_diag_batched",
    "naja_diag_noalloc",
    "naja_chr_timing",
    "naja_chr_kernel_times",
    "naja_kernel_basis",
    "naja_cube_center",
    "naja_extra_from_bounds",
    "naja_affine_hull",
    "naja_symmetric_eigen",
    "naja_precond_dikin",
    "naja_precond_dikin_batched",
    "naja_precond_ellipsoid_batched",
    "naja_precond_lowrank",
    "naja_precond_ellipsoid_device",
    "naja_precond_facet_device",
    "naja_precond_ellipsoid_large",
    "naja_precond_facet_large",
    "naja_analytic_center",
    "naja_lowrank_maps_device",
    "naja_lp_create",
    "naja_lp_destroy",
    "naja_lp_param_count",
    "naja_lp_get_params",
    "naja_lp_set_params",
    "naja_lp_forward",
    "naja_lp_loss_grad",
    "naja_lp_train_step",
    "naja_pilot_covariance",
    "naja_gemm_bf16",
    "naja_cond_run_create",
    "naja_cond_run_step",
    "naja_cond_run_sampling_ms",
    "naja_cond_run_destroy",
)


# int sink(const float* rows, int64_t n_rows, int n_cols, int64_t first_row, void* user)
SINK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p)


class NajaError(RuntimeError):
    """A failed library call (the reference throws std::runtime_error)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"naja error {code}: {msg}")
        self.code = code


class NajaInvalidArgument(NajaError, ValueError):
    """Shape / argument error (the reference throws std::invalid_argument)."""


class ChrParams(ctypes.Structure):
    _fields_ = [
        ("nspc", ctypes.c_int),
        ("thinning", ctypes.c_int),
        ("nchains", ctypes.c_int),
        ("tpb", ctypes.c_int),
        ("seed", ctypes.c_uint64),
        ("pair_prob", ctypes.c_double),
        ("resync_interval", ctypes.c_int),
        ("ksparse_prob", ctypes.c_double),
        ("ksparse_k", ctypes.c_int),
        ("pair_mode", ctypes.c_int),
        ("pair_i", ctypes.c_void_p),
        ("pair_j", ctypes.c_void_p),
        ("pair_c", ctypes.c_void_p),
        ("pair_s", ctypes.c_void_p),
        ("n_pairs", ctypes.c_int),
        ("chains_per_group", ctypes.c_int),
        ("move_log", ctypes.c_void_p),
        ("nlog", ctypes.c_int64),
    ]


_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libnaja_amd.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"naja_amd native library not built: {path} (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    lib.naja_last_error.restype = ctypes.c_char_p
    lib.naja_version.restype = i32
    lib.naja_chr_sample.argtypes = [vp, vp, i32, i32, vp, ctypes.POINTER(ChrParams), vp, vp, vp]
    lib.naja_chr_sample.restype = i32
    lib.naja_chr_sample_batched.argtypes = [vp, vp, i32, i32, i32, vp, vp, vp, vp,
                                            ctypes.POINTER(ChrParams), vp, vp, vp]
    lib.naja_chr_sample_batched.restype = i32
    P = ctypes.POINTER(ChrParams)
    lib.naja_chr_sample_batched_rounded.argtypes = [vp, vp, i32, i32, i32, vp, vp, vp, vp, vp, vp, P, vp, vp, vp]
    lib.naja_chr_sample_batched_rounded.restype = i32
    lib.naja_chr_plan_create.argtypes = [vp, vp, i32, i32, i32, vp, vp, vp, vp, vp, P, vp, ctypes.POINTER(vp)]
    lib.naja_chr_plan_create.restype = i32
    lib.naja_chr_plan_set_rounding.argtypes = [vp, vp, vp]
    lib.naja_chr_plan_set_rounding.restype = i32
    lib.naja_chr_plan_autotune.argtypes = [vp, vp, vp, vp]
    lib.naja_chr_plan_autotune.restype = i32
    lib.naja_chr_plan_shape.argtypes = [vp, vp]
    lib.naja_chr_plan_shape.restype = i32
    lib.naja_chr_plan_layout.argtypes = [vp, vp]
    lib.naja_chr_plan_layout.restype = i32
    lib.naja_chr_plan_launch.argtypes = [vp, vp, vp, vp, vp]
    lib.naja_chr_plan_launch.restype = i32
    lib.naja_chr_plan_destroy.argtypes = [vp]
    lib.naja_chr_plan_destroy.restype = i32
    lib.naja_chr_sample_streamed.argtypes = [vp, vp, i32, i32, vp, ctypes.POINTER(ChrParams), i32, vp, vp, i32,
                                             SINK, vp, vp]
    lib.naja_chr_sample_streamed.restype = i32
    lib.naja_backmap.argtypes = [vp, vp, vp, vp, i32, i32, i64, vp]
    lib.naja_backmap.restype = i32
    lib.naja_diag.argtypes = [vp, i32, i32, i32, i32, vp, vp, vp, vp]
    lib.naja_diag.restype = i32
    lib.naja_diag_batched.argtypes = [vp, i32, i32, i32, i32, i32, vp, vp, vp, vp, vp]
    lib.naja_diag_batched.restype = i32
    lib.naja_diag_noalloc.argtypes = [vp, i32, i32, i32, i32, vp, vp, vp, vp]
    lib.naja_diag_noalloc.restype = i32
    f64 = ctypes.c_double
    lib.naja_kernel_basis.argtypes = [vp, i32, i32, f64, vp, vp]
    lib.naja_chr_timing.argtypes = [i32]
    lib.naja_chr_timing.restype = i32
    lib.naja_chr_kernel_times.argtypes = [vp, vp, i32, vp]
    lib.naja_chr_kernel_times.restype = i32
    lib.naja_cube_center.argtypes = [vp, vp, i32, i32, vp, vp, vp, i32]
    lib.naja_cube_center.restype = i32
    lib.naja_extra_from_bounds.argtypes = [vp, vp, i32, i32, vp, vp, vp, vp, f64, vp, vp, vp]
    lib.naja_extra_from_bounds.restype = i32
    lib.naja_kernel_basis.restype = i32
    lib.naja_affine_hull.argtypes = [vp, vp, vp, i32, i32, f64, f64, vp, vp, vp]
    lib.naja_affine_hull.restype = i32
    lib.naja_symmetric_eigen.argtypes = [vp, i32, vp, vp]
    lib.naja_symmetric_eigen.restype = i32
    lib.naja_precond_dikin.argtypes = [vp, vp, vp, i32, i32, i32, f64, f64, f64, vp, vp, vp, vp, vp]
    lib.naja_precond_dikin.restype = i32
    lib.naja_precond_dikin_batched.argtypes = [vp, vp, i32, i32, i32, vp, vp, vp, vp, i32, f64, f64, f64, vp, vp,
                                               i32]
    lib.naja_precond_dikin_batched.restype = i32
    lib.naja_precond_ellipsoid_batched.argtypes = [vp, vp, i32, i32, i32, vp, vp, vp, vp, vp, i32]
    lib.naja_precond_ellipsoid_batched.restype = i32
    lib.naja_precond_lowrank.argtypes = [vp, i32, i32, i32, i32, vp, i32]
    lib.naja_precond_lowrank.restype = i32
    lib.naja_precond_ellipsoid_device.argtypes = [vp, vp, i32, i32, i32, vp, vp, vp, vp, vp, vp]
    lib.naja_precond_ellipsoid_device.restype = i32
    lib.naja_precond_facet_device.argtypes = [vp, vp, i32, i32, i32, vp, vp, vp, vp, vp, vp, i32, i32, ctypes.c_float,
                                               vp, vp]
    lib.naja_precond_facet_device.restype = i32
    lib.naja_precond_ellipsoid_large.argtypes = [vp, vp, i32, i32, vp, vp, vp]
    lib.naja_precond_ellipsoid_large.restype = i32
    lib.naja_precond_facet_large.argtypes = [vp, vp, i32, i32, vp, ctypes.c_float, vp, vp]
    lib.naja_precond_facet_large.restype = i32
    lib.naja_analytic_center.argtypes = [vp, vp, i32, i32, vp, i32, f64, vp, vp, vp]
    lib.naja_analytic_center.restype = i32
    lib.naja_lowrank_maps_device.argtypes = [vp, vp, i32, i32, i32, i32, vp, vp]
    lib.naja_lowrank_maps_device.restype = i32
    f32 = ctypes.c_float
    lib.naja_lp_create.argtypes = [i32, i32, i32, i32, i32, i32, ctypes.c_uint64, f32, ctypes.POINTER(vp)]
    lib.naja_lp_destroy.argtypes = [vp]
    lib.naja_lp_param_count.argtypes = [vp, ctypes.POINTER(i64)]
    lib.naja_lp_get_params.argtypes = [vp, vp, vp]
    lib.naja_lp_set_params.argtypes = [vp, vp, vp]
    lib.naja_lp_forward.argtypes = [vp, vp, i32, vp, vp]
    lib.naja_lp_loss_grad.argtypes = [vp, vp, vp, i32, vp, vp, vp]
    lib.naja_lp_train_step.argtypes = [vp, vp, vp, i32, f32, vp, vp]
    lib.naja_pilot_covariance.argtypes = [vp, i32, i64, i32, f64, vp, vp, vp]
    lib.naja_gemm_bf16.argtypes = [i32, i32, i32, i32, i32, vp, i64, vp, i64, vp, i64, f32, vp, vp, i64, vp]
    lib.naja_cond_run_create.argtypes = [vp, vp, i32, i32, i32, vp, vp, vp, f32, f32, ctypes.POINTER(ChrParams), vp,
                                         vp, vp, vp, vp, vp, vp, vp, i32, vp, ctypes.POINTER(vp)]
    lib.naja_cond_run_step.argtypes = [vp, vp]
    lib.naja_cond_run_sampling_ms.argtypes = [vp, ctypes.POINTER(f32)]
    lib.naja_cond_run_destroy.argtypes = [vp]
    for name in ("naja_lp_create", "naja_lp_destroy", "naja_lp_param_count", "naja_lp_get_params",
                 "naja_lp_set_params", "naja_lp_forward", "naja_lp_loss_grad", "naja_lp_train_step",
                 "naja_pilot_covariance", "naja_gemm_bf16", "naja_cond_run_create", "naja_cond_run_step",
                 "naja_cond_run_sampling_ms", "naja_cond_run_destroy"):
        getattr(lib, name).restype = i32
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc == NAJA_OK:
        return
    msg = _lib.naja_last_error().decode(errors="replace")
    if rc == NAJA_ERR_INVALID_ARGUMENT:
        raise NajaInvalidArgument(rc, msg)
    raise NajaError(rc, msg)


def _torch():
    import torch

    return torch


def _stream_ptr(stream) -> int:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def _dev(t, dtype):
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        t = torch.as_tensor(t)
    if not t.is_cuda:
        t = t.cuda()
    return t.to(dtype).contiguous()


def _colmajor_f32(A):
    """(m, d) tensor -> device FP32 buffer laid out column-major (lda = m)."""
    torch = _torch()
    A = _dev(A, torch.float32)
    return A.t().contiguous()


def chr_params(nspc, thinning, nchains, tpb_ss=-1, seed=0, pair_prob=0.0, resync_interval=0,
               ksparse_prob=0.0, ksparse_k=8, pair_mode=1, pair=None, chains_per_group=1,
               move_log=None, nlog=0) -> ChrParams:
    p = ChrParams()
    p.nspc, p.thinning, p.nchains, p.tpb = int(nspc), int(thinning), int(nchains), int(tpb_ss)
    p.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    p.pair_prob, p.resync_interval = float(pair_prob), int(resync_interval)
    p.ksparse_prob, p.ksparse_k, p.pair_mode = float(ksparse_prob), int(ksparse_k), int(pair_mode)
    if pair is not None:
        pi, pj, pc, ps = pair
        p.pair_i, p.pair_j = pi.data_ptr(), pj.data_ptr()
        p.pair_c, p.pair_s = pc.data_ptr(), ps.data_ptr()
        p.n_pairs = int(pi.numel())
    p.chains_per_group = int(chains_per_group)
    if move_log is not None:
        p.move_log = move_log.data_ptr()
        p.nlog = int(nlog)
    return p


def _pair_tensors(pair_i, pair_j, pair_c, pair_s):
    if pair_i is None:
        return None
    torch = _torch()
    return (_dev(pair_i, torch.int32), _dev(pair_j, torch.int32),
            _dev(pair_c, torch.float32), _dev(pair_s, torch.float32))


def coordinate_hit_and_run(A, b, X0, nspc, thinning, nchains, tpb_ss=-1, seed=0, pair_prob=0.0,
                           resync_interval=0, ksparse_prob=0.0, ksparse_k=8, pair_mode=1,
                           pair_i=None, pair_j=None, pair_c=None, pair_s=None,
                           chains_per_group=1, return_final=False, move_log_steps=0, stream=None):
    """naja::gpu::CoordinateHitAndRun (gpusamplers.h:30-48, chr.cu:13-70).

    Returns samples (nchains*nspc, d) float32 on the device; with
    return_final also the final states (nchains, d); with move_log_steps > 0
    also the (type, i, j) move log of chain 0 as an int32 (steps, 3) tensor.
    """
    torch = _torch()
    lib = load_library()
    A_cm = _colmajor_f32(A)
    d, m = A_cm.shape
    b_d = _dev(b, torch.float64)
    X = _dev(X0, torch.float32)
    if X.dim() == 1:
        X = X.unsqueeze(0).expand(nchains, d).contiguous()
    if tuple(X.shape) != (nchains, d):
        raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT, f"X0 shape {tuple(X.shape)} != ({nchains}, {d})")
    if b_d.numel() != m:
        raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT, f"b has {b_d.numel()} entries, A has {m} rows")
    pair = _pair_tensors(pair_i, pair_j, pair_c, pair_s)
    samples = torch.empty((nchains * nspc, d), dtype=torch.float32, device=A_cm.device)
    final = torch.empty((nchains, d), dtype=torch.float32, device=A_cm.device) if return_final else None
    mlog = None
    if move_log_steps > 0:
        mlog = torch.full((move_log_steps, 3), -7, dtype=torch.int32, device=A_cm.device)
    p = chr_params(nspc, thinning, nchains, tpb_ss, seed, pair_prob, resync_interval, ksparse_prob,
                   ksparse_k, pair_mode, pair, chains_per_group, mlog, move_log_steps)
    rc = lib.naja_chr_sample(A_cm.data_ptr(), b_d.data_ptr(), m, d, X.data_ptr(), ctypes.byref(p),
                             samples.data_ptr(), final.data_ptr() if final is not None else None,
                             _stream_ptr(stream))
    _check(rc)
    out = [samples]
    if return_final:
        out.append(final)
    if mlog is not None:
        out.append(mlog)
    return out[0] if len(out) == 1 else tuple(out)


def _batch_inputs(A, b, extra_A, extra_b):
    if isinstance(A, VariantBatch):
        return A.A_cm, A.b, A.d, A.m0, A.ks, A.Ae, A.be
    torch = _torch()
    A_cm = _colmajor_f32(A)
    d, m0 = A_cm.shape
    b_d = _dev(b, torch.float64)
    ks = [int(e.shape[0]) for e in extra_A]
    if all(not isinstance(e, torch.Tensor) for e in extra_A):
        # Host blocks: transpose and concatenate on the host, one upload.
        import numpy as np

        blocks = [np.asarray(e, dtype=np.float32).reshape(-1, d).T.ravel() for e in extra_A if len(e) > 0]
        Ae = torch.from_numpy(np.concatenate(blocks)).cuda() if blocks else None
        rhs = [np.asarray(f, dtype=np.float64).reshape(-1) for f in extra_b if len(f) > 0]
        be = torch.from_numpy(np.concatenate(rhs)).cuda() if blocks else None
        return A_cm, b_d, d, m0, ks, Ae, be
    blocks = [_colmajor_f32(e).reshape(-1) for e in extra_A if e.shape[0] > 0]
    Ae = torch.cat(blocks) if blocks else None
    be = torch.cat([_dev(f, torch.float64).reshape(-1) for f in extra_b if f.numel() > 0]) if blocks else None
    return A_cm, b_d, d, m0, ks, Ae, be


class VariantBatch:
    """A base polytope A y <= b and its conditioned variants [A; E_v] y <=
    [b; f_v], uploaded once in the batched sampler's layout (A column-major
    FP32, b FP64, the variants' extra rows as column-major blocks back to
    back). Pass it as `A` (with extra_A / extra_b None) to the batched
    sampler and the GPU roundings, so repeated launches upload nothing."""

    def __init__(self, A, b, extra_A, extra_b):
        self.A_cm, self.b, self.d, self.m0, self.ks, self.Ae, self.be = _batch_inputs(A, b, extra_A, extra_b)

    @property
    def n_variants(self) -> int:
        return len(self.ks)

    def slice(self, lo: int, hi: int) -> "VariantBatch":
        """Variants lo..hi-1 (views of the device buffers)."""
        out = VariantBatch.__new__(VariantBatch)
        out.A_cm, out.b, out.d, out.m0 = self.A_cm, self.b, self.d, self.m0
        r0, r1 = sum(self.ks[:lo]), sum(self.ks[:hi])
        out.ks = list(self.ks[lo:hi])
        out.Ae = self.Ae[r0 * self.d:r1 * self.d] if (self.Ae is not None and r1 > r0) else None
        out.be = self.be[r0:r1] if (self.be is not None and r1 > r0) else None
        return out

    def take(self, indices) -> "VariantBatch":
        """The variants at `indices`, in that order (device buffers gathered)."""
        torch = _torch()
        out = VariantBatch.__new__(VariantBatch)
        out.A_cm, out.b, out.d, out.m0 = self.A_cm, self.b, self.d, self.m0
        starts = [0]
        for k in self.ks:
            starts.append(starts[-1] + k)
        idx = [int(i) for i in indices]
        out.ks = [self.ks[i] for i in idx]
        a_parts = [self.Ae[starts[i] * self.d:starts[i + 1] * self.d] for i in idx if self.ks[i] > 0]
        b_parts = [self.be[starts[i]:starts[i + 1]] for i in idx if self.ks[i] > 0]
        out.Ae = torch.cat(a_parts) if a_parts else None
        out.be = torch.cat(b_parts) if b_parts else None
        return out

    def repeat(self, times: int) -> "VariantBatch":
        """The variant list repeated `times` times (device buffers concatenated)."""
        torch = _torch()
        out = VariantBatch.__new__(VariantBatch)
        out.A_cm, out.b, out.d, out.m0 = self.A_cm, self.b, self.d, self.m0
        out.ks = list(self.ks) * times
        out.Ae = torch.cat([self.Ae] * times) if self.Ae is not None else None
        out.be = torch.cat([self.be] * times) if self.be is not None else None
        return out


def coordinate_hit_and_run_batched(A, b, extra_A, extra_b, X0, nspc, thinning, nchains, tpb_ss=-1,
                                   seed=0, pair_prob=0.0, resync_interval=0, ksparse_prob=0.0,
                                   ksparse_k=8, pair_mode=1, chains_per_group=1, P=None, c=None,
                                   return_final=False, stream=None):
    """All conditioned variants [A; E_v] in one launch.

    extra_A / extra_b: lists (one entry per variant) of (k_v, d) and (k_v,)
    arrays, k_v >= 0; or A a VariantBatch (already on the device) with b,
    extra_A, extra_b None. X0: (n_variants, nchains, d) starts. With a rounding
    transform P (n_variants, d, d) and c (n_variants, d), variant v samples z
    with y = c_v + P_v z, X0 holds z-space starts and the samples are returned
    in y. Returns samples (n_variants, nchains*nspc, d).
    """
    torch = _torch()
    lib = load_library()
    A_cm, b_d, d, m0, ks, Ae, be = _batch_inputs(A, b, extra_A, extra_b)
    nv = len(ks)
    X = _dev(X0, torch.float32)
    if tuple(X.shape) != (nv, nchains, d):
        raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT, f"X0 shape {tuple(X.shape)} != ({nv}, {nchains}, {d})")
    Pd = _dev(P, torch.float32) if P is not None else None
    cd = _dev(c, torch.float32) if c is not None else None
    samples = torch.empty((nv, nchains * nspc, d), dtype=torch.float32, device=A_cm.device)
    final = torch.empty((nv, nchains, d), dtype=torch.float32, device=A_cm.device) if return_final else None
    karr = (ctypes.c_int * nv)(*ks)
    p = chr_params(nspc, thinning, nchains, tpb_ss, seed, pair_prob, resync_interval, ksparse_prob,
                   ksparse_k, pair_mode, None, chains_per_group)
    rc = lib.naja_chr_sample_batched_rounded(
        A_cm.data_ptr(), b_d.data_ptr(), m0, d, nv, karr, Ae.data_ptr() if Ae is not None else None,
        be.data_ptr() if be is not None else None, Pd.data_ptr() if Pd is not None else None,
        cd.data_ptr() if cd is not None else None, X.data_ptr(), ctypes.byref(p), samples.data_ptr(),
        final.data_ptr() if final is not None else None, _stream_ptr(stream))
    _check(rc)
    return (samples, final) if return_final else samples


class ChrPlan:
    """A packed batch that launches without allocating (naja_chr_plan_*).
    A may be a VariantBatch (b, extra_A, extra_b None).

    Keeps device copies of every input alive for the plan's lifetime. launch()
    is safe to capture in a torch.cuda.CUDAGraph (plain kernel launches on the
    current stream).
    """

    def __init__(self, A, b, nspc, thinning, nchains, extra_A=None, extra_b=None, P=None, c=None,
                 tpb_ss=-1, seed=0, pair_prob=0.0, resync_interval=0, ksparse_prob=0.0, ksparse_k=8,
                 pair_mode=1, chains_per_group=1, stream=None):
        torch = _torch()
        self._lib = load_library()
        if not isinstance(A, VariantBatch):  # a VariantBatch carries its extra rows
            extra_A = extra_A if extra_A is not None else [torch.zeros((0, A.shape[1]))]
            extra_b = extra_b if extra_b is not None else [torch.zeros(0)]
        self._A, self._b, self.d, self.m0, ks, self._Ae, self._be = _batch_inputs(A, b, extra_A, extra_b)
        self.n_variants = len(ks)
        self.nchains, self.nspc = nchains, nspc
        self._P = _dev(P, torch.float32) if P is not None else None
        self._c = _dev(c, torch.float32) if c is not None else None
        self._karr = (ctypes.c_int * self.n_variants)(*ks)
        self._params = chr_params(nspc, thinning, nchains, tpb_ss, seed, pair_prob, resync_interval,
                                  ksparse_prob, ksparse_k, pair_mode, None, chains_per_group)
        h = ctypes.c_void_p()
        ptr = lambda t: t.data_ptr() if t is not None else None  # noqa: E731
        _check(self._lib.naja_chr_plan_create(ptr(self._A), ptr(self._b), self.m0, self.d, self.n_variants,
                                              self._karr, ptr(self._Ae), ptr(self._be), ptr(self._P),
                                              ptr(self._c), ctypes.byref(self._params), _stream_ptr(stream),
                                              ctypes.byref(h)))
        self._h = h

    def set_rounding(self, P, c):
        """A new rounding for a plan created with one (same shapes): only
        A' = A P and b' = b - A c are formed again (naja_chr_plan_set_rounding)."""
        torch = _torch()
        if self._P is None:
            raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT, "the plan was created without a rounding")
        Pd, cd = _dev(P, torch.float32), _dev(c, torch.float32)
        if tuple(Pd.shape) != tuple(self._P.shape) or tuple(cd.shape) != tuple(self._c.shape):
            raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT,
                                      f"rounding shapes {tuple(Pd.shape)}, {tuple(cd.shape)} differ from the plan's")
        _check(self._lib.naja_chr_plan_set_rounding(self._h, Pd.data_ptr(), cd.data_ptr()))
        self._P, self._c = Pd, cd  # kept alive while the plan may transform again

    def samples_like(self):
        torch = _torch()
        return torch.empty((self.n_variants, self.nchains * self.nspc, self.d), dtype=torch.float32,
                           device=self._A.device)

    def autotune(self, X0, stream=None):
        shape = (ctypes.c_int * 3)()
        X = _dev(X0, _torch().float32)
        _check(self._lib.naja_chr_plan_autotune(self._h, X.data_ptr(), _stream_ptr(stream), shape))
        return tuple(shape)

    def shape(self):
        shape = (ctypes.c_int * 3)()
        _check(self._lib.naja_chr_plan_shape(self._h, shape))
        return tuple(shape)

    def layout(self):
        """(paired rows, chains per workgroup, scaled pairs) of the plan's
        launches: (1, 4, 1) is the paired LDS-column kernel on scaled pairs."""
        out = (ctypes.c_int * 3)()
        _check(self._lib.naja_chr_plan_layout(self._h, out))
        return tuple(out)

    def launch(self, X0, out, final=None, stream=None):
        _check(self._lib.naja_chr_plan_launch(self._h, X0.data_ptr(), out.data_ptr(),
                                              final.data_ptr() if final is not None else None,
                                              _stream_ptr(stream)))
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._lib.naja_chr_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class ConditionedRun:
    """Every conditioned variant of a VariantBatch sampled as one persistent
    hipGraph (naja_cond_run_*): facet-aligned rounding (completed from the
    base polytope's facets, maximum-volume exchange), rounding transform,
    batched CHR from z = 0 with the backmap to y, batched ESS / split-R-hat.
    Creation runs one step eagerly and captures the next; step() replays the
    graph and returns (samples (n_variants, nchains * nspc, d), ess_min,
    rhat_max) as device tensors owned by the run (overwritten by the next
    step). timed: event nodes around the CHR launch (sampling_ms)."""

    def __init__(self, batch, nspc, thinning, nchains, seed=0, resync_interval=0, theta=0.02, exchange_tol=1.02,
                 timed=False, stream=None):
        torch = _torch()
        lib = load_library()
        if not isinstance(batch, VariantBatch):
            raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT, "ConditionedRun takes a VariantBatch")
        self.batch = batch  # the run reads its device buffers
        nv, d = batch.n_variants, batch.d
        dev = batch.A_cm.device
        self.y0 = torch.zeros((nv, nchains, d), dtype=torch.float32, device=dev)
        self.c = torch.zeros((nv, d), dtype=torch.float32, device=dev)
        self.P = torch.empty((nv, d, d), dtype=torch.float32, device=dev)
        self.samples = torch.empty((nv, nchains * nspc, d), dtype=torch.float32, device=dev)
        self.ess = torch.empty((nv, d), dtype=torch.float32, device=dev)
        self.rhat = torch.empty((nv, d), dtype=torch.float32, device=dev)
        self.ess_min = torch.empty(nv, dtype=torch.float32, device=dev)
        self.rhat_max = torch.empty(nv, dtype=torch.float32, device=dev)
        self._p = chr_params(nspc, thinning, nchains, seed=seed, resync_interval=resync_interval)
        self._karr = (ctypes.c_int * nv)(*batch.ks)
        # A graph is captured on a stream of its own (not the legacy default
        # stream), after the work that made its inputs.
        self._torch_stream = stream if stream is not None else torch.cuda.Stream()
        self._stream = _stream_ptr(self._torch_stream)
        torch.cuda.current_stream().synchronize()
        h = ctypes.c_void_p()
        _check(lib.naja_cond_run_create(
            batch.A_cm.data_ptr(), batch.b.data_ptr(), batch.m0, d, nv, self._karr,
            batch.Ae.data_ptr() if batch.Ae is not None else None, batch.be.data_ptr() if batch.be is not None else None,
            float(theta), float(exchange_tol), ctypes.byref(self._p), self.y0.data_ptr(), self.c.data_ptr(),
            self.P.data_ptr(), self.samples.data_ptr(), self.ess.data_ptr(), self.rhat.data_ptr(),
            self.ess_min.data_ptr(), self.rhat_max.data_ptr(), int(bool(timed)), self._stream, ctypes.byref(h)))
        self._h = h

    def step(self, stream=None):
        _check(load_library().naja_cond_run_step(self._h, _stream_ptr(stream) if stream is not None else self._stream))
        return self.samples, self.ess_min, self.rhat_max

    def sampling_ms(self):
        ms = ctypes.c_float()
        _check(load_library().naja_cond_run_sampling_ms(self._h, ctypes.byref(ms)))
        return float(ms.value)

    def close(self):
        if getattr(self, "_h", None):
            _check(load_library().naja_cond_run_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def backmap(T, shift, Y, stream=None):
    """backmap_into (backmap.cu:21-48): V = Y T^T + shift, rows are samples."""
    torch = _torch()
    lib = load_library()
    T_d = _dev(T, torch.float32)
    sh = _dev(shift, torch.float32).reshape(-1)
    Y_d = _dev(Y, torch.float32)
    n, d = T_d.shape
    if Y_d.shape[1] != d:
        raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT,
                                  "backmap_into: transformation shape mismatch with samples")
    if sh.numel() != n:
        raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT,
                                  "backmap_into: shift length mismatch with transformation rows")
    V = torch.empty((Y_d.shape[0], n), dtype=torch.float32, device=T_d.device)
    _check(lib.naja_backmap(T_d.data_ptr(), sh.data_ptr(), Y_d.data_ptr(), V.data_ptr(), n, d,
                            Y_d.shape[0], _stream_ptr(stream)))
    return V


def coordinate_hit_and_run_backmap(A, b, X0, T, shift, nspc, thinning, nchains, **kw):
    """CoordinateHitAndRunBackmap (chr.cu:72-99): sample, then V = T y + shift."""
    Y = coordinate_hit_and_run(A, b, X0, nspc, thinning, nchains, **kw)
    return backmap(T, shift, Y, stream=kw.get("stream"))


def diagnostics(samples, nchains, layout=LAYOUT_INTERLEAVED, stream=None):
    """compute_ess + compute_split_rhat (mcmc.cpp:91-194) on the GPU.

    samples: (n_total, dim). Returns (ess[dim], rhat[dim], ess_chain[dim, nchains]).
    """
    torch = _torch()
    lib = load_library()
    S = _dev(samples, torch.float32)
    n_total, dim = S.shape
    ess = torch.empty(dim, dtype=torch.float32, device=S.device)
    rhat = torch.empty(dim, dtype=torch.float32, device=S.device)
    ess_chain = torch.empty((dim, nchains), dtype=torch.float32, device=S.device)
    _check(lib.naja_diag(S.data_ptr(), dim, n_total, nchains, layout, ess.data_ptr(), rhat.data_ptr(),
                         ess_chain.data_ptr(), _stream_ptr(stream)))
    return ess, rhat, ess_chain


def diagnostics_batched(samples, nchains, layout=LAYOUT_INTERLEAVED, stream=None):
    """diagnostics() of every variant of samples (n_variants, n_total, dim) in
    one set of launches. Returns (ess[v, dim], rhat[v, dim], ess_min[v], rhat_max[v])."""
    torch = _torch()
    lib = load_library()
    S = _dev(samples, torch.float32)
    nv, n_total, dim = S.shape
    ess = torch.empty((nv, dim), dtype=torch.float32, device=S.device)
    rhat = torch.empty((nv, dim), dtype=torch.float32, device=S.device)
    emin = torch.empty(nv, dtype=torch.float32, device=S.device)
    rmax = torch.empty(nv, dtype=torch.float32, device=S.device)
    _check(lib.naja_diag_batched(S.data_ptr(), nv, dim, n_total, nchains, layout, ess.data_ptr(), rhat.data_ptr(),
                                 emin.data_ptr(), rmax.data_ptr(), _stream_ptr(stream)))
    return ess, rhat, emin, rmax


def ellipsoid_rounding(A, b, extra_A, extra_b, x_c=None, stream=None):
    """Dikin-ellipsoid rounding of every variant [A; E_v] on the GPU:
    P (n_variants, d, d) FP32 with P P^T = H^{-1} (see naja_amd.precond for
    the FP64 host routine)."""
    torch = _torch()
    lib = load_library()
    A_cm, b_d, d, m0, ks, Ae, be = _batch_inputs(A, b, extra_A, extra_b)
    nv = len(ks)
    karr = (ctypes.c_int * nv)(*ks)
    xc = _dev(x_c, torch.float32) if x_c is not None else None
    P = torch.empty((nv, d, d), dtype=torch.float32, device=A_cm.device)
    _check(lib.naja_precond_ellipsoid_device(A_cm.data_ptr(), b_d.data_ptr(), m0, d, nv, karr,
                                             Ae.data_ptr() if Ae is not None else None,
                                             be.data_ptr() if be is not None else None,
                                             xc.data_ptr() if xc is not None else None, P.data_ptr(),
                                             _stream_ptr(stream)))
    return P


def ellipsoid_rounding_large(A, b, x_c=None, stream=None):
    """Dikin-ellipsoid rounding of one polytope A y <= b of any size on the
    GPU (FP64 Hessian by rocBLAS, blocked Cholesky): P (d, d) FP32, upper
    triangular, P P^T = H^{-1} at x_c (origin by default)."""
    torch = _torch()
    lib = load_library()
    A_cm = _colmajor_f32(A)
    d, m = A_cm.shape
    b_d = _dev(b, torch.float64)
    if b_d.numel() != m:
        raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT, f"b has {b_d.numel()} entries for {m} rows")
    xc = _dev(x_c, torch.float32).reshape(-1) if x_c is not None else None
    P = torch.empty((d, d), dtype=torch.float32, device=A_cm.device)
    _check(lib.naja_precond_ellipsoid_large(A_cm.data_ptr(), b_d.data_ptr(), m, d,
                                            xc.data_ptr() if xc is not None else None, P.data_ptr(),
                                            _stream_ptr(stream)))
    return P


def analytic_center(A, b, x0=None, max_iter=200, tol=1e-12, stream=None):
    """Analytic centre of A y <= b (damped Newton on the log barrier, FP64,
    on the GPU) from the interior point x0 (origin by default); returns
    (x (d,) FP32 device tensor, Newton steps)."""
    torch = _torch()
    lib = load_library()
    A_cm = _colmajor_f32(A)
    d, m = A_cm.shape
    b_d = _dev(b, torch.float64)
    x0d = _dev(x0, torch.float32).reshape(-1) if x0 is not None else None
    x = torch.empty(d, dtype=torch.float32, device=A_cm.device)
    it = ctypes.c_int(0)
    _check(lib.naja_analytic_center(A_cm.data_ptr(), b_d.data_ptr(), m, d, x0d.data_ptr() if x0d is not None else None,
                                    int(max_iter), float(tol), x.data_ptr(), ctypes.byref(it), _stream_ptr(stream)))
    return x, int(it.value)


def facet_rounding_large(A, b, theta=0.02, x_c=None, stream=None):
    """Facet-aligned rounding of one polytope A y <= b with at least d
    independent facets, at any size on the GPU: P (d, d) FP32 whose
    coordinate slabs are the d tightest independent facets at x_c."""
    torch = _torch()
    lib = load_library()
    A_cm = _colmajor_f32(A)
    d, m = A_cm.shape
    b_d = _dev(b, torch.float64)
    if b_d.numel() != m:
        raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT, f"b has {b_d.numel()} entries for {m} rows")
    xc = _dev(x_c, torch.float32).reshape(-1) if x_c is not None else None
    P = torch.empty((d, d), dtype=torch.float32, device=A_cm.device)
    _check(lib.naja_precond_facet_large(A_cm.data_ptr(), b_d.data_ptr(), m, d,
                                        xc.data_ptr() if xc is not None else None, float(theta), P.data_ptr(),
                                        _stream_ptr(stream)))
    return P


# How facet_rounding completes the directions the facets leave free
# (NAJA_FACET_COMPLETE_* in include/naja_amd.h).
FACET_COMPLETE_ROUNDING, FACET_COMPLETE_LOWER, FACET_COMPLETE_BASE_FACETS = 0, 1, 2


def facet_rounding(A, b, extra_A, extra_b, P0, theta, x_c=None, include_base=False, lower=False,
                   completion=None, exchange_tol=0.0, stream=None):
    """Facet-aligned rounding of every variant [A; E_v] on the GPU (FP64
    inside): P (n_variants, d, d) FP32 whose coordinate axes run along the
    edges of a parallelotope of the variant's binding facets. The directions
    the facets leave free are completed by `completion`:
    FACET_COMPLETE_ROUNDING (default) from the dual rows of the base rounding
    P0 (n_variants, d, d); FACET_COMPLETE_LOWER (or lower=True) from the
    lower Cholesky factor of P0 P0^T; FACET_COMPLETE_BASE_FACETS from the base
    polytope's own facets, orthonormalised in the free subspace (P0 unused,
    may be None). theta: one float or n_variants floats, the fraction of its
    norm a facet candidate must keep to be chosen (> 1: no facets).
    exchange_tol (0 = off, else >= 1): when the facets fill the basis, swap in
    a facet whose coefficient in the basis exceeds it (maximum-volume
    exchange)."""
    torch = _torch()
    lib = load_library()
    if completion is None:
        completion = FACET_COMPLETE_LOWER if lower else FACET_COMPLETE_ROUNDING
    A_cm, b_d, d, m0, ks, Ae, be = _batch_inputs(A, b, extra_A, extra_b)
    nv = len(ks)
    karr = (ctypes.c_int * nv)(*ks)
    P0d = None
    if completion != FACET_COMPLETE_BASE_FACETS or P0 is not None:
        P0d = _dev(P0, torch.float32).contiguous()
        if tuple(P0d.shape) != (nv, d, d):
            raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT, f"P0 {tuple(P0d.shape)}, expected ({nv}, {d}, {d})")
    th = torch.as_tensor(theta, dtype=torch.float32).reshape(-1)
    th = (th.expand(nv) if th.numel() == 1 else th).to(A_cm.device).contiguous()
    if th.numel() != nv:
        raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT, f"{th.numel()} theta values for {nv} variants")
    xc = _dev(x_c, torch.float32) if x_c is not None else None
    P = torch.empty((nv, d, d), dtype=torch.float32, device=A_cm.device)
    _check(lib.naja_precond_facet_device(A_cm.data_ptr(), b_d.data_ptr(), m0, d, nv, karr,
                                         Ae.data_ptr() if Ae is not None else None,
                                         be.data_ptr() if be is not None else None,
                                         xc.data_ptr() if xc is not None else None,
                                         P0d.data_ptr() if P0d is not None else None, th.data_ptr(),
                                         int(bool(include_base)), int(completion), float(exchange_tol), P.data_ptr(),
                                         _stream_ptr(stream)))
    return P


def lowrank_maps(P0, V, expand=True, stream=None):
    """P_v = P0_v (I + V_v V_v^T)^{+1/2} (expand) or ^{-1/2} on the GPU, FP64
    inside: P0 (n_variants, d, d) and V (n_variants, d, rank) FP32 CUDA
    tensors; returns P (n_variants, d, d) FP32 (see naja_amd.precond.lowrank
    for the FP64 host routine of the update alone)."""
    torch = _torch()
    lib = load_library()
    for name, t in (("P0", P0), ("V", V)):
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.dim() == 3):
            raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT, f"{name} must be a 3-d float32 CUDA tensor")
    nv, d, r = V.shape
    if tuple(P0.shape) != (nv, d, d):
        raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT, f"P0 {tuple(P0.shape)} does not match V {tuple(V.shape)}")
    P0c, Vc = P0.contiguous(), V.contiguous()
    P = torch.empty_like(P0c)
    _check(lib.naja_lowrank_maps_device(P0c.data_ptr(), Vc.data_ptr(), nv, d, r, int(bool(expand)), P.data_ptr(),
                                        _stream_ptr(stream)))
    return P


def coordinate_hit_and_run_streamed(A, b, X, nspc, thinning, nchains, chunk_nspc, sink, T=None, shift=None,
                                    tpb_ss=-1, seed=0, pair_prob=0.0, resync_interval=0, ksparse_prob=0.0,
                                    ksparse_k=8, pair_mode=1, chains_per_group=1, stream=None):
    """CoordinateHitAndRunStreamed (chr.cu:101-230): chunks of chunk_nspc
    samples per chain, each handed to sink(rows: np.ndarray (n_rows, n_cols),
    first_row: int) on the host as soon as it is copied back (rows in
    samples.npy order). X (nchains, d) is the start and is updated in place
    to the final state. With T, shift the chunks are backmapped first."""
    import numpy as np

    torch = _torch()
    lib = load_library()
    A_cm = _colmajor_f32(A)
    d, m = A_cm.shape
    b_d = _dev(b, torch.float64)
    if not (isinstance(X, torch.Tensor) and X.is_cuda and X.dtype == torch.float32 and X.is_contiguous()):
        raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT, "X must be a contiguous float32 CUDA tensor")
    Td = _dev(T, torch.float32) if T is not None else None
    sd = _dev(shift, torch.float32) if shift is not None else None
    n = int(Td.shape[0]) if Td is not None else 0
    failure = []

    def _sink(ptr, n_rows, n_cols, first_row, user):
        try:
            buf = (ctypes.c_float * (n_rows * n_cols)).from_address(ptr)
            sink(np.ctypeslib.as_array(buf).reshape(n_rows, n_cols).copy(), int(first_row))
            return 0
        except BaseException as e:  # surfaced after the call returns
            failure.append(e)
            return 1

    cb = SINK(_sink)
    p = chr_params(nspc, thinning, nchains, tpb_ss, seed, pair_prob, resync_interval, ksparse_prob, ksparse_k,
                   pair_mode, None, chains_per_group)
    rc = lib.naja_chr_sample_streamed(A_cm.data_ptr(), b_d.data_ptr(), m, d, X.data_ptr(), ctypes.byref(p),
                                      int(chunk_nspc), Td.data_ptr() if Td is not None else None,
                                      sd.data_ptr() if sd is not None else None, n, cb, None, _stream_ptr(stream))
    if failure:
        raise failure[0]
    _check(rc)


def chr_kernel_timing(enable: bool) -> None:
    """Record HIP events around every CHR step-kernel launch (measurement)."""
    _check(load_library().naja_chr_timing(1 if enable else 0))


def chr_kernel_times(max_launches: int = 4096):
    """[(work items, ms)] of the launches recorded since the last call, oldest first."""
    items = (ctypes.c_int * max_launches)()
    ms = (ctypes.c_float * max_launches)()
    n = ctypes.c_int(0)
    _check(load_library().naja_chr_kernel_times(items, ms, max_launches, ctypes.byref(n)))
    return [(int(items[i]), float(ms[i])) for i in range(n.value)]
