"""Learned preconditioner and pilot-chain statistics (ctypes over libnaja_amd.so).

`LearnedPreconditioner` owns an MLP on the GPU: extra-bound features ->
hidden1 -> hidden2 -> a dim x rank factor V. With the precision model its
whitening is P = (I + V V^T)^{-1/2}; with the covariance model it updates a
base rounding P0 as P0 (I + V V^T)^{1/2}. The sampler's rounding transform
applies the result (y = c + P z). Training, inference and the covariance of pilot chains run in
the library's own HIP kernels; tensors are torch device tensors.
"""
from __future__ import annotations

import ctypes

from . import NajaInvalidArgument, NAJA_ERR_INVALID_ARGUMENT, _check, _dev, _stream_ptr, _torch, load_library


class LearnedPreconditioner:
    PRECISION, COVARIANCE = 0, 1

    def __init__(self, n_features=300, dim=100, rank=32, hidden1=512, hidden2=512, seed=0, out_scale=64.0,
                 model=0):
        self.n_features, self.dim, self.rank = n_features, dim, rank
        self.hidden1, self.hidden2, self.out_scale = hidden1, hidden2, float(out_scale)
        self.model = int(model)
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib.naja_lp_create(n_features, dim, rank, hidden1, hidden2, self.model, ctypes.c_uint64(seed),
                                  self.out_scale, ctypes.byref(h)))
        self._h = h
        n = ctypes.c_int64()
        _check(lib.naja_lp_param_count(h, ctypes.byref(n)))
        self.n_params = int(n.value)

    def close(self):
        if self._h:
            _check(load_library().naja_lp_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _x(self, features):
        torch = _torch()
        X = _dev(features, torch.float32)
        if X.dim() != 2 or X.shape[1] != self.n_features:
            raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT,
                                      f"features shape {tuple(X.shape)}, expected (batch, {self.n_features})")
        return X

    def _s(self, S, batch):
        torch = _torch()
        S = _dev(S, torch.float32)
        if tuple(S.shape) != (batch, self.dim, self.dim):
            raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT,
                                      f"covariance shape {tuple(S.shape)}, expected ({batch}, {self.dim}, {self.dim})")
        return S

    def params(self, stream=None):
        torch = _torch()
        out = torch.empty(self.n_params, dtype=torch.float32, device="cuda")
        _check(load_library().naja_lp_get_params(self._h, out.data_ptr(), _stream_ptr(stream)))
        return out

    def set_params(self, p, stream=None):
        torch = _torch()
        p = _dev(p, torch.float32)
        if p.numel() != self.n_params:
            raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT, f"{p.numel()} parameters, expected {self.n_params}")
        _check(load_library().naja_lp_set_params(self._h, p.data_ptr(), _stream_ptr(stream)))

    def forward(self, features, stream=None):
        """V (batch, dim, rank)."""
        torch = _torch()
        X = self._x(features)
        V = torch.empty((X.shape[0], self.dim, self.rank), dtype=torch.float32, device=X.device)
        _check(load_library().naja_lp_forward(self._h, X.data_ptr(), X.shape[0], V.data_ptr(), _stream_ptr(stream)))
        return V

    def loss_grad(self, features, S, stream=None):
        """(per-condition loss (batch,), gradient of the batch mean (n_params,))."""
        torch = _torch()
        X = self._x(features)
        S = self._s(S, X.shape[0])
        loss = torch.empty(X.shape[0], dtype=torch.float32, device=X.device)
        g = torch.empty(self.n_params, dtype=torch.float32, device=X.device)
        _check(load_library().naja_lp_loss_grad(self._h, X.data_ptr(), S.data_ptr(), X.shape[0], loss.data_ptr(),
                                                g.data_ptr(), _stream_ptr(stream)))
        return loss, g

    def train_step(self, features, S, lr=1e-3, stream=None):
        torch = _torch()
        X = self._x(features)
        S = self._s(S, X.shape[0])
        loss = torch.empty(X.shape[0], dtype=torch.float32, device=X.device)
        _check(load_library().naja_lp_train_step(self._h, X.data_ptr(), S.data_ptr(), X.shape[0], float(lr),
                                                 loss.data_ptr(), _stream_ptr(stream)))
        return loss

    def split(self, p):
        """Views (W1, b1, W2, b2, W3, b3) of a parameter vector."""
        F, H1, H2, O = self.n_features, self.hidden1, self.hidden2, self.dim * self.rank
        sizes = [H1 * F, H1, H2 * H1, H2, O * H2, O]
        shapes = [(H1, F), (H1,), (H2, H1), (H2,), (O, H2), (O,)]
        out, o = [], 0
        for n, sh in zip(sizes, shapes):
            out.append(p[o:o + n].reshape(sh))
            o += n
        return out


def pilot_covariance(samples, scale=1.0, stream=None):
    """Per-variant covariance of samples (n_variants, rows, d): (S float32, mean float64)."""
    torch = _torch()
    X = _dev(samples, torch.float32)
    if X.dim() != 3:
        raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT, "samples must be (n_variants, rows, d)")
    nv, rows, d = X.shape
    S = torch.empty((nv, d, d), dtype=torch.float32, device=X.device)
    mean = torch.empty((nv, d), dtype=torch.float64, device=X.device)
    _check(load_library().naja_pilot_covariance(X.data_ptr(), nv, rows, d, float(scale), S.data_ptr(),
                                                mean.data_ptr(), _stream_ptr(stream)))
    return S, mean


def gemm_bf16(A, B, trans_a=False, trans_b=False, alpha=1.0, stream=None, row_sums=False, split_k=False):
    """alpha * op(A) op(B) on BF16 MFMA (FP32 accumulate). row_sums (trans_a
    only) also returns the sums over k of op(A) in FP32; split_k hands the
    product a workspace so that K may be split across workgroups."""
    torch = _torch()
    A = _dev(A, torch.float32)
    B = _dev(B, torch.float32)
    M, K = (A.shape[1], A.shape[0]) if trans_a else tuple(A.shape)
    K2, N = (B.shape[1], B.shape[0]) if trans_b else tuple(B.shape)
    if K != K2:
        raise NajaInvalidArgument(NAJA_ERR_INVALID_ARGUMENT, f"inner dimensions {K} != {K2}")
    C = torch.empty((M, N), dtype=torch.float32, device=A.device)
    rs = torch.empty(M, dtype=torch.float32, device=A.device) if row_sums else None
    ws = torch.empty(16 * (M * N + M), dtype=torch.float32, device=A.device) if split_k else None
    _check(load_library().naja_gemm_bf16(int(trans_a), int(trans_b), M, N, K, A.data_ptr(), A.shape[1],
                                         B.data_ptr(), B.shape[1], C.data_ptr(), N, float(alpha),
                                         rs.data_ptr() if rs is not None else None,
                                         ws.data_ptr() if ws is not None else None,
                                         ws.numel() if ws is not None else 0, _stream_ptr(stream)))
    return (C, rs) if row_sums else C
