// Internal interface of the ESS / split-R-hat kernels (diag.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace naja {

// Largest chain length the LDS-resident FFT handles (N/2 complex points in LDS).
int diag_max_chain_length();
// n_variants sample blocks of n_total x dim back to back; ess, rhat per
// variant and dim; optional per-variant min ESS / max R-hat.
void diag_launch(const float* samples, int dim, int n_total, int nchains, int layout, float* ess, float* rhat,
                 float* ess_chain, hipStream_t stream, int n_variants = 1, float* ess_min = nullptr,
                 float* rhat_max = nullptr);
// Scratch held by naja_diag_noalloc's static device buffer (per GPU).
constexpr size_t kDiagNoAllocBytes = (size_t)64 << 20;
// diag_launch for one sample set with all scratch in that static buffer
// (no allocation; NAJA_ERR_UNSUPPORTED beyond kDiagNoAllocBytes).
void diag_launch_noalloc(const float* samples, int dim, int n_total, int nchains, int layout, float* ess, float* rhat,
                         float* ess_chain, hipStream_t stream);

// Diagnostics of n_variants sample blocks with a caller-held workspace of
// diag_workspace_bytes (chains up to diag_max_chain_length samples, else
// NAJA_ERR_UNSUPPORTED), its twiddle tables written once by
// diag_workspace_init (synchronous); diag_launch_ws then issues kernels only:
// no allocation and no host copy (capturable in a graph).
size_t diag_workspace_bytes(int dim, int n_total, int nchains, int n_variants);
void diag_workspace_init(void* ws, int dim, int n_total, int nchains, int n_variants);
void diag_launch_ws(void* ws, const float* samples, int dim, int n_total, int nchains, int layout, float* ess,
                    float* rhat, hipStream_t stream, int n_variants, float* ess_min, float* rhat_max,
                    bool series_ready = false);
// The workspace's series array and its row stride: series_ready = true tells
// diag_launch_ws that the caller already wrote every (variant, chain, dim)
// series there, at ((v nchains + c) dim + d) * stride + s (interleaved
// samples only), so its transpose is skipped.
float* diag_workspace_series(void* ws);
int diag_series_stride(int n_total, int nchains);

}  // namespace naja
