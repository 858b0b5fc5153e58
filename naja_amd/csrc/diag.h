// Internal interface of the ESS / split-R-hat kernels (diag.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace naja {

// Largest chain length the LDS-resident FFT handles (N/2 complex points in LDS).
int diag_max_chain_length();
void diag_launch(const float* samples, int dim, int n_total, int nchains, int layout, float* ess, float* rhat,
                 float* ess_chain, hipStream_t stream);

}  // namespace naja
