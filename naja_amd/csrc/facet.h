// Internal interface of the facet-aligned rounding kernel (facet.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace naja {

// NAJA_ERR_UNSUPPORTED unless the kernel holds d columns and the candidates
// (facets plus the completion's) of the largest variant (kmax extra rows).
void facet_check_shape(int m0, int d, int kmax, int include_base, int completion);
// One workgroup per variant; eoff / ek the variants' extra-row offsets and
// counts on the device; status (n_variants ints, device) receives NAJA_OK or
// the variant's error; P0i: n_variants x d x d doubles of scratch unless the
// completion is NAJA_FACET_COMPLETE_BASE_FACETS. Launch only: no allocation,
// copy or synchronisation.
void facet_launch(const float* A0, const double* b0, int m0, int d, int n_variants, const float* Ae, const double* be,
                  const int64_t* eoff, const int* ek, const float* xc, const float* P0, const float* theta,
                  int include_base, int completion, float exchange_tol, double* P0i, float* P, int* status,
                  hipStream_t s);

}  // namespace naja
