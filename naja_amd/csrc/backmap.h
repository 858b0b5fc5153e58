// Internal interface of the MFMA backmap kernel (backmap.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace naja {

void backmap_launch(const float* T, const float* shift, const float* Y, float* V, int n, int d, int64_t ns,
                    hipStream_t stream);

}  // namespace naja
