// Internal interface of the MFMA backmap kernel (backmap.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace naja {

// Batch item z computes V_z = Y_z T_z^T + shift_z (strides in floats).
// ser (optional, the small-operand kernel only: backmap_series_ok): the
// outputs also in the diagnostics' series layout, ser[((z nchains + c) n +
// j) spcp + s] for row s nchains + c (ns a multiple of nchains).
void backmap_launch_batched(const float* T, const float* shift, const float* Y, float* V, int n, int d, int64_t ns,
                            int batch, int64_t sT, int64_t sShift, int64_t sY, int64_t sV, hipStream_t stream,
                            float* ser = nullptr, int nchains = 0, int spcp = 0);
bool backmap_series_ok(int n, int d, const float* Y, const float* V, int64_t sY, int64_t sV, int batch);
// True when backmap_launch_batched with V == Y (n == d, strides sY) runs the
// small-operand kernel, which may work in place: each workgroup reads a whole
// tile of rows before writing the same rows, and no two share a tile.
bool backmap_inplace_ok(int d, const float* Y, int64_t sY, int batch);
void backmap_launch(const float* T, const float* shift, const float* Y, float* V, int n, int d, int64_t ns,
                    hipStream_t stream);

}  // namespace naja
