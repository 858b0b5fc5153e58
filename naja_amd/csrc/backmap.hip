// Backmap V = T Y + shift as an FP32 MFMA GEMM with the shift fused into the
// epilogue. Replaces backmap_into (backmap.cu:21-48): cublasDgemm followed
// by add_shift_cols_kernel (backmap.cu:8-18).
//
// Operands are both K-contiguous: Y is (ns x d) sample rows, T is (n x d)
// reaction rows, V is (ns x n) backmapped sample rows, i.e.
//   V[s][r] = sum_k Y[s][k] * T[r][k] + shift[r].
// Workgroup tile 128 (samples) x 128 (reactions) x 32 (k), four wavefronts
// each owning a 64 x 64 sub-tile as 2 x 2 blocks of v_mfma_f32_32x32x2f32.
// Tiles are staged through LDS transposed to [k][row] so the 32 lanes that
// feed one MFMA operand read consecutive words.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "backmap.h"
#include "common.h"

namespace naja {

namespace {

constexpr int kBM = 128;
constexpr int kBN = 128;
constexpr int kBK = 32;
constexpr int kPad = 4;

typedef float floatx16 __attribute__((ext_vector_type(16)));

__global__ __launch_bounds__(256) void backmap_mfma_kernel(const float* __restrict__ T, const float* __restrict__ shift,
                                                          const float* __restrict__ Y, float* __restrict__ V, int n,
                                                          int d, int64_t ns, bool vec4, int64_t sT, int64_t sShift,
                                                          int64_t sY, int64_t sV)
{
    // Batch item blockIdx.z uses its own T, shift, Y and V (strides in floats).
    T += sT * blockIdx.z;
    shift += sShift * blockIdx.z;
    Y += sY * blockIdx.z;
    V += sV * blockIdx.z;
    __shared__ float Ys[kBK][kBM + kPad];
    __shared__ float Ts[kBK][kBN + kPad];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int64_t s0 = (int64_t)blockIdx.x * kBM;
    const int r0 = blockIdx.y * kBN;
    const int wm = (wave & 1) * 64;
    const int wn = (wave >> 1) * 64;

    floatx16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;

    // Loader: thread -> (row = tid/8 + 32p, k = 4*(tid%8) .. +3), p = 0..3.
    const int lrow = tid >> 3;
    const int lk = (tid & 7) * 4;

    for (int k0 = 0; k0 < d; k0 += kBK) {
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int row = lrow + 32 * p;
            const int64_t s = s0 + row;
            const int r = r0 + row;
            float yv[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            float tv[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            const int k = k0 + lk;
            if (vec4 && k + 3 < d) {
                if (s < ns) {
                    const float4 v4 = *reinterpret_cast<const float4*>(Y + s * d + k);
                    yv[0] = v4.x; yv[1] = v4.y; yv[2] = v4.z; yv[3] = v4.w;
                }
                if (r < n) {
                    const float4 v4 = *reinterpret_cast<const float4*>(T + (int64_t)r * d + k);
                    tv[0] = v4.x; tv[1] = v4.y; tv[2] = v4.z; tv[3] = v4.w;
                }
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    if (k + e < d) {
                        if (s < ns) yv[e] = Y[s * d + k + e];
                        if (r < n) tv[e] = T[(int64_t)r * d + k + e];
                    }
                }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                Ys[lk + e][row] = yv[e];
                Ts[lk + e][row] = tv[e];
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kBK; kk += 2) {
            const int kr = kk + (lane >> 5);
            float a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = Ys[kr][wm + i * 32 + (lane & 31)];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = Ts[kr][wn + j * 32 + (lane & 31)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }

    // D layout of the 32x32 block: lane holds column lane%32; register e
    // holds row 8*(e/4) + 4*(lane/32) + e%4.
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int r = r0 + wn + j * 32 + (lane & 31);
        if (r >= n) continue;
        const float sh = shift[r];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int64_t s = s0 + wm + i * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
                if (s < ns) V[s * n + r] = acc[i][j][e] + sh;
            }
        }
    }
}

}  // namespace

void backmap_launch_batched(const float* T, const float* shift, const float* Y, float* V, int n, int d, int64_t ns,
                            int batch, int64_t sT, int64_t sShift, int64_t sY, int64_t sV, hipStream_t stream)
{
    if (ns == 0 || batch == 0) return;
    const int64_t gx = (ns + kBM - 1) / kBM;
    NAJA_REQUIRE(gx <= 0x7FFFFFFF, NAJA_ERR_UNSUPPORTED, "too many sample rows for one backmap launch");
    NAJA_REQUIRE(batch <= 65535, NAJA_ERR_UNSUPPORTED, "too many backmap batch items (%d)", batch);
    dim3 grid((unsigned)gx, (unsigned)((n + kBN - 1) / kBN), (unsigned)batch);
    // 16-byte loads need d % 4 == 0 and 16-byte aligned operand bases.
    const bool vec4 = (d % 4 == 0) && (reinterpret_cast<uintptr_t>(T) % 16 == 0) &&
                      (reinterpret_cast<uintptr_t>(Y) % 16 == 0) && (sT % 4 == 0) && (sY % 4 == 0);
    hipLaunchKernelGGL(backmap_mfma_kernel, grid, dim3(256), 0, stream, T, shift, Y, V, n, d, ns, vec4, sT, sShift, sY,
                       sV);
    NAJA_HIP(hipGetLastError());
}

void backmap_launch(const float* T, const float* shift, const float* Y, float* V, int n, int d, int64_t ns,
                    hipStream_t stream)
{
    backmap_launch_batched(T, shift, Y, V, n, d, ns, 1, 0, 0, 0, 0, stream);
}

}  // namespace naja
