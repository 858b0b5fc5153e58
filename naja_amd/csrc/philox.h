// Philox4x32-10 counter RNG shared by the HIP kernels and the host library.
// Construction from Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3" (SC'11): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
// increments 0x9E3779B9 / 0xBB67AE85, ten rounds.
//
// It replaces the reference's per-chain cuRAND MRG32k3a states
// (chr.cu:43-48, chr.cu:556-560). Stream layout (fixed, so any launch shape
// gives the same numbers):
//   key          = (lo32(seed), hi32(seed) ^ variant)
//   move block   : counter (lo32(t), hi32(t), group, 0)
//   k-sparse     : counter (lo32(t), hi32(t), group, 1 + q/4), word q%4
//   alpha block  : counter (lo32(t), hi32(t), chain/4, 0xFFFFFFFF), word chain%4
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace naja {

constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;
constexpr uint32_t kAlphaBlock = 0xFFFFFFFFu;

struct U32x4 {
    uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ U32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                        uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += kPhiloxW0;
            k1 += kPhiloxW1;
        }
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    return U32x4{c0, c1, c2, c3};
}

// Index in [0, n) as the high word of w*n: integer-exact on host and device.
__host__ __device__ __forceinline__ uint32_t rng_index(uint32_t w, uint32_t n)
{
    return (uint32_t)(((uint64_t)w * (uint64_t)n) >> 32);
}

// Uniform in (0,1): an odd 24-bit integer times 2^-24 (exact in FP32).
__host__ __device__ __forceinline__ float rng_uniform(uint32_t w)
{
    return (float)(((w >> 9) << 1) | 1u) * 5.9604644775390625e-08f;
}

// w < threshold(p) selects an event of probability p (threshold capped at 2^32).
inline uint64_t rng_threshold(double p)
{
    if (!(p > 0.0)) return 0;
    if (p >= 1.0) return 4294967296ull;
    return (uint64_t)(p * 4294967296.0);
}

}  // namespace naja
