/*
 * ORACLE (test infrastructure only; never linked into the product library).
 *
 * Philox4x32-10 counter-based generator, written from the published
 * construction: J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11, section 5.
 *   round:  (c0,c1,c2,c3) -> (hi(M1*c2)^c1^k0, lo(M1*c2), hi(M0*c0)^c3^k1, lo(M0*c0))
 *   key schedule: k0 += 0x9E3779B9, k1 += 0xBB67AE85 between rounds.
 * Pinned by the Random123 known-answer vectors and by rocRAND's open-source
 * philox4x32_10 engine (tests/test_philox.py).
 *
 * The naja reference draws from cuRAND MRG32k3a (chr.cu:43-48, chr.cu:556-560),
 * a closed NVIDIA stream that is not reproduced. This build replaces it with
 * the counter layout below, shared by the oracle and the HIP kernel:
 *   key                      = (lo32(seed), hi32(seed) ^ variant)
 *   move block   (per group) : ctr = (lo32(t), hi32(t), group, 0)
 *       w0 -> move type, w1 -> coordinate i, w2 -> partner j / pair index,
 *       w3 -> bit 0 "which axis", bit 1 sign (pair mode 2); bit q sign of
 *             k-sparse entry q
 *   k-sparse blocks          : ctr = (lo32(t), hi32(t), group, 1 + q/4), word q%4
 *   alpha block  (per chain) : ctr = (lo32(t), hi32(t), chain/4, 0xFFFFFFFF), word chain%4
 */
#ifndef NAJA_ORACLE_PHILOX_H
#define NAJA_ORACLE_PHILOX_H
#include <stdint.h>

#define ORC_PHILOX_M0 0xD2511F53u
#define ORC_PHILOX_M1 0xCD9E8D57u
#define ORC_PHILOX_W0 0x9E3779B9u
#define ORC_PHILOX_W1 0xBB67AE85u
#define ORC_ALPHA_BLOCK 0xFFFFFFFFu

static inline void orc_philox4x32_10(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += ORC_PHILOX_W0; k1 += ORC_PHILOX_W1; }
        uint64_t p0 = (uint64_t)ORC_PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)ORC_PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Index in [0, n): the high word of w*n. Integer-only, so exact everywhere. */
static inline uint32_t orc_index(uint32_t w, uint32_t n)
{
    return (uint32_t)(((uint64_t)w * (uint64_t)n) >> 32);
}

/* Uniform in (0,1): odd 24-bit integer times 2^-24, exact in float. */
static inline float orc_uniform(uint32_t w)
{
    return (float)(((w >> 9) << 1) | 1u) * 5.9604644775390625e-08f;
}

/* Move-type threshold for probability p: w < floor(p * 2^32) (capped at 2^32). */
static inline uint64_t orc_threshold(double p)
{
    if (!(p > 0.0)) return 0;
    if (p >= 1.0) return 4294967296ull;
    return (uint64_t)(p * 4294967296.0);
}

#endif
