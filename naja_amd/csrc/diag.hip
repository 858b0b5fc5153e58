// GPU ESS and split-R-hat, replacing naja validate's host loops
// compute_ess / compute_split_rhat (mcmc.cpp:91-194) and their vendored
// kissfft dependency.
//
// ESS (mcmc.cpp:17-76): the samples are first gathered into contiguous
// series (one tiled transpose), then one workgroup per (variant, chain, dim)
// series of length n.
//   * mean summed in FP64, cast to float; series centred in float;
//   * autocovariance by a hand-written FFT: the real series of length
//     N = 2^ceil(log2(2n)) (zero padded; linear, not circular, correlation)
//     is packed as N/2 complex points z_j = x_2j + i x_2j+1 and transformed
//     in LDS by an in-place decimation-in-frequency FFT of radix-16 passes,
//     16 points per thread in registers (natural in, bit-reversed out;
//     twiddles from two small LDS tables); the real spectrum X_k is split
//     out, |X_k|^2 is
//     re-packed in place and inverted by a decimation-in-time FFT
//     (bit-reversed in, natural out), so no reordering pass is needed; the
//     interleaved result is the unnormalised autocorrelation, scaled by
//     1/(N n) as mcmc.cpp:49-53 scales kissfft's by 1/(2n n);
//   * truncated pair-sum tau (mcmc.cpp:66-72): the pair sums
//     rho_{2k-1} + rho_{2k} are formed in parallel, the truncation point is
//     the first negative pair (a workgroup minimum), and the pairs before it
//     are summed in parallel in FP64 (the reference adds them one by one in
//     float; the two agree to float rounding).
// Split-R-hat (mcmc.cpp:126-194): FP64 sums per (dim, half chain), taken by
// the first ESS pass while it reads each series, then the reference's
// B / W / var_hat formula per dim.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <vector>

#include "common.h"
#include "diag.h"

// Device memory of naja_diag_noalloc (see diag_launch_noalloc below): the
// per-call scratch, and the FFT twiddle tables of every transform size.
constexpr int kTwiddleEntries = 1024;
static __device__ __align__(256) unsigned char scratch[naja::kDiagNoAllocBytes];
static __device__ __align__(256) float2 scratch_twiddles[kTwiddleEntries];

namespace naja {

namespace {

// Threads per ESS workgroup: one per 16-point group of a radix-16 pass
// (H >= kRadix16MinH) or per 4-point group of a radix-4 pass (shorter
// transforms), between 64 and 1024.
constexpr int kDiagThreads = 1024;
constexpr int kRadix16MinH = 512;  // radix-16 passes from here (3 LDS passes at 512 points instead of 5)
int ess_threads(int H) { return std::max(64, std::min(kDiagThreads, H >= kRadix16MinH ? H / 16 : H / 4)); }
constexpr int kMaxHalf = 16384;  // N/2 complex points held in LDS (136 KB with the padding)

__device__ __forceinline__ int64_t sample_row(int layout, int s, int c, int nchains, int spc)
{
    return layout == 0 ? (int64_t)s * nchains + c : (int64_t)c * spc + s;
}

// Complex product with each component one multiply and one fused
// multiply-add (the library is built without FMA contraction elsewhere).
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x));
}

// Twiddles W^k = exp(-2 pi i k / N), N = 2H, from two small LDS tables:
// k = hi * B + lo, W^k = thi[hi] * tlo[lo] (B = 2^tb, both tables <= 128
// entries for H <= 16384), each entry rounded once from FP64 on the host.
// A pass over N' = N / 2^sh points reads W_{N'}^k = W_N^(k 2^sh) from the same
// tables.
struct Twiddles {
    const float2* tlo;
    const float2* thi;
    int tb;
    int sh;
    __device__ __forceinline__ float2 at(int k) const
    {
        const int e = k << sh;
        return cmul(thi[e >> tb], tlo[e & ((1 << tb) - 1)]);
    }
};

// The transform's LDS array with one float2 of padding after every 16
// (128 bytes): the power-of-two strides of the radix-16 passes then spread
// over all LDS banks instead of a few.
struct PaddedZ {
    float2* p;
    __device__ __forceinline__ float2& operator[](int i) const { return p[i + (i >> 4)]; }
    // Element m of the interleaved real view (x_0, y_0, x_1, y_1, ...).
    __device__ __forceinline__ float real(int m) const
    {
        const float2 v = (*this)[m >> 1];
        return (m & 1) ? v.y : v.x;
    }
};
__host__ __device__ constexpr size_t padded_lds_bytes(int H) { return sizeof(float2) * (size_t)(H + H / 16); }

int ess_threads(int H);

// Workgroups of the full-size redo pass, which loops over all series and
// redoes those the short pass could not settle: two rounds of what the CUs
// hold at once (by waves, 32 per CU, and by LDS, 160 KB per CU with about
// 4 KB of static arrays per workgroup). Dispatching one workgroup per series
// only to have most of them exit costs milliseconds for millions of short
// series.
unsigned ess_grid(int H, int64_t n_series)
{
    static const int n_cu = [] {
        int dev = 0, cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0)
            cu = 256;
        return cu;
    }();
    const int by_waves = std::max(1, 32 / (ess_threads(H) / 64));
    const int by_lds = std::max(1, (int)((160 * 1024) / (padded_lds_bytes(H) + 4096)));
    const int64_t g = (int64_t)n_cu * std::min(by_waves, by_lds) * 2;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(n_series, g));
}

// (x) * (-i) and (x) * (+i).
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }
__device__ __forceinline__ float2 mul_pi(float2 a) { return make_float2(-a.y, a.x); }
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// 4-point DFT in place, natural order (INV: the unnormalised inverse).
template <bool INV>
__device__ __forceinline__ void dft4(float2& u0, float2& u1, float2& u2, float2& u3)
{
    const float2 t0 = cadd(u0, u2), t1 = csub(u0, u2), t2 = cadd(u1, u3);
    const float2 t3 = INV ? mul_pi(csub(u1, u3)) : mul_mi(csub(u1, u3));
    u0 = cadd(t0, t2);
    u1 = cadd(t1, t3);
    u2 = csub(t0, t2);
    u3 = csub(t1, t3);
}

// x * W_16^k (conjugated for INV), k in {1, 2, 3, 4, 6, 9}.
template <bool INV, int K>
__device__ __forceinline__ float2 w16(float2 x)
{
    constexpr float c1 = 0.92387953251128674f, s1 = 0.38268343236508978f, h = 0.70710678118654752f;
    float2 w;
    if (K == 1) w = make_float2(c1, -s1);
    if (K == 2) w = make_float2(h, -h);
    if (K == 3) w = make_float2(s1, -c1);
    if (K == 4) return INV ? mul_pi(x) : mul_mi(x);
    if (K == 6) w = make_float2(-h, -h);
    if (K == 9) w = make_float2(-c1, s1);
    if (INV) w.y = -w.y;
    return cmul(x, w);
}

// 16-point DFT of x[0..15] (natural order) in registers as 4 x 4:
// DFT4 over m1 of x[4 m1 + m2], twiddle W_16^(m2 r1), DFT4 over m2.
// Output X_r ends in x[4 (r & 3) + (r >> 2)].
template <bool INV>
__device__ __forceinline__ void dft16(float2 (&x)[16])
{
#pragma unroll
    for (int m2 = 0; m2 < 4; ++m2) dft4<INV>(x[m2], x[4 + m2], x[8 + m2], x[12 + m2]);
    x[5] = w16<INV, 1>(x[5]);
    x[9] = w16<INV, 2>(x[9]);
    x[13] = w16<INV, 3>(x[13]);
    x[6] = w16<INV, 2>(x[6]);
    x[10] = w16<INV, 4>(x[10]);
    x[14] = w16<INV, 6>(x[14]);
    x[7] = w16<INV, 3>(x[7]);
    x[11] = w16<INV, 6>(x[11]);
    x[15] = w16<INV, 9>(x[15]);
#pragma unroll
    for (int r1 = 0; r1 < 4; ++r1) dft4<INV>(x[4 * r1], x[4 * r1 + 1], x[4 * r1 + 2], x[4 * r1 + 3]);
}

__host__ __device__ constexpr int bitrev4(int r) { return ((r & 1) << 3) | ((r & 2) << 1) | ((r & 4) >> 1) | ((r & 8) >> 3); }

// w[r] = W_L^(p0 r), r = 0..15, from w1 = W_L^p0 and w4 = W_L^(4 p0) (both
// from the tables; every other power is at most three products away).
__device__ __forceinline__ void twiddle_powers(float2 w1, float2 w4, float2 (&w)[16])
{
    w[0] = make_float2(1.0f, 0.0f);
    w[1] = w1;
    w[2] = cmul(w1, w1);
    w[3] = cmul(w[2], w1);
    w[4] = w4;
    w[8] = cmul(w4, w4);
    w[12] = cmul(w[8], w4);
#pragma unroll
    for (int b = 4; b < 16; b += 4)
#pragma unroll
        for (int k = 1; k < 4; ++k) w[b + k] = cmul(w[b], w[k]);
}

// One radix-16 pass of the in-place DIF transform over blocks of L = 16 q
// points (q = 2^logq): per block offset p0 < q, the DFT16 of the points
// p0 + m q, output r scaled by W_L^(p0 r) and stored at p0 + bitrev4(r) q.
// Passes at L = H, H/16, ... leave the spectrum in bit-reversed order.
//
// Padded offset of point base + m q from the padded base, for a group base
// = B 16 q + p0 (p0 < q): base + m q + ((base + m q) >> 4) = pad(base) +
// m q + ((m q) >> 4), because 16 q is a multiple of 16 and p0 never carries
// the low four bits of m q past a multiple of 16. The offsets depend on q
// only, so they are wave-uniform (scalar) and each point of a group is one
// add from the group's padded address.
__device__ __forceinline__ int pad_off(int m, int q) { return m * q + ((m * q) >> 4); }

// The group of block offset p0 (points at zb[pad_off(m, q)], m < 16, already
// in x; zb the padded address of the group's base): DFT16, twiddles
// W_L^(p0 r) (W_L^j = W_N^(j << sh)), bit-reversed stores.
__device__ __forceinline__ void dif16_group(float2* zb, float2 (&x)[16], int q, int p0, int sh, const Twiddles& tw)
{
    float2 w[16];
    dft16<false>(x);
    twiddle_powers(tw.at(p0 << sh), tw.at((4 * p0) << sh), w);
#pragma unroll
    for (int r = 0; r < 16; ++r) zb[pad_off(bitrev4(r), q)] = cmul(x[4 * (r & 3) + (r >> 2)], w[r]);
}

__device__ void dif16_pass(const PaddedZ& z, int H, int logq, int logN, const Twiddles& tw)
{
    const int q = 1 << logq;
    const int sh = logN - logq - 4;  // W_L^j = W_N^(j << sh)
    for (int g = threadIdx.x; g < H / 16; g += blockDim.x) {
        const int p0 = g & (q - 1);
        const int base = ((g >> logq) << (logq + 4)) + p0;
        float2* zb = z.p + base + (base >> 4);
        float2 x[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) x[m] = zb[pad_off(m, q)];
        dif16_group(zb, x, q, p0, sh, tw);
    }
}

// The inverse of dif16_pass (times 16): conjugate twiddles on the
// bit-reversed inputs, inverse DFT16, natural-order outputs.
__device__ void dit16_pass(const PaddedZ& z, int H, int logq, int logN, const Twiddles& tw)
{
    const int q = 1 << logq;
    const int sh = logN - logq - 4;
    for (int g = threadIdx.x; g < H / 16; g += blockDim.x) {
        const int p0 = g & (q - 1);
        const int base = ((g >> logq) << (logq + 4)) + p0;
        float2* zb = z.p + base + (base >> 4);
        float2 x[16], w[16];
        twiddle_powers(tw.at(p0 << sh), tw.at((4 * p0) << sh), w);
#pragma unroll
        for (int r = 0; r < 16; ++r) x[r] = cmul(zb[pad_off(bitrev4(r), q)], conjf2(w[r]));
        dft16<true>(x);
#pragma unroll
        for (int m = 0; m < 16; ++m) zb[pad_off(m, q)] = x[4 * (m & 3) + (m >> 2)];
    }
}

// Radix-4 passes of the same scheme for short transforms (H < 512, where
// H / 16 groups would leave most of a 64-thread workgroup idle): per block
// of L = 4 q points, DFT4 of p0 + m q, output r scaled by W_L^(p0 r) and
// stored at p0 + bitrev2(r) q (bitrev2 swaps outputs 1 and 2).
__device__ void dif4_pass(const PaddedZ& z, int H, int logq, int logN, const Twiddles& tw)
{
    const int q = 1 << logq;
    const int sh = logN - logq - 2;
    for (int g = threadIdx.x; g < H / 4; g += blockDim.x) {
        const int p0 = g & (q - 1);
        const int base = ((g >> logq) << (logq + 2)) + p0;
        float2 x0 = z[base], x1 = z[base + q], x2 = z[base + 2 * q], x3 = z[base + 3 * q];
        dft4<false>(x0, x1, x2, x3);
        const float2 w1 = tw.at(p0 << sh), w2 = cmul(w1, w1), w3 = cmul(w2, w1);
        z[base] = x0;
        z[base + 2 * q] = cmul(x1, w1);
        z[base + q] = cmul(x2, w2);
        z[base + 3 * q] = cmul(x3, w3);
    }
}

__device__ void dit4_pass(const PaddedZ& z, int H, int logq, int logN, const Twiddles& tw)
{
    const int q = 1 << logq;
    const int sh = logN - logq - 2;
    for (int g = threadIdx.x; g < H / 4; g += blockDim.x) {
        const int p0 = g & (q - 1);
        const int base = ((g >> logq) << (logq + 2)) + p0;
        const float2 w1 = tw.at(p0 << sh), w2 = cmul(w1, w1), w3 = cmul(w2, w1);
        float2 y0 = z[base], y1 = cmul(z[base + 2 * q], conjf2(w1)), y2 = cmul(z[base + q], conjf2(w2)),
               y3 = cmul(z[base + 3 * q], conjf2(w3));
        dft4<true>(y0, y1, y2, y3);
        z[base] = y0;
        z[base + q] = y1;
        z[base + 2 * q] = y2;
        z[base + 3 * q] = y3;
    }
}


// The last DIF pass (first DIT pass) when logH is not a multiple of 4:
// DFTs of 2^rem contiguous points, no twiddles; DIF stores output r at
// bitrev_rem(r), DIT reads input r from there.
template <bool INV>
__device__ void tail_pass(const PaddedZ& z, int H, int rem)
{
    const int L = 1 << rem;
    for (int g = threadIdx.x; g < H / L; g += blockDim.x) {
        const int o = g * L;
        if (rem == 1) {
            const float2 u = z[o], v = z[o + 1];
            z[o] = cadd(u, v);
            z[o + 1] = csub(u, v);
        } else if (rem == 2) {
            // bitrev2: slots 0, 1, 2, 3 hold outputs 0, 2, 1, 3
            float2 u0 = z[o], u1 = z[o + (INV ? 2 : 1)], u2 = z[o + (INV ? 1 : 2)], u3 = z[o + 3];
            dft4<INV>(u0, u1, u2, u3);
            z[o] = u0;
            z[o + (INV ? 1 : 2)] = u1;
            z[o + (INV ? 2 : 1)] = u2;
            z[o + 3] = u3;
        } else {
            // 8 points: DFT4 over m1 of x[2 m1 + m2], twiddle W_8^(m2 r1),
            // DFT2 over m2; output r1 + 4 r2; bitrev3 = {0, 4, 2, 6, 1, 5, 3, 7}.
            constexpr int br3[8] = {0, 4, 2, 6, 1, 5, 3, 7};
            float2 x[8];
#pragma unroll
            for (int m = 0; m < 8; ++m) x[m] = z[o + (INV ? br3[m] : m)];
            dft4<INV>(x[0], x[2], x[4], x[6]);
            dft4<INV>(x[1], x[3], x[5], x[7]);
            constexpr float h = 0.70710678118654752f;
            // x[2 r1 + 1] *= W_8^r1
            x[3] = cmul(x[3], make_float2(h, INV ? h : -h));
            x[5] = INV ? mul_pi(x[5]) : mul_mi(x[5]);
            x[7] = cmul(x[7], make_float2(-h, INV ? h : -h));
            float2 X[8];
#pragma unroll
            for (int r1 = 0; r1 < 4; ++r1) {
                X[r1] = cadd(x[2 * r1], x[2 * r1 + 1]);
                X[r1 + 4] = csub(x[2 * r1], x[2 * r1 + 1]);
            }
#pragma unroll
            for (int r = 0; r < 8; ++r) z[o + (INV ? r : br3[r])] = X[r];
        }
    }
}

// Points of the last DIF pass (first DIT pass) that the radix-16 or radix-4
// passes leave over: 2^tail_bits contiguous points.
__device__ __forceinline__ int tail_bits(int H, int logH) { return H >= kRadix16MinH ? (logH & 3) : (logH & 1); }

// In-place FFT over H points, decimation in frequency (natural in,
// bit-reversed out): radix-16 passes at block sizes H, H/16, ... while
// >= 16, then one radix-2/4/8 pass. Each pass holds 16 (or 2^rem) points
// per thread in registers, so an 8192-point transform takes 4 LDS passes.
// first_done: the caller ran the first radix-16 pass; skip_tail: the caller
// folds a radix-2 last pass into the split that follows.
__device__ void fft_dif(const PaddedZ& z, int H, int logH, const Twiddles& tw, bool first_done, bool skip_tail)
{
    const int logN = logH + 1;
    int logL = logH;
    if (H < kRadix16MinH) {
        for (; logL >= 2; logL -= 2) {
            dif4_pass(z, H, logL - 2, logN, tw);
            __syncthreads();
        }
        if (logL == 1 && !skip_tail) {
            tail_pass<false>(z, H, 1);
            __syncthreads();
        }
        return;
    }
    if (first_done) logL -= 4;
    for (; logL >= 4; logL -= 4) {
        dif16_pass(z, H, logL - 4, logN, tw);
        __syncthreads();
    }
    if (logL > 0 && !skip_tail) {
        tail_pass<false>(z, H, logL);
        __syncthreads();
    }
}

// Unnormalised inverse of fft_dif's output order (bit-reversed in, natural
// out): the passes of fft_dif undone in reverse order (skip_head: the
// caller ran the radix-2 first pass).
__device__ void ifft_dit(const PaddedZ& z, int H, int logH, const Twiddles& tw, bool skip_head)
{
    const int logN = logH + 1;
    if (H < kRadix16MinH) {
        const int rem2 = logH & 1;
        if (rem2 && !skip_head) {
            tail_pass<true>(z, H, 1);
            __syncthreads();
        }
        for (int logL = rem2 + 2; logL <= logH; logL += 2) {
            dit4_pass(z, H, logL - 2, logN, tw);
            __syncthreads();
        }
        return;
    }
    const int rem = logH & 3;
    if (rem > 0 && !skip_head) {
        tail_pass<true>(z, H, rem);
        __syncthreads();
    }
    for (int logL = rem + 4; logL <= logH; logL += 4) {
        dit16_pass(z, H, logL - 4, logN, tw);
        __syncthreads();
    }
}

__device__ __forceinline__ int bitrev(int x, int bits)
{
    return (int)(__builtin_bitreverse32((uint32_t)x) >> (32 - bits));
}

// Gathers every (variant, chain) series into rows of stride spcp (spc
// rounded up to 64 floats, so every row starts on a 256-byte boundary):
// out[((v * nchains + c) * dim + d) * spcp + s] = sample s of chain c, dim d.
// A workgroup moves kTs samples x td dims of one chain through LDS (row
// stride td + 1, so the column reads hit distinct banks) and writes, per dim,
// kTs consecutive samples (one aligned 256-byte run). Block ids run over
// (chain, sample tile) so that the chains of 8 consecutive sample tiles are
// 8 apart: they land on the same XCD (blocks go to XCDs round robin) close
// in time, and the cache lines their interleaved sample rows share are read
// from HBM once.
constexpr int kTs = 64;
constexpr int kTd = 128;

// A plain store that zeroes the redo count (a memset would not be safe
// inside a captured graph).
__global__ void zero_count_kernel(int* __restrict__ count)
{
    if (threadIdx.x == 0) *count = 0;
}

__global__ __launch_bounds__(256) void series_transpose_kernel(const float* __restrict__ samples, int dim,
                                                              int nchains, int spc, int spcp, int layout,
                                                              int64_t vstride, int s_tiles, int d_tiles,
                                                              float* __restrict__ out, int* __restrict__ redo)
{
    __shared__ float tile[kTs * (kTd + 1)];
    const int64_t b = blockIdx.x;
    // The ESS short pass that follows appends to this list (count first): a
    // plain store here, not a memset, keeps the launch sequence capturable.
    if (b == 0 && threadIdx.x == 0 && redo) *redo = 0;
    const int64_t grp = b / (8 * nchains);
    const int within = (int)(b % (8 * nchains));
    const int c = within / 8;
    const int s_groups = (s_tiles + 7) / 8;
    const int st = (int)(grp % s_groups) * 8 + within % 8;
    const int64_t rest = grp / s_groups;
    const int dt = (int)(rest % d_tiles);
    const int v = (int)(rest / d_tiles);
    if (st >= s_tiles) return;  // padding blocks of the last sample group
    const int s0 = st * kTs, d0 = dt * kTd;
    const int tdv = min(kTd, dim - d0);
    const float* src = samples + (int64_t)v * vstride;
    // Reads: lanes along dims of one sample row (runs of tdv floats), two
    // rows per pass of the 256 threads; all loads issued before the LDS
    // stores.
    constexpr int kRows = kTs / (256 / kTd);
    const int dr = threadIdx.x % kTd, sr = threadIdx.x / kTd;
    float vals[kRows];
#pragma unroll
    for (int k = 0; k < kRows; ++k) {
        const int ss = s0 + sr + k * (256 / kTd);
        vals[k] = (dr < tdv && ss < spc) ? src[sample_row(layout, ss, c, nchains, spc) * dim + d0 + dr] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < kRows; ++k) tile[(sr + k * (256 / kTd)) * (kTd + 1) + dr] = vals[k];
    __syncthreads();
    // Writes: lanes along samples of one series.
    const int s = threadIdx.x % kTs;
    const int ss = s0 + s;
    if (ss >= spc) return;  // no barrier follows
    for (int d = threadIdx.x / kTs; d < tdv; d += 256 / kTs)
        out[(((int64_t)v * nchains + c) * dim + d0 + d) * spcp + ss] = tile[s * (kTd + 1) + d];
}

// Split of the packed spectrum for the pair {k, H-k}: Z_k = zk, Z_{H-k} = zb,
// wk = W^k, wb = W^(H-k) (W^H = -1 for k = 0). X_k = E_k + W^k O_k with
// E = (Z_k + conj Z_{H-k})/2, O = (Z_k - conj Z_{H-k})/(2i), P = |X|^2, and
// the re-packed Z'_k = A_k + i B_k, A_k = P_k + P_{H-k},
// B_k = (P_k - P_{H-k}) e^{+2 pi i k/N} = (P_k - P_{H-k}) conj(W^k); likewise
// Z'_{H-k}.
__device__ __forceinline__ void split_pair(float2 zk, float2 zb, float2 wk, float2 wb, float2& ok, float2& ob)
{
    const float2 Ek = make_float2(0.5f * (zk.x + zb.x), 0.5f * (zk.y - zb.y));
    const float2 Ok = make_float2(0.5f * (zk.y + zb.y), -0.5f * (zk.x - zb.x));
    const float2 Eb = make_float2(0.5f * (zb.x + zk.x), 0.5f * (zb.y - zk.y));
    const float2 Ob = make_float2(0.5f * (zb.y + zk.y), -0.5f * (zb.x - zk.x));
    const float2 Xk = cadd(Ek, cmul(wk, Ok));
    const float2 Xb = cadd(Eb, cmul(wb, Ob));
    const float Pk = Xk.x * Xk.x + Xk.y * Xk.y;
    const float Pb = Xb.x * Xb.x + Xb.y * Xb.y;
    const float Dk = Pk - Pb, Db = Pb - Pk;
    ok = make_float2((Pk + Pb) + Dk * wk.y, Dk * wk.x);
    ob = make_float2((Pb + Pk) + Db * wb.y, Db * wb.x);
}

// ESS of one contiguous series per workgroup (blockIdx.x = (v * nchains + c) * dim + d).
// mode 0: the full transform (N >= 2n, every lag). mode 1: a shorter
// transform, N < 2n: zero padding to N keeps the correlation exact for lags
// k <= N - n, so the result is final when the truncation point (the first
// negative pair, lags 2k-1 and 2k) lies within them; otherwise the series is
// marked -1 for mode 2, which redoes only the marked series at full size.
// Workgroup b takes series b, b + gridDim.x, ... one at a time, with the
// twiddle tables loaded once: the short and single passes run one workgroup
// per series, the full-size redo pass (mode 2) a bounded grid that skips the
// series the short pass settled.
__global__ __launch_bounds__(kDiagThreads) void ess_series_kernel(const float* __restrict__ series, int dim,
                                                                  int nchains, int spc, int spcp, int logH,
                                                                  const float2* __restrict__ tw_tables, int tb,
                                                                  int tw_shift, int mode,
                                                                  float* __restrict__ ess_chain, int half,
                                                                  double* __restrict__ mom, int64_t n_series,
                                                                  int* __restrict__ redo)
{
    extern __shared__ float2 z_lds[];
    const PaddedZ z{z_lds};
    __shared__ float2 tlo_s[128], thi_s[128];
    __shared__ double red[5][kDiagThreads / 64];
    __shared__ float meanf_s;
    __shared__ int kfirst[kDiagThreads / 64];
    __shared__ double psum[kDiagThreads / 64];

    const int n = spc;
    const int H = 1 << logH;
    const int N = 2 * H;
    const int tid = threadIdx.x;
    const int nlo = 1 << tb, nhi = (H << tw_shift) >> tb;
    for (int i = tid; i < nlo; i += blockDim.x) tlo_s[i] = tw_tables[i];
    for (int i = tid; i < nhi; i += blockDim.x) thi_s[i] = tw_tables[nlo + i];
    const Twiddles tw{tlo_s, thi_s, tb, tw_shift};
    __syncthreads();

    // Every early return below is uniform across the workgroup.
    auto one = [&](const int64_t sid) {
        const int d = (int)(sid % dim);
        const int vc = (int)(sid / dim);
        const int c = vc % nchains;
        const int v = vc / nchains;
        float* out = ess_chain + (int64_t)v * dim * nchains + (int64_t)d * nchains + c;
        const float* x = series + sid * spcp;

        if (mode == 2 && *out >= 0.0f) return;  // settled by the short pass
        if (n < 4) {
            if (tid == 0) *out = (float)n;
            return;
        }

        // One pass over the series: raw values into the packed complex layout
        // z_j = x_2j + i x_2j+1 (zero padded), the FP64 sum alongside
        // (mcmc.cpp:25-27), and the split-R-hat moments of its two halves
        // (sum and sum of squares of samples [0, half) and [half, 2 half) in FP64,
        // mcmc.cpp:150-165), written for the R-hat pass by the first ESS pass.
        // A thread owns at most kPerThread points (H <= 16 * blockDim); their
        // loads are all issued before any of them is used. With exactly 16
        // (radix-16 passes, H = 16 * blockDim) thread tid holds the points
        // tid + m H/16, which is its group of the first radix-16 pass: the
        // points stay in registers, are centred there and go through that pass
        // before anything is written to LDS.
        const bool regs_first = H >= kRadix16MinH && 16 * (int)blockDim.x == H;
        double part[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // total, half-0 sum, sq, half-1 sum, sq
        constexpr int kPerThread = 16;
        const int nq = (H + (int)blockDim.x - 1) / (int)blockDim.x;  // <= kPerThread, uniform
        float2 xv[kPerThread];
#pragma unroll
        for (int q = 0; q < kPerThread; ++q) {
            xv[q] = make_float2(0.0f, 0.0f);
            if (q < nq) {
                const int s0 = 2 * (tid + q * (int)blockDim.x);
                if (s0 + 1 < n)
                    xv[q] = *reinterpret_cast<const float2*>(x + s0);  // rows are 256-byte aligned, s0 even
                else if (s0 < n)
                    xv[q].x = x[s0];
            }
        }
#pragma unroll
        for (int q = 0; q < kPerThread; ++q) {
            const int j = tid + q * (int)blockDim.x;
            if (q >= nq || j >= H) continue;
            const int s0 = 2 * j, s1 = 2 * j + 1;
            const float x0 = xv[q].x, x1 = xv[q].y;
            const double d0 = (double)x0, d1 = (double)x1;
            part[0] += d0 + d1;
            if (s0 < 2 * half) {
                const int h = s0 >= half;
                part[1 + 2 * h] += d0;
                part[2 + 2 * h] += d0 * d0;
            }
            if (s1 < 2 * half) {
                const int h = s1 >= half;
                part[1 + 2 * h] += d1;
                part[2 + 2 * h] += d1 * d1;
            }
            if (!regs_first) z[j] = make_float2(x0, x1);
        }
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            for (int off = 32; off > 0; off >>= 1) part[q] += __shfl_xor(part[q], off);
            if ((tid & 63) == 0) red[q][tid >> 6] = part[q];
        }
        __syncthreads();
        if (tid < 5) {
            double sum = 0.0;
            for (int w = 0; w < (int)(blockDim.x / 64); ++w) sum += red[tid][w];
            if (tid == 0) meanf_s = (float)(sum / n);
            else if (mom && mode != 2) {
                // mom[v][2c + h][d][moment]
                const int h = (tid - 1) >> 1, q = (tid - 1) & 1;
                mom[(((int64_t)v * 2 * nchains + 2 * c + h) * dim + d) * 2 + q] = sum;
            }
        }
        __syncthreads();
        const float meanf = meanf_s;
        // Centre in float; the padding stays zero.
        if (regs_first) {
            float2 xr[16];
#pragma unroll
            for (int m = 0; m < 16; ++m) {
                const int j = tid + m * (int)blockDim.x;
                xr[m] = xv[m];
                if (2 * j < n) xr[m].x -= meanf;
                if (2 * j + 1 < n) xr[m].y -= meanf;
            }
            // first pass: L = H, q = H/16, p0 = base = tid, W_L^j = W_N^(2j)
            dif16_group(z.p + tid + (tid >> 4), xr, H / 16, tid, 1, tw);
        } else {
            for (int j = tid; j < H; j += blockDim.x) {
                const int s0 = 2 * j, s1 = 2 * j + 1;
                float2 p = z[j];
                if (s0 < n) p.x -= meanf;
                if (s1 < n) p.y -= meanf;
                z[j] = p;
            }
        }
        __syncthreads();
        // A radix-2 last pass is folded into the split below.
        const bool fuse2 = tail_bits(H, logH) == 1;
        fft_dif(z, H, logH, tw, regs_first, fuse2);  // Z_k now sits at z[bitrev(k)] (after the radix-2 stage)

        // Split the real spectrum, form P = |X|^2 and re-pack it for the
        // half-size inverse (split_pair), written back to z[bitrev(k)] (the
        // slot it came from, which is where the DIT inverse wants it).
        if (fuse2) {
            // The DIF radix-2 last pass, the split and the DIT radix-2 first
            // pass in one sweep. Item k (0 <= k <= H/4) owns the slot pairs
            // a = bitrev(k)/2 and b = bitrev(H/2 - k)/2: after the radix-2
            // stage on them, slots 2a, 2a+1, 2b, 2b+1 hold Z_k, Z_{k+H/2},
            // Z_{H/2-k} and Z_{H-k}, two split pairs {k, H-k} and
            // {H/2+k, H/2-k} (k = 0: b = a with pairs {0} and {H/2};
            // k = H/4: b = a with the one pair {H/4, 3H/4}). The twiddles
            // come from W^k: W^(H-k) = -conj W^k, W^(H/2+k) = -i W^k,
            // W^(H/2-k) = -i conj W^k.
            // Threads walk the slot pairs in order rather than k (whose
            // slots are scattered over the LDS banks): a = bitrev(k) / 2 is k
            // reversed in logH - 1 bits, and reversing H/2 - k = -k mod H/2
            // gives b = a ^ (2^msb(a) - 1) for a > 0. Item 0 is k = 0 (a = 0),
            // item 1 is k = H/4 (a = 1); item i >= 2 takes the pair {c, c ^
            // (2^msb(c) - 1)} with c = i' + 2^msb(i'), i' = i - 1 (each pair
            // of [2^j, 2^(j+1)) once), and its even member as a, which is the
            // one with k < H/4 (k's top two bits are a's low bit and 0).
            for (int it = tid; it <= H / 4; it += blockDim.x) {
                int a = it, b = it;
                if (it >= 2) {
                    const int ip = it - 1;
                    const int hb = 1 << (31 - __builtin_clz(ip));
                    const int c = ip + hb, cp = c ^ (2 * hb - 1);
                    a = (c & 1) ? cp : c;
                    b = (c & 1) ? c : cp;
                }
                const int k = bitrev(2 * a, logH);
                const float2 ua = z[2 * a], va = z[2 * a + 1], ub = z[2 * b], vb = z[2 * b + 1];
                const float2 Zk = cadd(ua, va), Zkh = csub(ua, va), Zm = cadd(ub, vb), Zmh = csub(ub, vb);
                const float2 w = tw.at(k);
                const float2 wmc = make_float2(-w.x, w.y);   // W^(H-k)
                const float2 wh = make_float2(w.y, -w.x);    // W^(H/2+k)
                const float2 whc = make_float2(-w.y, -w.x);  // W^(H/2-k)
                float2 o_k, o_mh, o_kh, o_m;  // Z' of k, H-k, H/2+k, H/2-k
                split_pair(Zk, k == 0 ? Zk : Zmh, w, wmc, o_k, o_mh);
                split_pair(Zkh, k == 0 ? Zkh : Zm, wh, whc, o_kh, o_m);
                // DIT radix-2 first pass on each slot pair.
                z[2 * a] = cadd(o_k, o_kh);
                z[2 * a + 1] = csub(o_k, o_kh);
                if (b != a) {
                    z[2 * b] = cadd(o_m, o_mh);
                    z[2 * b + 1] = csub(o_m, o_mh);
                }
            }
        } else {
            // Each thread owns the pair {k, H-k}, so no other thread touches
            // those two slots. Threads walk the slots rather than k, so that
            // the z[pk] accesses of a wavefront are adjacent: the even slots
            // hold every k < H/2 (the top bit of k is the low bit of its slot)
            // and slot 1 holds k = H/2.
            for (int i = tid; i <= H / 2; i += blockDim.x) {
                const int pk = (i < H / 2) ? 2 * i : 1;
                const int k = bitrev(pk, logH);
                const int kb = (H - k) & (H - 1);
                const int pb = bitrev(kb, logH);
                // The partner of k = 0 is X_H, with W^H = -1.
                const float2 wk = tw.at(k);
                const float2 wb = (k == 0) ? make_float2(-1.0f, 0.0f) : tw.at(H - k);
                float2 ok, ob;
                split_pair(z[pk], z[pb], wk, wb, ok, ob);
                z[pk] = ok;
                if (k != 0 && kb != k) z[pb] = ob;
            }
        }
        __syncthreads();
        ifft_dit(z, H, logH, tw, fuse2);

        // acov[k] = raw[k] / (N n); raw is the interleaved real view of z.
        const float norm = 1.0f / ((float)N * (float)n);
        const float var = z.real(0) * norm;
        const int kend = n / 2;  // pairs k = 1 .. kend-1
        if (var < 1e-30f) {
            if (tid == 0) *out = 1.0f;
            return;  // uniform across the workgroup: no barrier follows for anyone
        }
        // Pairs whose lags are exact in this transform: 2k <= N - n.
        const int klim = (mode == 1) ? min(kend, (N - n) / 2 + 1) : kend;
        int kneg = klim;
        for (int k = 1 + tid; k < klim; k += blockDim.x) {
            const float pair_sum = (z.real(2 * k - 1) * norm) / var + (z.real(2 * k) * norm) / var;
            if (pair_sum < 0.0f) {
                kneg = k;
                break;
            }
        }
        for (int off = 32; off > 0; off >>= 1) kneg = min(kneg, __shfl_xor(kneg, off));
        if ((tid & 63) == 0) kfirst[tid >> 6] = kneg;
        __syncthreads();
        int K = klim;
        for (int w = 0; w < (int)(blockDim.x / 64); ++w) K = min(K, kfirst[w]);
        if (K == klim && klim < kend) {
            // Truncation beyond the exact lags: redo at full size (listed for
            // the redo pass, which then visits only these series).
            if (tid == 0) {
                *out = -1.0f;
                if (redo) redo[1 + atomicAdd(redo, 1)] = (int)sid;
            }
            return;
        }
        double acc = 0.0;
        for (int k = 1 + tid; k < K; k += blockDim.x)
            acc += (double)((z.real(2 * k - 1) * norm) / var + (z.real(2 * k) * norm) / var);
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        if ((tid & 63) == 0) psum[tid >> 6] = acc;
        __syncthreads();
        if (tid == 0) {
            double sum = 0.0;
            for (int w = 0; w < (int)(blockDim.x / 64); ++w) sum += psum[w];
            const float tau = fmaxf(1.0f + 2.0f * (float)sum, 1.0f);
            *out = (float)n / tau;
        }
    };

    if (mode == 2 && redo) {
        // The series the short pass listed (their order is immaterial: each
        // result depends on its series only).
        const int cnt = *redo;
        for (int i = blockIdx.x; i < cnt; i += gridDim.x) {
            one(redo[1 + i]);
            __syncthreads();  // LDS is reused by the next series
        }
        return;
    }
    for (int64_t sid = blockIdx.x; sid < n_series; sid += gridDim.x) {
        one(sid);
        __syncthreads();  // LDS is reused by the next series
    }
}

__global__ void ess_min_kernel(const float* __restrict__ ess_chain, int dim, int nchains, float* __restrict__ ess)
{
    ess_chain += (int64_t)blockIdx.y * dim * nchains;
    ess += (int64_t)blockIdx.y * dim;
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= dim) return;
    float m = 3.402823466e+38f;
    for (int c = 0; c < nchains; ++c) m = fminf(m, ess_chain[(int64_t)d * nchains + c]);
    ess[d] = m;
}

// mcmc.cpp:150-186 per dim.
__global__ void rhat_final_kernel(const double* __restrict__ mom, int dim, int nchains, int half,
                                  float* __restrict__ rhat)
{
    mom += (int64_t)blockIdx.y * 2 * nchains * dim * 2;
    rhat += (int64_t)blockIdx.y * dim;
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= dim) return;
    const int m = 2 * nchains;
    double grand = 0.0;
    for (int j = 0; j < m; ++j) grand += mom[((int64_t)j * dim + d) * 2] / half;
    const double gm = grand / m;
    double B = 0.0, W = 0.0;
    for (int j = 0; j < m; ++j) {
        const double S = mom[((int64_t)j * dim + d) * 2];
        const double S2 = mom[((int64_t)j * dim + d) * 2 + 1];
        const double mean = S / half;
        double var = (S2 / half) - mean * mean;
        var = var * half / (half - 1);
        const double df = mean - gm;
        B += df * df;
        W += var;
    }
    B *= (double)half / (m - 1);
    W /= m;
    const double var_hat = ((half - 1.0) / half) * W + B / half;
    rhat[d] = (float)((W > 1e-30) ? sqrt(var_hat / W) : 1.0);
}

// Per variant: the minimum over dims of ess and the maximum of rhat.
__global__ __launch_bounds__(256) void variant_summary_kernel(const float* __restrict__ ess,
                                                             const float* __restrict__ rhat, int dim,
                                                             float* __restrict__ ess_min, float* __restrict__ rhat_max)
{
    __shared__ float smin[256], smax[256];
    const int v = blockIdx.x, tid = threadIdx.x;
    float mn = 3.402823466e+38f, mx = -3.402823466e+38f;
    for (int d = tid; d < dim; d += 256) {
        mn = fminf(mn, ess[(int64_t)v * dim + d]);
        mx = fmaxf(mx, rhat[(int64_t)v * dim + d]);
    }
    smin[tid] = mn;
    smax[tid] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            smin[tid] = fminf(smin[tid], smin[tid + s]);
            smax[tid] = fmaxf(smax[tid], smax[tid + s]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        if (ess_min) ess_min[v] = smin[0];
        if (rhat_max) rhat_max[v] = smax[0];
    }
}

}  // namespace

namespace {

// ---------------------------------------------------------------------------
// Chains longer than the LDS transform holds (N/2 > kMaxHalf complex points):
// the same computation with the transforms in global memory, series in
// batches. Per series: FP64 moments and the float mean (long_moments_kernel),
// the centred packed pairs (long_pack_kernel), the forward complex FFT of
// length H by Stockham passes of radix 4 (one radix-2 pass when log2 H is
// odd), natural order in and out (stockham_kernel; Govindaraju et al., "High
// performance discrete Fourier transforms on graphics processors", SC 2008),
// the real-spectrum split and power-spectrum re-pack of split_pair
// (long_split_kernel), the inverse passes, and the truncated pair sum
// (long_trunc_kernel) with the short kernel's arithmetic.
// ---------------------------------------------------------------------------
constexpr int kLongThreads = 256;

__global__ __launch_bounds__(kLongThreads) void long_moments_kernel(const float* __restrict__ series, int dim,
                                                                   int nchains, int n, int spcp, int half,
                                                                   int64_t s0, float* __restrict__ meanf,
                                                                   double* __restrict__ mom)
{
    __shared__ double red[5][kLongThreads / 64];
    const int64_t sid = s0 + blockIdx.x;
    const float* x = series + sid * spcp;
    double part[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int t = threadIdx.x; t < n; t += kLongThreads) {
        const double v = (double)x[t];
        part[0] += v;
        if (t < 2 * half) {
            const int h = t >= half;
            part[1 + 2 * h] += v;
            part[2 + 2 * h] += v * v;
        }
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        for (int off = 32; off > 0; off >>= 1) part[q] += __shfl_xor(part[q], off);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = part[q];
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        double sum = 0.0;
        for (int w = 0; w < kLongThreads / 64; ++w) sum += red[threadIdx.x][w];
        const int d = (int)(sid % dim);
        const int vc = (int)(sid / dim);
        const int c = vc % nchains, v = vc / nchains;
        if (threadIdx.x == 0) {
            meanf[blockIdx.x] = (float)(sum / n);
        } else {
            const int h = (threadIdx.x - 1) >> 1, q = (threadIdx.x - 1) & 1;
            mom[(((int64_t)v * 2 * nchains + 2 * c + h) * dim + d) * 2 + q] = sum;
        }
    }
}

// z_j = (x_2j - mean) + i (x_2j+1 - mean), zero beyond the series, j < H.
__global__ void long_pack_kernel(const float* __restrict__ series, int n, int spcp, int H, int64_t s0,
                                 const float* __restrict__ meanf, float2* __restrict__ z)
{
    const int b = blockIdx.y;
    const float* x = series + (s0 + b) * spcp;
    const float mu = meanf[b];
    float2* zb = z + (int64_t)b * H;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < H; j += gridDim.x * blockDim.x) {
        const int t0 = 2 * j, t1 = 2 * j + 1;
        zb[j] = make_float2(t0 < n ? x[t0] - mu : 0.0f, t1 < n ? x[t1] - mu : 0.0f);
    }
}

// One Stockham pass of radix R over a batch of length-H transforms: butterfly
// j (< H/R) reads src[j + r H/R], multiplies by W_{Ns R}^{r (j mod Ns)}
// (from tw[t] = W_H^t, conjugated for the inverse), transforms, and writes
// dst[(j / Ns) Ns R + (j mod Ns) + r Ns].
template <int R>
__global__ void stockham_kernel(const float2* __restrict__ src, float2* __restrict__ dst, int H, int Ns,
                                const float2* __restrict__ tw, int inverse)
{
    const int b = blockIdx.y;
    const float2* sb = src + (int64_t)b * H;
    float2* db = dst + (int64_t)b * H;
    const int nbf = H / R;
    const int step = H / (Ns * R);  // W_{Ns R}^e = W_H^(e step)
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nbf; j += gridDim.x * blockDim.x) {
        const int jm = j % Ns;
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float2 a = sb[j + r * nbf];
            if (r > 0) {
                float2 w = tw[r * jm * step];
                if (inverse) w.y = -w.y;
                a = cmul(a, w);
            }
            v[r] = a;
        }
        float2 o[R];
        if (R == 2) {
            o[0] = cadd(v[0], v[1]);
            o[1] = csub(v[0], v[1]);
        } else {
            // 4-point DFT; the forward kernel is W_4 = -i, the inverse +i.
            const float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
            const float2 a2 = cadd(v[1], v[3]), d3 = csub(v[1], v[3]);
            const float2 b3 = inverse ? make_float2(-d3.y, d3.x) : make_float2(d3.y, -d3.x);  // (+/-i)(v1 - v3)
            o[0] = cadd(a0, a2);
            o[1 % R] = cadd(a1, b3);
            o[2 % R] = csub(a0, a2);
            o[3 % R] = csub(a1, b3);
        }
        const int base = (j / Ns) * Ns * R + jm;
#pragma unroll
        for (int r = 0; r < R; ++r) db[base + r * Ns] = o[r];
    }
}

// Y_k and Y_{H-k} from Z_k, Z_{H-k} (split_pair), k in [0, H/2]; wN[k] = W_N^k.
__global__ void long_split_kernel(const float2* __restrict__ Z, float2* __restrict__ Y, int H,
                                  const float2* __restrict__ wN)
{
    const int b = blockIdx.y;
    const float2* zb = Z + (int64_t)b * H;
    float2* yb = Y + (int64_t)b * H;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k <= H / 2; k += gridDim.x * blockDim.x) {
        const int kb = (H - k) & (H - 1);
        const float2 wk = wN[k];
        const float2 wb = (k == 0) ? make_float2(-1.0f, 0.0f) : wN[H - k];  // the partner of 0 is X_H
        float2 ok, ob;
        split_pair(zb[k], zb[kb], wk, wb, ok, ob);
        yb[k] = ok;
        if (k != 0 && kb != k) yb[kb] = ob;
    }
}

// The truncated pair sum of one series (one workgroup) from the inverse
// transform R (interleaved real view: lag m at R[m/2].x or .y), with the
// short kernel's arithmetic.
__global__ __launch_bounds__(kLongThreads) void long_trunc_kernel(const float2* __restrict__ Rz, int H, int n,
                                                                 int dim, int nchains, int64_t s0,
                                                                 float* __restrict__ ess_chain)
{
    __shared__ int kfirst[kLongThreads / 64];
    __shared__ double psum[kLongThreads / 64];
    const int b = blockIdx.x;
    const int64_t sid = s0 + b;
    const int d = (int)(sid % dim);
    const int vc = (int)(sid / dim);
    const int c = vc % nchains, v = vc / nchains;
    float* out = ess_chain + (int64_t)v * dim * nchains + (int64_t)d * nchains + c;
    const float* r = reinterpret_cast<const float*>(Rz + (int64_t)b * H);
    const int N = 2 * H;
    const float norm = 1.0f / ((float)N * (float)n);
    const float var = r[0] * norm;
    if (var < 1e-30f) {
        if (threadIdx.x == 0) *out = 1.0f;
        return;
    }
    const int kend = n / 2;
    int kneg = kend;
    for (int k = 1 + threadIdx.x; k < kend; k += kLongThreads) {
        const float pair_sum = (r[2 * k - 1] * norm) / var + (r[2 * k] * norm) / var;
        if (pair_sum < 0.0f) {
            kneg = k;
            break;
        }
    }
    for (int off = 32; off > 0; off >>= 1) kneg = min(kneg, __shfl_xor(kneg, off));
    if ((threadIdx.x & 63) == 0) kfirst[threadIdx.x >> 6] = kneg;
    __syncthreads();
    int K = kend;
    for (int w = 0; w < kLongThreads / 64; ++w) K = min(K, kfirst[w]);
    double acc = 0.0;
    for (int k = 1 + threadIdx.x; k < K; k += kLongThreads)
        acc += (double)((r[2 * k - 1] * norm) / var + (r[2 * k] * norm) / var);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) psum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int w = 0; w < kLongThreads / 64; ++w) sum += psum[w];
        const float tau = fmaxf(1.0f + 2.0f * (float)sum, 1.0f);
        *out = (float)n / tau;
    }
}

}  // namespace

int diag_max_chain_length() { return kMaxHalf; }

namespace {

// Sizes and launch geometry of one diagnostics call.
struct DiagPlan {
    int spc, logH, H, N, tb, half, spcp, s_tiles, d_tiles;
    int64_t n_series, t_blocks, vstride;
    unsigned nv;
    size_t b_ser, b_mom, b_ec;  // 256-byte rounded scratch regions
    size_t b_redo;              // the redo list of the two-pass ESS: count, then series ids
    size_t main_bytes() const { return b_ser + b_mom + b_ec + b_redo; }
};

size_t round256(size_t b) { return (b + 255) / 256 * 256; }

DiagPlan diag_plan(int dim, int n_total, int nchains, int n_variants, bool need_ec, bool allow_long = false)
{
    DiagPlan p{};
    p.spc = n_total / nchains;
    p.logH = 1;
    while ((1 << (p.logH + 1)) < 2 * p.spc) ++p.logH;  // N = 2H >= 2 spc
    p.H = 1 << p.logH;
    NAJA_REQUIRE(allow_long || p.H <= kMaxHalf, NAJA_ERR_UNSUPPORTED,
                 "chains of %d samples exceed the LDS FFT limit (%d)", p.spc, kMaxHalf);
    NAJA_REQUIRE(p.logH <= 28, NAJA_ERR_UNSUPPORTED, "chains of %d samples are too long", p.spc);
    NAJA_REQUIRE(n_variants >= 1 && n_variants <= 65535, NAJA_ERR_UNSUPPORTED, "%d variants (1..65535)", n_variants);
    p.N = 2 * p.H;
    p.tb = (p.logH + 1) / 2;
    p.half = p.spc / 2;
    p.nv = (unsigned)n_variants;
    p.vstride = (int64_t)n_total * dim;
    p.n_series = (int64_t)dim * nchains * p.nv;
    p.spcp = (p.spc + kTs - 1) / kTs * kTs;  // series row stride (256-byte aligned rows)
    p.s_tiles = (p.spc + kTs - 1) / kTs;
    p.d_tiles = (dim + kTd - 1) / kTd;
    p.t_blocks = (int64_t)((p.s_tiles + 7) / 8) * 8 * nchains * p.d_tiles * p.nv;
    NAJA_REQUIRE(p.t_blocks < (int64_t)1 << 31 && p.n_series < (int64_t)1 << 31, NAJA_ERR_UNSUPPORTED,
                 "too many series for one diagnostics launch");
    p.b_ser = round256(sizeof(float) * (size_t)p.n_series * p.spcp);
    p.b_mom = round256(sizeof(double) * 2 * (size_t)dim * 2 * nchains * p.nv);
    p.b_ec = need_ec ? round256(sizeof(float) * (size_t)dim * nchains * p.nv) : 0;
    p.b_redo = round256(sizeof(int) * (1 + (size_t)p.n_series));
    return p;
}

// Twiddle tables of the transform with H = 2^logH complex points, N = 2H:
// tlo[l] = W^l (l < 2^tb), then thi[h] = W^(h 2^tb) (h < H / 2^tb), each
// rounded once from FP64.
int twiddle_count(int logH)
{
    const int tb = (logH + 1) / 2;
    return (1 << tb) + ((1 << logH) >> tb);
}

void twiddle_tables(int logH, float2* out)
{
    const int tb = (logH + 1) / 2;
    const int nlo = 1 << tb, nhi = (1 << logH) >> tb;
    const double N = (double)(2 << logH);
    for (int l = 0; l < nlo; ++l) {
        const double ang = -2.0 * M_PI * (double)l / N;
        out[l] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    for (int h = 0; h < nhi; ++h) {
        const double ang = -2.0 * M_PI * (double)h * (double)nlo / N;
        out[nlo + h] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
}

// The launches of one diagnostics call. mem holds the series, the moments
// and, when ess_chain is NULL, the per-chain ESS (b_ser + b_mom + b_ec
// bytes); tw the twiddle tables of logH. Nothing in mem needs to start
// zeroed: every kernel writes what it later reads.
void diag_run(const DiagPlan& p, const float* samples, int dim, int nchains, int layout, float* ess, float* rhat,
              float* ess_chain, char* mem, const float2* tw, hipStream_t stream, float* ess_min, float* rhat_max,
              bool series_ready = false)
{
    float* ser = reinterpret_cast<float*>(mem);
    double* mom = reinterpret_cast<double*>(mem + p.b_ser);
    float* ec = ess_chain ? ess_chain : reinterpret_cast<float*>(mem + p.b_ser + p.b_mom);
    int* redo = reinterpret_cast<int*>(mem + p.b_ser + p.b_mom + p.b_ec);
    const int spc = p.spc, H = p.H, logH = p.logH, N = p.N, tb = p.tb, half = p.half, spcp = p.spcp;
    const int64_t n_series = p.n_series;
    const unsigned nv = p.nv;

    if (series_ready) {  // the producer wrote the series layout: only the redo count starts at zero
        hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(64), 0, stream, redo);
    } else {
        hipLaunchKernelGGL(series_transpose_kernel, dim3((unsigned)p.t_blocks), dim3(256), 0, stream, samples, dim,
                           nchains, spc, spcp, layout, p.vstride, p.s_tiles, p.d_tiles, ser, redo);
    }
    NAJA_HIP(hipGetLastError());
    // A half-size pass first when it leaves at least 1/8 of the chain as exact
    // lags (N/2 - n >= n/8); the full pass then redoes only the series whose
    // truncation point lies beyond them.
    const bool two_pass = N / 2 - spc >= spc / 8;
    if (two_pass) {
        hipLaunchKernelGGL(ess_series_kernel, dim3((unsigned)n_series), dim3(ess_threads(H / 2)),
                           padded_lds_bytes(H / 2), stream,
                           ser, dim, nchains, spc, spcp, logH - 1, tw, tb, 1, 1, ec, half, mom, n_series, redo);
        NAJA_HIP(hipGetLastError());
        hipLaunchKernelGGL(ess_series_kernel, dim3(ess_grid(H, n_series)), dim3(ess_threads(H)), padded_lds_bytes(H),
                           stream, ser, dim, nchains, spc, spcp, logH, tw, tb, 0, 2, ec, half, mom, n_series, redo);
    } else {
        hipLaunchKernelGGL(ess_series_kernel, dim3((unsigned)n_series), dim3(ess_threads(H)), padded_lds_bytes(H),
                           stream, ser, dim, nchains, spc, spcp, logH, tw, tb, 0, 0, ec, half, mom, n_series, redo);
    }
    NAJA_HIP(hipGetLastError());
    hipLaunchKernelGGL(ess_min_kernel, dim3((dim + 255) / 256, nv), dim3(256), 0, stream, ec, dim, nchains, ess);
    NAJA_HIP(hipGetLastError());
    hipLaunchKernelGGL(rhat_final_kernel, dim3((dim + 255) / 256, nv), dim3(256), 0, stream, mom, dim, nchains, half,
                       rhat);
    NAJA_HIP(hipGetLastError());
    if (ess_min || rhat_max) {
        hipLaunchKernelGGL(variant_summary_kernel, dim3(nv), dim3(256), 0, stream, ess, rhat, dim, ess_min, rhat_max);
        NAJA_HIP(hipGetLastError());
    }
}

// Allocation-free diagnostics: the series, moments and per-chain ESS live in
// a file-scope device array (one copy per GPU, held until the process ends),
// the twiddle tables of every transform size in a second one. Host code
// never uses these names as addresses: each device's addresses come from
// hipGetSymbolAddress at set-up.
constexpr int kMinLogH = 2, kMaxLogH = 14;  // H = 4 (4-sample chains) .. kMaxHalf
static_assert((1 << kMaxLogH) == kMaxHalf, "twiddle tables cover every transform size");
static_assert(kTwiddleEntries >= 886, "sum of twiddle_count(logH) over kMinLogH..kMaxLogH");

struct NoAllocDevice {
    bool ready = false;
    char* scratch = nullptr;       // this device's address of `scratch`
    float2* twiddles = nullptr;    // this device's address of `scratch_twiddles`
    int offset[kMaxLogH + 1] = {}; // first entry of each transform size's tables
};
std::mutex g_noalloc_mu;
std::vector<NoAllocDevice> g_noalloc;

const NoAllocDevice& noalloc_device()
{
    int dev = 0;
    NAJA_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if ((int)g_noalloc.size() <= dev) g_noalloc.resize((size_t)dev + 1);
    NoAllocDevice& st = g_noalloc[(size_t)dev];
    if (!st.ready) {
        void* ptr = nullptr;
        NAJA_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        void* tptr = nullptr;
        NAJA_HIP(hipGetSymbolAddress(&tptr, HIP_SYMBOL(scratch_twiddles)));
        std::vector<float2> tables;
        for (int lh = kMinLogH; lh <= kMaxLogH; ++lh) {
            st.offset[lh] = (int)tables.size();
            tables.resize(tables.size() + (size_t)twiddle_count(lh));
            twiddle_tables(lh, tables.data() + st.offset[lh]);
        }
        NAJA_REQUIRE((int)tables.size() <= kTwiddleEntries, NAJA_ERR_RUNTIME, "twiddle tables exceed their array");
        // Synchronous upload, once per device (the first call on a device
        // must not be under stream capture).
        NAJA_HIP(hipMemcpy(tptr, tables.data(), sizeof(float2) * tables.size(), hipMemcpyHostToDevice));
        st.scratch = static_cast<char*>(ptr);
        st.twiddles = static_cast<float2*>(tptr);
        st.ready = true;
    }
    return st;
}

}  // namespace

// Chains beyond the LDS transform: the transforms in global memory (the
// long_* kernels), series in batches whose two complex buffers stay within
// about 2 GB.
void diag_launch_long(const DiagPlan& p, const float* samples, int dim, int nchains, int layout, float* ess,
                      float* rhat, float* ess_chain, hipStream_t stream, float* ess_min, float* rhat_max)
{
    const int H = p.H, N = p.N, spc = p.spc;
    std::vector<float2> tw_h((size_t)2 * H);  // W_H^t, then W_N^k, t, k < H
    for (int t = 0; t < H; ++t) {
        const double a = -2.0 * M_PI * (double)t / (double)H;
        tw_h[(size_t)t] = make_float2((float)std::cos(a), (float)std::sin(a));
        const double c = -2.0 * M_PI * (double)t / (double)N;
        tw_h[(size_t)H + t] = make_float2((float)std::cos(c), (float)std::sin(c));
    }
    const int64_t nb_max = std::max<int64_t>(1, std::min<int64_t>(p.n_series, ((int64_t)1 << 30) / ((int64_t)H * 8)));
    const size_t b_tw = round256(sizeof(float2) * tw_h.size());
    const size_t b_buf = round256(sizeof(float2) * (size_t)nb_max * H);
    const size_t b_mean = round256(sizeof(float) * (size_t)nb_max);
    const size_t b_main = p.b_ser + p.b_mom + p.b_ec;
    char* mem = nullptr;
    scratch_alloc((void**)&mem, b_main + b_tw + 2 * b_buf + b_mean, stream);
    float* ser = reinterpret_cast<float*>(mem);
    double* mom = reinterpret_cast<double*>(mem + p.b_ser);
    float* ec = ess_chain ? ess_chain : reinterpret_cast<float*>(mem + p.b_ser + p.b_mom);
    float2* tw = reinterpret_cast<float2*>(mem + b_main);
    float2* bufA = reinterpret_cast<float2*>(mem + b_main + b_tw);
    float2* bufB = reinterpret_cast<float2*>(mem + b_main + b_tw + b_buf);
    float* meanf = reinterpret_cast<float*>(mem + b_main + b_tw + 2 * b_buf);
    NAJA_HIP(hipMemcpyAsync(tw, tw_h.data(), sizeof(float2) * tw_h.size(), hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(series_transpose_kernel, dim3((unsigned)p.t_blocks), dim3(256), 0, stream, samples, dim,
                       nchains, spc, p.spcp, layout, p.vstride, p.s_tiles, p.d_tiles, ser, nullptr);
    NAJA_HIP(hipGetLastError());
    const float2* twH = tw;
    const float2* twN = tw + H;
    for (int64_t s0 = 0; s0 < p.n_series; s0 += nb_max) {
        const int bn = (int)std::min<int64_t>(nb_max, p.n_series - s0);
        const dim3 g2((unsigned)std::min(64, std::max(1, H / 4 / 256)), (unsigned)bn);
        hipLaunchKernelGGL(long_moments_kernel, dim3(bn), dim3(kLongThreads), 0, stream, ser, dim, nchains, spc,
                           p.spcp, p.half, s0, meanf, mom);
        hipLaunchKernelGGL(long_pack_kernel, g2, dim3(256), 0, stream, ser, spc, p.spcp, H, s0, meanf, bufA);
        NAJA_HIP(hipGetLastError());
        float2* cur = bufA;
        float2* oth = bufB;
        auto passes = [&](int inverse) {
            int Ns = 1;
            if (p.logH & 1) {
                hipLaunchKernelGGL(stockham_kernel<2>, g2, dim3(256), 0, stream, cur, oth, H, Ns, twH, inverse);
                std::swap(cur, oth);
                Ns *= 2;
            }
            for (; Ns < H; Ns *= 4) {
                hipLaunchKernelGGL(stockham_kernel<4>, g2, dim3(256), 0, stream, cur, oth, H, Ns, twH, inverse);
                std::swap(cur, oth);
            }
            NAJA_HIP(hipGetLastError());
        };
        passes(0);
        hipLaunchKernelGGL(long_split_kernel, g2, dim3(256), 0, stream, cur, oth, H, twN);
        std::swap(cur, oth);
        passes(1);
        hipLaunchKernelGGL(long_trunc_kernel, dim3(bn), dim3(kLongThreads), 0, stream, cur, H, spc, dim, nchains, s0,
                           ec);
        NAJA_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(ess_min_kernel, dim3((dim + 255) / 256, p.nv), dim3(256), 0, stream, ec, dim, nchains, ess);
    hipLaunchKernelGGL(rhat_final_kernel, dim3((dim + 255) / 256, p.nv), dim3(256), 0, stream, mom, dim, nchains,
                       p.half, rhat);
    NAJA_HIP(hipGetLastError());
    if (ess_min || rhat_max) {
        hipLaunchKernelGGL(variant_summary_kernel, dim3(p.nv), dim3(256), 0, stream, ess, rhat, dim, ess_min, rhat_max);
        NAJA_HIP(hipGetLastError());
    }
    scratch_free(mem, stream);
}

void diag_launch(const float* samples, int dim, int n_total, int nchains, int layout, float* ess, float* rhat,
                 float* ess_chain, hipStream_t stream, int n_variants, float* ess_min, float* rhat_max)
{
    const DiagPlan p = diag_plan(dim, n_total, nchains, n_variants, ess_chain == nullptr, true);
    if (p.H > kMaxHalf) {
        diag_launch_long(p, samples, dim, nchains, layout, ess, rhat, ess_chain, stream, ess_min, rhat_max);
        return;
    }
    std::vector<float2> tw_h((size_t)twiddle_count(p.logH));
    twiddle_tables(p.logH, tw_h.data());
    // One scratch allocation for the series, the moments, the per-chain ESS
    // (when the caller passes none) and the twiddle tables: each pool call
    // costs tens of microseconds of host time.
    const size_t b_tw = round256(sizeof(float2) * tw_h.size());
    const size_t b_main = p.main_bytes();
    char* mem = nullptr;
    scratch_alloc((void**)&mem, b_main + b_tw, stream);
    float2* tw = reinterpret_cast<float2*>(mem + b_main);
    NAJA_HIP(hipMemcpyAsync(tw, tw_h.data(), sizeof(float2) * tw_h.size(), hipMemcpyHostToDevice, stream));
    diag_run(p, samples, dim, nchains, layout, ess, rhat, ess_chain, mem, tw, stream, ess_min, rhat_max);
    scratch_free(mem, stream);
}

// Diagnostics with a caller-held workspace: the series, moments, per-chain
// ESS and twiddle tables in `ws` (diag_workspace_bytes), the tables written by
// diag_workspace_init once. A run issues kernels only (no allocation, no copy
// from the host), so it can be captured in a graph.
size_t diag_workspace_bytes(int dim, int n_total, int nchains, int n_variants)
{
    const DiagPlan p = diag_plan(dim, n_total, nchains, n_variants, true);
    return p.main_bytes() + round256(sizeof(float2) * (size_t)twiddle_count(p.logH));
}

void diag_workspace_init(void* ws, int dim, int n_total, int nchains, int n_variants)
{
    const DiagPlan p = diag_plan(dim, n_total, nchains, n_variants, true);
    std::vector<float2> tw_h((size_t)twiddle_count(p.logH));
    twiddle_tables(p.logH, tw_h.data());
    char* tw = static_cast<char*>(ws) + p.main_bytes();
    NAJA_HIP(hipMemcpy(tw, tw_h.data(), sizeof(float2) * tw_h.size(), hipMemcpyHostToDevice));
}

void diag_launch_ws(void* ws, const float* samples, int dim, int n_total, int nchains, int layout, float* ess,
                    float* rhat, hipStream_t stream, int n_variants, float* ess_min, float* rhat_max,
                    bool series_ready)
{
    const DiagPlan p = diag_plan(dim, n_total, nchains, n_variants, true);
    char* mem = static_cast<char*>(ws);
    const float2* tw = reinterpret_cast<const float2*>(mem + p.main_bytes());
    diag_run(p, samples, dim, nchains, layout, ess, rhat, nullptr, mem, tw, stream, ess_min, rhat_max, series_ready);
}

float* diag_workspace_series(void* ws) { return static_cast<float*>(ws); }

int diag_series_stride(int n_total, int nchains)
{
    const int spc = n_total / nchains;
    return (spc + kTs - 1) / kTs * kTs;
}

void diag_launch_noalloc(const float* samples, int dim, int n_total, int nchains, int layout, float* ess, float* rhat,
                         float* ess_chain, hipStream_t stream)
{
    const DiagPlan p = diag_plan(dim, n_total, nchains, 1, ess_chain == nullptr);
    const size_t need = p.main_bytes();
    NAJA_REQUIRE(need <= kDiagNoAllocBytes, NAJA_ERR_UNSUPPORTED,
                 "naja_diag_noalloc: %d dims x %d chains x %d samples need %zu bytes of scratch; the static buffer "
                 "holds %zu (about %d x 4 bytes per (dim, chain) series, rounded up to 64 samples)",
                 dim, nchains, p.spc, need, (size_t)kDiagNoAllocBytes, p.spcp);
    const NoAllocDevice& st = noalloc_device();
    diag_run(p, samples, dim, nchains, layout, ess, rhat, ess_chain, st.scratch, st.twiddles + st.offset[p.logH],
             stream, nullptr, nullptr);
}

}  // namespace naja
