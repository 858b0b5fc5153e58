// Device code of the CHR step kernel, included by the chr_inst_*.hip
// translation units (each instantiates a subset of launch shapes, so the
// shapes compile in parallel). The design notes are in chr_kernel.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <type_traits>

#include "chr_kernel.h"
#include "philox.h"

namespace naja {
namespace chr_detail {

constexpr int kMoveCoord = 0;
constexpr int kMovePair = 1;
constexpr int kMoveKsparse = 2;
constexpr int kKmax = 32;

struct Move {
    int type;
    int i;
    int j;
    float ci;
    float cj;
    uint32_t w3;
};

__device__ __forceinline__ Move draw_move(const ChrArgs& a, uint32_t k0, uint32_t k1, uint64_t t, uint32_t group)
{
    const U32x4 w = philox4x32_10((uint32_t)t, (uint32_t)(t >> 32), group, 0u, k0, k1);
    Move mv;
    mv.w3 = w.w;
    mv.type = kMoveCoord;
    mv.i = (int)rng_index(w.y, (uint32_t)a.d);
    mv.j = -1;
    mv.ci = 1.0f;
    mv.cj = 0.0f;
    const bool ks = (uint64_t)w.x < a.thr_k;
    const bool pr = !ks && a.has_pair && (uint64_t)w.x < a.thr_kp;
    if (ks) {
        mv.type = kMoveKsparse;
    } else if (pr && a.d >= 2) {
        if (a.pair_mode == 1) {
            const int r = (int)rng_index(w.z, (uint32_t)(a.d - 1));
            mv.type = kMovePair;
            mv.j = (r >= mv.i) ? r + 1 : r;
            mv.ci = 0.70710678118654752440f;
            mv.cj = -0.70710678118654752440f;
        } else if (a.pair_mode == 2 && a.n_pairs > 0) {
            const int p = (int)rng_index(w.z, (uint32_t)a.n_pairs);
            const float c = a.pair_c[p];
            const float s = a.pair_s[p];
            mv.type = kMovePair;
            mv.i = a.pair_i[p];
            mv.j = a.pair_j[p];
            float ci = ((w.w & 1u) == 0u) ? c : -s;
            float cj = ((w.w & 1u) == 0u) ? s : c;
            if ((w.w >> 1) & 1u) {
                ci = -ci;
                cj = -cj;
            }
            mv.ci = ci;
            mv.cj = cj;
        }
    }
    return mv;
}

// k-sparse entry q: index from block (t, group, 1 + q/4), sign from bit q of w3.
__device__ __forceinline__ int ksparse_index(const ChrArgs& a, uint32_t k0, uint32_t k1, uint64_t t,
                                             uint32_t group, int q)
{
    const U32x4 v = philox4x32_10((uint32_t)t, (uint32_t)(t >> 32), group, 1u + (uint32_t)(q >> 2), k0, k1);
    const uint32_t word = (q & 3) == 0 ? v.x : ((q & 3) == 1 ? v.y : ((q & 3) == 2 ? v.z : v.w));
    return (int)rng_index(word, (uint32_t)a.d);
}

template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v)
{
    return __builtin_bit_cast(float,
                              __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

__device__ __forceinline__ float readlane_f32(float v, int lane)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, 0xF, 0xF, false);
}

__device__ __forceinline__ uint32_t umin32(uint32_t a, uint32_t b) { return a < b ? a : b; }
__device__ __forceinline__ int32_t smin32(int32_t a, int32_t b) { return a < b ? a : b; }

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_mov(uint32_t v)
{
    // bound_ctrl set and full masks: every lane of these patterns has a
    // source, and the move folds into the consuming min as a DPP operand.
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}

// DPP row broadcast (row_bcast:15 = 0x142, row_bcast:31 = 0x143, gfx9):
// the rows in ROW_MASK take lane 15 (31) of the row below; the other rows get
// IDENT, the identity of the consuming min, so the move folds into it.
template <int CTRL, int ROW_MASK, uint32_t IDENT>
__device__ __forceinline__ uint32_t dpp_bcast(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)IDENT, (int)v, CTRL, ROW_MASK, 0xF, false);
}

// Unsigned and signed minimum over the 64 lanes of a wavefront: DPP within
// each row of 16 (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror,
// row_mirror), then across rows with row_bcast:15 (rows 1 and 3 take the row
// below) and row_bcast:31 (rows 2 and 3 take lane 31), which leaves the wave
// minimum in lane 63. Every step is a DPP operand of a min: three VALU per
// value across rows.
__device__ __forceinline__ void wave_min2_dpp(uint32_t& hb, int32_t& lb)
{
    uint32_t l = (uint32_t)lb;
    hb = umin32(hb, dpp_mov<0xB1>(hb));
    l = (uint32_t)smin32((int32_t)l, (int32_t)dpp_mov<0xB1>(l));
    hb = umin32(hb, dpp_mov<0x4E>(hb));
    l = (uint32_t)smin32((int32_t)l, (int32_t)dpp_mov<0x4E>(l));
    hb = umin32(hb, dpp_mov<0x141>(hb));
    l = (uint32_t)smin32((int32_t)l, (int32_t)dpp_mov<0x141>(l));
    hb = umin32(hb, dpp_mov<0x140>(hb));
    l = (uint32_t)smin32((int32_t)l, (int32_t)dpp_mov<0x140>(l));
    hb = umin32(hb, dpp_bcast<0x142, 0xA, 0xFFFFFFFFu>(hb));
    l = (uint32_t)smin32((int32_t)l, (int32_t)dpp_bcast<0x142, 0xA, 0x7FFFFFFFu>(l));
    hb = umin32(hb, dpp_bcast<0x143, 0xC, 0xFFFFFFFFu>(hb));
    l = (uint32_t)smin32((int32_t)l, (int32_t)dpp_bcast<0x143, 0xC, 0x7FFFFFFFu>(l));
    lb = (int32_t)l;
}

// The same minima, returned wave-uniform (v_readlane of lane 63).
__device__ __forceinline__ void wave_min2(uint32_t& hb, int32_t& lb)
{
    wave_min2_dpp(hb, lb);
    hb = (uint32_t)__builtin_amdgcn_readlane((int)hb, 63);
    lb = __builtin_amdgcn_readlane(lb, 63);
}

// Unsigned and signed minima of v[0..N) within the lane, as a tree of
// three-input minima (ceil(log3 N) dependent levels instead of N - 1).
template <int N>
__device__ __forceinline__ void lane_min2(const uint32_t (&v)[N], uint32_t& hb, int32_t& lb)
{
    uint32_t u[N];
    int32_t t[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        u[i] = v[i];
        t[i] = (int32_t)v[i];
    }
#pragma unroll
    for (int st = 1; st < N; st *= 3)
#pragma unroll
        for (int i = 0; i < N; i += 3 * st) {
            if (i + st < N) {
                u[i] = umin32(u[i], u[i + st]);
                t[i] = smin32(t[i], t[i + st]);
            }
            if (i + 2 * st < N) {
                u[i] = umin32(u[i], u[i + 2 * st]);
                t[i] = smin32(t[i], t[i + 2 * st]);
            }
        }
    hb = u[0];
    lb = t[0];
}

// Chord ends in the direct form q = s / ae:
// hi = the least q with the sign bit clear and lo = the greatest q with the
// sign bit set. For IEEE floats these are an unsigned and a signed integer
// minimum over the bit patterns, so no branch, select or final reciprocal is
// needed: ae = 0 gives +-inf and 0/0 a NaN, which never win against a real
// facet. In exact arithmetic hi = 1/max(ae/s) and lo = 1/min(ae/s), the
// reference's ends (chr.cu:416-438), each rounded once here instead of twice.
// Every slack is stored clamped: a value that FP32 rounding left at or below
// zero, after an update or a resync, is replaced by +0, so a point just
// outside a facet is treated as on it (the FP64 reference has no such
// clamp). The clamp is a signed integer maximum with 0 over the bit pattern
// (one instruction): negative values and -0 become +0, positive values pass,
// and so does a NaN with the sign clear, whose quotient is a NaN that never
// wins either minimum (no finite input produces one).
__device__ __forceinline__ float clamp_slack(float s)
{
    return __builtin_bit_cast(float, max(__builtin_bit_cast(int32_t, s), 0));
}

// The scaled-pair form's clamp to [0, 1] (ChrArgs::scaled): on an FMA's
// result it is the instruction's own output clamp (no extra instruction).
// Negative values and -0 become +0, as with clamp_slack; the pair's slacks
// stay within [0, 1/2] (S >= 2 (b + bn)), so the upper end never applies to
// a real row, and a padding row's +inf becomes 1.
__device__ __forceinline__ float clamp_unit(float s) { return __builtin_amdgcn_fmed3f(s, 0.0f, 1.0f); }

__device__ __forceinline__ uint32_t chord_bits(float ae, float s)
{
    return __builtin_bit_cast(uint32_t, s / ae);
}

// Coordinate moves use the packed reciprocal column rinv = RN(1/A[r,i]):
// q = RN(s * rinv), one multiply instead of an IEEE division.
__device__ __forceinline__ uint32_t chord_bits_rcp(float rinv, float s)
{
    return __builtin_bit_cast(uint32_t, s * rinv);
}

// Lane l of every wavefront draws the move and the alpha uniforms of step
// tb + l (VALU Philox, about three VALU instructions per step amortised);
// each step then reads its draw with readlane.
template <int G>
struct MoveBatch {
    int tij;     // type << 30 | i
    int j;       // partner coordinate, or the k-sparse sign word w3
    float ci;
    float cj;
    float u[G];  // alpha uniform of chain c0 + g
};

// CO: the launch has no pair or k-sparse moves (every draw is a coordinate
// move), so only the index is kept; the draws are the same.
// CO: tij is the index times col_scale (the column's byte offset when the
// caller passes its column stride, else 1).
template <int G, bool CO>
__device__ __forceinline__ void gen_batch(const ChrArgs& a, uint32_t k0, uint32_t k1, uint64_t tb, uint32_t group,
                                          int c0, MoveBatch<G>& mb, int col_scale = 1)
{
    const uint64_t t = tb + (uint64_t)(threadIdx.x & 63);
    if (CO) {
        const U32x4 w = philox4x32_10((uint32_t)t, (uint32_t)(t >> 32), group, 0u, k0, k1);
        mb.tij = (int)rng_index(w.y, (uint32_t)a.d) * col_scale;
    } else {
        const Move mv = draw_move(a, k0, k1, t, group);
        mb.tij = (mv.type << 30) | mv.i;
        mb.j = (mv.type == kMoveKsparse) ? (int)mv.w3 : mv.j;
        mb.ci = mv.ci;
        mb.cj = mv.cj;
    }
    uint32_t ublk[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const uint32_t chain = (uint32_t)(c0 + g);
        if (g == 0 || (chain & 3u) == 0u) {
            const U32x4 u = philox4x32_10((uint32_t)t, (uint32_t)(t >> 32), chain >> 2, kAlphaBlock, k0, k1);
            ublk[0] = u.x;
            ublk[1] = u.y;
            ublk[2] = u.z;
            ublk[3] = u.w;
        }
        const uint32_t w = ublk[chain & 3u];
        mb.u[g] = rng_uniform(w);
    }
}

__device__ __forceinline__ int readlane_i32(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

template <int G, bool CO>
__device__ __forceinline__ Move read_move(const MoveBatch<G>& mb, int lane)
{
    Move mv;
    const int tij = readlane_i32(mb.tij, lane);
    if (CO) {
        mv.type = kMoveCoord;
        mv.i = tij;
        mv.j = -1;
        mv.w3 = 0u;
        mv.ci = 1.0f;
        mv.cj = 0.0f;
        return mv;
    }
    mv.type = (int)((uint32_t)tij >> 30);
    mv.i = tij & 0x3FFFFFFF;
    const int jv = readlane_i32(mb.j, lane);
    mv.j = jv;
    mv.w3 = (uint32_t)jv;
    mv.ci = readlane_f32(mb.ci, lane);
    mv.cj = readlane_f32(mb.cj, lane);
    return mv;
}

// Raw buffer loads with a descriptor built by hand: base address, stride 0,
// num_records (bytes; loads beyond it read 0) and the gfx9 dword-3 format
// word (32-bit data). The loads are the LLVM intrinsics, named as functions.
typedef int chr_i32x4 __attribute__((ext_vector_type(4)));
typedef float chr_f32x2 __attribute__((ext_vector_type(2)));
__device__ float chr_raw_load_f32(chr_i32x4 rsrc, int voffset, int soffset,
                                  int aux) __asm("llvm.amdgcn.raw.buffer.load.f32");
__device__ chr_f32x2 chr_raw_load_v2f32(chr_i32x4 rsrc, int voffset, int soffset,
                                        int aux) __asm("llvm.amdgcn.raw.buffer.load.v2f32");
typedef float chr_f32x4 __attribute__((ext_vector_type(4)));
__device__ chr_f32x4 chr_raw_load_v4f32(chr_i32x4 rsrc, int voffset, int soffset,
                                        int aux) __asm("llvm.amdgcn.raw.buffer.load.v4f32");
__device__ __forceinline__ chr_i32x4 chr_make_rsrc(const void* p, uint32_t bytes)
{
    const uint64_t addr = (uint64_t)p;
    chr_i32x4 r;
    // Wave-uniform by construction; readfirstlane keeps it in SGPRs.
    r[0] = __builtin_amdgcn_readfirstlane((int)(uint32_t)addr);
    r[1] = __builtin_amdgcn_readfirstlane((int)(uint32_t)(addr >> 32));
    r[2] = __builtin_amdgcn_readfirstlane((int)bytes);
    r[3] = 0x00020000;
    return r;
}

// PF: prefetch distance in steps. The columns of step t + PF are loaded
// during step t into one of PF + 1 rotating register sets, so each load has
// PF steps to arrive (moves do not depend on the state).
// XR > 0 (one wavefront): the chain states live in VGPRs, coordinate c of
// chain g in lane c % 64 of xv[g][c / 64] (d <= 64 XR), and every lane
// applies the update of its own coordinates, so a step reads and writes no
// LDS and needs no exec-masked store; XR = 0 keeps the states in LDS.
// PAIRED: packed row r stands for two constraints, a_r y <= b_r and
// -a_r y <= bn_r (a row and its negation), with two slacks per slot: the
// column loads and the cache footprint are those of one row. Every value is
// the one the unpaired layout computes for the negated row: its chord ratio
// s / (-ae) = -(s / ae) and RN(1 / -a) = -RN(1 / a) (rounding is symmetric),
// its update fmaf(-alpha, -ae, s) = fmaf(alpha, ae, s), its FP64 resync
// bn minus the negated FMA chain. A row without a partner has bn = NaN: its second slack
// stays NaN and a NaN ratio never wins either minimum.
// NC > 1 (one-wavefront coordinate-only chains with VGPR states): a workgroup
// runs NC chains of one variant, one per wavefront and otherwise independent,
// with the variant's packed A copied into LDS once, so each step reads its
// column A[:, i] from LDS and only 1/A[:, i] through the caches.
// SC (paired kernels on scaled pairs, ChrArgs::scaled): slacks are clamped
// to [0, 1] by the slack update's own FMA clamp instead of a separate
// maximum; padding rows are read like real rows (A = 0, b = bn = +inf,
// reciprocal +inf): their clamped slacks are 1, their ratios 1 * (+-inf) or
// 1 / (+-0), infinities that never bound the chord, as the unscaled
// layout's inf / 0 from an out-of-range read.
template <int W, int G, int R, bool CO = false, int PF = 1, int XR = 0, bool PAIRED = false, int NC = 1,
          bool SC = false>
__global__ __launch_bounds__(64 * W * NC) void chr_step_kernel(ChrArgs a)
{
    static_assert(PF >= 1 && PF <= 4, "prefetch distance 1 to 4");
    static_assert(XR == 0 || W == 1, "register-resident states need one wavefront");
    static_assert(NC == 1 || (W == 1 && G == 1 && CO && XR > 0), "LDS-resident columns: one-wavefront coordinate chains");
    static_assert(!SC || PAIRED, "scaled pairs need the paired-row layout");
    constexpr int T = 64 * W;
    extern __shared__ __align__(16) float xs[];  // G x d current states (XR = 0); A columns (NC > 1)
    __shared__ uint2 part[2][G][W];
    __shared__ int ks_idx[kKmax];
    __shared__ float ks_coef[kKmax];

    // NC > 1: wavefront w of the workgroup runs item NC * blockIdx.x + w, and
    // tid is the thread's index within its chain.
    const int chain_slot = NC > 1 ? __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6) : 0;
    const int tid = NC > 1 ? (int)threadIdx.x & 63 : (int)threadIdx.x;
    const int lane = tid & 63;
    // readfirstlane makes the wave index provably wave-uniform.
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const ChrItem item = a.items[(int64_t)blockIdx.x * NC + chain_slot];
    const ChrVariant V = a.variants[item.variant];
    const int d = a.d;
    const int lda = V.lda;  // == T * R: every slot of every lane is a real or a padding row
    // Row layout: lane tid holds its rows in runs of VEC consecutive rows,
    // run q at rows q*VEC*T + tid*VEC .. + VEC-1, so one column's slots are
    // R/VEC vector loads per lane and a wavefront's load covers VEC*T
    // consecutive rows. Padding rows (A = 0, b = +inf) fill every run past m.
    constexpr int VEC = (R % 4 == 0) ? 4 : ((R % 2 == 0) ? 2 : 1);
    constexpr int XS = XR > 0 ? XR : 1;
    auto row_of = [&](int k) { return (k / VEC) * (VEC * T) + tid * VEC + (k % VEC); };
    const float* Av = a.A + V.a_off;
    const int geff = min(G, a.nchains - item.c0);
    const uint32_t group = (uint32_t)(item.c0 / G);
    const uint32_t k0 = a.seed_lo;
    const uint32_t k1 = a.seed_hi ^ (uint32_t)item.variant;
    // NC > 1: the workgroup's chains belong to one variant (items come in
    // groups of NC); its packed A is copied into LDS once, column c at
    // xs[c * lda] (the packed block's own stride, so one byte offset
    // addresses a column in both), the rows past m as zeros.
    if constexpr (NC > 1) {
        for (int e = (int)threadIdx.x; e < lda * d; e += 64 * NC) {
            const int c = e / lda, r = e % lda;
            xs[e] = r < V.m ? Av[(int64_t)c * lda + r] : 0.0f;
        }
        __syncthreads();
    }
    // CO with VGPR states: a draw carries its column's byte offset i * 4 lda,
    // so a step forms no column address (the buffer loads take it as their
    // scalar offset, the LDS reads add it to the lane's row offset) and the
    // state update compares it with the lanes' coordinate offsets.
    constexpr bool kColOff = CO && XR > 0;
    const int col_scale = kColOff ? 4 * lda : 1;

    float xv[G][XS];
    if constexpr (XR > 0) {
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int q = 0; q < XR; ++q) {
                const int c = q * 64 + lane;
                xv[g][q] = (g < geff && c < d) ? a.y0[(V.y0_col + item.c0 + g) * (int64_t)d + c] : 0.0f;
            }
    } else {
        for (int g = 0; g < G; ++g) {
            for (int c = tid; c < d; c += T) {
                xs[g * d + c] = (g < geff) ? a.y0[(V.y0_col + item.c0 + g) * (int64_t)d + c] : 0.0f;
            }
        }
    }

    // Slot k of lane tid holds row row_of(k); sn its partner's slack (PAIRED).
    constexpr int RP = PAIRED ? R : 1;
    float s[G][R];
    float sn[G][RP];
    float cs[PF + 1][2][R];  // rotating column sets: step t uses set (t - t_begin) % (PF + 1)
    // Steps t0 .. t0 + nsteps - 1 (t0 > 0 continues an earlier launch).
    const int64_t t_begin = a.t0;
    const int64_t t_end = a.t0 + a.nsteps;
    MoveBatch<G> mb;
    gen_batch<G, CO>(a, k0, k1, (uint64_t)(t_begin & ~(int64_t)63), group, item.c0, mb, col_scale);
    const int L0 = (int)(t_begin & 63);
    // (A[:,i], 1/A[:,i]) for a coordinate move, (A[:,i], A[:,j]) for a pair
    // move; a k-sparse move (or no move) loads a dummy pair. One buffer
    // descriptor per packed array over the variant's whole block: the column
    // is the instruction's SGPR offset, the lane's run of rows its VGPR
    // offset, fixed for the launch. Runs wholly past the real rows (m rounded
    // up to VEC; the rows past m in memory are padding) get an offset beyond
    // any block (bit 31; the launch keeps blocks under 2 GiB), so they read 0
    // without touching memory: their slack is +inf, so their chord ratio is
    // inf * 0 or inf / 0 and never bounds the chord. On scaled pairs (slacks
    // clamped to 1, so a zero reciprocal would give a zero ratio) they read
    // their padding rows' +inf reciprocals (and, without LDS columns, their
    // zero A entries) from memory.
    const uint32_t blk_bytes = 4u * (uint32_t)lda * (uint32_t)d;
    const chr_i32x4 rsA = chr_make_rsrc(a.A + V.a_off, blk_bytes);
    const chr_i32x4 rsI = chr_make_rsrc(a.Ainv + V.a_off, blk_bytes);
    const int col_stride = 4 * lda;  // bytes from one column to the next
    const int mpad = (V.m + VEC - 1) / VEC * VEC;
    int voff[R / VEC];
    int loff[R / VEC];  // LDS byte offsets of the lane's runs (NC > 1)
#pragma unroll
    for (int q = 0; q < R / VEC; ++q) {
        const int r0 = q * VEC * T + tid * VEC;
        voff[q] = (SC || r0 < mpad) ? 4 * r0 : (int)0x80000000u;
        loff[q] = 4 * r0;
    }
    auto load_cols = [&](const Move& m, bool real_move, float (&xa)[R], float (&xb)[R]) {
        const int mt = CO ? kMoveCoord : m.type;
        const bool real = real_move && mt != kMoveKsparse;
        const bool pair = real && mt == kMovePair;
        const int si = real ? (kColOff ? m.i : m.i * col_stride) : 0;
        const int sj = pair ? m.j * col_stride : si;
        const chr_i32x4 rj = pair ? rsA : rsI;
        if constexpr (NC > 1) {
            const char* colp = reinterpret_cast<const char*>(xs) + si;
#pragma unroll
            for (int q = 0; q < R / VEC; ++q) {
                if (VEC == 4) {
                    const float4 u = *reinterpret_cast<const float4*>(colp + loff[q]);
                    const chr_f32x4 w = chr_raw_load_v4f32(rsI, voff[q], si, 0);
                    xa[q * 4] = u.x;
                    xa[q * 4 + 1] = u.y;
                    xa[q * 4 + 2] = u.z;
                    xa[q * 4 + 3] = u.w;
#pragma unroll
                    for (int e = 0; e < 4; ++e) xb[q * 4 + e] = w[e];
                } else if (VEC == 2) {
                    const float2 u = *reinterpret_cast<const float2*>(colp + loff[q]);
                    const chr_f32x2 w = chr_raw_load_v2f32(rsI, voff[q], si, 0);
                    xa[q * 2] = u.x;
                    xa[q * 2 + 1] = u.y;
                    xb[q * 2] = w[0];
                    xb[q * 2 + 1] = w[1];
                } else {
                    xa[q] = *reinterpret_cast<const float*>(colp + loff[q]);
                    xb[q] = chr_raw_load_f32(rsI, voff[q], si, 0);
                }
            }
            return;
        }
#pragma unroll
        for (int q = 0; q < R / VEC; ++q) {
            if (VEC == 4) {
                const chr_f32x4 u = chr_raw_load_v4f32(rsA, voff[q], si, 0);
                const chr_f32x4 w = chr_raw_load_v4f32(rj, voff[q], sj, 0);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    xa[q * 4 + e] = u[e];
                    xb[q * 4 + e] = w[e];
                }
            } else if (VEC == 2) {
                const chr_f32x2 u = chr_raw_load_v2f32(rsA, voff[q], si, 0);
                const chr_f32x2 w = chr_raw_load_v2f32(rj, voff[q], sj, 0);
                xa[q * 2] = u[0];
                xa[q * 2 + 1] = u[1];
                xb[q * 2] = w[0];
                xb[q * 2 + 1] = w[1];
            } else {
                xa[q] = chr_raw_load_f32(rsA, voff[q], si, 0);
                xb[q] = chr_raw_load_f32(rj, voff[q], sj, 0);
            }
        }
    };
    // Draw FIFO: the move and alpha uniforms of steps t .. t + PF - 1 (entry
    // 0 is the current step), each read from the batch holding it at the
    // step that requests its columns, PF steps ahead.
    Move fm[PF];
    float fu[PF][G];
    fm[0] = read_move<G, CO>(mb, L0);
#pragma unroll
    for (int g = 0; g < G; ++g) fu[0][g] = readlane_f32(mb.u[g], L0);
    load_cols(fm[0], true, cs[0][0], cs[0][1]);
#pragma unroll
    for (int k = 1; k < PF; ++k) {
        fm[k] = fm[0];
#pragma unroll
        for (int g = 0; g < G; ++g) fu[k][g] = 0.0f;
        if (t_begin + k < t_end) {
            const int Lk = (L0 + k) & 63;
            if (Lk == 0) gen_batch<G, CO>(a, k0, k1, (uint64_t)(t_begin + k), group, item.c0, mb, col_scale);
            fm[k] = read_move<G, CO>(mb, Lk);
#pragma unroll
            for (int g = 0; g < G; ++g) fu[k][g] = readlane_f32(mb.u[g], Lk);
        }
        load_cols(fm[k], t_begin + k < t_end, cs[k][0], cs[k][1]);
    }

    float last_i[G], last_j[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        last_i[g] = 0.0f;
        last_j[g] = 0.0f;
    }
    int prev_i = -1, prev_j = -1;
    int buf = 0;
    // Events by step offset u from t_begin: the sample of the state after
    // step next_save (u % thin == 0) is stored at the start of step
    // next_save + 1 (or after the last step), the FP64 slacks are recomputed
    // at the start of step next_resync, and a new batch of 64 draws is made
    // at the start of step next_batch, the first whose draw PF steps ahead
    // opens one; each step tests only next_event, the earliest of these.
    int64_t next_save = 0, next_resync = 0;
    int next_batch = (64 - ((L0 + PF) & 63)) & 63;
    int next_event = 0;
    int64_t sample_idx = 0;

    // Stores the current states (the state after the last step) as sample
    // sample_idx.
    auto save_sample = [&]() {
        const int64_t col = sample_idx * a.nchains + item.c0;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if (g < geff) {
                float* dst = a.samples + V.sample_off + (col + g) * (int64_t)d;
                // Streaming stores: the samples are not read again by this
                // launch, so they should not displace the columns of A in L2.
                if constexpr (XR > 0) {
#pragma unroll
                    for (int q = 0; q < XR; ++q)
                        if (q * 64 + lane < d) __builtin_nontemporal_store(xv[g][q], dst + q * 64 + lane);
                } else {
                    for (int c = tid; c < d; c += T) {
                        float v = xs[g * d + c];
                        if (c == prev_i) v = last_i[g];
                        if (c == prev_j) v = last_j[g];
                        __builtin_nontemporal_store(v, dst + c);
                    }
                }
            }
        }
        ++sample_idx;
    };

    // x_c <- fmaf(alpha, coef, x_c) for chain g, in the lane that holds c.
    // With kColOff, c is the coordinate's column offset.
    auto xr_update = [&](int g, int c, float alpha_g, float coef) {
#pragma unroll
        for (int q = 0; q < XS; ++q) {
            const float nv = fmaf(alpha_g, coef, xv[g][q]);
            xv[g][q] = ((q * 64 + lane) * col_scale == c) ? nv : xv[g][q];
        }
    };

    // One CHR step on the columns (ca, cb), prefetching the columns of step
    // t + PF into (na, nb). The register sets rotate between consecutive
    // steps (the loop below is unrolled by PF + 1), so the prefetched columns
    // are never copied and their loads stay in flight across PF steps.
    // Move logging (a test hook) is decided once per workgroup and runs in
    // the step's event path.
    const bool log_on = a.move_log != nullptr && item.variant == 0 && item.c0 == 0;
    // PF_ALL: every step of the caller's round has a step PF ahead to prefetch
    // for (the main loop), so the per-step bound test is dropped there.
    // Steps are counted from t_begin in 32 bits (a launch runs fewer than
    // 2^31); the 64-bit step number is formed only for a new draw batch.
    const int nst = (int)(t_end - t_begin);
    auto step = [&](auto pf_all, int u, float (&ca)[R], float (&cb)[R], float (&na)[R], float (&nb)[R]) {
        constexpr bool kPfAll = decltype(pf_all)::value;
        const Move mv = fm[0];
        const int mtype = CO ? kMoveCoord : mv.type;
        // Events are rare (a sample every thin steps, a draw batch every 64,
        // a resync every resync steps): the hint keeps the step's common path
        // straight-line, with the event path out of line.
        if (__builtin_expect(u == next_event, 0)) {
            if (u == next_save + 1) {
                save_sample();
                next_save += a.thin;
            }
            if (u == next_resync) {
                // s = b - A y in FP64 (chr.cu:386-395); also the initial slack.
                // The partner row's FMA chain over -a gives exactly -acc (a zero
                // sum stays +0 in both: RN(x + y) is -0 only when x and y are),
                // so its slack is bn - (acc != 0 ? -acc : +0), no second chain.
                double acc[G][R];
    #pragma unroll
                for (int k = 0; k < R; ++k)
    #pragma unroll
                    for (int g = 0; g < G; ++g) acc[g][k] = 0.0;
                const float* pc = Av;
                auto add_col = [&](const double (&xc)[G]) {
    #pragma unroll
                    for (int k = 0; k < R; ++k) {
                        const double av = (double)pc[row_of(k)];
    #pragma unroll
                        for (int g = 0; g < G; ++g) acc[g][k] = fma(av, xc[g], acc[g][k]);
                    }
                    pc += lda;
                };
                if constexpr (XR > 0) {
    #pragma unroll
                    for (int q = 0; q < XR; ++q)
                        for (int l = 0; l < 64 && q * 64 + l < d; ++l) {
                            double xc[G];
    #pragma unroll
                            for (int g = 0; g < G; ++g) xc[g] = (double)readlane_f32(xv[g][q], l);
                            add_col(xc);
                        }
                } else {
                    __syncthreads();
                    for (int c = 0; c < d; ++c) {
                        double xc[G];
    #pragma unroll
                        for (int g = 0; g < G; ++g) xc[g] = (double)xs[g * d + c];
                        add_col(xc);
                    }
                }
                const double* bv = a.b + V.b_off;
    #pragma unroll
                for (int k = 0; k < R; ++k) {
                    const double bb = bv[row_of(k)];
    #pragma unroll
                    for (int g = 0; g < G; ++g)
                        s[g][k] = SC ? clamp_unit((float)(bb - acc[g][k])) : clamp_slack((float)(bb - acc[g][k]));
                    if constexpr (PAIRED) {
                        const double bn = a.bn[V.b_off + row_of(k)];
    #pragma unroll
                        for (int g = 0; g < G; ++g) {
                            const float v = (float)(bn - (acc[g][k] != 0.0 ? -acc[g][k] : 0.0));
                            sn[g][k] = SC ? clamp_unit(v) : clamp_slack(v);
                        }
                    }
                }
                prev_i = -1;
                prev_j = -1;
                // No resync: a step the launch never reaches (fewer than 2^31 steps).
                next_resync = a.resync > 0 ? u + (int64_t)a.resync : (int64_t)0x7FFFFFFF;
            }
            if (u == next_batch) {
                gen_batch<G, CO>(a, k0, k1, (uint64_t)(t_begin + u + PF), group, item.c0, mb, col_scale);
                next_batch += 64;
            }
            // Move logging is an event of every step while it lasts.
            if (log_on && u < (int)a.nlog && tid == 0) {
                const int64_t tl = u;
                a.move_log[tl * 3 + 0] = mtype;
                a.move_log[tl * 3 + 1] = kColOff ? mv.i / col_scale : mv.i;
                a.move_log[tl * 3 + 2] = (mtype == kMovePair) ? mv.j : -1;
            }
            int64_t e = next_resync < next_save + 1 ? next_resync : next_save + 1;
            if (next_batch < e) e = next_batch;
            if (log_on && u + 1 < (int)a.nlog && u + 1 < e) e = u + 1;
            next_event = e < 0x7FFFFFFF ? (int)e : 0x7FFFFFFF;
        }

        // Alpha uniforms of step t, and the draw of step t + PF (from the
        // batch that holds it; the FIFO already holds the draws of the steps
        // before it).
        float U[G];
#pragma unroll
        for (int g = 0; g < G; ++g) U[g] = fu[0][g];
        Move nf = mv;  // the move of step t + PF
        float nu[G];
#pragma unroll
        for (int g = 0; g < G; ++g) nu[g] = 0.0f;
        // Launches run fewer than 2^31 steps: 32-bit differences compare on
        // the scalar unit.
        const bool has_pf = kPfAll || nst - u > PF;
        if (has_pf) {
            const int Lp = (L0 + u + PF) & 63;
            nf = read_move<G, CO>(mb, Lp);
#pragma unroll
            for (int g = 0; g < G; ++g) nu[g] = readlane_f32(mb.u[g], Lp);
        }

        // Prefetch the columns of step t + PF. Every step issues exactly 2R/VEC
        // loads, so the in-order vmcnt wait for this step's columns never
        // waits on the later prefetches.
        load_cols(nf, has_pf, na, nb);

        // ae = (A u)_r for this lane's rows.
        float ae[R];
        const int kq = a.ks_k;
        if (mtype == kMoveCoord) {
#pragma unroll
            for (int k = 0; k < R; ++k) ae[k] = ca[k];
        } else if (mtype == kMovePair) {
#pragma unroll
            for (int k = 0; k < R; ++k) ae[k] = fmaf(ca[k], mv.ci, cb[k] * mv.cj);
        } else {
            // Rare escape move: entry q's index and coefficient go through LDS.
            __syncthreads();
            if (tid < kq) {
                ks_idx[tid] = ksparse_index(a, k0, k1, (uint64_t)(t_begin + u), group, tid);
                ks_coef[tid] = ((mv.w3 >> tid) & 1u) ? -a.ks_mag : a.ks_mag;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < R; ++k) ae[k] = 0.0f;
            for (int q = 0; q < kq; ++q) {
                const float* pq = Av + (int64_t)ks_idx[q] * lda;
                const float cq = ks_coef[q];
#pragma unroll
                for (int k = 0; k < R; ++k) ae[k] = fmaf(pq[row_of(k)], cq, ae[k]);
            }
        }

        // Pre-update coordinates (LDS states). With one wavefront its own LDS
        // write of the previous update is ordered before this read, so the
        // read is issued here, early; with several, registers carry the
        // previous update because the owner's LDS write of it is not yet
        // ordered.
        float oi[G], oj[G];
        if (XR == 0 && mtype != kMoveKsparse) {
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (W == 1) {
                    oi[g] = xs[g * d + mv.i];
                    oj[g] = (mtype == kMovePair) ? xs[g * d + mv.j] : 0.0f;
                } else {
                    oi[g] = (mv.i == prev_j) ? last_j[g] : ((mv.i == prev_i) ? last_i[g] : xs[g * d + mv.i]);
                    oj[g] = 0.0f;
                    if (mtype == kMovePair)
                        oj[g] = (mv.j == prev_j) ? last_j[g] : ((mv.j == prev_i) ? last_i[g] : xs[g * d + mv.j]);
                }
            }
        }

        // Chord ends per chain: every ratio of the lane, reduced as a tree,
        // then across the wavefront. One wavefront forms alpha in lane 63,
        // where the minima end, and reads that one value back; several
        // exchange their minima through LDS first.
        constexpr int NQ = PAIRED ? 2 * R : R;
        uint32_t rhi[G];
        int32_t rlo[G];
        float alpha[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            uint32_t qv[NQ];
            if (mtype == kMoveCoord) {
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    qv[k] = chord_bits_rcp(cb[k], s[g][k]);
                    if constexpr (PAIRED) qv[R + k] = chord_bits_rcp(-cb[k], sn[g][k]);
                }
            } else {
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    qv[k] = chord_bits(ae[k], s[g][k]);
                    if constexpr (PAIRED) qv[R + k] = chord_bits(-ae[k], sn[g][k]);
                }
            }
            uint32_t hb;
            int32_t lb;
            lane_min2(qv, hb, lb);
            if constexpr (W == 1) {
                wave_min2_dpp(hb, lb);
                const float hi = __builtin_bit_cast(float, hb);
                const float lo = __builtin_bit_cast(float, lb);
                alpha[g] = readlane_f32(fmaf(U[g], hi - lo, lo), 63);
            } else {
                wave_min2(hb, lb);
                rhi[g] = hb;
                rlo[g] = lb;
                if (lane == 0) part[buf][g][wave] = make_uint2(hb, (uint32_t)lb);
            }
        }

        if constexpr (W > 1) {
            __syncthreads();
#pragma unroll
            for (int g = 0; g < G; ++g) {
                uint32_t hb = rhi[g];
                int32_t lb = rlo[g];
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    const uint2 p = part[buf][g][w];
                    hb = umin32(hb, p.x);
                    lb = smin32(lb, (int32_t)p.y);
                }
                const float hi = __builtin_bit_cast(float, hb);
                const float lo = __builtin_bit_cast(float, lb);
                alpha[g] = fmaf(U[g], hi - lo, lo);
            }
        }

        // Fused rank-1 slack update s -= alpha * ae (chr.cu:464-482); the
        // partner row's is sn -= alpha * (-ae).
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const float nal = -alpha[g];
#pragma unroll
            for (int k = 0; k < R; ++k)
                s[g][k] = SC ? clamp_unit(fmaf(nal, ae[k], s[g][k])) : clamp_slack(fmaf(nal, ae[k], s[g][k]));
            if constexpr (PAIRED) {
#pragma unroll
                for (int k = 0; k < R; ++k)
                    sn[g][k] = SC ? clamp_unit(fmaf(alpha[g], ae[k], sn[g][k]))
                                  : clamp_slack(fmaf(alpha[g], ae[k], sn[g][k]));
            }
        }

        if constexpr (XR > 0) {
            // Same roundings as the LDS form: x_i = fmaf(alpha, ci, x_i), then
            // x_j = fmaf(alpha, cj, x_j), the second on the first's result
            // when j == i; k-sparse entries in order.
            if (mtype != kMoveKsparse) {
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    xr_update(g, mv.i, alpha[g], mv.ci);
                    if (mtype == kMovePair) xr_update(g, mv.j, alpha[g], mv.cj);
                }
            } else {
                for (int q = 0; q < kq; ++q) {
                    const int c = __builtin_amdgcn_readfirstlane(ks_idx[q]);
                    const float cq = ks_coef[q];
#pragma unroll
                    for (int g = 0; g < G; ++g) xr_update(g, c, alpha[g], cq);
                }
            }
        } else if (mtype != kMoveKsparse) {
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const float ni = fmaf(alpha[g], mv.ci, oi[g]);
                last_i[g] = ni;
                if (mtype == kMovePair) {
                    const float base = (mv.j == mv.i) ? ni : oj[g];
                    last_j[g] = fmaf(alpha[g], mv.cj, base);
                }
            }
            prev_i = mv.i;
            prev_j = (mtype == kMovePair) ? mv.j : -1;
            if (tid == 0) {
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    xs[g * d + mv.i] = last_i[g];
                    if (mtype == kMovePair) xs[g * d + mv.j] = last_j[g];
                }
            }
        } else {
            __syncthreads();
            if (tid == 0) {
#pragma unroll
                for (int g = 0; g < G; ++g)
                    for (int q = 0; q < kq; ++q)
                        xs[g * d + ks_idx[q]] = fmaf(alpha[g], ks_coef[q], xs[g * d + ks_idx[q]]);
            }
            __syncthreads();
            prev_i = -1;
            prev_j = -1;
        }


#pragma unroll
        for (int k = 0; k + 1 < PF; ++k) {
            fm[k] = fm[k + 1];
#pragma unroll
            for (int g = 0; g < G; ++g) fu[k][g] = fu[k + 1][g];
        }
        fm[PF - 1] = nf;
#pragma unroll
        for (int g = 0; g < G; ++g) fu[PF - 1][g] = nu[g];
        buf ^= 1;
    };
    // Step t + j of a round computes on set j and prefetches into set
    // (j + PF) % (PF + 1), the one the step before it has finished with.
    // Rounds whose every step u + j (j <= PF) has u + j + PF < nst, then
    // at most one checked round, then the last steps.
    int u = 0;
    for (; u + 2 * PF < nst; u += PF + 1) {
#pragma unroll
        for (int j = 0; j <= PF; ++j)
            step(std::true_type{}, u + j, cs[j][0], cs[j][1], cs[(j + PF) % (PF + 1)][0],
                 cs[(j + PF) % (PF + 1)][1]);
    }
    if (u + PF < nst) {
#pragma unroll
        for (int j = 0; j <= PF; ++j)
            step(std::false_type{}, u + j, cs[j][0], cs[j][1], cs[(j + PF) % (PF + 1)][0],
                 cs[(j + PF) % (PF + 1)][1]);
        u += PF + 1;
    }
#pragma unroll
    for (int j = 0; j < PF; ++j)
        if (u + j < nst)
            step(std::false_type{}, u + j, cs[j][0], cs[j][1], cs[(j + PF) % (PF + 1)][0],
                 cs[(j + PF) % (PF + 1)][1]);
    if ((int64_t)nst == next_save + 1) save_sample();  // the last step's sample

    if constexpr (XR > 0) {
        if (a.y_final) {
#pragma unroll
            for (int g = 0; g < G; ++g)
#pragma unroll
                for (int q = 0; q < XR; ++q)
                    if (g < geff && q * 64 + lane < d)
                        a.y_final[(V.y0_col + item.c0 + g) * (int64_t)d + q * 64 + lane] = xv[g][q];
        }
    } else {
        __syncthreads();
        if (a.y_final) {
            for (int g = 0; g < geff; ++g)
                for (int c = tid; c < d; c += T)
                    a.y_final[(V.y0_col + item.c0 + g) * (int64_t)d + c] = xs[g * d + c];
        }
    }
}

// Large polytopes (more rows than the register-resident kernel above holds):
// one chain per workgroup of W wavefronts, R rows per lane, a register
// budget of 3R: the slack s and two column arrays that swap roles every step
// (the loop is unrolled by two). X holds this step's column A[:, i],
// prefetched during the previous step; a pair move loads its partner
// A[:, j] into Y at the start of the step and combines in place, X = ci X +
// cj Y; then the next step's A[:, i] is prefetched into Y and stays in
// flight during the chord, the reductions and the update. Buffer loads:
// the variant's block is the resource, the lane's row the one VGPR offset,
// the column / row-slot offset a wave-uniform SGPR, so the R loads of a
// column cost no address registers. No reciprocal column: a coordinate move
// forms RN(1/A[r,i]) with a correctly rounded division, which keeps the chord
// bits of the packed-reciprocal kernel (q = RN(s * RN(1/A))) while reading
// one column per step from memory instead of two. The FP64 resync keeps one
// accumulator at a time (same c-ascending FMA order as above).
template <int W, int R>
__global__ __launch_bounds__(64 * W) void chr_big_kernel(ChrArgs a)
{
    constexpr int T = 64 * W;
    extern __shared__ float xs[];  // d: the chain's state
    __shared__ uint2 part[2][W];
    __shared__ int ks_idx[kKmax];
    __shared__ float ks_coef[kKmax];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const ChrItem item = a.items[blockIdx.x];
    const ChrVariant V = a.variants[item.variant];
    const int d = a.d;
    const int lda = V.lda;  // == T * R
    const chr_i32x4 rsrc = chr_make_rsrc(a.A + V.a_off, (uint32_t)(4 * (int64_t)lda * d));
    const int voff = 4 * tid;
    auto ld = [&](int col, int k) -> float { return chr_raw_load_f32(rsrc, voff, 4 * (col * lda + k * T), 0); };
    // b of the block through a buffer too (64-bit global addresses per row
    // slot would be hoisted out of the step loop and held in VGPRs).
    const chr_i32x4 rsrc_b = chr_make_rsrc(a.b + V.b_off, (uint32_t)(8 * lda));
    auto ldb = [&](int k) -> double { return __builtin_bit_cast(double, chr_raw_load_v2f32(rsrc_b, 8 * tid, 8 * k * T, 0)); };
    const uint32_t group = (uint32_t)item.c0;  // one chain per group
    const uint32_t k0 = a.seed_lo;
    const uint32_t k1 = a.seed_hi ^ (uint32_t)item.variant;
    const int64_t t_begin = a.t0;
    const int64_t t_end = a.t0 + a.nsteps;

    for (int c = tid; c < d; c += T) xs[c] = a.y0[(V.y0_col + item.c0) * (int64_t)d + c];

    float s[R], X0[R], X1[R];
    MoveBatch<1> mb;
    gen_batch<1, false>(a, k0, k1, (uint64_t)(t_begin & ~(int64_t)63), group, item.c0, mb);
    Move mv = read_move<1, false>(mb, (int)(t_begin & 63));
    {
        const int ci = (mv.type == kMoveKsparse) ? 0 : mv.i;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            X0[k] = ld(ci, k);
            X1[k] = 0.0f;
        }
    }
    float last_i = 0.0f, last_j = 0.0f;
    int prev_i = -1, prev_j = -1;
    int buf = 0;
    int until_save = 0;
    int until_resync = 0;
    int64_t sample_idx = 0;

    const bool log_on = a.move_log != nullptr && item.variant == 0 && item.c0 == 0;
    auto step = [&](int64_t t, float (&X)[R], float (&Y)[R]) {
        if (until_resync == 0) {
            __syncthreads();
#pragma unroll
            for (int k = 0; k < R; ++k) {
                double acc = 0.0;
                for (int c = 0; c < d; ++c) acc = fma((double)ld(c, k), (double)xs[c], acc);
                s[k] = clamp_slack((float)(ldb(k) - acc));
            }
            prev_i = -1;
            prev_j = -1;
            // No resync: a count the launch never reaches (fewer than 2^31 steps).
            until_resync = a.resync > 0 ? a.resync : 0x7FFFFFFF;
        }

        if (log_on) {
            if ((int)(t - t_begin) < (int)a.nlog && tid == 0) {
                const int64_t tl = t - t_begin;
                a.move_log[tl * 3 + 0] = mv.type;
                a.move_log[tl * 3 + 1] = mv.i;
                a.move_log[tl * 3 + 2] = (mv.type == kMovePair) ? mv.j : -1;
            }
        }

        const int L = (int)(t & 63);
        const float U = readlane_f32(mb.u[0], L);
        Move nx = mv;
        const bool has_next = t + 1 < t_end;
        if (has_next) {
            if (L == 63) gen_batch<1, false>(a, k0, k1, (uint64_t)(t + 1), group, item.c0, mb);
            nx = read_move<1, false>(mb, (L + 1) & 63);
        }

        // This step's direction column in X.
        const int kq = a.ks_k;
        if (mv.type == kMovePair) {
#pragma unroll
            for (int k = 0; k < R; ++k) Y[k] = ld(mv.j, k);
#pragma unroll
            for (int k = 0; k < R; ++k) X[k] = fmaf(X[k], mv.ci, Y[k] * mv.cj);
        } else if (mv.type == kMoveKsparse) {
            __syncthreads();
            if (tid < kq) {
                ks_idx[tid] = ksparse_index(a, k0, k1, (uint64_t)t, group, tid);
                ks_coef[tid] = ((mv.w3 >> tid) & 1u) ? -a.ks_mag : a.ks_mag;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < R; ++k) X[k] = 0.0f;
            for (int q = 0; q < kq; ++q) {
                const int col = __builtin_amdgcn_readfirstlane(ks_idx[q]);
                const float cq = ks_coef[q];
#pragma unroll
                for (int k = 0; k < R; ++k) X[k] = fmaf(ld(col, k), cq, X[k]);
            }
        }
        // The next step's first column, in flight for the rest of this step.
        {
            const int ni = (has_next && nx.type != kMoveKsparse) ? nx.i : 0;
#pragma unroll
            for (int k = 0; k < R; ++k) Y[k] = ld(ni, k);
        }

        uint32_t hb = 0xFFFFFFFFu;
        int32_t lb = 0x7FFFFFFF;
        if (mv.type == kMoveCoord) {
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const uint32_t qb = chord_bits_rcp(1.0f / X[k], s[k]);
                hb = umin32(hb, qb);
                lb = smin32(lb, (int32_t)qb);
            }
        } else {
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const uint32_t qb = chord_bits(X[k], s[k]);
                hb = umin32(hb, qb);
                lb = smin32(lb, (int32_t)qb);
            }
        }
        wave_min2(hb, lb);
        if (lane == 0) part[buf][wave] = make_uint2(hb, (uint32_t)lb);

        float oi = 0.0f, oj = 0.0f;
        if (mv.type != kMoveKsparse) {
            oi = (mv.i == prev_j) ? last_j : ((mv.i == prev_i) ? last_i : xs[mv.i]);
            if (mv.type == kMovePair) oj = (mv.j == prev_j) ? last_j : ((mv.j == prev_i) ? last_i : xs[mv.j]);
        }
        __syncthreads();
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint2 p = part[buf][w];
            hb = umin32(hb, p.x);
            lb = smin32(lb, (int32_t)p.y);
        }
        const float hi = __builtin_bit_cast(float, hb);
        const float lo = __builtin_bit_cast(float, lb);
        const float alpha = fmaf(U, hi - lo, lo);

        const float nal = -alpha;
#pragma unroll
        for (int k = 0; k < R; ++k) s[k] = clamp_slack(fmaf(nal, X[k], s[k]));

        if (mv.type != kMoveKsparse) {
            const float ni = fmaf(alpha, mv.ci, oi);
            last_i = ni;
            if (mv.type == kMovePair) {
                const float base = (mv.j == mv.i) ? ni : oj;
                last_j = fmaf(alpha, mv.cj, base);
            }
            prev_i = mv.i;
            prev_j = (mv.type == kMovePair) ? mv.j : -1;
            if (tid == 0) {
                xs[mv.i] = last_i;
                if (mv.type == kMovePair) xs[mv.j] = last_j;
            }
        } else {
            __syncthreads();
            if (tid == 0)
                for (int q = 0; q < kq; ++q) xs[ks_idx[q]] = fmaf(alpha, ks_coef[q], xs[ks_idx[q]]);
            __syncthreads();
            prev_i = -1;
            prev_j = -1;
        }

        if (until_save == 0) {
            const int64_t col = sample_idx * a.nchains + item.c0;
            float* dst = a.samples + V.sample_off + col * (int64_t)d;
            for (int c = tid; c < d; c += T) {
                float v = xs[c];
                if (c == prev_i) v = last_i;
                if (c == prev_j) v = last_j;
                __builtin_nontemporal_store(v, dst + c);  // streaming, as above
            }
            ++sample_idx;
            until_save = a.thin;
        }
        --until_save;
        --until_resync;
        mv = nx;
        buf ^= 1;
    };
    int64_t t = t_begin;
    for (; t + 1 < t_end; t += 2) {
        step(t, X0, X1);
        step(t + 1, X1, X0);
    }
    if (t < t_end) step(t, X0, X1);

    __syncthreads();
    if (a.y_final)
        for (int c = tid; c < d; c += T) a.y_final[(V.y0_col + item.c0) * (int64_t)d + c] = xs[c];
}

}  // namespace chr_detail
}  // namespace naja
