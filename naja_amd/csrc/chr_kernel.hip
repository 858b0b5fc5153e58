// Multi-chain coordinate hit-and-run step kernel for gfx950.
//
// Replaces the reference's chrKernel (chr.cu:233-490), launchChrKernel
// (chr.cu:499-553), initPrngStatesCHR (chr.cu:556-560) and the slack-init
// DGEMM of CoordinateHitAndRun (chr.cu:34-40).
//
// Layout and schedule (one workgroup = one group of G chains that share the
// move draw; W wavefronts of 64 lanes):
//   * row r of the polytope belongs to lane tid = r mod 64W, slot r / 64W;
//     the slack s_r of every chain of the group lives in that lane's VGPRs
//     for the whole run (R slots, G chains), so the per-step slack traffic of
//     the reference (read twice, written once in global memory, chr.cu:402,
//     chr.cu:481) never leaves the register file;
//   * the chosen column A[:, i] (and A[:, j] for pair moves) is read once per
//     step per workgroup and broadcast from registers to all G chains; the
//     move for step t+1 depends only on the Philox counter, so its columns are
//     prefetched while step t computes;
//   * the chord's max/min of inverse distances is a DPP wavefront reduction
//     plus one LDS exchange and one barrier per step (double-buffered
//     partials), instead of the reference's two CUB block reductions and
//     three barriers;
//   * x lives in LDS; the lane-0 owner applies the 1-2 coordinate update, and
//     every lane carries the last update in registers so no second barrier is
//     needed before the next read;
//   * the slack is FP32 and is recomputed in FP64 as s = b - A y at t = 0 and
//     every resync_interval steps (the reference's resync, chr.cu:386-395).
// Arithmetic is the oracle's, operation for operation (explicit fmaf, IEEE
// division), so GPU and oracle trajectories agree bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "chr_kernel.h"
#include "common.h"
#include "philox.h"

namespace naja {

namespace {

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

// Max and min over the 64 lanes of a wavefront: DPP within each row of 16
// (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror),
// then the four row results through readlane.
__device__ __forceinline__ void wave_maxmin(float& vmax, float& vmin)
{
    vmax = fmaxf(vmax, dpp_f32<0xB1>(vmax));
    vmin = fminf(vmin, dpp_f32<0xB1>(vmin));
    vmax = fmaxf(vmax, dpp_f32<0x4E>(vmax));
    vmin = fminf(vmin, dpp_f32<0x4E>(vmin));
    vmax = fmaxf(vmax, dpp_f32<0x141>(vmax));
    vmin = fminf(vmin, dpp_f32<0x141>(vmin));
    vmax = fmaxf(vmax, dpp_f32<0x140>(vmax));
    vmin = fminf(vmin, dpp_f32<0x140>(vmin));
    const float a0 = readlane_f32(vmax, 0), a1 = readlane_f32(vmax, 16);
    const float a2 = readlane_f32(vmax, 32), a3 = readlane_f32(vmax, 48);
    const float b0 = readlane_f32(vmin, 0), b1 = readlane_f32(vmin, 16);
    const float b2 = readlane_f32(vmin, 32), b3 = readlane_f32(vmin, 48);
    vmax = fmaxf(fmaxf(a0, a1), fmaxf(a2, a3));
    vmin = fminf(fminf(b0, b1), fminf(b2, b3));
}

// chr.cu:416-426: s == 0 -> +-inf by the sign of ae (0 if ae == 0); NaN -> 0.
// A slack that FP32 rounding left negative takes the same branch, so a point
// just outside a facet is treated as on it (the FP64 reference has no clamp).
__device__ __forceinline__ float inverse_distance(float ae, float s)
{
    float q;
    if (s <= 0.0f) {
        q = (ae > 0.0f) ? INFINITY : ((ae < 0.0f) ? -INFINITY : 0.0f);
    } else {
        q = ae / s;
        q = __builtin_isnan(q) ? 0.0f : q;
    }
    return q;
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

template <int G>
__device__ __forceinline__ void gen_batch(const ChrArgs& a, uint32_t k0, uint32_t k1, uint64_t tb, uint32_t group,
                                          int c0, MoveBatch<G>& mb)
{
    const uint64_t t = tb + (uint64_t)(threadIdx.x & 63);
    const Move mv = draw_move(a, k0, k1, t, group);
    mb.tij = (mv.type << 30) | mv.i;
    mb.j = (mv.type == kMoveKsparse) ? (int)mv.w3 : mv.j;
    mb.ci = mv.ci;
    mb.cj = mv.cj;
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

template <int G>
__device__ __forceinline__ Move read_move(const MoveBatch<G>& mb, int lane)
{
    Move mv;
    const int tij = readlane_i32(mb.tij, lane);
    mv.type = (int)((uint32_t)tij >> 30);
    mv.i = tij & 0x3FFFFFFF;
    const int jv = readlane_i32(mb.j, lane);
    mv.j = jv;
    mv.w3 = (uint32_t)jv;
    mv.ci = readlane_f32(mb.ci, lane);
    mv.cj = readlane_f32(mb.cj, lane);
    return mv;
}

template <int W, int G, int R>
__global__ __launch_bounds__(64 * W) void chr_step_kernel(ChrArgs a)
{
    constexpr int T = 64 * W;
    extern __shared__ float xs[];  // G x d current states
    __shared__ float2 part[2][G][W];
    __shared__ int ks_idx[kKmax];
    __shared__ float ks_coef[kKmax];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const ChrItem item = a.items[blockIdx.x];
    const ChrVariant V = a.variants[item.variant];
    const int d = a.d;
    const int lda = V.lda;
    const float* Av = a.A + V.a_off;
    const int geff = min(G, a.nchains - item.c0);
    const uint32_t group = (uint32_t)(item.c0 / G);
    const uint32_t k0 = a.seed_lo;
    const uint32_t k1 = a.seed_hi ^ (uint32_t)item.variant;
    const bool log_moves = (a.move_log != nullptr) && item.variant == 0 && item.c0 == 0 && tid == 0;

    for (int g = 0; g < G; ++g) {
        for (int c = tid; c < d; c += T) {
            xs[g * d + c] = (g < geff) ? a.y0[(V.y0_col + item.c0 + g) * (int64_t)d + c] : 0.0f;
        }
    }

    // Slot k of this wave covers rows k*T + wave*64 + [0, 64); whole slots
    // beyond lda are skipped with a wave-uniform test.
    const int wrow = wave * 64;
    int nslot = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) nslot += (k * T + wrow < lda) ? 1 : 0;

    float s[G][R];
    float ca[R], cb[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        ca[k] = 0.0f;
        cb[k] = 0.0f;
#pragma unroll
        for (int g = 0; g < G; ++g) s[g][k] = __builtin_huge_valf();
    }

    MoveBatch<G> mb;
    gen_batch<G>(a, k0, k1, 0, group, item.c0, mb);
    Move mv = read_move<G>(mb, 0);
    {
        const float* pi = Av + (int64_t)mv.i * lda + wrow + lane;
        const float* pj = Av + (int64_t)(mv.j < 0 ? 0 : mv.j) * lda + wrow + lane;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (k < nslot) {
                if (mv.type != kMoveKsparse) ca[k] = pi[k * T];
                if (mv.type == kMovePair) cb[k] = pj[k * T];
            }
        }
    }

    float last_i[G], last_j[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        last_i[g] = 0.0f;
        last_j[g] = 0.0f;
    }
    int prev_i = -1, prev_j = -1;
    int buf = 0;
    int until_save = 0;    // steps until the next saved sample (t % thin == 0)
    int until_resync = 0;  // steps until the next FP64 slack computation
    int64_t sample_idx = 0;
    const int64_t nsteps = a.nsteps;

    for (int64_t t = 0; t < nsteps; ++t) {
        if (until_resync == 0) {
            // s = b - A y in FP64 (chr.cu:386-395); also the initial slack.
            __syncthreads();
            double acc[G][R];
#pragma unroll
            for (int k = 0; k < R; ++k)
#pragma unroll
                for (int g = 0; g < G; ++g) acc[g][k] = 0.0;
            const float* pc = Av + wrow + lane;
#pragma unroll 2
            for (int c = 0; c < d; ++c) {
                double xv[G];
#pragma unroll
                for (int g = 0; g < G; ++g) xv[g] = (double)xs[g * d + c];
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    if (k < nslot) {
                        const double av = (double)pc[k * T];
#pragma unroll
                        for (int g = 0; g < G; ++g) acc[g][k] = fma(av, xv[g], acc[g][k]);
                    }
                }
                pc += lda;
            }
            const double* bv = a.b + V.b_off + wrow + lane;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                if (k < nslot) {
                    const double bb = bv[k * T];
#pragma unroll
                    for (int g = 0; g < G; ++g) s[g][k] = (float)(bb - acc[g][k]);
                }
            }
            prev_i = -1;
            prev_j = -1;
            until_resync = a.resync > 0 ? a.resync : -1;
        }

        if (log_moves && t < a.nlog) {
            a.move_log[t * 3 + 0] = mv.type;
            a.move_log[t * 3 + 1] = mv.i;
            a.move_log[t * 3 + 2] = (mv.type == kMovePair) ? mv.j : -1;
        }

        // Alpha uniforms of step t, then the draw of step t+1 (a new batch of
        // 64 steps when t+1 crosses a batch boundary).
        const int L = (int)(t & 63);
        float U[G];
#pragma unroll
        for (int g = 0; g < G; ++g) U[g] = readlane_f32(mb.u[g], L);
        Move nx = mv;
        const bool has_next = t + 1 < nsteps;
        if (has_next) {
            if (L == 63) gen_batch<G>(a, k0, k1, (uint64_t)(t + 1), group, item.c0, mb);
            nx = read_move<G>(mb, (L + 1) & 63);
        }

        // Prefetch the columns of step t+1 (moves do not depend on state).
        float na[R], nb[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            na[k] = 0.0f;
            nb[k] = 0.0f;
        }
        if (has_next && nx.type != kMoveKsparse) {
            const float* pi = Av + (int64_t)nx.i * lda + wrow + lane;
#pragma unroll
            for (int k = 0; k < R; ++k)
                if (k < nslot) na[k] = pi[k * T];
            if (nx.type == kMovePair) {
                const float* pj = Av + (int64_t)nx.j * lda + wrow + lane;
#pragma unroll
                for (int k = 0; k < R; ++k)
                    if (k < nslot) nb[k] = pj[k * T];
            }
        }

        // ae = (A u)_r for this lane's rows.
        float ae[R];
        const int kq = a.ks_k;
        if (mv.type == kMoveCoord) {
#pragma unroll
            for (int k = 0; k < R; ++k) ae[k] = ca[k];
        } else if (mv.type == kMovePair) {
#pragma unroll
            for (int k = 0; k < R; ++k) ae[k] = fmaf(ca[k], mv.ci, cb[k] * mv.cj);
        } else {
            // Rare escape move: entry q's index and coefficient go through LDS.
            __syncthreads();
            if (tid < kq) {
                ks_idx[tid] = ksparse_index(a, k0, k1, (uint64_t)t, group, tid);
                ks_coef[tid] = ((mv.w3 >> tid) & 1u) ? -a.ks_mag : a.ks_mag;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < R; ++k) {
                float acc = 0.0f;
                if (k < nslot) {
                    for (int q = 0; q < kq; ++q)
                        acc = fmaf(Av[(int64_t)ks_idx[q] * lda + k * T + wrow + lane], ks_coef[q], acc);
                }
                ae[k] = acc;
            }
        }

        // Chord: max / min of inverse distances per chain.
        float rmax[G], rmin[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            float vmax = -3.402823466e+38f, vmin = 3.402823466e+38f;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                if (k < nslot) {
                    const float q = inverse_distance(ae[k], s[g][k]);
                    vmax = fmaxf(vmax, q);
                    vmin = fminf(vmin, q);
                }
            }
            wave_maxmin(vmax, vmin);
            rmax[g] = vmax;
            rmin[g] = vmin;
            if (W > 1 && lane == 0) part[buf][g][wave] = make_float2(vmax, vmin);
        }

        // Pre-update coordinates (registers carry the previous step's update
        // because the owner's LDS write of it is not yet ordered).
        float oi[G], oj[G];
        if (mv.type != kMoveKsparse) {
#pragma unroll
            for (int g = 0; g < G; ++g) {
                oi[g] = (mv.i == prev_j) ? last_j[g] : ((mv.i == prev_i) ? last_i[g] : xs[g * d + mv.i]);
                oj[g] = 0.0f;
                if (mv.type == kMovePair)
                    oj[g] = (mv.j == prev_j) ? last_j[g] : ((mv.j == prev_i) ? last_i[g] : xs[g * d + mv.j]);
            }
        }

        if (W > 1) __syncthreads();

        float alpha[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            float vmax = rmax[g], vmin = rmin[g];
            if (W > 1) {
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    const float2 p = part[buf][g][w];
                    vmax = fmaxf(vmax, p.x);
                    vmin = fminf(vmin, p.y);
                }
            }
            const float lo = 1.0f / vmin;
            const float hi = 1.0f / vmax;
            alpha[g] = fmaf(U[g], hi - lo, lo);
        }

        // Fused rank-1 slack update s -= alpha * ae (chr.cu:464-482).
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const float nal = -alpha[g];
#pragma unroll
            for (int k = 0; k < R; ++k) s[g][k] = fmaf(nal, ae[k], s[g][k]);
        }

        if (mv.type != kMoveKsparse) {
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const float ni = fmaf(alpha[g], mv.ci, oi[g]);
                last_i[g] = ni;
                if (mv.type == kMovePair) {
                    const float base = (mv.j == mv.i) ? ni : oj[g];
                    last_j[g] = fmaf(alpha[g], mv.cj, base);
                }
            }
            prev_i = mv.i;
            prev_j = (mv.type == kMovePair) ? mv.j : -1;
            if (tid == 0) {
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    xs[g * d + mv.i] = last_i[g];
                    if (mv.type == kMovePair) xs[g * d + mv.j] = last_j[g];
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

        if (until_save == 0) {
            const int64_t col = sample_idx * a.nchains + item.c0;
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (g < geff) {
                    float* dst = a.samples + V.sample_off + (col + g) * (int64_t)d;
                    for (int c = tid; c < d; c += T) {
                        float v = xs[g * d + c];
                        if (c == prev_i) v = last_i[g];
                        if (c == prev_j) v = last_j[g];
                        dst[c] = v;
                    }
                }
            }
            ++sample_idx;
            until_save = a.thin;
        }
        --until_save;
        if (until_resync > 0) --until_resync;

        mv = nx;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            ca[k] = na[k];
            cb[k] = nb[k];
        }
        buf ^= 1;
    }

    __syncthreads();
    if (a.y_final) {
        for (int g = 0; g < geff; ++g)
            for (int c = tid; c < d; c += T) a.y_final[(V.y0_col + item.c0 + g) * (int64_t)d + c] = xs[g * d + c];
    }
}

// Packs variant v: rows [A; E_v; 0...] with lda = V.lda, b = [b; f_v; +inf...].
__global__ void chr_pack_kernel(const float* A, const double* b, int m0, int d, const float* A_extra,
                                const double* b_extra, const ChrVariant* variants, const int64_t* extra_off,
                                const int* extra_k, float* A_pack, double* b_pack)
{
    const int v = blockIdx.y;
    const ChrVariant V = variants[v];
    const int kv = extra_k[v];
    const int64_t eoff = extra_off[v];
    const int64_t total = (int64_t)V.lda * d;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(e % V.lda);
        const int c = (int)(e / V.lda);
        float val = 0.0f;
        if (r < m0) val = A[(int64_t)c * m0 + r];
        else if (r < m0 + kv) val = A_extra[eoff * d + (int64_t)c * kv + (r - m0)];
        A_pack[V.a_off + e] = val;
    }
    if (blockIdx.x == 0) {
        for (int r = threadIdx.x; r < V.lda; r += blockDim.x) {
            double bb = __builtin_huge_val();
            if (r < m0) bb = b[r];
            else if (r < m0 + kv) bb = b_extra[eoff + (r - m0)];
            b_pack[V.b_off + r] = bb;
        }
    }
}

using KernelFn = void (*)(ChrArgs);

template <int W, int G, int R>
constexpr KernelFn kfn()
{
    return chr_step_kernel<W, G, R>;
}

struct KernelEntry {
    int W, G, R;
    KernelFn fn;
};

// Compiled launch shapes. Rows per lane R bounds m <= 64 W R.
const KernelEntry kKernels[] = {
    {1, 1, 1, kfn<1, 1, 1>()},   {1, 1, 2, kfn<1, 1, 2>()},   {1, 1, 4, kfn<1, 1, 4>()},
    {1, 1, 8, kfn<1, 1, 8>()},   {1, 1, 16, kfn<1, 1, 16>()}, {4, 1, 1, kfn<4, 1, 1>()},
    {4, 1, 2, kfn<4, 1, 2>()},   {4, 1, 4, kfn<4, 1, 4>()},   {4, 1, 8, kfn<4, 1, 8>()},
    {4, 1, 16, kfn<4, 1, 16>()}, {8, 1, 2, kfn<8, 1, 2>()},   {8, 1, 4, kfn<8, 1, 4>()},
    {8, 1, 8, kfn<8, 1, 8>()},   {8, 1, 16, kfn<8, 1, 16>()}, {16, 1, 1, kfn<16, 1, 1>()},
    {16, 1, 2, kfn<16, 1, 2>()}, {16, 1, 4, kfn<16, 1, 4>()}, {16, 1, 6, kfn<16, 1, 6>()},
    {16, 1, 8, kfn<16, 1, 8>()}, {16, 1, 12, kfn<16, 1, 12>()}, {16, 1, 16, kfn<16, 1, 16>()},
    {16, 1, 24, kfn<16, 1, 24>()}, {16, 1, 32, kfn<16, 1, 32>()},
    {1, 4, 1, kfn<1, 4, 1>()},   {1, 4, 2, kfn<1, 4, 2>()},   {1, 4, 4, kfn<1, 4, 4>()},
    {1, 4, 8, kfn<1, 4, 8>()},   {4, 4, 1, kfn<4, 4, 1>()},   {4, 4, 2, kfn<4, 4, 2>()},
    {4, 4, 4, kfn<4, 4, 4>()},   {4, 4, 8, kfn<4, 4, 8>()},   {16, 4, 1, kfn<16, 4, 1>()},
    {16, 4, 2, kfn<16, 4, 2>()}, {16, 4, 4, kfn<16, 4, 4>()},
};

}  // namespace

int chr_pick_waves(int lda_max, int tpb)
{
    if (tpb > 0) {
        const int w = std::max(1, tpb / 64);
        return w >= 16 ? 16 : (w >= 8 ? 8 : (w >= 4 ? 4 : 1));
    }
    // One wavefront (no barrier) while each lane keeps at most 8 rows;
    // otherwise widen the workgroup so each lane keeps about six rows.
    if (lda_max <= 64 * 8) return 1;
    if (lda_max <= 256 * 8) return 4;
    if (lda_max <= 512 * 6) return 8;
    return 16;
}

bool chr_shape_available(int W, int G, int lda_max, int* R_out)
{
    int best = -1;
    for (const KernelEntry& e : kKernels) {
        if (e.W == W && e.G == G && e.R * 64 * W >= lda_max) {
            if (best < 0 || e.R < best) best = e.R;
        }
    }
    if (R_out) *R_out = best;
    return best > 0;
}

void chr_launch(const ChrArgs& args, int n_items, int W, int G, int R, hipStream_t stream)
{
    KernelFn fn = nullptr;
    for (const KernelEntry& e : kKernels)
        if (e.W == W && e.G == G && e.R == R) fn = e.fn;
    NAJA_REQUIRE(fn != nullptr, NAJA_ERR_UNSUPPORTED, "no compiled CHR kernel for W=%d G=%d R=%d", W, G, R);
    const size_t lds = sizeof(float) * (size_t)G * (size_t)args.d;
    NAJA_REQUIRE(lds <= 96 * 1024, NAJA_ERR_UNSUPPORTED, "d=%d with %d chains per group exceeds the LDS state budget",
                 args.d, G);
    hipLaunchKernelGGL(fn, dim3(n_items), dim3(64 * W), lds, stream, args);
    NAJA_HIP(hipGetLastError());
}

void chr_pack(const float* A, const double* b, int m0, int d, const float* A_extra, const double* b_extra,
              const ChrVariant* variants_dev, const int64_t* extra_off_dev, const int* extra_k_dev, int n_variants,
              float* A_pack, double* b_pack, int lda_max, hipStream_t stream)
{
    const int64_t per = (int64_t)lda_max * d;
    int gx = (int)std::min<int64_t>((per + 255) / 256, 4096);
    if (gx < 1) gx = 1;
    hipLaunchKernelGGL(chr_pack_kernel, dim3(gx, n_variants), dim3(256), 0, stream, A, b, m0, d, A_extra, b_extra,
                       variants_dev, extra_off_dev, extra_k_dev, A_pack, b_pack);
    NAJA_HIP(hipGetLastError());
}

}  // namespace naja
