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
//   * the chord ends (direct form q = s/ae; the least non-negative and the
//     greatest negative q are an unsigned and a signed integer minimum over
//     the float bits) are DPP wavefront reductions plus one LDS exchange and
//     one barrier per step (double-buffered partials), instead of the
//     reference's two CUB block reductions, three barriers and two final
//     reciprocals;
//   * one-wavefront launches with d <= 128 keep x in VGPRs (coordinate c in
//     lane c % 64 of slot c / 64) and every lane updates its own coordinates:
//     a step touches no LDS; otherwise x lives in LDS, the lane-0 owner
//     applies the 1-2 coordinate update, and every lane carries the last
//     update in registers so no second barrier is needed before the next read;
//   * the columns are buffer loads over the variant's whole block, the
//     column in the instruction's SGPR offset and each lane's run of rows in
//     a fixed VGPR offset (runs past the real rows read 0 from no memory);
//   * the slack is FP32 and is recomputed in FP64 as s = b - A y at t = 0 and
//     every resync_interval steps (the reference's resync, chr.cu:386-395).
// Arithmetic is the oracle's, operation for operation (explicit fmaf, IEEE
// division), so GPU and oracle trajectories agree bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <vector>

#include "chr_kernel.h"
#include "chr_kernel_impl.h"
#include "common.h"

namespace naja {

namespace {

// Scale of a pair of width w = b + bn in the scaled-pair form: the power of
// two S = 2^(e+1) >= 2 w for w = m 2^e, m in [0.5, 1). Dividing a pair's rows
// and right-hand sides by S and multiplying its reciprocals by S is exact
// (barring FP32 underflow), so every slack and chord ratio is the unscaled
// one times 1/S or exactly equal, and the slacks, at most w, stay within
// [0, 1/2].
__device__ __forceinline__ double pair_scale(double w)
{
    int e = 0;
    (void)frexp(w, &e);
    return ldexp(1.0, e + 1);
}

// Packs variant v: rows [A; E_v; 0...] with lda = V.lda, b = [b; f_v; +inf...],
// and the elementwise IEEE reciprocal of the packed block.
// PAIRED: a block whose second half negates its first (paired[0] for A,
// paired[1 + v] for E_v) packs only its first half, row r with bn_r = the b
// of row r + k/2; the rows of an unpaired block are packed alone with
// bn = NaN; padding rows get b = bn = +inf and a reciprocal of 0 (their
// ratios are then NaN, which never bound the chord). scaled (every block
// paired): the scaled-pair form (pair_scale), padding reciprocals +inf (the
// clamped padding slack is 1, its ratio +-inf never bounds the chord).
template <bool PAIRED>
__global__ void chr_pack_kernel(const float* A, const double* b, int m0, int d, const float* A_extra,
                                const double* b_extra, const ChrVariant* variants, const int64_t* extra_off,
                                const int* extra_k, const int* paired, float* A_pack, float* Ainv_pack,
                                double* b_pack, double* bn_pack, int scaled)
{
    const int v = blockIdx.y;
    const ChrVariant V = variants[v];
    const int kv = extra_k[v];
    const int64_t eoff = extra_off[v];
    const bool pb = PAIRED && paired[0] != 0, pe = PAIRED && paired[1 + v] != 0;
    const int sb = pb ? m0 / 2 : m0;  // packed rows of the base block
    const int se = pe ? kv / 2 : kv;  // packed rows of the extra block
    // Right-hand sides of packed row r (padding: +inf, +inf).
    auto rhs = [&](int r, double& bb, double& bn) {
        bb = __builtin_huge_val();
        bn = __builtin_huge_val();
        if (r < sb) {
            bb = b[r];
            bn = pb ? b[r + m0 / 2] : __builtin_nan("");
        } else if (r < sb + se) {
            bb = b_extra[eoff + (r - sb)];
            bn = pe ? b_extra[eoff + (r - sb) + kv / 2] : __builtin_nan("");
        }
    };
    const int64_t total = (int64_t)V.lda * d;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(e % V.lda);
        const int c = (int)(e / V.lda);
        float val = 0.0f;
        const bool real = r < sb + se;
        if (r < sb) val = A[(int64_t)c * m0 + r];
        else if (real) val = A_extra[eoff * d + (int64_t)c * kv + (r - sb)];
        float inv = (PAIRED && !real) ? 0.0f : 1.0f / val;
        if (PAIRED && scaled) {
            if (real) {
                double bb, bn;
                rhs(r, bb, bn);
                const double S = pair_scale(bb + bn);
                val *= (float)(1.0 / S);
                inv *= (float)S;
            } else {
                inv = __builtin_huge_valf();
            }
        }
        A_pack[V.a_off + e] = val;
        Ainv_pack[V.a_off + e] = inv;
    }
    if (blockIdx.x == 0) {
        for (int r = threadIdx.x; r < V.lda; r += blockDim.x) {
            double bb, bn;
            rhs(r, bb, bn);
            if (PAIRED && scaled && r < sb + se) {
                const double S = pair_scale(bb + bn);
                bb /= S;
                bn /= S;
            }
            b_pack[V.b_off + r] = bb;
            if (PAIRED) bn_pack[V.b_off + r] = bn;
        }
    }
}

// Row pairing of the blocks: flags[0] for the base block (m0 rows), flags[1 + v]
// for variant v's extra block; the host sets them to 1 and this clears a flag
// when its block has an odd row count or a row r + k/2 that is not the
// negation of row r (compared as values, so +0 and -0 pair; NaN never does).
__global__ void chr_pair_check_kernel(const float* A, int m0, int d, const float* A_extra,
                                      const int64_t* extra_off, const int* extra_k, int* flags)
{
    const int blk = blockIdx.y;
    const int k = blk == 0 ? m0 : extra_k[blk - 1];
    if (k == 0) return;
    const float* M = blk == 0 ? A : A_extra + extra_off[blk - 1] * d;
    if (k % 2 != 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) flags[blk] = 0;
        return;
    }
    const int h = k / 2;
    const int64_t total = (int64_t)h * d;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(e % h);
        const int c = (int)(e / h);
        const float x = M[(int64_t)c * k + r], y = M[(int64_t)c * k + r + h];
        if (!(y == -x)) flags[blk] = 0;
    }
}

// Rounding transform of the packed blocks: variant v samples z with
// y = c_v + P_v z, so its rows become A' = A P_v (FP32, fmaf over k in
// order) and b' = b - A c_v (FP64). Padding rows stay A' = 0, b' = +inf.
// Also writes the IEEE reciprocal of A'. PAIRED: bn' = bn - (-A) c_v by the
// FMA chain over the negated row (what the unpaired layout computes for it),
// and padding rows (r >= V.m) get a reciprocal of 0. scaled: the outputs in
// the scaled-pair form, S from the untransformed pair width b + bn (the
// shift c moves b and bn by opposite amounts), padding reciprocals +inf.
template <bool PAIRED>
__global__ void chr_transform_kernel(const float* __restrict__ Asrc, const double* __restrict__ bsrc,
                                     const double* __restrict__ bnsrc, const ChrVariant* __restrict__ variants, int d,
                                     const float* __restrict__ P, const float* __restrict__ c,
                                     float* __restrict__ Adst, float* __restrict__ Ainv, double* __restrict__ bdst,
                                     double* __restrict__ bndst, int scaled)
{
    const int v = blockIdx.y;
    const ChrVariant V = variants[v];
    const float* A = Asrc + V.a_off;
    const float* Pv = P + (int64_t)v * d * d;
    // Each thread owns one row r and a block of kJB columns, so every load of
    // A[k, r] feeds kJB products; the per-element order (k ascending, fmaf)
    // is unchanged. A wavefront shares its column block (lda is a multiple of
    // 64), so the P reads are uniform: readfirstlane states it, and they
    // become scalar loads.
    constexpr int kJB = 8;
    const int njb = (d + kJB - 1) / kJB;
    const int64_t total = (int64_t)V.lda * njb;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(e % V.lda);
        const int j0 = __builtin_amdgcn_readfirstlane((int)(e / V.lda) * kJB);
        const int nj = min(kJB, d - j0);
        float acc[kJB];
#pragma unroll
        for (int q = 0; q < kJB; ++q) acc[q] = 0.0f;
        if (nj == kJB) {
#pragma unroll 4
            for (int k = 0; k < d; ++k) {
                const float a = A[(int64_t)k * V.lda + r];
                const float* pk = Pv + (int64_t)k * d + j0;
#pragma unroll
                for (int q = 0; q < kJB; ++q) acc[q] = fmaf(a, pk[q], acc[q]);
            }
        } else {
            for (int k = 0; k < d; ++k) {
                const float a = A[(int64_t)k * V.lda + r];
                const float* pk = Pv + (int64_t)k * d + j0;
#pragma unroll
                for (int q = 0; q < kJB; ++q)
                    if (q < nj) acc[q] = fmaf(a, pk[q], acc[q]);
            }
        }
        const bool pad = PAIRED && r >= V.m;
        float sdn = 1.0f, sup = 1.0f;  // 1 / S and S of a scaled pair
        if (PAIRED && scaled && !pad) {
            const double S = pair_scale(bsrc[V.b_off + r] + bnsrc[V.b_off + r]);
            sdn = (float)(1.0 / S);
            sup = (float)S;
        }
#pragma unroll
        for (int q = 0; q < kJB; ++q)
            if (q < nj) {
                const int64_t o = V.a_off + (int64_t)(j0 + q) * V.lda + r;
                Adst[o] = acc[q] * sdn;
                Ainv[o] = pad ? (scaled ? __builtin_huge_valf() : 0.0f) : (1.0f / acc[q]) * sup;
            }
    }
    if (blockIdx.x == 0) {
        const float* cv = c + (int64_t)v * d;
        for (int r = threadIdx.x; r < V.lda; r += blockDim.x) {
            double acc = 0.0, accn = 0.0;
            for (int k = 0; k < d; ++k) {
                const double av = (double)A[(int64_t)k * V.lda + r];
                acc = fma(av, (double)cv[k], acc);
                if (PAIRED) accn = fma(-av, (double)cv[k], accn);
            }
            double bo = bsrc[V.b_off + r] - acc, bno = PAIRED ? bnsrc[V.b_off + r] - accn : 0.0;
            if (PAIRED && scaled && r < V.m) {
                const double S = pair_scale(bsrc[V.b_off + r] + bnsrc[V.b_off + r]);
                bo /= S;
                bno /= S;
            }
            bdst[V.b_off + r] = bo;
            if (PAIRED) bndst[V.b_off + r] = bno;
        }
    }
}

}  // namespace

void chr_transform(const float* A_pack, const double* b_pack, const double* bn_pack, const ChrVariant* variants_dev,
                   int n_variants, int d, int lda, const float* P, const float* c, float* A_out, float* Ainv_out,
                   double* b_out, double* bn_out, hipStream_t stream, bool scaled)
{
    NAJA_REQUIRE(!scaled || bn_pack, NAJA_ERR_INVALID_ARGUMENT, "scaled pairs need the paired-row layout");
    const int64_t per = (int64_t)lda * ((d + 7) / 8);
    int gx = (int)std::min<int64_t>((per + 255) / 256, 4096);
    if (gx < 1) gx = 1;
    if (bn_pack)
        hipLaunchKernelGGL(chr_transform_kernel<true>, dim3(gx, n_variants), dim3(256), 0, stream, A_pack, b_pack,
                           bn_pack, variants_dev, d, P, c, A_out, Ainv_out, b_out, bn_out, scaled ? 1 : 0);
    else
        hipLaunchKernelGGL(chr_transform_kernel<false>, dim3(gx, n_variants), dim3(256), 0, stream, A_pack, b_pack,
                           nullptr, variants_dev, d, P, c, A_out, Ainv_out, b_out, nullptr, 0);
    NAJA_HIP(hipGetLastError());
}

void chr_pair_check(const float* A, int m0, int d, const float* A_extra, const int64_t* extra_off_dev,
                    const int* extra_k_dev, int n_variants, int* flags_dev, hipStream_t stream)
{
    const int64_t per = (int64_t)(m0 / 2 + 1) * d;
    int gx = (int)std::min<int64_t>((per + 255) / 256, 1024);
    if (gx < 1) gx = 1;
    hipLaunchKernelGGL(chr_pair_check_kernel, dim3(gx, n_variants + 1), dim3(256), 0, stream, A, m0, d, A_extra,
                       extra_off_dev, extra_k_dev, flags_dev);
    NAJA_HIP(hipGetLastError());
}

namespace {

const ChrKernelTable& all_kernels()
{
    static const ChrKernelTable tables[] = {chr_table_w1(), chr_table_w4a(), chr_table_w4b(), chr_table_w8(),
                                            chr_table_w16(), chr_table_g4(), chr_table_big(), chr_table_w1p(),
                                            chr_table_w1l(), chr_table_w4p(), chr_table_w8p(), chr_table_w1s(),
                                            chr_table_w4s(), chr_table_w8s()};
    static std::vector<ChrKernelEntry> flat;
    static ChrKernelTable all{nullptr, 0};
    if (!all.e) {
        for (const ChrKernelTable& t : tables) flat.insert(flat.end(), t.e, t.e + t.n);
        all = ChrKernelTable{flat.data(), (int)flat.size()};
    }
    return all;
}

}  // namespace

// The launch functions of shape (W, G, R), merged over every table that
// compiles some of them.
ChrKernelEntry chr_entry(int W, int G, int R)
{
    const ChrKernelTable& t = all_kernels();
    ChrKernelEntry m{W, G, R, nullptr};
    for (int q = 0; q < t.n; ++q) {
        const ChrKernelEntry& e = t.e[q];
        if (e.W != W || e.G != G || e.R != R) continue;
        for (auto f : {&ChrKernelEntry::fn, &ChrKernelEntry::fn_coord, &ChrKernelEntry::fn_x2,
                       &ChrKernelEntry::fn_coord_x2, &ChrKernelEntry::fn_p, &ChrKernelEntry::fn_coord_p,
                       &ChrKernelEntry::fn_x2_p, &ChrKernelEntry::fn_coord_x2_p, &ChrKernelEntry::fn_coord_x2_l4,
                       &ChrKernelEntry::fn_coord_x2_p_l4, &ChrKernelEntry::fn_ps, &ChrKernelEntry::fn_coord_ps,
                       &ChrKernelEntry::fn_x2_ps, &ChrKernelEntry::fn_coord_x2_ps, &ChrKernelEntry::fn_coord_x2_ps_l4})
            if (e.*f) m.*f = e.*f;
    }
    return m;
}

// Smallest compiled R with 64 W R >= rows (with the paired-row layout's
// kernels when paired); -1 if none.
int chr_smallest_r(int W, int G, int rows, bool paired)
{
    const ChrKernelTable& t = all_kernels();
    int best = -1;
    for (int q = 0; q < t.n; ++q) {
        const ChrKernelEntry& e = t.e[q];
        const bool has = paired ? e.fn_p != nullptr : e.fn != nullptr;
        if (has && e.W == W && e.G == G && e.R * 64 * W >= rows && (best < 0 || e.R < best)) best = e.R;
    }
    return best;
}

ChrShape chr_choose_shape(int m_max, int G, int tpb, int n_items, bool paired)
{
    // A step costs (rows per lane x ~15 VALU + ~100 VALU of per-wave chord,
    // alpha and bookkeeping work) x wavefronts per SIMD. One wavefront needs
    // no barrier while it holds <= 8 rows per lane; beyond that four
    // wavefronts (one per SIMD) with up to 32 rows per lane amortise the
    // per-wave work best; eight when the rows would not fit in VGPRs.
    const int rows = (m_max + 63) / 64 * 64;
    int W = rows <= 64 * 8 ? 1 : (rows <= 256 * 32 ? 4 : 8);
    // A launch of few chain groups leaves most SIMDs idle: eight wavefronts
    // per group then shorten each step without costing parallelism (the
    // 4-chain genome-scale run: 1.21 -> 1.10 us per step).
    if (W == 4 && n_items > 0 && n_items * 8 <= 1024 && chr_smallest_r(8, G, rows, paired) > 0) W = 8;
    if (tpb > 0) {
        // The reference's --tpb (threads per chain) is a preference here.
        const int w = std::max(1, tpb / 64);
        const int Wp = w >= 16 ? 16 : (w >= 8 ? 8 : (w >= 4 ? 4 : 1));
        if (chr_smallest_r(Wp, G, rows, paired) > 0) W = Wp;
    }
    int R = chr_smallest_r(W, G, rows, paired);
    if (R < 0) {
        for (int w : {1, 4, 8, 16}) {
            R = chr_smallest_r(w, G, rows, paired);
            if (R > 0) {
                W = w;
                break;
            }
        }
    }
    if (R < 0 && paired) return ChrShape{0, G, 0};
    NAJA_REQUIRE(R > 0, NAJA_ERR_UNSUPPORTED, "no compiled CHR kernel holds %d rows with %d chains per group", m_max, G);
    return ChrShape{W, G, R};
}

namespace {

// Launch timing (off unless enabled): a pair of HIP events on the launch's
// stream around every CHR step-kernel launch.
struct TimedLaunch {
    hipEvent_t start, stop;
    int items;
};
std::mutex g_timing_mu;
bool g_timing = false;
std::vector<TimedLaunch> g_timed;

}  // namespace

void chr_timing_enable(bool on)
{
    std::lock_guard<std::mutex> g(g_timing_mu);
    g_timing = on;
}

int chr_timing_collect(int* items, float* ms, int max)
{
    std::vector<TimedLaunch> recs;
    {
        std::lock_guard<std::mutex> g(g_timing_mu);
        recs.swap(g_timed);
    }
    int n = 0;
    for (TimedLaunch& r : recs) {
        if (n < max) {
            NAJA_HIP(hipEventSynchronize(r.stop));
            float t = 0.0f;
            NAJA_HIP(hipEventElapsedTime(&t, r.start, r.stop));
            if (items) items[n] = r.items;
            if (ms) ms[n] = t;
            ++n;
        }
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    return n;
}

size_t chr_lds_columns_bytes(const ChrShape& sh, int d)
{
    // Columns at the packed block's stride, 64 W R rows (every slot of every
    // lane), so a column's byte offset is the same in LDS and in the block.
    return sizeof(float) * (size_t)d * (size_t)(64 * sh.W * sh.R);
}

bool chr_has_lds_kernel(const ChrShape& sh, bool paired)
{
    const ChrKernelEntry e = chr_entry(sh.W, sh.G, sh.R);
    return (paired ? e.fn_coord_x2_p_l4 : e.fn_coord_x2_l4) != nullptr;
}

namespace {

// The kernel of a launch: coordinate-only twins when the moves allow,
// VGPR-state twins when d does, the LDS-column kernel for nc > 1; with
// scaled, the scaled-pair twin of that choice (null if not compiled).
ChrKernelFn pick_kernel(const ChrKernelEntry& e, bool paired, bool coord_only, bool reg_state, int nc, bool scaled)
{
    if (nc > 1) return paired ? (scaled ? e.fn_coord_x2_ps_l4 : e.fn_coord_x2_p_l4) : e.fn_coord_x2_l4;
    if (paired) {
        const ChrKernelFn f = scaled ? e.fn_ps : e.fn_p, fc = scaled ? e.fn_coord_ps : e.fn_coord_p;
        const ChrKernelFn x = scaled ? e.fn_x2_ps : e.fn_x2_p, xc = scaled ? e.fn_coord_x2_ps : e.fn_coord_x2_p;
        // The scaled twins exist for the same choices as the unscaled ones
        // or not at all: a choice falls back only along the unscaled path.
        ChrKernelFn fn = coord_only && e.fn_coord_p ? fc : f;
        const ChrKernelFn fx = coord_only && e.fn_coord_x2_p ? xc : x;
        if (reg_state && (coord_only && e.fn_coord_x2_p ? e.fn_coord_x2_p : e.fn_x2_p)) fn = fx;
        return fn;
    }
    ChrKernelFn fn = coord_only && e.fn_coord ? e.fn_coord : e.fn;
    const ChrKernelFn fx = coord_only && e.fn_coord_x2 ? e.fn_coord_x2 : e.fn_x2;
    if (reg_state && fx) fn = fx;
    return fn;
}

}  // namespace

bool chr_has_scaled_kernel(const ChrShape& sh, bool coord_only, bool reg_state, int nc)
{
    return pick_kernel(chr_entry(sh.W, sh.G, sh.R), true, coord_only, reg_state, nc, true) != nullptr;
}

void chr_launch(const ChrArgs& args, int n_items, const ChrShape& sh, hipStream_t stream, int nc, size_t lds_bytes)
{
    const ChrKernelEntry e = chr_entry(sh.W, sh.G, sh.R);
    const bool coord_only = args.thr_k == 0 && !args.has_pair;
    const bool reg_state = args.d <= kChrRegStateMaxD;
    const bool paired = args.bn != nullptr;
    NAJA_REQUIRE(!args.scaled || paired, NAJA_ERR_INVALID_ARGUMENT, "scaled pairs need the paired-row layout");
    ChrKernelFn fn = nullptr;
    if (paired) {
        fn = pick_kernel(e, true, coord_only, reg_state, 1, args.scaled != 0);
    } else {
        fn = coord_only && e.fn_coord ? e.fn_coord : e.fn;
        const ChrKernelFn fx = coord_only && e.fn_coord_x2 ? e.fn_coord_x2 : e.fn_x2;
        if (reg_state && fx) fn = fx;
    }
    if (nc > 1) {
        NAJA_REQUIRE(nc == kChrLdsChains && coord_only && reg_state && n_items % nc == 0, NAJA_ERR_INVALID_ARGUMENT,
                     "an LDS-column launch needs coordinate moves, d <= %d and items in groups of %d",
                     kChrRegStateMaxD, kChrLdsChains);
        fn = pick_kernel(e, paired, coord_only, reg_state, nc, args.scaled != 0);
    }
    NAJA_REQUIRE(fn != nullptr, NAJA_ERR_UNSUPPORTED, "no compiled CHR kernel for W=%d G=%d R=%d%s%s", sh.W, sh.G, sh.R,
                 paired ? " (paired rows)" : "", nc > 1 ? " (LDS columns)" : "");
    NAJA_REQUIRE(args.nsteps >= 0 && args.nsteps < ((int64_t)1 << 31) && args.nlog < ((int64_t)1 << 31),
                 NAJA_ERR_UNSUPPORTED, "%lld steps in one launch (at most 2^31 - 1; the streamed API splits longer runs)",
                 (long long)args.nsteps);
    // The register kernels address a variant's block with 32-bit byte
    // offsets and mark padding runs with bit 31.
    NAJA_REQUIRE(sh.big() || 4.0 * sh.rows() * (double)args.d < 2147483648.0, NAJA_ERR_UNSUPPORTED,
                 "a packed block of %d rows x %d columns exceeds 2 GiB", sh.rows(), args.d);
    const size_t lds = nc > 1 ? lds_bytes : sh.lds_bytes(args.d);
    NAJA_REQUIRE(nc > 1 || lds <= 96u * 1024u, NAJA_ERR_UNSUPPORTED,
                 "d=%d with %d chains per group exceeds the LDS state budget", args.d, sh.G);
    NAJA_REQUIRE(lds <= 156u * 1024u, NAJA_ERR_UNSUPPORTED, "%zu bytes of LDS columns exceed the budget", lds);
    bool timed;
    {
        std::lock_guard<std::mutex> g(g_timing_mu);
        timed = g_timing;
    }
    if (timed) {  // a captured launch is timed by the graph's own event nodes, if at all
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        NAJA_HIP(hipStreamIsCapturing(stream, &cap));
        timed = cap == hipStreamCaptureStatusNone;
    }
    TimedLaunch rec{nullptr, nullptr, n_items};
    if (timed) {
        NAJA_HIP(hipEventCreate(&rec.start));
        NAJA_HIP(hipEventCreate(&rec.stop));
        NAJA_HIP(hipEventRecord(rec.start, stream));
    }
    hipLaunchKernelGGL(fn, dim3(n_items / nc), dim3(64 * sh.W * nc), lds, stream, args);
    NAJA_HIP(hipGetLastError());
    if (timed) {
        NAJA_HIP(hipEventRecord(rec.stop, stream));
        std::lock_guard<std::mutex> g(g_timing_mu);
        g_timed.push_back(rec);
    }
}

void chr_pack(const float* A, const double* b, int m0, int d, const float* A_extra, const double* b_extra,
              const ChrVariant* variants_dev, const int64_t* extra_off_dev, const int* extra_k_dev, int n_variants,
              const int* paired_dev, float* A_pack, float* Ainv_pack, double* b_pack, double* bn_pack, int lda_max,
              hipStream_t stream, bool scaled)
{
    NAJA_REQUIRE(!scaled || bn_pack, NAJA_ERR_INVALID_ARGUMENT, "scaled pairs need the paired-row layout");
    const int64_t per = (int64_t)lda_max * d;
    int gx = (int)std::min<int64_t>((per + 255) / 256, 4096);
    if (gx < 1) gx = 1;
    if (bn_pack)
        hipLaunchKernelGGL(chr_pack_kernel<true>, dim3(gx, n_variants), dim3(256), 0, stream, A, b, m0, d, A_extra,
                           b_extra, variants_dev, extra_off_dev, extra_k_dev, paired_dev, A_pack, Ainv_pack, b_pack,
                           bn_pack, scaled ? 1 : 0);
    else
        hipLaunchKernelGGL(chr_pack_kernel<false>, dim3(gx, n_variants), dim3(256), 0, stream, A, b, m0, d, A_extra,
                           b_extra, variants_dev, extra_off_dev, extra_k_dev, nullptr, A_pack, Ainv_pack, b_pack,
                           nullptr, 0);
    NAJA_HIP(hipGetLastError());
}

}  // namespace naja
