// Facet-aligned rounding on the GPU: for each conditioned variant
// [A0; E_v] y <= [b0; f_v], a rounding y = c + P z whose coordinate axes run
// along the edges of a parallelotope built from the variant's binding
// facets.
//
// Coordinate hit-and-run resamples one coordinate of z per step. On an
// axis-aligned box that is an exact draw of that coordinate; on the same box
// rotated it is not, however well the box is whitened. A rounding that
// aligns the axes with the facets that bound the body therefore mixes far
// faster than one that only matches its second moments (the Dikin ellipsoid
// P0, P0 P0^T = H^-1, whose orientation is that of a Cholesky factor).
//
// The dual rows u_r = a_r / s_r (s_r the slack at the centre x_c) of the
// candidate facets are chosen greedily by modified Gram-Schmidt: at every
// step the candidate with the largest residual norm, first among the
// variant's extra rows (and, on request, its base rows) while that candidate
// keeps at least theta_v of its own norm. The directions the facets leave
// free are completed in one of three ways:
//   * from the dual rows of a base rounding P0 (the rows of P0^-1), chosen by
//     the same greedy residual rule;
//   * from the rows of L^-1 for the lower Cholesky factor L of P0 P0^T (the
//     same ellipsoid with its axes in triangular order);
//   * from the base polytope's own facets, orthonormalised: the base dual
//     rows' residuals against everything chosen so far, the largest first,
//     each normalised and kept as it is (FP64 classical Gram-Schmidt, twice).
//     These rows lie in the free subspace V (orthogonal to the chosen
//     facets), so a move along a chosen facet's axis stays in V's complement,
//     a free move stays in V, and the free axes are the base facets' normals
//     made orthonormal in the order of their residuals: on a slice of a box,
//     as near to the box's own axes as an orthonormal basis of V gets.
// With N the d chosen rows, P = N^-1: every chosen facet is a coordinate slab
// |z_k| <= ~1. A diagonal rescaling of P would not change the chain in y,
// and neither would a permutation (up to which coordinate a draw picks), so
// only the choice of rows matters. theta_v > 1 takes no facet rows.
//
// Exchange (when the chosen facets fill the basis): every facet candidate is
// expressed in the basis, C = U N^-1, and while some |C_ik| exceeds the
// tolerance the basis row k is replaced by candidate i (one rank-1 update of
// C per exchange). Each exchange multiplies |det N| by |C_ik|, so the
// parallelotope {|N y| <= 1} shrinks towards the smallest one spanned by the
// variant's facets (the maximum-volume submatrix rule of Goreinov,
// Oseledets, Savostyanov, Tyrtyshnikov and Zamarashkin, "How to find a good
// submatrix", 2010), and the facets left out cut it less deeply.
//
// Per variant, one workgroup of 512 threads: P0^-1 and N^-1 by Gauss-Jordan
// with partial pivoting in LDS (FP64); the selection on FP32 residuals, one
// candidate per thread held in registers (so a Gram-Schmidt pass is a dot
// product and an update against the current direction broadcast from LDS);
// the chosen rows are re-formed in FP64 for the inverse; the exchange on
// FP32 coefficient rows, one per thread.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/naja_amd.h"
#include "common.h"
#include "facet.h"

namespace naja {
namespace {

constexpr int kFacetThreads = 512;  // one candidate row per thread
constexpr int kFacetMaxD = 128;
static_assert(kFacetMaxD <= 128 && kFacetThreads % 128 == 0, "the elimination maps 128 columns x row groups");

// Block-wide argmax of (value, index), larger value first, then smaller index.
__device__ void block_argmax(float& val, int& idx, float* rv, int* ri)
{
    for (int off = 32; off > 0; off >>= 1) {
        const float ov = __shfl_xor(val, off);
        const int oi = __shfl_xor(idx, off);
        if (ov > val || (ov == val && oi < idx)) {
            val = ov;
            idx = oi;
        }
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        rv[threadIdx.x >> 6] = val;
        ri[threadIdx.x >> 6] = idx;
    }
    __syncthreads();
    val = rv[0];
    idx = ri[0];
    for (int w = 1; w < kFacetThreads / 64; ++w)
        if (rv[w] > val || (rv[w] == val && ri[w] < idx)) {
            val = rv[w];
            idx = ri[w];
        }
}

// In-place inverse of the d x d row-major FP64 matrix M (LDS) by Gauss-Jordan
// with partial pivoting (d <= 128). col: d doubles of LDS; piv: d ints of
// LDS. Returns false (uniformly) when a pivot is zero or not finite. The
// pivot search runs on the first wavefront (two rows per lane).
__device__ bool gj_inverse(double* M, int d, double* col, int* piv, int* flag)
{
    const int tid = threadIdx.x;
    for (int k = 0; k < d; ++k) {
        if (tid < 64) {
            double best = -1.0;
            int p = 0x7fffffff;
            for (int i = k + tid; i < d; i += 64) {
                const double a = fabs(M[i * d + k]);
                if (a > best) {
                    best = a;
                    p = i;
                }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double ob = __shfl_xor(best, off);
                const int op = __shfl_xor(p, off);
                if (ob > best || (ob == best && op < p)) {
                    best = ob;
                    p = op;
                }
            }
            if (tid == 0) {
                piv[k] = p;
                *flag = (best > 0.0 && isfinite(best)) ? 0 : 1;
            }
        }
        __syncthreads();
        if (*flag) return false;
        const int p = piv[k];
        const double inv = 1.0 / M[p * d + k];
        // Column k after the interchange of rows k and p.
        for (int i = tid; i < d; i += kFacetThreads) col[i] = M[(i == k ? p : (i == p ? k : i)) * d + k];
        __syncthreads();
        // Interchange rows k and p, scaling the new row k.
        for (int j = tid; j < d; j += kFacetThreads) {
            const double rk = M[k * d + j], rp = M[p * d + j];
            M[k * d + j] = (j == k) ? inv : rp * inv;
            if (p != k) M[p * d + j] = rk;
        }
        __syncthreads();
        // Thread (ig, j) = (tid / 128, tid % 128) updates column j of rows
        // ig, ig + 4, ...: no integer division per element (d <= 128). Rows
        // go in batches of kGjBatch whose LDS reads are all issued before any
        // write (the rows are independent), so the reads' latency is paid
        // once per batch rather than once per row.
        {
            constexpr int kGjBatch = 8;
            constexpr int kStride = kFacetThreads / 128;
            const int j = tid & 127;
            if (j < d) {
                const double mkj = M[k * d + j];
                for (int i0 = tid >> 7; i0 < d; i0 += kGjBatch * kStride) {
                    double f[kGjBatch], m[kGjBatch];
#pragma unroll
                    for (int b = 0; b < kGjBatch; ++b) {
                        const int i = i0 + b * kStride;
                        f[b] = i < d ? col[i] : 0.0;
                        m[b] = i < d ? M[i * d + j] : 0.0;
                    }
#pragma unroll
                    for (int b = 0; b < kGjBatch; ++b) {
                        const int i = i0 + b * kStride;
                        if (i < d && i != k) M[i * d + j] = (j == k) ? -f[b] * inv : fma(-f[b], mkj, m[b]);
                    }
                }
            }
        }
        __syncthreads();
    }
    // Undo the row interchanges as column interchanges, last first.
    for (int k = d - 1; k >= 0; --k) {
        const int p = piv[k];
        if (p != k)
            for (int i = tid; i < d; i += kFacetThreads) {
                const double t = M[i * d + k];
                M[i * d + k] = M[i * d + p];
                M[i * d + p] = t;
            }
        __syncthreads();
    }
    return true;
}

// One workgroup per variant, one candidate per thread: thread i holds the
// FP32 residual of candidate i in registers (COLS >= d columns, the rest 0).
// P0i: per-variant global scratch for the FP64 rows of P0^-1 (completions 0
// and 1).
template <int COLS>
__global__ __launch_bounds__(kFacetThreads) void facet_kernel(
    const float* __restrict__ A0, const double* __restrict__ b0, int m0, int d, const float* __restrict__ Ae,
    const double* __restrict__ be, const int64_t* __restrict__ eoff, const int* __restrict__ ek,
    const float* __restrict__ xc, const float* __restrict__ P0, const float* __restrict__ theta, int include_base,
    int completion, float exchange_tol, double* __restrict__ P0iall, float* __restrict__ P, int* __restrict__ status)
{
    extern __shared__ double sh[];
    double* M = sh;                                 // d x d: P0, then N (or Q, then N)
    double* col = M + d * d;                        // d
    double* sl = col + d;                           // kFacetThreads: slack of each candidate with a row
    float* q = reinterpret_cast<float*>(sl + kFacetThreads);  // COLS: the current direction
    int* sel = reinterpret_cast<int*>(q + COLS);    // d: chosen candidates in order
    int* piv = sel + d;                             // d
    __shared__ float rv[kFacetThreads / 64];
    __shared__ int ri[kFacetThreads / 64];
    __shared__ int flag;
    __shared__ double xs[kFacetMaxD];
    __shared__ double dots[kFacetMaxD];
    __shared__ double upd[kFacetThreads / 128][kFacetMaxD];

    const int v = blockIdx.x, tid = threadIdx.x;
    const int k = ek[v];
    const int64_t off = eoff[v];
    const bool base_facets = completion == NAJA_FACET_COMPLETE_BASE_FACETS;
    const int nf = (include_base ? m0 : 0) + k;  // facet candidates
    const int nr = nf + (base_facets ? m0 : d);   // + the completion candidates
    double* P0i = P0iall + (int64_t)v * d * d;
    const float th = theta[v];
    const float th2 = th * th;
    auto arow = [&](int r, int c) -> double {
        return r < m0 ? (double)A0[(int64_t)c * m0 + r] : (double)Ae[off * d + (int64_t)c * k + (r - m0)];
    };
    auto facet_row = [&](int i) { return include_base ? i : m0 + i; };
    // Polytope row of candidate i (facets, then base rows as completions), or -1 for a P0^-1 row.
    auto poly_row = [&](int i) { return i < nf ? facet_row(i) : (base_facets ? i - nf : -1); };

    for (int c = tid; c < d; c += kFacetThreads) xs[c] = xc ? (double)xc[(int64_t)v * d + c] : 0.0;
    if (tid == 0) flag = 0;
    __syncthreads();
    if (tid < nr && poly_row(tid) >= 0) {
        const int r = poly_row(tid);
        double ax = 0.0;
        for (int c = 0; c < d; ++c) ax = fma(arow(r, c), xs[c], ax);
        const double s = (r < m0 ? b0[r] : be[off + (r - m0)]) - ax;
        if (!(s > 0.0)) flag = 1;
        sl[tid] = s;
    }
    if (!base_facets) {
        const float* P0v = P0 + (int64_t)v * d * d;
        if (completion == NAJA_FACET_COMPLETE_LOWER) {
            // The same ellipsoid with its axes in triangular order: C = P0 P0^T in
            // FP64 (row-major; P0 read through the L1/L2 caches), then its lower
            // Cholesky factor in place.
            for (int e = tid; e < d * d; e += kFacetThreads) {
                const int i = e / d, j = e % d;
                double c = 0.0;
                if (j <= i)
                    for (int t = 0; t < d; ++t)
                        c = fma((double)P0v[(int64_t)i * d + t], (double)P0v[(int64_t)j * d + t], c);
                M[e] = c;
            }
            __syncthreads();
            for (int j = 0; j < d; ++j) {
                if (tid == 0) {
                    double s = M[j * d + j];
                    for (int t = 0; t < j; ++t) s -= M[j * d + t] * M[j * d + t];
                    if (!(s > 0.0)) flag = 2;
                    M[j * d + j] = sqrt(s > 0.0 ? s : 1.0);
                }
                __syncthreads();
                const double ljj = M[j * d + j];
                for (int i = j + 1 + tid; i < d; i += kFacetThreads) {
                    double s = M[i * d + j];
                    for (int t = 0; t < j; ++t) s -= M[i * d + t] * M[j * d + t];
                    M[i * d + j] = s / ljj;
                }
                __syncthreads();
            }
            for (int e = tid; e < d * d; e += kFacetThreads)
                if (e % d > e / d) M[e] = 0.0;
        } else {
            for (int e = tid; e < d * d; e += kFacetThreads) M[e] = (double)P0v[e];
        }
    }
    __syncthreads();
    if (flag) {
        if (tid == 0) status[v] = flag == 1 ? NAJA_ERR_INVALID_ARGUMENT : NAJA_ERR_RUNTIME;
        return;
    }
    if (!base_facets) {
        if (!gj_inverse(M, d, col, piv, &flag)) {  // M = P0^-1: its rows complete the basis
            if (tid == 0) status[v] = NAJA_ERR_RUNTIME;
            return;
        }
        for (int e = tid; e < d * d; e += kFacetThreads) P0i[e] = M[e];
    }
    // FP64 dual row of candidate i (after P0^-1 is in P0i).
    auto cand64 = [&](int i, int c) -> double {
        const int r = poly_row(i);
        return r >= 0 ? arow(r, c) / sl[i] : P0i[(int64_t)(i - nf) * d + c];
    };

    // This thread's candidate.
    float r[COLS];
    float n0 = 0.0f;
    const bool live = tid < nr;
    const bool is_facet = tid < nf;
    if (live && poly_row(tid) >= 0) {
        const int rr = poly_row(tid);
        const double w = 1.0 / sl[tid];
#pragma unroll
        for (int c = 0; c < COLS; ++c) r[c] = c < d ? (float)(arow(rr, c) * w) : 0.0f;
    } else if (live) {
#pragma unroll
        for (int c = 0; c < COLS; ++c) r[c] = c < d ? (float)M[(tid - nf) * d + c] : 0.0f;
    } else {
#pragma unroll
        for (int c = 0; c < COLS; ++c) r[c] = 0.0f;
    }
#pragma unroll
    for (int c = 0; c < COLS; ++c) n0 = fmaf(r[c], r[c], n0);
    bool taken = !live;
    float n2 = n0;

    // Greedy modified Gram-Schmidt over the candidates.
    for (int t = 0; t < d; ++t) {
        if (t > 0 && !taken) {
            float cq = 0.0f;
#pragma unroll
            for (int c = 0; c < COLS; ++c) cq = fmaf(r[c], q[c], cq);
            n2 = 0.0f;
#pragma unroll
            for (int c = 0; c < COLS; ++c) {
                r[c] = fmaf(-cq, q[c], r[c]);
                n2 = fmaf(r[c], r[c], n2);
            }
        }
        float vf = -1.0f, vb = -1.0f;
        int i_f = 0x7fffffff, i_b = 0x7fffffff;
        if (!taken) {
            if (is_facet) {
                if (n2 >= th2 * n0 && n2 > 1e-12f * n0) {
                    vf = n2;
                    i_f = tid;
                }
            } else if (!base_facets || n2 > 1e-12f * n0) {
                vb = n2;
                i_b = tid;
            }
        }
        block_argmax(vf, i_f, rv, ri);
        block_argmax(vb, i_b, rv, ri);
        const bool from_f = vf > 0.0f;
        const int p = from_f ? i_f : i_b;
        const float best = from_f ? vf : vb;
        if (p == 0x7fffffff || !(best > 0.0f)) {
            if (tid == 0) status[v] = NAJA_ERR_RUNTIME;
            return;
        }
        __syncthreads();  // every thread has read q for this pass
        if (tid == p) {
            const float inv = 1.0f / sqrtf(best);
#pragma unroll
            for (int c = 0; c < COLS; ++c) q[c] = r[c] * inv;
            taken = true;
            sel[t] = p;
        }
        __syncthreads();
    }

    // Exchange, when the chosen facets fill the basis: C = U N^-1 for the facet
    // candidates (FP32 rows, one per thread, summed in column order), then
    // rank-1 updates while the largest |C_ik| exceeds the tolerance. N^-1
    // (FP64, LDS M) follows each exchange by the same rank-1 form
    // (Sherman-Morrison: row k of N becomes u_i, g = u_i N^-1, N^-1 -=
    // N^-1 e_k (g - e_k) / g_k), so it is P at the end.
    int nfac = 0;
    for (int t = 0; t < d; ++t) nfac += sel[t] < nf;
    bool have_inverse = false;
    if (exchange_tol > 0.0f) {
        if (nfac == d) {
            for (int e = tid; e < d * d; e += kFacetThreads) M[e] = cand64(sel[e / d], e % d);
            __syncthreads();
            if (!gj_inverse(M, d, col, piv, &flag)) {
                if (tid == 0) status[v] = NAJA_ERR_RUNTIME;
                return;
            }
#pragma unroll
            for (int c = 0; c < COLS; ++c) r[c] = 0.0f;
            if (is_facet) {
                // The candidate's dual row comes from global memory in batches
                // of eight entries, all loads issued before the first is used.
                constexpr int kUB = 8;
                const int rr = poly_row(tid);
                const double st = sl[tid];
                for (int j0 = 0; j0 < d; j0 += kUB) {
                    double a8[kUB];
#pragma unroll
                    for (int b = 0; b < kUB; ++b) a8[b] = j0 + b < d ? arow(rr, j0 + b) : 0.0;
#pragma unroll
                    for (int b = 0; b < kUB; ++b) {
                        const int j = j0 + b;
                        if (j < d) {
                            const float uj = (float)(a8[b] / st);
#pragma unroll
                            for (int c = 0; c < COLS; ++c)
                                if (c < d) r[c] = r[c] + uj * (float)M[j * d + c];
                        }
                    }
                }
            }
            for (int it = 0; it < 4 * d; ++it) {
                float best = -1.0f;
                int bi = 0x7fffffff;
                if (is_facet)
#pragma unroll
                    for (int c = 0; c < COLS; ++c) {
                        const float a = fabsf(r[c]);
                        if (c < d && a > best) {
                            best = a;
                            bi = tid * kFacetMaxD + c;
                        }
                    }
                block_argmax(best, bi, rv, ri);
                if (!(best > exchange_tol)) break;
                const int pi = bi / kFacetMaxD, pk = bi % kFacetMaxD;
                __syncthreads();  // rv / ri read; q free
                if (tid == pi)
#pragma unroll
                    for (int c = 0; c < COLS; ++c) q[c] = r[c];
                __syncthreads();
                if (is_facet) {
                    float cjk = 0.0f;
#pragma unroll
                    for (int c = 0; c < COLS; ++c) cjk = c == pk ? r[c] : cjk;
                    const float f = cjk / q[pk];
#pragma unroll
                    for (int c = 0; c < COLS; ++c) r[c] = r[c] - f * (q[c] - (c == pk ? 1.0f : 0.0f));
                }
                if (tid == 0) sel[pk] = pi;
                // N^-1 after row pk of N becomes u_pi: col = u_pi (FP64),
                // xs = column pk of N^-1, dots = g = u_pi N^-1 (four threads
                // per column, as in the completion below).
                for (int c = tid; c < d; c += kFacetThreads) {
                    col[c] = cand64(pi, c);
                    xs[c] = M[c * d + pk];
                }
                __syncthreads();
                {
                    const int c = tid & 127, g = tid >> 7;
                    if (c < d) {
                        double w0 = 0.0, w1 = 0.0;
                        int j = g;
                        for (; j + 4 < d; j += 8) {
                            w0 = fma(col[j], M[j * d + c], w0);
                            w1 = fma(col[j + 4], M[(j + 4) * d + c], w1);
                        }
                        if (j < d) w0 = fma(col[j], M[j * d + c], w0);
                        upd[g][c] = w0 + w1;
                    }
                }
                __syncthreads();
                for (int c = tid; c < d; c += kFacetThreads)
                    dots[c] = (upd[0][c] + upd[1][c]) + (upd[2][c] + upd[3][c]);
                __syncthreads();
                {
                    const double gk = dots[pk];
                    for (int e = tid; e < d * d; e += kFacetThreads) {
                        const int i = e / d, c = e - i * d;
                        M[e] -= xs[i] * ((dots[c] - (c == pk ? 1.0 : 0.0)) / gk);
                    }
                }
                __syncthreads();
            }
            have_inverse = true;
        }
    }

    if (base_facets && nfac < d) {
        // Q (LDS M): orthonormal rows of everything chosen, in order, by
        // classical Gram-Schmidt twice in FP64; a completion row of N is its
        // Q row, a facet row of N its dual row.
        for (int t = 0; t < d; ++t) {
            const int i = sel[t];
            for (int c = tid; c < d; c += kFacetThreads) col[c] = cand64(i, c);
            __syncthreads();
            for (int pass = 0; pass < 2; ++pass) {
                // dots[j] = Q_j . col: four threads per row j (t < 128), each
                // over the columns c = part (mod 4) with two accumulators, the
                // four partial sums added over the quad.
                if (tid < 4 * t) {
                    const int j = tid >> 2, part = tid & 3;
                    double s0 = 0.0, s1 = 0.0;
                    int c = part;
                    for (; c + 4 < d; c += 8) {
                        s0 = fma(M[j * d + c], col[c], s0);
                        s1 = fma(M[j * d + c + 4], col[c + 4], s1);
                    }
                    if (c < d) s0 = fma(M[j * d + c], col[c], s0);
                    double s = s0 + s1;
                    s += __shfl_xor(s, 1);
                    s += __shfl_xor(s, 2);
                    if (part == 0) dots[j] = s;
                }
                __syncthreads();
                // col -= sum_j dots[j] Q_j: thread (g, c) = (tid / 128, tid %
                // 128) sums the rows j = g (mod 4), then column c adds the four.
                {
                    const int c = tid & 127, g = tid >> 7;
                    if (c < d) {
                        double w0 = 0.0, w1 = 0.0;
                        int j = g;
                        for (; j + 4 < t; j += 8) {
                            w0 = fma(dots[j], M[j * d + c], w0);
                            w1 = fma(dots[j + 4], M[(j + 4) * d + c], w1);
                        }
                        if (j < t) w0 = fma(dots[j], M[j * d + c], w0);
                        upd[g][c] = w0 + w1;
                    }
                }
                __syncthreads();
                for (int c = tid; c < d; c += kFacetThreads)
                    col[c] = col[c] - ((upd[0][c] + upd[1][c]) + (upd[2][c] + upd[3][c]));
                __syncthreads();
            }
            if (tid < 64) {
                double s = 0.0;
                for (int c = tid; c < d; c += 64) s = fma(col[c], col[c], s);
                for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
                if (tid == 0) dots[kFacetMaxD - 1] = s;
            }
            __syncthreads();
            const double nrm = sqrt(dots[kFacetMaxD - 1]);
            if (!(nrm > 0.0)) {
                if (tid == 0) status[v] = NAJA_ERR_RUNTIME;
                return;
            }
            for (int c = tid; c < d; c += kFacetThreads) M[t * d + c] = col[c] / nrm;
            __syncthreads();
        }
        for (int e = tid; e < d * d; e += kFacetThreads)
            if (sel[e / d] < nf) M[e] = cand64(sel[e / d], e % d);
    } else if (!have_inverse) {
        // N (the chosen dual rows, FP64).
        for (int e = tid; e < d * d; e += kFacetThreads) M[e] = cand64(sel[e / d], e % d);
    }
    __syncthreads();
    if (!have_inverse && !gj_inverse(M, d, col, piv, &flag)) {
        if (tid == 0) status[v] = NAJA_ERR_RUNTIME;
        return;
    }
    float* Pv = P + (int64_t)v * d * d;
    for (int e = tid; e < d * d; e += kFacetThreads) Pv[e] = (float)M[e];
    if (tid == 0) status[v] = NAJA_OK;
}

size_t facet_lds_bytes(int d, int cols)
{
    return sizeof(double) * ((size_t)d * d + d + kFacetThreads) + sizeof(float) * (size_t)cols +
           sizeof(int) * 2 * (size_t)d;
}

}  // namespace

int facet_columns(int d) { return d <= 32 ? 32 : (d <= 64 ? 64 : (d <= 112 ? 112 : 128)); }

void facet_check_shape(int m0, int d, int kmax, int include_base, int completion)
{
    const bool base_facets = completion == NAJA_FACET_COMPLETE_BASE_FACETS;
    const int nr_max = (include_base ? m0 : 0) + kmax + (base_facets ? m0 : d);
    const size_t lds = facet_lds_bytes(d, facet_columns(d));
    NAJA_REQUIRE(d <= kFacetMaxD && nr_max <= kFacetThreads && lds <= 150 * 1024, NAJA_ERR_UNSUPPORTED,
                 "facet rounding: d=%d with %d candidate rows exceeds the kernel (d <= %d, %d candidates)", d, nr_max,
                 kFacetMaxD, kFacetThreads);
}

void facet_launch(const float* A0, const double* b0, int m0, int d, int n_variants, const float* Ae, const double* be,
                  const int64_t* eoff, const int* ek, const float* xc, const float* P0, const float* theta,
                  int include_base, int completion, float exchange_tol, double* P0i, float* P, int* status,
                  hipStream_t s)
{
    const int cols = facet_columns(d);
    auto kern = cols == 32    ? facet_kernel<32>
                : cols == 64  ? facet_kernel<64>
                : cols == 112 ? facet_kernel<112>
                              : facet_kernel<128>;  // 113..128 columns spill to scratch
    hipLaunchKernelGGL(kern, dim3(n_variants), dim3(kFacetThreads), facet_lds_bytes(d, cols), s, A0, b0, m0, d, Ae,
                       be, eoff, ek, xc, P0, theta, include_base ? 1 : 0, completion, exchange_tol, P0i, P, status);
    NAJA_HIP(hipGetLastError());
}

}  // namespace naja

extern "C" int naja_precond_facet_device(const float* A0, const double* b0, int m0, int d, int n_variants,
                                         const int* extra_rows, const float* extra_A, const double* extra_b,
                                         const float* x_c, const float* P0, const float* theta, int include_base,
                                         int completion, float exchange_tol, float* P, void* stream)
{
    using namespace naja;
    return guarded([&]() {
        const bool base_facets = completion == NAJA_FACET_COMPLETE_BASE_FACETS;
        NAJA_REQUIRE(A0 && b0 && theta && P && m0 > 0 && d > 0 && n_variants > 0 && extra_rows &&
                         (P0 || base_facets),
                     NAJA_ERR_INVALID_ARGUMENT, "naja_precond_facet_device: bad arguments");
        NAJA_REQUIRE(completion == NAJA_FACET_COMPLETE_ROUNDING || completion == NAJA_FACET_COMPLETE_LOWER ||
                         base_facets,
                     NAJA_ERR_INVALID_ARGUMENT, "naja_precond_facet_device: unknown completion %d", completion);
        NAJA_REQUIRE(!(base_facets && include_base), NAJA_ERR_INVALID_ARGUMENT,
                     "naja_precond_facet_device: base rows cannot be both facet candidates and the completion");
        NAJA_REQUIRE(exchange_tol == 0.0f || exchange_tol >= 1.0f, NAJA_ERR_INVALID_ARGUMENT,
                     "naja_precond_facet_device: exchange tolerance %g (0 = off, else >= 1)", (double)exchange_tol);
        std::vector<int64_t> off(n_variants);
        int64_t tot = 0;
        int kmax = 0;
        for (int v = 0; v < n_variants; ++v) {
            NAJA_REQUIRE(extra_rows[v] >= 0, NAJA_ERR_INVALID_ARGUMENT, "variant %d has %d extra rows", v,
                         extra_rows[v]);
            off[v] = tot;
            tot += extra_rows[v];
            kmax = std::max(kmax, extra_rows[v]);
        }
        NAJA_REQUIRE(tot == 0 || (extra_A && extra_b), NAJA_ERR_INVALID_ARGUMENT,
                     "naja_precond_facet_device: extra rows without extra_A/extra_b");
        facet_check_shape(m0, d, kmax, include_base, completion);
        hipStream_t s = (hipStream_t)stream;
        const size_t b_off = (sizeof(int64_t) * n_variants + 255) / 256 * 256;
        const size_t b_k = (sizeof(int) * n_variants + 255) / 256 * 256;
        const size_t b_P0i = base_facets ? 0 : sizeof(double) * (size_t)n_variants * d * d;
        HeldBuf& hb = held_buf();
        std::lock_guard<std::mutex> hold(hb.mu);
        char* mem = held_reserve(hb, b_off + 2 * b_k + b_P0i);
        int64_t* eoff = reinterpret_cast<int64_t*>(mem);
        int* ek = reinterpret_cast<int*>(mem + b_off);
        int* status = reinterpret_cast<int*>(mem + b_off + b_k);
        double* P0i = reinterpret_cast<double*>(mem + b_off + 2 * b_k);
        NAJA_HIP(hipMemcpyAsync(eoff, off.data(), sizeof(int64_t) * n_variants, hipMemcpyHostToDevice, s));
        NAJA_HIP(hipMemcpyAsync(ek, extra_rows, sizeof(int) * n_variants, hipMemcpyHostToDevice, s));
        facet_launch(A0, b0, m0, d, n_variants, extra_A, extra_b, eoff, ek, x_c, P0, theta, include_base, completion,
                     exchange_tol, P0i, P, status, s);
        int* st = held_host(hb, (size_t)n_variants);
        NAJA_HIP(hipMemcpyAsync(st, status, sizeof(int) * n_variants, hipMemcpyDeviceToHost, s));
        NAJA_HIP(hipStreamSynchronize(s));
        for (int v = 0; v < n_variants; ++v) {
            NAJA_REQUIRE(st[v] != NAJA_ERR_INVALID_ARGUMENT, NAJA_ERR_INVALID_ARGUMENT,
                         "facet rounding of variant %d: the centre is not interior", v);
            NAJA_REQUIRE(st[v] == NAJA_OK, NAJA_ERR_RUNTIME,
                         "facet rounding of variant %d: a singular base rounding or no independent completion", v);
        }
    });
}
