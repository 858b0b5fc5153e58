// Learned preconditioner: a 3-layer MLP (features -> H1 -> H2 -> dim x rank
// factor V) trained on the GPU with hand-written BF16 MFMA kernels.
//
//   z1 = x W1^T + b1, h1 = relu(z1)
//   z2 = h1 W2^T + b2, h2 = relu(z2)
//   o  = h2 W3^T + b3, V = out_scale * reshape(o, dim, rank)
//
// Two covariance models are trained by the Gaussian negative log likelihood
// of pilot covariances S (see nll_kernel), averaged over the batch:
// * precision: Pi = I + V V^T in units of the base polytope; its whitening
//   is P = (I + V V^T)^{-1/2};
// * covariance: C = tau (I + V V^T) in the frame of a base rounding, tau
//   profiled out (scale-free); its update of the base is P = (I + V V^T)^{1/2}.
//
// Every matrix product runs on v_mfma_f32_32x32x8bf16_1k: operands are
// rounded to BF16 (round-to-nearest-even) as they are staged into LDS and
// accumulated in FP32. Weights are kept as FP32 masters for Adam and as BF16
// copies that the products read. The loss and its gradient are FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/naja_amd.h"
#include "common.h"

namespace naja {
namespace mlp {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef unsigned short bf16_t;

__device__ __forceinline__ bf16_t f2bf(float x)
{
    unsigned u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__device__ __forceinline__ bf16_t to_bf(float x) { return f2bf(x); }
__device__ __forceinline__ bf16_t to_bf(bf16_t x) { return x; }

enum { EPI_NONE = 0, EPI_BIAS = 1, EPI_BIAS_RELU = 2, EPI_MASK = 3 };

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 4;

// One thread's share (8 elements) of a 64-row x 32-k operand tile, held in
// registers between its global load and its BF16 store into LDS, so that the
// next tile's loads are in flight while the current one is multiplied.
// KC (k contiguous in memory, element (row, k) at P[row*ld + k]): row tid / 4,
// k (tid % 4) * 8 + j. Otherwise (row contiguous, P[k*ld + row]): k tid / 8,
// row (tid % 8) * 8 + j. `vec`: the host found P 16-byte aligned and ld a
// multiple of 16 bytes' worth of elements, so a whole in-bounds segment of 8
// is loaded as 16-byte vectors; partial segments are loaded element-wise.
template <typename T, bool KC>
struct TileRegs {
    float v[8];

    __device__ __forceinline__ void load(const T* __restrict__ P, int64_t ld, int row0, int nrows, int k0, int kend,
                                         bool vec, int tid)
    {
        const int row = KC ? row0 + tid / 4 : row0 + (tid % 8) * 8;
        const int k = KC ? k0 + (tid % 4) * 8 : k0 + tid / 8;
        const T* q = KC ? P + (int64_t)row * ld + k : P + (int64_t)k * ld + row;
        const bool whole = KC ? (row < nrows && k + 8 <= kend) : (k < kend && row + 8 <= nrows);
        if (vec && whole) {
            if (sizeof(T) == 4) {
                const float4 x0 = *reinterpret_cast<const float4*>(q);
                const float4 x1 = *reinterpret_cast<const float4*>(q + 4);
                v[0] = x0.x, v[1] = x0.y, v[2] = x0.z, v[3] = x0.w;
                v[4] = x1.x, v[5] = x1.y, v[6] = x1.z, v[7] = x1.w;
            } else {
                const uint4 x = *reinterpret_cast<const uint4*>(q);
                const unsigned w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    v[2 * j] = __uint_as_float(w[j] << 16);
                    v[2 * j + 1] = __uint_as_float(w[j] & 0xffff0000u);
                }
            }
            return;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const bool in = KC ? (row < nrows && k + j < kend) : (k < kend && row + j < nrows);
            v[j] = in ? elem(q + j) : 0.0f;
        }
    }

    __device__ __forceinline__ static float elem(const float* q) { return *q; }
    __device__ __forceinline__ static float elem(const bf16_t* q) { return __uint_as_float((unsigned)*q << 16); }

    __device__ __forceinline__ void store(bf16_t (*S)[LDK], int tid) const
    {
        if (KC) {
            const int r = tid / 4, kk = (tid % 4) * 8;
            unsigned w[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) w[j] = (unsigned)f2bf(v[2 * j]) | ((unsigned)f2bf(v[2 * j + 1]) << 16);
            *reinterpret_cast<uint2*>(&S[r][kk]) = make_uint2(w[0], w[1]);
            *reinterpret_cast<uint2*>(&S[r][kk + 4]) = make_uint2(w[2], w[3]);
        } else {
            const int k = tid / 8, r = (tid % 8) * 8;
#pragma unroll
            for (int j = 0; j < 8; ++j) S[r + j][k] = f2bf(v[j]);
        }
    }
};

// C(m, n) = alpha * (sum_k A(m, k) B(k, n) [+ bias(n)]) [relu] [* (mask(m, n) > 0)].
// A(m, k) = A[m*lda + k] if AK (k contiguous) else A[k*lda + m];
// B(k, n) = B[n*ldb + k] if BKC else B[k*ldb + n].
// 64 x 64 output tile per workgroup, 4 waves of 32 x 32, K staged 32 at a time
// (the next 32 loaded into registers while the current 32 are multiplied).
// Split K: workgroup z of grid.z covers k in [z kps, (z + 1) kps) and, with
// `part`, writes its raw tile to part[z][m][n] for gemm_reduce_kernel (the
// epilogue then runs there). rowsum (A stored m-contiguous only): the FP32
// sums over k of A(m, k), taken from the unrounded operand while it is staged
// by the workgroups of the first column block (a bias gradient folded into the
// weight-gradient product); per split in part[splits*M*N + z*M + m] when split.
template <typename TA, typename TB, bool AK, bool BKC, int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(int M, int N, int K, const TA* __restrict__ A, int64_t lda,
                                                   const TB* __restrict__ B, int64_t ldb, float* __restrict__ C,
                                                   int64_t ldc, float alpha, const float* __restrict__ bias,
                                                   const float* __restrict__ mask, int64_t ldm, int kps,
                                                   float* __restrict__ part, float* __restrict__ rowsum, int vec_a,
                                                   int vec_b)
{
    __shared__ __align__(16) bf16_t As[BM][LDK];
    __shared__ __align__(16) bf16_t Bs[BN][LDK];
    __shared__ float rsh[BK][BM];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN, z = blockIdx.z;
    const int kbeg = z * kps, kend = min(K, kbeg + kps);
    const bool do_rs = !AK && rowsum != nullptr && blockIdx.x == 0;
    float rs[8];  // row sums of this thread's 8 rows over its k (tid / 8 of every tile)
#pragma unroll
    for (int j = 0; j < 8; ++j) rs[j] = 0.0f;
    floatx16 acc;
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
    TileRegs<TA, AK> ra;
    TileRegs<TB, BKC> rb;
    if (kbeg < kend) {
        ra.load(A, lda, m0, M, kbeg, kend, vec_a, tid);
        rb.load(B, ldb, n0, N, kbeg, kend, vec_b, tid);
    }
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        if (do_rs) {
#pragma unroll
            for (int j = 0; j < 8; ++j) rs[j] += ra.v[j];
        }
        ra.store(As, tid);
        rb.store(Bs, tid);
        __syncthreads();
        if (k0 + BK < kend) {
            ra.load(A, lda, m0, M, k0 + BK, kend, vec_a, tid);
            rb.load(B, ldb, n0, N, k0 + BK, kend, vec_b, tid);
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 8) {
            // Lanes 0-31 carry k = kk..kk+3 of rows/columns 0-31, lanes 32-63 k = kk+4..kk+7.
            const int kb = kk + (lane >> 5) * 4;
            const uint2 ua = *reinterpret_cast<const uint2*>(&As[wm + (lane & 31)][kb]);
            const uint2 ub = *reinterpret_cast<const uint2*>(&Bs[wn + (lane & 31)][kb]);
            bf16x4 a, b;
            a[0] = (short)(ua.x & 0xffffu);
            a[1] = (short)(ua.x >> 16);
            a[2] = (short)(ua.y & 0xffffu);
            a[3] = (short)(ua.y >> 16);
            b[0] = (short)(ub.x & 0xffffu);
            b[1] = (short)(ub.x >> 16);
            b[2] = (short)(ub.y & 0xffffu);
            b[3] = (short)(ub.y >> 16);
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    if (do_rs) {
#pragma unroll
        for (int j = 0; j < 8; ++j) rsh[tid / 8][(tid % 8) * 8 + j] = rs[j];
        __syncthreads();
        if (tid < BM && m0 + tid < M) {
            float sum = 0.0f;
            for (int q = 0; q < BK; ++q) sum += rsh[q][tid];
            if (part)
                part[(int64_t)gridDim.z * M * N + (int64_t)z * M + m0 + tid] = sum;
            else
                rowsum[m0 + tid] = sum;
        }
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int r = m0 + wm + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
        const int c = n0 + wn + (lane & 31);
        if (r < M && c < N) {
            float v = acc[e];
            if (part) {
                part[((int64_t)z * M + r) * N + c] = v;
                continue;
            }
            if (EPI == EPI_BIAS || EPI == EPI_BIAS_RELU) v += bias[c];
            v *= alpha;
            if (EPI == EPI_BIAS_RELU) v = fmaxf(v, 0.0f);
            if (EPI == EPI_MASK) v = (mask[(int64_t)r * ldm + c] > 0.0f) ? v : 0.0f;
            C[(int64_t)r * ldc + c] = v;
        }
    }
}

// The split-K sum in split order, then the epilogue of gemm_kernel; and the
// row sums' split sum.
template <int EPI>
__global__ void gemm_reduce_kernel(const float* __restrict__ part, int splits, int M, int N, float* __restrict__ C,
                                   int64_t ldc, float alpha, const float* __restrict__ bias,
                                   const float* __restrict__ mask, int64_t ldm, float* __restrict__ rowsum)
{
    const int64_t MN = (int64_t)M * N;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < MN; e += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / N), c = (int)(e % N);
        float v = part[e];
        for (int z = 1; z < splits; ++z) v += part[(int64_t)z * MN + e];
        if (EPI == EPI_BIAS || EPI == EPI_BIAS_RELU) v += bias[c];
        v *= alpha;
        if (EPI == EPI_BIAS_RELU) v = fmaxf(v, 0.0f);
        if (EPI == EPI_MASK) v = (mask[(int64_t)r * ldm + c] > 0.0f) ? v : 0.0f;
        C[(int64_t)r * ldc + c] = v;
    }
    if (rowsum)
        for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < M; m += (int64_t)gridDim.x * blockDim.x) {
            float v = part[splits * MN + m];
            for (int z = 1; z < splits; ++z) v += part[splits * MN + (int64_t)z * M + m];
            rowsum[m] = v;
        }
}

// Launch: with `part` (part_floats floats of scratch), K is split while the
// launch has fewer than 512 workgroups, every split keeps at least 64 of K,
// and the partial tiles (and row sums) fit.
template <typename TA, typename TB, bool AK, bool BKC, int EPI>
void gemm(hipStream_t s, int M, int N, int K, const TA* A, int64_t lda, const TB* B, int64_t ldb, float* C,
          int64_t ldc, float alpha = 1.0f, const float* bias = nullptr, const float* mask = nullptr, int64_t ldm = 0,
          float* rowsum = nullptr, float* part = nullptr, size_t part_floats = 0)
{
    const int gx = (N + BN - 1) / BN, gy = (M + BM - 1) / BM;
    int splits = 1;
    if (part) {
        auto fits = [&](int sp) { return (size_t)sp * M * N + (size_t)sp * M <= part_floats; };
        while (gx * gy * splits < 512 && K / (2 * splits) >= 64 && fits(2 * splits)) splits *= 2;
    }
    const int kps = (((K + splits - 1) / splits + BK - 1) / BK) * BK;
    splits = (K + kps - 1) / kps;
    auto vec = [](const void* p, int64_t ld, size_t elem) {
        return (int)(((uintptr_t)p % 16) == 0 && ((size_t)ld * elem) % 16 == 0);
    };
    const int va = vec(A, lda, sizeof(TA)), vb = vec(B, ldb, sizeof(TB));
    if (splits <= 1) {
        hipLaunchKernelGGL((gemm_kernel<TA, TB, AK, BKC, EPI>), dim3(gx, gy, 1), dim3(256), 0, s, M, N, K, A, lda, B,
                           ldb, C, ldc, alpha, bias, mask, ldm, K, nullptr, rowsum, va, vb);
        NAJA_HIP(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL((gemm_kernel<TA, TB, AK, BKC, EPI>), dim3(gx, gy, splits), dim3(256), 0, s, M, N, K, A, lda,
                       B, ldb, C, ldc, alpha, bias, mask, ldm, kps, part, rowsum, va, vb);
    NAJA_HIP(hipGetLastError());
    const int64_t MN = (int64_t)M * N;
    hipLaunchKernelGGL(gemm_reduce_kernel<EPI>, dim3((unsigned)std::min<int64_t>((MN + 255) / 256, 4096)), dim3(256),
                       0, s, part, splits, M, N, C, ldc, alpha, bias, mask, ldm, rowsum);
    NAJA_HIP(hipGetLastError());
}

// Adam (Kingma & Ba, 2015) on the FP32 masters; refreshes the BF16 copies.
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                            float* __restrict__ m2, bf16_t* __restrict__ pbf, int64_t n, float lr, float beta1,
                            float beta2, float eps, float c1, float c2)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = beta1 * m1[i] + (1.0f - beta1) * gi;
    const float b = beta2 * m2[i] + (1.0f - beta2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float mhat = a * c1, vhat = b * c2;
    const float w = p[i] - lr * mhat / (sqrtf(vhat) + eps);
    p[i] = w;
    pbf[i] = f2bf(w);
}

__global__ void to_bf16_kernel(const float* __restrict__ p, bf16_t* __restrict__ pbf, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pbf[i] = f2bf(p[i]);
}

// One workgroup per condition b, FP64 throughout. V = out_scale * O[b]
// (dim x R), S = S[b] (dim x dim), M = I + V^T V = L L^T (Cholesky),
// M^{-1} = L^{-T} L^{-1} (L inverted column by column).
//
// MODEL 0, precision model Pi = I + V V^T (S in units of the base polytope):
//   loss = tr(V^T S V) - logdet M,  dL/dV = 2 S V - 2 V M^{-1}.
// MODEL 1, scale-free covariance model C = tau (I + V V^T), tau profiled out
// (tau = q / dim):
//   q = tr(C~^{-1} S) = tr S - tr(M^{-1} G), G = V^T S V,
//   loss = dim log(q / dim) + logdet M,
//   dL/dV = (dim / q) (2 V M^{-1} G M^{-1} - 2 S V M^{-1}) + 2 V M^{-1}.
// dO[b] = out_scale * grad_scale * dL/dV.
template <int R, int MODEL>
__global__ __launch_bounds__(256) void nll_kernel(const float* __restrict__ O, int64_t ldo,
                                                  const float* __restrict__ S, int dim, float out_scale,
                                                  float grad_scale, float* __restrict__ loss, float* __restrict__ dO)
{
    extern __shared__ double lds[];
    double* V = lds;                      // dim x R
    double* SV = V + (size_t)dim * R;     // dim x R
    double* L = SV + (size_t)dim * R;     // R x R (M, then its Cholesky factor)
    double* Li = L + R * R;               // R x R, L^{-1}
    double* Mi = Li + R * R;              // R x R, M^{-1}
    double* G = Mi + R * R;               // R x R, V^T S V        (MODEL 1)
    double* T = G + R * R;                // R x R, M^{-1} G       (MODEL 1)
    double* K = T + R * R;                // R x R, M^{-1} G M^{-1} (MODEL 1)
    __shared__ double red[256];
    __shared__ double red2[256];
    const int b = blockIdx.x, tid = threadIdx.x;
    const float* Ob = O + (int64_t)b * ldo;
    const float* Sb = S + (int64_t)b * dim * dim;
    for (int i = tid; i < dim * R; i += 256) V[i] = (double)out_scale * (double)Ob[i];
    __syncthreads();
    double tr = 0.0;
    for (int i = tid; i < dim * R; i += 256) {
        const int r = i / R, c = i % R;
        const float* Srow = Sb + (int64_t)r * dim;
        double s = 0.0;
        for (int k = 0; k < dim; ++k) s += (double)Srow[k] * V[k * R + c];
        SV[i] = s;
        tr += V[i] * s;
    }
    double trs = 0.0;
    if (MODEL == 1)
        for (int i = tid; i < dim; i += 256) trs += (double)Sb[(int64_t)i * dim + i];
    for (int i = tid; i < R * R; i += 256) {
        const int a = i / R, c = i % R;
        double s = (a == c) ? 1.0 : 0.0;
        for (int k = 0; k < dim; ++k) s += V[k * R + a] * V[k * R + c];
        L[i] = s;
    }
    red[tid] = tr;
    red2[tid] = trs;
    __syncthreads();
    if (MODEL == 1) {
        for (int i = tid; i < R * R; i += 256) {
            const int a = i / R, c = i % R;
            double s = 0.0;
            for (int k = 0; k < dim; ++k) s += V[k * R + a] * SV[k * R + c];
            G[i] = s;
        }
    }
    for (int j = 0; j < R; ++j) {
        if (tid == 0) {
            double s = L[j * R + j];
            for (int k = 0; k < j; ++k) s -= L[j * R + k] * L[j * R + k];
            L[j * R + j] = sqrt(s);
        }
        __syncthreads();
        const int i = j + 1 + tid;
        if (i < R) {
            double s = L[i * R + j];
            for (int k = 0; k < j; ++k) s -= L[i * R + k] * L[j * R + k];
            L[i * R + j] = s / L[j * R + j];
        }
        __syncthreads();
    }
    if (tid < R) {  // column tid of L^{-1}
        const int c = tid;
        for (int i = 0; i < c; ++i) Li[i * R + c] = 0.0;
        for (int i = c; i < R; ++i) {
            double s = (i == c) ? 1.0 : 0.0;
            for (int k = c; k < i; ++k) s -= L[i * R + k] * Li[k * R + c];
            Li[i * R + c] = s / L[i * R + i];
        }
    }
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            red[tid] += red[tid + s];
            red2[tid] += red2[tid + s];
        }
        __syncthreads();
    }
    double logdet = 0.0;
    for (int j = 0; j < R; ++j) logdet += 2.0 * log(L[j * R + j]);
    for (int i = tid; i < R * R; i += 256) {
        const int a = i / R, c = i % R;
        double s = 0.0;
        for (int k = (a > c ? a : c); k < R; ++k) s += Li[k * R + a] * Li[k * R + c];
        Mi[i] = s;
    }
    __syncthreads();
    const double gs = (double)out_scale * (double)grad_scale;
    float* out = dO + (int64_t)b * ldo;
    if (MODEL == 0) {
        if (tid == 0) loss[b] = (float)(red[0] - logdet);
        for (int i = tid; i < dim * R; i += 256) {
            const int r = i / R, c = i % R;
            double x = 0.0;
            for (int k = 0; k < R; ++k) x += V[r * R + k] * Mi[k * R + c];
            out[i] = (float)(gs * (2.0 * SV[i] - 2.0 * x));
        }
    } else {
        for (int i = tid; i < R * R; i += 256) {
            const int a = i / R, c = i % R;
            double s = 0.0;
            for (int k = 0; k < R; ++k) s += Mi[a * R + k] * G[k * R + c];
            T[i] = s;
        }
        __syncthreads();
        double tmg = 0.0;
        for (int i = tid; i < R * R; i += 256) {
            const int a = i / R, c = i % R;
            double s = 0.0;
            for (int k = 0; k < R; ++k) s += T[a * R + k] * Mi[k * R + c];
            K[i] = s;
            if (a == c) tmg += T[i];
        }
        __syncthreads();  // red[] is free again: the sums above were read before
        red[tid] = tmg;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s) red[tid] += red[tid + s];
            __syncthreads();
        }
        const double q = red2[0] - red[0];
        const double fd = (double)dim / q;
        if (tid == 0) loss[b] = (float)((double)dim * log(q / (double)dim) + logdet);
        for (int i = tid; i < dim * R; i += 256) {
            const int r = i / R, c = i % R;
            double vk = 0.0, svm = 0.0, vm = 0.0;
            for (int k = 0; k < R; ++k) {
                vk += V[r * R + k] * K[k * R + c];
                svm += SV[r * R + k] * Mi[k * R + c];
                vm += V[r * R + k] * Mi[k * R + c];
            }
            out[i] = (float)(gs * (fd * (2.0 * vk - 2.0 * svm) + 2.0 * vm));
        }
    }
}

// Per-variant sample mean (FP64): samples (nv, rows, d).
__global__ void mean_kernel(const float* __restrict__ X, int64_t rows, int d, double* __restrict__ mean)
{
    const int v = blockIdx.y;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d) return;
    const float* Xv = X + (int64_t)v * rows * d;
    double s = 0.0;
    for (int64_t r = 0; r < rows; ++r) s += (double)Xv[r * d + j];
    mean[(int64_t)v * d + j] = s / (double)rows;
}

// S[v] = scale / (rows - 1) * sum_r (x_r - mean)(x_r - mean)^T, 32 x 32
// output tile per workgroup, FP64 accumulation, 64 centred rows staged per pass.
__global__ __launch_bounds__(256) void cov_kernel(const float* __restrict__ X, int64_t rows, int d,
                                                  const double* __restrict__ mean, double scale,
                                                  float* __restrict__ S)
{
    __shared__ double xi[64][33];
    __shared__ double xj[64][33];
    const int v = blockIdx.z, ti = blockIdx.y * 32, tj = blockIdx.x * 32, tid = threadIdx.x;
    const float* Xv = X + (int64_t)v * rows * d;
    const double* mu = mean + (int64_t)v * d;
    const int c = tid & 31, r0 = tid >> 5;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t s0 = 0; s0 < rows; s0 += 64) {
        for (int i = tid; i < 64 * 32; i += 256) {
            const int rr = i >> 5, cc = i & 31;
            const int64_t row = s0 + rr;
            double a = 0.0, bb = 0.0;
            if (row < rows) {
                if (ti + cc < d) a = (double)Xv[row * d + ti + cc] - mu[ti + cc];
                if (tj + cc < d) bb = (double)Xv[row * d + tj + cc] - mu[tj + cc];
            }
            xi[rr][cc] = a;
            xj[rr][cc] = bb;
        }
        __syncthreads();
        for (int rr = 0; rr < 64; ++rr) {
            const double bj = xj[rr][c];
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] += xi[rr][r0 + 8 * q] * bj;
        }
        __syncthreads();
    }
    const double f = scale / (double)(rows - 1);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = ti + r0 + 8 * q, j = tj + c;
        if (i < d && j < d) S[((int64_t)v * d + i) * d + j] = (float)(acc[q] * f);
    }
}

}  // namespace mlp
}  // namespace naja

using namespace naja;
using namespace naja::mlp;

struct naja_lp {
    int F, dim, rank, H1, H2, O;
    int model;
    float out_scale;
    int64_t n_params;
    int64_t off_w1, off_b1, off_w2, off_b2, off_w3, off_b3;
    float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    bf16_t* pbf = nullptr;
    int step = 0;
    // activations for `cap` rows
    int cap = 0;
    float *h1 = nullptr, *h2 = nullptr, *o = nullptr, *d1 = nullptr, *d2 = nullptr, *dO = nullptr, *lossb = nullptr;
    // split-K partial tiles and row sums of the products, part_floats floats
    float* part = nullptr;
    size_t part_floats = 0;
};

namespace {

constexpr int kMaxNllDim = 160;  // LDS: (2 * dim + 6 * 32) * 32 doubles <= 128 KB

void free_acts(naja_lp* h)
{
    for (float** q : {&h->h1, &h->h2, &h->o, &h->d1, &h->d2, &h->dO, &h->lossb, &h->part}) {
        if (*q) (void)hipFree(*q);
        *q = nullptr;
    }
    h->cap = 0;
    h->part_floats = 0;
}

void ensure_acts(naja_lp* h, int batch)
{
    if (batch <= h->cap) return;
    NAJA_HIP(hipDeviceSynchronize());
    free_acts(h);
    const int cap = std::max(batch, 256);
    NAJA_HIP(hipMalloc(&h->h1, sizeof(float) * cap * h->H1));
    NAJA_HIP(hipMalloc(&h->h2, sizeof(float) * cap * h->H2));
    NAJA_HIP(hipMalloc(&h->o, sizeof(float) * cap * h->O));
    NAJA_HIP(hipMalloc(&h->d1, sizeof(float) * cap * h->H1));
    NAJA_HIP(hipMalloc(&h->d2, sizeof(float) * cap * h->H2));
    NAJA_HIP(hipMalloc(&h->dO, sizeof(float) * cap * h->O));
    NAJA_HIP(hipMalloc(&h->lossb, sizeof(float) * cap));
    // Up to 16 K-splits of the batch-row products and their row sums.
    const size_t part = 16 * ((size_t)cap * std::max(h->H1, h->H2) + std::max({cap, h->O, h->H1, h->H2}));
    NAJA_HIP(hipMalloc(&h->part, sizeof(float) * part));
    h->part_floats = part;
    h->cap = cap;
}

const bf16_t* wbf(const naja_lp* h, int64_t off) { return h->pbf + off; }

// Layers 1-2 into h1, h2; layer 3 into out3 = alpha3 * (h2 W3^T + b3).
void forward(naja_lp* h, const float* X, int B, hipStream_t s, float* out3, float alpha3)
{
    float* P = h->part;
    const size_t PF = h->part_floats;
    gemm<float, bf16_t, true, true, EPI_BIAS_RELU>(s, B, h->H1, h->F, X, h->F, wbf(h, h->off_w1), h->F, h->h1, h->H1,
                                                   1.0f, h->p + h->off_b1, nullptr, 0, nullptr, P, PF);
    gemm<float, bf16_t, true, true, EPI_BIAS_RELU>(s, B, h->H2, h->H1, h->h1, h->H1, wbf(h, h->off_w2), h->H1, h->h2,
                                                   h->H2, 1.0f, h->p + h->off_b2, nullptr, 0, nullptr, P, PF);
    gemm<float, bf16_t, true, true, EPI_BIAS>(s, B, h->O, h->H2, h->h2, h->H2, wbf(h, h->off_w3), h->H2, out3, h->O,
                                              alpha3, h->p + h->off_b3, nullptr, 0, nullptr, P, PF);
}

void loss_and_backward(naja_lp* h, const float* X, const float* S, int B, float* loss, hipStream_t s)
{
    const size_t lds = sizeof(double) * (2 * (size_t)h->dim * 32 + 6 * 32 * 32);
    auto fn = (h->model == NAJA_LP_COVARIANCE) ? nll_kernel<32, 1> : nll_kernel<32, 0>;
    hipLaunchKernelGGL(fn, dim3(B), dim3(256), lds, s, h->o, (int64_t)h->O, S, h->dim, h->out_scale, 1.0f / (float)B,
                       loss ? loss : h->lossb, h->dO);
    NAJA_HIP(hipGetLastError());
    float* g = h->g;
    float* P = h->part;
    const size_t PF = h->part_floats;
    // Layer 3: dW3 = dO^T h2 with db3 = the row sums of dO^T, d2 = (dO W3) * relu'(z2).
    gemm<float, float, false, false, EPI_NONE>(s, h->O, h->H2, B, h->dO, h->O, h->h2, h->H2, g + h->off_w3, h->H2,
                                               1.0f, nullptr, nullptr, 0, g + h->off_b3, P, PF);
    gemm<float, bf16_t, true, false, EPI_MASK>(s, B, h->H2, h->O, h->dO, h->O, wbf(h, h->off_w3), h->H2, h->d2, h->H2,
                                               1.0f, nullptr, h->h2, h->H2, nullptr, P, PF);
    // Layer 2.
    gemm<float, float, false, false, EPI_NONE>(s, h->H2, h->H1, B, h->d2, h->H2, h->h1, h->H1, g + h->off_w2, h->H1,
                                               1.0f, nullptr, nullptr, 0, g + h->off_b2, P, PF);
    gemm<float, bf16_t, true, false, EPI_MASK>(s, B, h->H1, h->H2, h->d2, h->H2, wbf(h, h->off_w2), h->H1, h->d1, h->H1,
                                               1.0f, nullptr, h->h1, h->H1, nullptr, P, PF);
    // Layer 1.
    gemm<float, float, false, false, EPI_NONE>(s, h->H1, h->F, B, h->d1, h->H1, X, h->F, g + h->off_w1, h->F, 1.0f,
                                               nullptr, nullptr, 0, g + h->off_b1, P, PF);
}

void check_batch(const naja_lp* h, const float* X, int B)
{
    NAJA_REQUIRE(h && X && B > 0, NAJA_ERR_INVALID_ARGUMENT, "learned preconditioner: null handle/features or batch %d",
                 B);
}

// splitmix64 stream for the initial weights.
struct SplitMix {
    uint64_t s;
    double uniform()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    }
};

}  // namespace

extern "C" int naja_lp_create(int n_features, int dim, int rank, int hidden1, int hidden2, int model, uint64_t seed,
                              float out_scale, naja_lp** out)
{
    return guarded([&]() {
        NAJA_REQUIRE(out, NAJA_ERR_INVALID_ARGUMENT, "naja_lp_create: null output");
        NAJA_REQUIRE(n_features > 0 && dim > 0 && hidden1 > 0 && hidden2 > 0, NAJA_ERR_INVALID_ARGUMENT,
                     "naja_lp_create: sizes must be positive");
        NAJA_REQUIRE(rank == 32, NAJA_ERR_UNSUPPORTED, "naja_lp_create: rank %d (compiled for 32)", rank);
        NAJA_REQUIRE(model == NAJA_LP_PRECISION || model == NAJA_LP_COVARIANCE, NAJA_ERR_INVALID_ARGUMENT,
                     "naja_lp_create: unknown model %d", model);
        NAJA_REQUIRE(dim >= rank && dim <= kMaxNllDim, NAJA_ERR_UNSUPPORTED,
                     "naja_lp_create: dim %d outside [%d, %d]", dim, rank, kMaxNllDim);
        auto* h = new naja_lp();
        h->F = n_features;
        h->dim = dim;
        h->rank = rank;
        h->H1 = hidden1;
        h->H2 = hidden2;
        h->O = dim * rank;
        h->out_scale = out_scale;
        h->model = model;
        h->off_w1 = 0;
        h->off_b1 = h->off_w1 + (int64_t)hidden1 * n_features;
        h->off_w2 = h->off_b1 + hidden1;
        h->off_b2 = h->off_w2 + (int64_t)hidden2 * hidden1;
        h->off_w3 = h->off_b2 + hidden2;
        h->off_b3 = h->off_w3 + (int64_t)h->O * hidden2;
        h->n_params = h->off_b3 + h->O;
        // He-uniform for the ReLU layers; the output layer starts 100x
        // smaller so the initial map is close to the identity (V ~ 0).
        std::vector<float> init(h->n_params, 0.0f);
        SplitMix rng{seed};
        auto fill = [&](int64_t off, int64_t count, double limit) {
            for (int64_t i = 0; i < count; ++i) init[off + i] = (float)((2.0 * rng.uniform() - 1.0) * limit);
        };
        fill(h->off_w1, (int64_t)hidden1 * n_features, std::sqrt(6.0 / n_features));
        fill(h->off_w2, (int64_t)hidden2 * hidden1, std::sqrt(6.0 / hidden1));
        fill(h->off_w3, (int64_t)h->O * hidden2, 0.01 * std::sqrt(6.0 / hidden2));
        try {
            NAJA_HIP(hipMalloc(&h->p, sizeof(float) * h->n_params));
            NAJA_HIP(hipMalloc(&h->g, sizeof(float) * h->n_params));
            NAJA_HIP(hipMalloc(&h->m1, sizeof(float) * h->n_params));
            NAJA_HIP(hipMalloc(&h->m2, sizeof(float) * h->n_params));
            NAJA_HIP(hipMalloc(&h->pbf, sizeof(bf16_t) * h->n_params));
            NAJA_HIP(hipMemcpy(h->p, init.data(), sizeof(float) * h->n_params, hipMemcpyHostToDevice));
            NAJA_HIP(hipMemset(h->g, 0, sizeof(float) * h->n_params));
            NAJA_HIP(hipMemset(h->m1, 0, sizeof(float) * h->n_params));
            NAJA_HIP(hipMemset(h->m2, 0, sizeof(float) * h->n_params));
            hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((h->n_params + 255) / 256)), dim3(256), 0, nullptr,
                               h->p, h->pbf, h->n_params);
            NAJA_HIP(hipGetLastError());
            NAJA_HIP(hipDeviceSynchronize());
        } catch (...) {
            for (void* q : {(void*)h->p, (void*)h->g, (void*)h->m1, (void*)h->m2, (void*)h->pbf})
                if (q) (void)hipFree(q);
            delete h;
            throw;
        }
        *out = h;
    });
}

extern "C" int naja_lp_destroy(naja_lp* h)
{
    return guarded([&]() {
        if (!h) return;
        NAJA_HIP(hipDeviceSynchronize());
        free_acts(h);
        for (void* q : {(void*)h->p, (void*)h->g, (void*)h->m1, (void*)h->m2, (void*)h->pbf})
            if (q) (void)hipFree(q);
        delete h;
    });
}

extern "C" int naja_lp_param_count(const naja_lp* h, int64_t* n)
{
    return guarded([&]() {
        NAJA_REQUIRE(h && n, NAJA_ERR_INVALID_ARGUMENT, "naja_lp_param_count: null argument");
        *n = h->n_params;
    });
}

extern "C" int naja_lp_get_params(const naja_lp* h, float* dst, void* stream)
{
    return guarded([&]() {
        NAJA_REQUIRE(h && dst, NAJA_ERR_INVALID_ARGUMENT, "naja_lp_get_params: null argument");
        NAJA_HIP(hipMemcpyAsync(dst, h->p, sizeof(float) * h->n_params, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    });
}

extern "C" int naja_lp_set_params(naja_lp* h, const float* src, void* stream)
{
    return guarded([&]() {
        NAJA_REQUIRE(h && src, NAJA_ERR_INVALID_ARGUMENT, "naja_lp_set_params: null argument");
        hipStream_t s = (hipStream_t)stream;
        NAJA_HIP(hipMemcpyAsync(h->p, src, sizeof(float) * h->n_params, hipMemcpyDeviceToDevice, s));
        NAJA_HIP(hipMemsetAsync(h->m1, 0, sizeof(float) * h->n_params, s));
        NAJA_HIP(hipMemsetAsync(h->m2, 0, sizeof(float) * h->n_params, s));
        hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((h->n_params + 255) / 256)), dim3(256), 0, s, h->p, h->pbf,
                           h->n_params);
        NAJA_HIP(hipGetLastError());
        h->step = 0;
    });
}

extern "C" int naja_lp_forward(naja_lp* h, const float* features, int batch, float* V, void* stream)
{
    return guarded([&]() {
        check_batch(h, features, batch);
        NAJA_REQUIRE(V, NAJA_ERR_INVALID_ARGUMENT, "naja_lp_forward: null output");
        ensure_acts(h, batch);
        forward(h, features, batch, (hipStream_t)stream, V, h->out_scale);
    });
}

extern "C" int naja_lp_loss_grad(naja_lp* h, const float* features, const float* S, int batch, float* loss,
                                 float* grad, void* stream)
{
    return guarded([&]() {
        check_batch(h, features, batch);
        NAJA_REQUIRE(S, NAJA_ERR_INVALID_ARGUMENT, "naja_lp_loss_grad: null covariance");
        ensure_acts(h, batch);
        hipStream_t s = (hipStream_t)stream;
        forward(h, features, batch, s, h->o, 1.0f);
        loss_and_backward(h, features, S, batch, loss, s);
        if (grad) NAJA_HIP(hipMemcpyAsync(grad, h->g, sizeof(float) * h->n_params, hipMemcpyDeviceToDevice, s));
    });
}

extern "C" int naja_lp_train_step(naja_lp* h, const float* features, const float* S, int batch, float lr,
                                  float* loss, void* stream)
{
    return guarded([&]() {
        check_batch(h, features, batch);
        NAJA_REQUIRE(S, NAJA_ERR_INVALID_ARGUMENT, "naja_lp_train_step: null covariance");
        NAJA_REQUIRE(lr > 0.0f, NAJA_ERR_INVALID_ARGUMENT, "naja_lp_train_step: learning rate must be > 0");
        ensure_acts(h, batch);
        hipStream_t s = (hipStream_t)stream;
        forward(h, features, batch, s, h->o, 1.0f);
        loss_and_backward(h, features, S, batch, loss, s);
        h->step += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, h->step)));
        const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, h->step)));
        hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((h->n_params + 255) / 256)), dim3(256), 0, s, h->p, h->g,
                           h->m1, h->m2, h->pbf, h->n_params, lr, b1, b2, 1e-8f, c1, c2);
        NAJA_HIP(hipGetLastError());
    });
}

extern "C" int naja_pilot_covariance(const float* samples, int n_variants, int64_t rows, int d, double scale,
                                     float* S, double* mean, void* stream)
{
    return guarded([&]() {
        NAJA_REQUIRE(samples && S && mean && n_variants > 0 && d > 0, NAJA_ERR_INVALID_ARGUMENT,
                     "naja_pilot_covariance: bad arguments");
        NAJA_REQUIRE(rows >= 2, NAJA_ERR_INVALID_ARGUMENT, "naja_pilot_covariance: need at least 2 samples");
        NAJA_REQUIRE(n_variants <= 65535, NAJA_ERR_UNSUPPORTED, "naja_pilot_covariance: at most 65535 variants");
        hipStream_t s = (hipStream_t)stream;
        hipLaunchKernelGGL(mean_kernel, dim3((d + 127) / 128, n_variants), dim3(128), 0, s, samples, rows, d, mean);
        const int t = (d + 31) / 32;
        hipLaunchKernelGGL(cov_kernel, dim3(t, t, n_variants), dim3(256), 0, s, samples, rows, d, mean, scale, S);
        NAJA_HIP(hipGetLastError());
    });
}

extern "C" int naja_gemm_bf16(int trans_a, int trans_b, int M, int N, int K, const float* A, int64_t lda,
                              const float* B, int64_t ldb, float* C, int64_t ldc, float alpha, float* rowsum,
                              float* workspace, int64_t workspace_floats, void* stream)
{
    return guarded([&]() {
        NAJA_REQUIRE(A && B && C && M > 0 && N > 0 && K > 0, NAJA_ERR_INVALID_ARGUMENT, "naja_gemm_bf16: bad arguments");
        NAJA_REQUIRE(!rowsum || trans_a, NAJA_ERR_INVALID_ARGUMENT, "naja_gemm_bf16: row sums need A stored K x M");
        NAJA_REQUIRE(workspace_floats >= 0 && (workspace || workspace_floats == 0), NAJA_ERR_INVALID_ARGUMENT,
                     "naja_gemm_bf16: bad workspace");
        hipStream_t s = (hipStream_t)stream;
        const size_t wf = (size_t)workspace_floats;
        // op(A) is M x K: row-major A (trans_a = 0) has k contiguous; op(B) is
        // K x N: row-major B (trans_b = 0) has n contiguous.
        if (!trans_a && !trans_b)
            gemm<float, float, true, false, EPI_NONE>(s, M, N, K, A, lda, B, ldb, C, ldc, alpha, nullptr, nullptr, 0,
                                                      nullptr, workspace, wf);
        else if (!trans_a && trans_b)
            gemm<float, float, true, true, EPI_NONE>(s, M, N, K, A, lda, B, ldb, C, ldc, alpha, nullptr, nullptr, 0,
                                                     nullptr, workspace, wf);
        else if (trans_a && !trans_b)
            gemm<float, float, false, false, EPI_NONE>(s, M, N, K, A, lda, B, ldb, C, ldc, alpha, nullptr, nullptr, 0,
                                                       rowsum, workspace, wf);
        else
            gemm<float, float, false, true, EPI_NONE>(s, M, N, K, A, lda, B, ldb, C, ldc, alpha, nullptr, nullptr, 0,
                                                      rowsum, workspace, wf);
    });
}

// ---------------------------------------------------------------------------
// Dikin-ellipsoid rounding on the GPU (FP64): one workgroup per variant of a
// ragged batch [A0; E_v] y <= [b0; f_v]. H = sum_r u_r u_r^T, u_r = a_r / s_r
// at x_c: rows staged kEllRows at a time in LDS (read down the column-major
// blocks), each thread accumulating up to three 4 x 4 tiles of the lower
// triangle in registers; then H = L L^T in place, M = L^{-1} column by column
// into the free upper triangle (diagonal kept apart), P = M^T written as FP32.
// The host routine of the same name in precond.cpp does the same in FP64.
// ---------------------------------------------------------------------------
namespace naja {
namespace mlp {

constexpr int kEllRows = 16;
constexpr int kEllTiles = 3;  // 4 x 4 tiles of H per thread: d <= 4 * 38 = 152

// Doubles of the H region: H itself, or one staged row chunk when larger.
__host__ __device__ inline size_t ellipsoid_h_doubles(int d)
{
    const size_t dpad = (size_t)((d + 3) & ~3);
    const size_t hd = (size_t)d * d, ch = (size_t)kEllRows * dpad;
    return hd > ch ? hd : ch;
}

__global__ __launch_bounds__(256) void ellipsoid_kernel(const float* __restrict__ A0, const double* __restrict__ b0,
                                                        int m0, int d, const float* __restrict__ Ae,
                                                        const double* __restrict__ be,
                                                        const int64_t* __restrict__ eoff,
                                                        const int* __restrict__ ek, const float* __restrict__ xc,
                                                        float* __restrict__ P, int* __restrict__ status)
{
    extern __shared__ double sh[];
    double* H = sh;                                // d x d (first the staged row chunks)
    double* Ub = sh;                               // kEllRows x dpad
    double* Md = sh + ellipsoid_h_doubles(d);      // d: diagonal of M
    double* w = Md + d;                            // m_v: 1 / s_r
    const int v = blockIdx.x, tid = threadIdx.x;
    const int k = ek[v];
    const int64_t off = eoff[v];
    const int m = m0 + k;
    const float* x = xc ? xc + (int64_t)v * d : nullptr;
    // Row r, column c of [A0; E_v] (both blocks column-major).
    auto arow = [&](int r, int c) -> double {
        return r < m0 ? (double)A0[(int64_t)c * m0 + r] : (double)Ae[off * d + (int64_t)c * k + (r - m0)];
    };
    __shared__ int bad;
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int r = tid; r < m; r += 256) {
        double ax = 0.0;
        if (x)
            for (int c = 0; c < d; ++c) ax = fma(arow(r, c), (double)x[c], ax);
        const double s = (r < m0 ? b0[r] : be[off + (r - m0)]) - ax;
        if (!(s > 0.0)) bad = 1;
        w[r] = 1.0 / s;
    }
    __syncthreads();
    if (bad) {
        if (tid == 0) status[v] = NAJA_ERR_INVALID_ARGUMENT;
        return;
    }
    const int dpad = (d + 3) & ~3;
    const int nt = dpad / 4;
    const int ntiles = nt * (nt + 1) / 2;
    int ti[kEllTiles], tj[kEllTiles];
#pragma unroll
    for (int q = 0; q < kEllTiles; ++q) {
        // Tile t -> (I, J), I >= J, row-major over the lower triangle of tiles.
        const int t = tid + 256 * q;
        ti[q] = -1;
        tj[q] = 0;
        if (t < ntiles) {
            int I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
            while ((I + 1) * (I + 2) / 2 <= t) ++I;
            while (I * (I + 1) / 2 > t) --I;
            ti[q] = I;
            tj[q] = t - I * (I + 1) / 2;
        }
    }
    double acc[kEllTiles][4][4];
#pragma unroll
    for (int q = 0; q < kEllTiles; ++q)
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) acc[q][x][y] = 0.0;
    for (int r0 = 0; r0 < m; r0 += kEllRows) {
        const int nr = min(kEllRows, m - r0);
        __syncthreads();  // the previous chunk has been read
        for (int e = tid; e < kEllRows * dpad; e += 256) {
            const int rr = e % kEllRows, c = e / kEllRows;  // rows fastest: reads run down a column
            Ub[rr * dpad + c] = (rr < nr && c < d) ? arow(r0 + rr, c) * w[r0 + rr] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kEllTiles; ++q) {
            if (ti[q] < 0) continue;
            const double* ui = Ub + 4 * ti[q];
            const double* uj = Ub + 4 * tj[q];
            for (int rr = 0; rr < nr; ++rr) {
                double a4[4], b4[4];
#pragma unroll
                for (int x = 0; x < 4; ++x) {
                    a4[x] = ui[rr * dpad + x];
                    b4[x] = uj[rr * dpad + x];
                }
#pragma unroll
                for (int x = 0; x < 4; ++x)
#pragma unroll
                    for (int y = 0; y < 4; ++y) acc[q][x][y] = fma(a4[x], b4[y], acc[q][x][y]);
            }
        }
    }
    __syncthreads();  // every chunk read: H may now overwrite them
#pragma unroll
    for (int q = 0; q < kEllTiles; ++q) {
        if (ti[q] < 0) continue;
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                const int i = 4 * ti[q] + x, j = 4 * tj[q] + y;
                if (i < d && j <= i) H[(size_t)i * d + j] = acc[q][x][y];
            }
    }
    __syncthreads();
    __shared__ int singular;
    if (tid == 0) singular = 0;
    for (int j = 0; j < d; ++j) {
        if (tid == 0) {
            double s = H[(size_t)j * d + j];
            for (int q = 0; q < j; ++q) s -= H[(size_t)j * d + q] * H[(size_t)j * d + q];
            if (!(s > 0.0)) singular = 1;
            H[(size_t)j * d + j] = sqrt(s > 0.0 ? s : 1.0);
        }
        __syncthreads();
        const double ljj = H[(size_t)j * d + j];
        for (int i = j + 1 + tid; i < d; i += 256) {
            double s = H[(size_t)i * d + j];
            for (int q = 0; q < j; ++q) s -= H[(size_t)i * d + q] * H[(size_t)j * d + q];
            H[(size_t)i * d + j] = s / ljj;
        }
        __syncthreads();
    }
    if (singular) {
        if (tid == 0) status[v] = NAJA_ERR_INVALID_ARGUMENT;
        return;
    }
    // Column c of M = L^{-1}: M[i][c] (i > c) at H[c][i], M[c][c] at Md[c].
    for (int c = tid; c < d; c += 256) {
        Md[c] = 1.0 / H[(size_t)c * d + c];
        for (int i = c + 1; i < d; ++i) {
            double s = H[(size_t)i * d + c] * Md[c];
            for (int q = c + 1; q < i; ++q) s += H[(size_t)i * d + q] * H[(size_t)c * d + q];
            H[(size_t)c * d + i] = -s / H[(size_t)i * d + i];
        }
    }
    __syncthreads();
    float* Pv = P + (int64_t)v * d * d;
    for (int e = tid; e < d * d; e += 256) {
        const int r = e / d, c = e % d;
        const double val = (c == r) ? Md[r] : (c > r ? H[(size_t)r * d + c] : 0.0);
        Pv[e] = (float)val;
    }
    if (tid == 0) status[v] = NAJA_OK;
}

}  // namespace mlp
}  // namespace naja

extern "C" int naja_precond_ellipsoid_device(const float* A0, const double* b0, int m0, int d, int n_variants,
                                             const int* extra_rows, const float* extra_A, const double* extra_b,
                                             const float* x_c, float* P, void* stream)
{
    return guarded([&]() {
        NAJA_REQUIRE(A0 && b0 && P && m0 > 0 && d > 0 && n_variants > 0 && extra_rows, NAJA_ERR_INVALID_ARGUMENT,
                     "naja_precond_ellipsoid_device: bad arguments");
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
                     "naja_precond_ellipsoid_device: extra rows without extra_A/extra_b");
        const size_t lds = sizeof(double) * (mlp::ellipsoid_h_doubles(d) + d + (size_t)(m0 + kmax));
        NAJA_REQUIRE(lds <= 150 * 1024 && d <= 4 * 38, NAJA_ERR_UNSUPPORTED,
                     "naja_precond_ellipsoid_device: d=%d with %d rows exceeds the LDS budget (use the host routine)",
                     d, m0 + kmax);
        hipStream_t s = (hipStream_t)stream;
        // Offsets, row counts and per-variant status in the device's small
        // buffer, held until the status has been read back.
        const size_t b_off = (sizeof(int64_t) * n_variants + 255) / 256 * 256;
        const size_t b_k = (sizeof(int) * n_variants + 255) / 256 * 256;
        HeldBuf& sb = held_buf();
        std::lock_guard<std::mutex> hold(sb.mu);
        char* mem = held_reserve(sb, b_off + 2 * b_k);
        int64_t* eoff = reinterpret_cast<int64_t*>(mem);
        int* ek = reinterpret_cast<int*>(mem + b_off);
        int* status = reinterpret_cast<int*>(mem + b_off + b_k);
        NAJA_HIP(hipMemcpyAsync(eoff, off.data(), sizeof(int64_t) * n_variants, hipMemcpyHostToDevice, s));
        NAJA_HIP(hipMemcpyAsync(ek, extra_rows, sizeof(int) * n_variants, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(mlp::ellipsoid_kernel, dim3(n_variants), dim3(256), lds, s, A0, b0, m0, d, extra_A, extra_b,
                           eoff, ek, x_c, P, status);
        NAJA_HIP(hipGetLastError());
        int* st = held_host(sb, (size_t)n_variants);
        NAJA_HIP(hipMemcpyAsync(st, status, sizeof(int) * n_variants, hipMemcpyDeviceToHost, s));
        NAJA_HIP(hipStreamSynchronize(s));
        for (int v = 0; v < n_variants; ++v)
            NAJA_REQUIRE(st[v] == NAJA_OK, NAJA_ERR_INVALID_ARGUMENT,
                         "Dikin ellipsoid of variant %d: the centre is not interior or the barrier Hessian is singular",
                         v);
    });
}

// ---------------------------------------------------------------------------
// Learned map on the GPU: P_v = P0_v (I + V_v V_v^T)^{p}, p = +1/2 (expand)
// or -1/2, FP64 inside, one workgroup per variant. With G = V^T V (r x r),
// (I + V V^T)^{p} = I + V M V^T, M = f(G), f(l) = ((1 + l)^p - 1) / l on
// G's eigenvalues, i.e. with S = (I + G)^{1/2}:
//   expand:  M = (I + S)^{-1};   shrink:  M = -S^{-1} (I + S)^{-1}.
// S and S^{-1} come from the coupled Newton-Schulz iteration (Higham,
// Functions of Matrices, 2008, eq. 6.35) on A = (I + G) / c with
// c = ||I + G||_F, so A's eigenvalues lie in (0, 1]:
//   T = (3 I - Z Y) / 2,  Y <- Y T,  Z <- T Z;   Y -> A^{1/2}, Z -> A^{-1/2},
// stopped once ||I - Z Y||_F <= 1e-12 or its quadratic decrease stalls at
// rounding level (a variant that meets neither within kNsMaxIter iterations
// reports NAJA_ERR_RUNTIME). (I + S)^{-1} by
// Gauss-Jordan without pivoting (I + S is SPD with eigenvalues >= 2). Then
// Q = M V^T (r x d), W = P0 V (d x r) and P = P0 + W Q, written as FP32.
// The host routine naja_precond_lowrank forms the same matrix function from
// an eigendecomposition of G.
// ---------------------------------------------------------------------------
namespace naja {
namespace mlp {

constexpr int kLrThreads = 256;
constexpr int kNsMaxIter = 100;

__device__ double lr_block_sum(double x, double* red)
{
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < kLrThreads / 64; ++w) s += red[w];
    return s;
}

// C = A B for r x r row-major LDS matrices; entries e = tid + 256 q go to out[q].
__device__ __forceinline__ void lr_matmul(const double* A, const double* B, int r, double out[4])
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int e = threadIdx.x + kLrThreads * q;
        double s = 0.0;
        if (e < r * r) {
            const int i = e / r, j = e % r;
            for (int t = 0; t < r; ++t) s = fma(A[i * r + t], B[t * r + j], s);
        }
        out[q] = s;
    }
}

__global__ __launch_bounds__(kLrThreads) void lowrank_map_kernel(const float* __restrict__ P0,
                                                                 const float* __restrict__ V, int d, int r, int expand,
                                                                 float* __restrict__ P, int* __restrict__ status)
{
    extern __shared__ double lr_sh[];
    __shared__ double red[kLrThreads / 64];
    const int v = blockIdx.x, tid = threadIdx.x;
    const int rr = r * r;
    double* Vs = lr_sh;            // d x r
    double* M = Vs + d * r;        // r x r
    double* R2 = M + rr;           // max(5 r^2, 2 d r)
    double* Y = R2;
    double* Z = Y + rr;
    double* T = Z + rr;
    double* B = T + rr;            // I + S, then reduced to I
    double* Bi = B + rr;           // (I + S)^{-1}
    const float* Vv = V + (int64_t)v * d * r;
    const float* P0v = P0 + (int64_t)v * d * d;

    for (int e = tid; e < d * r; e += kLrThreads) Vs[e] = (double)Vv[e];
    __syncthreads();
    // A = I + G and its Frobenius norm.
    double fro = 0.0;
    for (int e = tid; e < rr; e += kLrThreads) {
        const int i = e / r, j = e % r;
        double s = (i == j) ? 1.0 : 0.0;
        for (int k = 0; k < d; ++k) s = fma(Vs[k * r + i], Vs[k * r + j], s);
        T[e] = s;
        fro = fma(s, s, fro);
    }
    const double c = sqrt(lr_block_sum(fro, red));
    for (int e = tid; e < rr; e += kLrThreads) {
        Y[e] = T[e] / c;
        Z[e] = (e / r == e % r) ? 1.0 : 0.0;
    }
    __syncthreads();

    bool converged = false;
    double prev = 1.0;
    for (int it = 0; it < kNsMaxIter; ++it) {
        double zy[4];
        lr_matmul(Z, Y, r, zy);
        double err = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + kLrThreads * q;
            if (e < rr) {
                const double dlt = ((e / r == e % r) ? 1.0 : 0.0) - zy[q];
                err = fma(dlt, dlt, err);
                T[e] = 0.5 * (((e / r == e % r) ? 3.0 : 0.0) - zy[q]);
            }
        }
        // The sum ends with a barrier: T is visible to every thread. Stop at
        // 1e-12, or once the quadratic convergence stalls at rounding level
        // (the residual no longer halves after falling below 1e-6).
        const double res = sqrt(lr_block_sum(err, red));
        if (!(res == res) || res > 1e300) break;
        if (res <= 1e-12 || (prev < 1e-6 && res > 0.5 * prev)) {
            converged = true;
            break;
        }
        prev = res;
        double yn[4], zn[4];
        lr_matmul(Y, T, r, yn);
        lr_matmul(T, Z, r, zn);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + kLrThreads * q;
            if (e < rr) {
                Y[e] = yn[q];
                Z[e] = zn[q];
            }
        }
        __syncthreads();
    }
    if (!converged) {
        if (tid == 0) status[v] = NAJA_ERR_RUNTIME;
        return;  // uniform across the workgroup
    }
    // S = sqrt(c) Y, S^{-1} = Z / sqrt(c). Gauss-Jordan on [I + S | I].
    const double sc = sqrt(c);
    for (int e = tid; e < rr; e += kLrThreads) {
        const bool diag = (e / r == e % r);
        B[e] = (diag ? 1.0 : 0.0) + sc * Y[e];
        Bi[e] = diag ? 1.0 : 0.0;
    }
    __syncthreads();
    __shared__ double rowp[64];  // pivot row of B / pivot, then of Bi / pivot
    __shared__ double fcol[32];
    for (int p = 0; p < r; ++p) {
        const double piv = B[p * r + p];
        if (tid < r) {
            rowp[tid] = B[p * r + tid] / piv;
            rowp[r + tid] = Bi[p * r + tid] / piv;
            fcol[tid] = B[tid * r + p];
        }
        __syncthreads();
        for (int e = tid; e < rr; e += kLrThreads) {
            const int i = e / r, j = e % r;
            if (i == p) {
                B[e] = rowp[j];
                Bi[e] = rowp[r + j];
            } else {
                B[e] = fma(-fcol[i], rowp[j], B[e]);
                Bi[e] = fma(-fcol[i], rowp[r + j], Bi[e]);
            }
        }
        __syncthreads();
    }
    if (expand) {
        for (int e = tid; e < rr; e += kLrThreads) M[e] = Bi[e];
    } else {
        double m[4];
        lr_matmul(Z, Bi, r, m);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + kLrThreads * q;
            if (e < rr) M[e] = -m[q] / sc;
        }
    }
    __syncthreads();
    double* Q = R2;          // r x d
    double* W = R2 + r * d;  // d x r
    for (int e = tid; e < r * d; e += kLrThreads) {
        const int t = e / d, b = e % d;
        double s = 0.0;
        for (int u = 0; u < r; ++u) s = fma(M[t * r + u], Vs[b * r + u], s);
        Q[e] = s;
    }
    for (int e = tid; e < d * r; e += kLrThreads) {
        const int a = e / r, t = e % r;
        double s = 0.0;
        for (int k = 0; k < d; ++k) s = fma((double)P0v[(int64_t)a * d + k], Vs[k * r + t], s);
        W[e] = s;
    }
    __syncthreads();
    float* Pv = P + (int64_t)v * d * d;
    for (int e = tid; e < d * d; e += kLrThreads) {
        const int a = e / d, b = e % d;
        double s = (double)P0v[e];
        for (int t = 0; t < r; ++t) s = fma(W[a * r + t], Q[t * d + b], s);
        Pv[e] = (float)s;
    }
    if (tid == 0) status[v] = NAJA_OK;
}

}  // namespace mlp
}  // namespace naja

extern "C" int naja_lowrank_maps_device(const float* P0, const float* V, int n_variants, int d, int rank, int expand,
                                        float* P, void* stream)
{
    return guarded([&]() {
        NAJA_REQUIRE(P0 && V && P && n_variants > 0 && rank > 0 && rank <= d, NAJA_ERR_INVALID_ARGUMENT,
                     "naja_lowrank_maps_device: bad arguments");
        NAJA_REQUIRE(rank <= 32 && d <= 160, NAJA_ERR_UNSUPPORTED,
                     "naja_lowrank_maps_device: rank %d, d %d (rank <= 32, d <= 160)", rank, d);
        NAJA_REQUIRE(P != P0, NAJA_ERR_INVALID_ARGUMENT, "naja_lowrank_maps_device: P must not alias P0");
        const size_t rr = (size_t)rank * rank;
        const size_t lds = sizeof(double) * ((size_t)d * rank + rr + std::max(5 * rr, 2 * (size_t)d * rank));
        hipStream_t s = (hipStream_t)stream;
        HeldBuf& sb = held_buf();
        std::lock_guard<std::mutex> hold(sb.mu);
        int* status = reinterpret_cast<int*>(held_reserve(sb, sizeof(int) * n_variants));
        hipLaunchKernelGGL(mlp::lowrank_map_kernel, dim3(n_variants), dim3(mlp::kLrThreads), lds, s, P0, V, d, rank,
                           expand, P, status);
        NAJA_HIP(hipGetLastError());
        int* st = held_host(sb, (size_t)n_variants);
        NAJA_HIP(hipMemcpyAsync(st, status, sizeof(int) * n_variants, hipMemcpyDeviceToHost, s));
        NAJA_HIP(hipStreamSynchronize(s));
        for (int v = 0; v < n_variants; ++v)
            NAJA_REQUIRE(st[v] == NAJA_OK, NAJA_ERR_RUNTIME,
                         "low-rank map of variant %d: the Newton-Schulz square root did not converge", v);
    });
}
