// Roundings of one polytope of any size on the GPU, the path beyond the
// batched kernels' LDS budgets (genome scale d = 512, large scale d = 2400
// with m = 21200 rows). FP64 throughout.
//
// Dikin ellipsoid (naja_precond_ellipsoid_large):
//
//   s = b - A x_c,  As = diag(1/s) A            (scaling kernel, m x d)
//   H = As^T As                                  (rocBLAS dsyrk: FP64 MFMA)
//   H = L L^T                                    (blocked right-looking Cholesky:
//        64 x 64 diagonal blocks factored in LDS by cholesky_block_kernel,
//        the panel below by dtrsm, the trailing matrix by dsyrk)
//   P = L^{-T}                                   (dtrsm against the identity)
//
// P P^T = H^{-1}: y = x_c + P z whitens the barrier Hessian at x_c. The GEMM
// work (2 m d^2 flops for H, about d^3 for the factor and the inverse) goes to
// rocBLAS as plain library products; the factorisation's sequential part and
// the scaling and layout kernels are this library's.
//
// Facet-aligned rounding (naja_precond_facet_large): the dual rows
// u_r = a_r / s_r taken in order of tightness ||u_r|| (the closest facets
// first), each accepted while its residual against the accepted ones keeps
// at least theta of its norm, until d are accepted; P = N^-1 for the accepted
// rows N. Blocks of 128 candidates: two classical Gram-Schmidt passes against
// the accepted orthonormal basis Q by rocBLAS dgemm (the coefficients kept in
// a lower-triangular R with N = R Q^T), then the selection within the block
// by a Cholesky with rejection on the block's Gram matrix in LDS
// (block_select_kernel), the new directions by dtrsm; finally P = Q R^-1.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <vector>

#include "../../include/naja_amd.h"
#include "common.h"

namespace naja {
namespace {

constexpr int kCholNb = 64;  // diagonal block size of the blocked Cholesky

#define NAJA_ROCBLAS(call)                                                                        \
    do {                                                                                          \
        rocblas_status st_ = (call);                                                              \
        NAJA_REQUIRE(st_ == rocblas_status_success, NAJA_ERR_RUNTIME, "%s failed: %s", #call,     \
                     rocblas_status_to_string(st_));                                              \
    } while (0)

// One rocBLAS handle per device, created on first use and kept; a call holds
// its device's lock while it uses the handle.
struct BlasSlot {
    std::mutex mu;
    rocblas_handle h = nullptr;
};

BlasSlot& blas_slot()
{
    static BlasSlot slots[64];
    int dev = 0;
    NAJA_HIP(hipGetDevice(&dev));
    NAJA_REQUIRE(dev >= 0 && dev < 64, NAJA_ERR_UNSUPPORTED, "device %d", dev);
    return slots[dev];
}

// As[i + j m] = A[i + j m] / s_i with s_i = b_i - a_i . x_c (FP64); a slack
// that is not positive sets *bad.
__global__ void slack_kernel(const float* __restrict__ A, const double* __restrict__ b, int m, int d,
                             const float* __restrict__ xc, double* __restrict__ winv, int* __restrict__ bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double ax = 0.0;
    if (xc)
        for (int j = 0; j < d; ++j) ax = fma((double)A[(int64_t)j * m + i], (double)xc[j], ax);
    const double s = b[i] - ax;
    if (!(s > 0.0)) *bad = 1;
    winv[i] = 1.0 / s;
}

__global__ void scale_rows_kernel(const float* __restrict__ A, const double* __restrict__ winv, int64_t m, int64_t n,
                                  double* __restrict__ As)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m * n; e += (int64_t)gridDim.x * blockDim.x)
        As[e] = (double)A[e] * winv[e % m];
}

// Lower Cholesky factor, in place, of the jb x jb diagonal block at H + j0 (1 + ld)
// (column-major, leading dimension ld), in LDS: column by column, the
// diagonal by one thread, the column below in parallel. A pivot that is not
// positive sets *bad.
__global__ __launch_bounds__(256) void cholesky_block_kernel(double* __restrict__ H, int ld, int j0, int jb,
                                                             int* __restrict__ bad)
{
    __shared__ double L[kCholNb * kCholNb];
    __shared__ int fail;
    const int tid = threadIdx.x;
    double* D = H + (int64_t)j0 * ld + j0;
    for (int e = tid; e < jb * jb; e += 256) L[e] = D[(int64_t)(e / jb) * ld + (e % jb)];  // L[col * jb + row]
    if (tid == 0) fail = 0;
    __syncthreads();
    for (int j = 0; j < jb; ++j) {
        if (tid == 0) {
            double s = L[j * jb + j];
            for (int t = 0; t < j; ++t) s -= L[t * jb + j] * L[t * jb + j];
            if (!(s > 0.0)) fail = 1;
            L[j * jb + j] = sqrt(s > 0.0 ? s : 1.0);
        }
        __syncthreads();
        const double ljj = L[j * jb + j];
        for (int i = j + 1 + tid; i < jb; i += 256) {
            double s = L[j * jb + i];
            for (int t = 0; t < j; ++t) s -= L[t * jb + i] * L[t * jb + j];
            L[j * jb + i] = s / ljj;
        }
        __syncthreads();
    }
    for (int e = tid; e < jb * jb; e += 256) {
        const int c = e / jb, r = e % jb;
        if (r >= c) D[(int64_t)c * ld + r] = L[e];
    }
    if (tid == 0 && fail) *bad = 1;
}

__global__ void identity_kernel(double* __restrict__ X, int d)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < (int64_t)d * d;
         e += (int64_t)gridDim.x * blockDim.x)
        X[e] = (e % d == e / d) ? 1.0 : 0.0;
}

// P (row-major FP32) from X (column-major FP64): P[i][j] = X[i + j d].
__global__ void to_row_major_kernel(const double* __restrict__ X, int d, float* __restrict__ P)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < (int64_t)d * d;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / d, j = e % d;
        P[e] = (float)X[j * d + i];
    }
}


// Tightness of each row as a facet candidate: ||a_r|| / s_r (winv = 1/s).
__global__ void tightness_kernel(const float* __restrict__ A, const double* __restrict__ winv, int m, int d,
                                 double* __restrict__ tight)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double n2 = 0.0;
    for (int j = 0; j < d; ++j) {
        const double a = (double)A[(int64_t)j * m + i];
        n2 = fma(a, a, n2);
    }
    tight[i] = sqrt(n2) * winv[i];
}

// U[:, j] = a_{rows[j]} / s_{rows[j]} (d x B column-major) for j < nb.
__global__ void gather_rows_kernel(const float* __restrict__ A, const double* __restrict__ winv, int m, int d,
                                   const int* __restrict__ rows, int nb, double* __restrict__ U)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < (int64_t)nb * d;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(e / d), c = (int)(e % d);
        const int r = rows[j];
        U[e] = (double)A[(int64_t)c * m + r] * winv[r];
    }
}

// Greedy selection within a block from its Gram matrix G = U^T U (nb x nb,
// column-major, U already orthogonal to the accepted basis): a right-looking
// Cholesky that skips candidate j when its residual norm^2 G_jj falls below
// theta^2 n0_j^2 (n0_j its norm before any projection), and stops once
// `room` candidates are accepted. Outputs: acc (count), pos[p] (block index of
// the p-th accepted), Lb (nb x nb, column p = new direction p): Lb[i + p nb]
// is the coefficient of candidate i on direction p (zero for i before pos[p]).
__global__ __launch_bounds__(1024) void block_select_kernel(const double* __restrict__ Gg, int nb,
                                                            const double* __restrict__ n0, double theta2, int room,
                                                            int* __restrict__ acc_out, int* __restrict__ pos,
                                                            double* __restrict__ Lb)
{
    extern __shared__ double G[];  // nb x nb, column-major
    __shared__ int acc;
    __shared__ int accepted_j;
    const int tid = threadIdx.x;
    for (int e = tid; e < nb * nb; e += blockDim.x) G[e] = Gg[e];
    for (int e = tid; e < nb * nb; e += blockDim.x) Lb[e] = 0.0;
    if (tid == 0) acc = 0;
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        __syncthreads();  // every thread has read the previous decision
        if (tid == 0) {
            const double r2 = G[j * nb + j];
            accepted_j = (acc < room && r2 > 0.0 && r2 >= theta2 * n0[j] * n0[j]) ? 1 : 0;
        }
        __syncthreads();
        if (!accepted_j) continue;  // uniform
        const int p = acc;
        const double ljj = sqrt(G[j * nb + j]);
        // Column p of the factor over the later candidates: l_i = G_ij / ljj.
        for (int i = j + 1 + tid; i < nb; i += blockDim.x) G[j * nb + i] /= ljj;  // G[i][j] stored at j*nb+i
        __syncthreads();
        // Trailing update G_ik -= l_i l_k (k <= i, both after j).
        const int rest = nb - j - 1;
        for (int e = tid; e < rest * rest; e += blockDim.x) {
            const int i = j + 1 + e / rest, kk = j + 1 + e % rest;
            if (kk <= i) G[kk * nb + i] -= G[j * nb + i] * G[j * nb + kk];
        }
        // Record: new direction p has coefficient ljj on candidate j and l_i on later i.
        if (tid == 0) {
            pos[p] = j;
            Lb[(int64_t)p * nb + j] = ljj;
        }
        for (int i = j + 1 + tid; i < nb; i += blockDim.x) Lb[(int64_t)p * nb + i] = G[j * nb + i];
        __syncthreads();
        if (tid == 0) acc = p + 1;
        __syncthreads();
    }
    if (tid == 0) *acc_out = acc;
}

// Gathers the accepted columns of U (d x nb) into Qb (d x kb) and the lower
// triangular block factor Lsel (kb x kb, column-major) with
// Lsel[q + p kb] = Lb[pos[q] + p nb] (q >= p): U[:, pos] = Qb Lsel^T after
// the solve Qb = U[:, pos] Lsel^{-T}.
__global__ void gather_accepted_kernel(const double* __restrict__ U, const double* __restrict__ Lb,
                                       const int* __restrict__ pos, int d, int nb, int kb, double* __restrict__ Qb,
                                       double* __restrict__ Lsel)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < (int64_t)kb * d;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int p = (int)(e / d), c = (int)(e % d);
        Qb[e] = U[(int64_t)pos[p] * d + c];
    }
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < (int64_t)kb * kb;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int p = (int)(e / kb), q = (int)(e % kb);
        Lsel[e] = q >= p ? Lb[(int64_t)p * nb + pos[q]] : 0.0;
    }
}

// Rows k0 .. k0 + kb - 1 of the coefficient matrix R (d x d, column-major,
// lower triangular): R[k0 + q, i] = C[i, pos[q]] (i < k0; the coefficients on
// the earlier directions) and R[k0 + q, k0 + p] = Lsel[q + p kb].
__global__ void record_coeffs_kernel(const double* __restrict__ C, int k0, int nb, const int* __restrict__ pos,
                                     const double* __restrict__ Lsel, int kb, int d, double* __restrict__ R)
{
    const int64_t tot = (int64_t)kb * (k0 + kb);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        const int q = (int)(e % kb), i = (int)(e / kb);
        double v;
        if (i < k0)
            v = C[(int64_t)pos[q] * k0 + i];  // C is k0 x nb column-major
        else
            v = Lsel[q + (int64_t)(i - k0) * kb];
        R[(int64_t)i * d + k0 + q] = v;
    }
}

__global__ void add_kernel(double* __restrict__ x, const double* __restrict__ y, int64_t n)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        x[e] += y[e];
}


// H = L L^T in place (lower triangle of the column-major d x d matrix H):
// blocked right-looking, 64 x 64 diagonal blocks factored in LDS, the panel
// below by dtrsm, the trailing matrix by dsyrk. A pivot that is not positive
// sets *bad.
void blocked_cholesky(rocblas_handle h, double* H, int d, int* bad, hipStream_t s)
{
    const double one = 1.0, minus_one = -1.0;
    for (int j0 = 0; j0 < d; j0 += kCholNb) {
        const int jb = std::min(kCholNb, d - j0);
        hipLaunchKernelGGL(cholesky_block_kernel, dim3(1), dim3(256), 0, s, H, d, j0, jb, bad);
        NAJA_HIP(hipGetLastError());
        const int rest = d - j0 - jb;
        if (rest > 0) {
            double* L11 = H + (int64_t)j0 * d + j0;
            double* L21 = H + (int64_t)j0 * d + j0 + jb;
            double* H22 = H + (int64_t)(j0 + jb) * d + j0 + jb;
            NAJA_ROCBLAS(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                                       rocblas_diagonal_non_unit, rest, jb, &one, L11, d, L21, d));
            NAJA_ROCBLAS(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_none, rest, jb, &minus_one, L21, d,
                                       &one, H22, d));
        }
    }
}

// x <- x + t dx with t from the Newton decrement lambda^2 = -g . dx (on the
// device, read by one thread): the damped step 1 / (1 + lambda) keeps x
// inside the polytope for the self-concordant log barrier; full steps once
// lambda < 1/4.
__global__ void newton_update_kernel(const double* __restrict__ g, const double* __restrict__ dx, int d,
                                     double* __restrict__ x, float* __restrict__ xf, double* __restrict__ lam2_out)
{
    __shared__ double part[256];
    __shared__ double tstep;
    double acc = 0.0;
    for (int j = threadIdx.x; j < d; j += 256) acc = fma(-g[j], dx[j], acc);
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double lam2 = part[0] > 0.0 ? part[0] : 0.0;
        *lam2_out = lam2;
        tstep = lam2 < 0.0625 ? 1.0 : 1.0 / (1.0 + sqrt(lam2));
    }
    __syncthreads();
    for (int j = threadIdx.x; j < d; j += 256) {
        x[j] = fma(tstep, dx[j], x[j]);
        xf[j] = (float)x[j];
    }
}

// 1/s for s = b - A x at an FP64 iterate x (the FP32 copy would put a floor
// of FP32 rounding under the Newton decrement).
__global__ void slack64_kernel(const float* __restrict__ A, const double* __restrict__ b, int m, int d,
                               const double* __restrict__ x, double* __restrict__ winv, int* __restrict__ bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double ax = 0.0;
    for (int j = 0; j < d; ++j) ax = fma((double)A[(int64_t)j * m + i], x[j], ax);
    const double s = b[i] - ax;
    if (!(s > 0.0)) *bad = 1;
    winv[i] = 1.0 / s;
}

__global__ void fill_kernel(double* __restrict__ v, double value, int64_t n)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        v[e] = value;
}

__global__ void widen_kernel(const float* __restrict__ src, double* __restrict__ x, float* __restrict__ xf, int d)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < d) {
        x[j] = src ? (double)src[j] : 0.0;
        xf[j] = src ? src[j] : 0.0f;
    }
}

unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 65535); }

}  // namespace
}  // namespace naja

extern "C" int naja_precond_ellipsoid_large(const float* A, const double* b, int m, int d, const float* x_c,
                                            float* P, void* stream)
{
    using namespace naja;
    return guarded([&]() {
        NAJA_REQUIRE(A && b && P && m > 0 && d > 0, NAJA_ERR_INVALID_ARGUMENT,
                     "naja_precond_ellipsoid_large: bad arguments");
        NAJA_REQUIRE((int64_t)m * d < ((int64_t)1 << 31), NAJA_ERR_UNSUPPORTED,
                     "naja_precond_ellipsoid_large: m * d must stay below 2^31");
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        const size_t b_w = ((size_t)m * sizeof(double) + 255) / 256 * 256;
        const size_t b_as = (size_t)m * d * sizeof(double);
        const size_t b_h = (size_t)d * d * sizeof(double);
        char* mem = nullptr;
        scratch_alloc((void**)&mem, 256 + b_w + b_as + 2 * b_h, s);
        int* bad = reinterpret_cast<int*>(mem);
        double* winv = reinterpret_cast<double*>(mem + 256);
        double* As = reinterpret_cast<double*>(mem + 256 + b_w);
        double* H = reinterpret_cast<double*>(mem + 256 + b_w + b_as);
        double* X = H + (size_t)d * d;
        NAJA_HIP(hipMemsetAsync(bad, 0, sizeof(int) * 2, s));
        hipLaunchKernelGGL(slack_kernel, dim3((m + 255) / 256), dim3(256), 0, s, A, b, m, d, x_c, winv, bad);
        NAJA_HIP(hipGetLastError());
        hipLaunchKernelGGL(scale_rows_kernel, dim3(grid_for((int64_t)m * d)), dim3(256), 0, s, A, winv, (int64_t)m,
                           (int64_t)d, As);
        NAJA_HIP(hipGetLastError());

        BlasSlot& slot = blas_slot();
        std::lock_guard<std::mutex> hold(slot.mu);
        if (!slot.h) NAJA_ROCBLAS(rocblas_create_handle(&slot.h));
        rocblas_handle h = slot.h;
        NAJA_ROCBLAS(rocblas_set_stream(h, s));
        NAJA_ROCBLAS(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host));
        const double one = 1.0, zero = 0.0;
        // H = As^T As (lower triangle).
        NAJA_ROCBLAS(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_transpose, d, m, &one, As, m, &zero, H, d));
        blocked_cholesky(h, H, d, bad + 1, s);
        // X = L^{-T}: solve L^T X = I.
        hipLaunchKernelGGL(identity_kernel, dim3(grid_for((int64_t)d * d)), dim3(256), 0, s, X, d);
        NAJA_HIP(hipGetLastError());
        NAJA_ROCBLAS(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_transpose,
                                   rocblas_diagonal_non_unit, d, d, &one, H, d, X, d));
        hipLaunchKernelGGL(to_row_major_kernel, dim3(grid_for((int64_t)d * d)), dim3(256), 0, s, X, d, P);
        NAJA_HIP(hipGetLastError());
        int flags[2] = {0, 0};
        NAJA_HIP(hipMemcpyAsync(flags, bad, sizeof(flags), hipMemcpyDeviceToHost, s));
        scratch_free(mem, s);
        NAJA_HIP(hipStreamSynchronize(s));
        NAJA_REQUIRE(flags[0] == 0, NAJA_ERR_INVALID_ARGUMENT, "Dikin ellipsoid: the centre is not interior");
        NAJA_REQUIRE(flags[1] == 0, NAJA_ERR_RUNTIME, "Dikin ellipsoid: the barrier Hessian is not positive definite");
    });
}

extern "C" int naja_precond_facet_large(const float* A, const double* b, int m, int d, const float* x_c, float theta,
                                        float* P, void* stream)
{
    using namespace naja;
    return guarded([&]() {
        NAJA_REQUIRE(A && b && P && m > 0 && d > 0 && theta >= 0.0f, NAJA_ERR_INVALID_ARGUMENT,
                     "naja_precond_facet_large: bad arguments");
        NAJA_REQUIRE((int64_t)m * d < ((int64_t)1 << 31) && m >= d, NAJA_ERR_UNSUPPORTED,
                     "naja_precond_facet_large: needs d <= m and m * d < 2^31");
        constexpr int B = 128;  // candidates per block
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        auto up = [](size_t x) { return (x + 255) / 256 * 256; };
        const size_t b_w = up(sizeof(double) * m), b_t = up(sizeof(double) * m), b_o = up(sizeof(int) * m);
        const size_t b_dd = up(sizeof(double) * (size_t)d * d), b_db = up(sizeof(double) * (size_t)d * B);
        const size_t b_bb = up(sizeof(double) * B * B), b_pos = up(sizeof(int) * B);
        char* mem = nullptr;
        scratch_alloc((void**)&mem, 256 + b_w + 2 * b_t + b_o + 2 * b_dd + 3 * b_db + 2 * b_bb + b_pos, s);
        char* cur = mem;
        auto take = [&](size_t bytes) {
            char* p = cur;
            cur += bytes;
            return p;
        };
        int* flags = reinterpret_cast<int*>(take(256));  // [0] exterior centre, [1] accepted count
        double* winv = reinterpret_cast<double*>(take(b_w));
        double* tight = reinterpret_cast<double*>(take(b_t));
        double* tight_sorted = reinterpret_cast<double*>(take(b_t));
        int* order = reinterpret_cast<int*>(take(b_o));
        double* Q = reinterpret_cast<double*>(take(b_dd));
        double* R = reinterpret_cast<double*>(take(b_dd));
        double* U = reinterpret_cast<double*>(take(b_db));
        double* C = reinterpret_cast<double*>(take(b_db));
        double* T = reinterpret_cast<double*>(take(b_db));
        double* G = reinterpret_cast<double*>(take(b_bb));
        double* Lb = reinterpret_cast<double*>(take(b_bb));
        int* pos = reinterpret_cast<int*>(take(b_pos));
        double* Lsel = G;  // G is consumed by the selection before Lsel is formed

        NAJA_HIP(hipMemsetAsync(flags, 0, 256, s));
        hipLaunchKernelGGL(slack_kernel, dim3((m + 255) / 256), dim3(256), 0, s, A, b, m, d, x_c, winv, flags);
        NAJA_HIP(hipGetLastError());
        hipLaunchKernelGGL(tightness_kernel, dim3((m + 255) / 256), dim3(256), 0, s, A, winv, m, d, tight);
        NAJA_HIP(hipGetLastError());
        std::vector<double> th((size_t)m);
        int bad = 0;
        NAJA_HIP(hipMemcpyAsync(th.data(), tight, sizeof(double) * m, hipMemcpyDeviceToHost, s));
        NAJA_HIP(hipMemcpyAsync(&bad, flags, sizeof(int), hipMemcpyDeviceToHost, s));
        NAJA_HIP(hipStreamSynchronize(s));
        if (bad) {
            scratch_free(mem, s);
            NAJA_REQUIRE(false, NAJA_ERR_INVALID_ARGUMENT, "facet rounding: the centre is not interior");
        }
        // Candidates in order of tightness (ties by row index).
        std::vector<int> ord((size_t)m);
        for (int i = 0; i < m; ++i) ord[(size_t)i] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int c) { return th[(size_t)a] > th[(size_t)c]; });
        std::vector<double> ths((size_t)m);
        for (int i = 0; i < m; ++i) ths[(size_t)i] = th[(size_t)ord[(size_t)i]];
        NAJA_HIP(hipMemcpyAsync(order, ord.data(), sizeof(int) * m, hipMemcpyHostToDevice, s));
        NAJA_HIP(hipMemcpyAsync(tight_sorted, ths.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
        NAJA_HIP(hipMemsetAsync(R, 0, sizeof(double) * (size_t)d * d, s));

        BlasSlot& slot = blas_slot();
        std::lock_guard<std::mutex> hold(slot.mu);
        if (!slot.h) NAJA_ROCBLAS(rocblas_create_handle(&slot.h));
        rocblas_handle h = slot.h;
        NAJA_ROCBLAS(rocblas_set_stream(h, s));
        NAJA_ROCBLAS(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host));
        const double one = 1.0, zero = 0.0, minus_one = -1.0;
        const double theta2 = (double)theta * theta;
        int k = 0;
        for (int start = 0; start < m && k < d; start += B) {
            const int nb = std::min(B, m - start);
            hipLaunchKernelGGL(gather_rows_kernel, dim3(grid_for((int64_t)nb * d)), dim3(256), 0, s, A, winv, m, d,
                               order + start, nb, U);
            NAJA_HIP(hipGetLastError());
            if (k > 0) {
                // Two classical Gram-Schmidt passes against the accepted basis
                // Q[:, :k]; C accumulates the coefficients.
                NAJA_ROCBLAS(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, k, nb, d, &one, Q, d,
                                           U, d, &zero, C, k));
                NAJA_ROCBLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, d, nb, k, &minus_one, Q,
                                           d, C, k, &one, U, d));
                NAJA_ROCBLAS(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, k, nb, d, &one, Q, d,
                                           U, d, &zero, T, k));
                NAJA_ROCBLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, d, nb, k, &minus_one, Q,
                                           d, T, k, &one, U, d));
                hipLaunchKernelGGL(add_kernel, dim3(grid_for((int64_t)k * nb)), dim3(256), 0, s, C, T, (int64_t)k * nb);
                NAJA_HIP(hipGetLastError());
            }
            NAJA_ROCBLAS(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_transpose, nb, d, &one, U, d, &zero, G,
                                       nb));
            hipLaunchKernelGGL(block_select_kernel, dim3(1), dim3(1024), sizeof(double) * nb * nb, s, G, nb,
                               tight_sorted + start, theta2, d - k, flags + 1, pos, Lb);
            NAJA_HIP(hipGetLastError());
            int kb = 0;
            NAJA_HIP(hipMemcpyAsync(&kb, flags + 1, sizeof(int), hipMemcpyDeviceToHost, s));
            NAJA_HIP(hipStreamSynchronize(s));
            NAJA_REQUIRE(kb >= 0 && kb <= nb && k + kb <= d, NAJA_ERR_RUNTIME, "facet selection: bad block count %d",
                         kb);
            if (kb == 0) continue;
            double* Qb = Q + (int64_t)k * d;
            hipLaunchKernelGGL(gather_accepted_kernel, dim3(grid_for((int64_t)kb * d)), dim3(256), 0, s, U, Lb, pos,
                               d, nb, kb, Qb, Lsel);
            NAJA_HIP(hipGetLastError());
            NAJA_ROCBLAS(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                                       rocblas_diagonal_non_unit, d, kb, &one, Lsel, kb, Qb, d));
            hipLaunchKernelGGL(record_coeffs_kernel, dim3(grid_for((int64_t)kb * (k + kb))), dim3(256), 0, s, C, k, nb,
                               pos, Lsel, kb, d, R);
            NAJA_HIP(hipGetLastError());
            k += kb;
        }
        if (k < d) {
            scratch_free(mem, s);
            NAJA_REQUIRE(false, NAJA_ERR_RUNTIME,
                         "facet rounding: only %d independent facets for d = %d (use a base rounding to complete)", k, d);
        }
        // P = N^{-1} = Q R^{-1} for N = R Q^T (the accepted dual rows).
        NAJA_ROCBLAS(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_none,
                                   rocblas_diagonal_non_unit, d, d, &one, R, d, Q, d));
        hipLaunchKernelGGL(to_row_major_kernel, dim3(grid_for((int64_t)d * d)), dim3(256), 0, s, Q, d, P);
        NAJA_HIP(hipGetLastError());
        scratch_free(mem, s);
        NAJA_HIP(hipStreamSynchronize(s));
    });
}

extern "C" int naja_analytic_center(const float* A, const double* b, int m, int d, const float* x0, int max_iter,
                                    double tol, float* x_out, int* iters_out, void* stream)
{
    using namespace naja;
    return guarded([&]() {
        NAJA_REQUIRE(A && b && x_out && m > 0 && d > 0 && max_iter > 0 && tol > 0.0, NAJA_ERR_INVALID_ARGUMENT,
                     "naja_analytic_center: bad arguments");
        NAJA_REQUIRE((int64_t)m * d < ((int64_t)1 << 31) && m > d, NAJA_ERR_UNSUPPORTED,
                     "naja_analytic_center: needs m > d and m * d < 2^31");
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        char* mem = nullptr;
        try {
            auto up = [](size_t x) { return (x + 255) / 256 * 256; };
            const size_t b_w = up(sizeof(double) * m), b_as = up(sizeof(double) * (size_t)m * d);
            const size_t b_h = up(sizeof(double) * (size_t)d * d), b_d = up(sizeof(double) * d);
            scratch_alloc((void**)&mem, 256 + 2 * b_w + b_as + b_h + 3 * b_d + up(sizeof(float) * d), s);
            char* cur = mem;
            auto take = [&](size_t bytes) {
                char* p = cur;
                cur += bytes;
                return p;
            };
            int* flags = reinterpret_cast<int*>(take(128));
            double* lam2 = reinterpret_cast<double*>(take(128));
            double* winv = reinterpret_cast<double*>(take(b_w));
            double* ones = reinterpret_cast<double*>(take(b_w));
            double* As = reinterpret_cast<double*>(take(b_as));
            double* H = reinterpret_cast<double*>(take(b_h));
            double* g = reinterpret_cast<double*>(take(b_d));
            double* dx = reinterpret_cast<double*>(take(b_d));
            double* x = reinterpret_cast<double*>(take(b_d));
            float* xf = reinterpret_cast<float*>(take(up(sizeof(float) * d)));
            NAJA_HIP(hipMemsetAsync(flags, 0, 128, s));
            hipLaunchKernelGGL(fill_kernel, dim3(grid_for(m)), dim3(256), 0, s, ones, 1.0, (int64_t)m);
            hipLaunchKernelGGL(widen_kernel, dim3((d + 255) / 256), dim3(256), 0, s, x0, x, xf, d);
            NAJA_HIP(hipGetLastError());
            BlasSlot& slot = blas_slot();
            std::lock_guard<std::mutex> hold(slot.mu);
            if (!slot.h) NAJA_ROCBLAS(rocblas_create_handle(&slot.h));
            rocblas_handle h = slot.h;
            NAJA_ROCBLAS(rocblas_set_stream(h, s));
            NAJA_ROCBLAS(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host));
            const double one = 1.0, zero = 0.0, minus_one = -1.0;
            int it = 0;
            double l2 = 1.0;
            int f[2] = {0, 0};
            for (; it < max_iter; ++it) {
                // s = b - A x, gradient g = A^T (1/s), Hessian H = A^T diag(1/s^2) A.
                hipLaunchKernelGGL(slack64_kernel, dim3((m + 255) / 256), dim3(256), 0, s, A, b, m, d, x, winv, flags);
                hipLaunchKernelGGL(scale_rows_kernel, dim3(grid_for((int64_t)m * d)), dim3(256), 0, s, A, winv,
                                   (int64_t)m, (int64_t)d, As);
                NAJA_HIP(hipGetLastError());
                NAJA_ROCBLAS(rocblas_dgemv(h, rocblas_operation_transpose, m, d, &one, As, m, ones, 1, &zero, g, 1));
                NAJA_ROCBLAS(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_transpose, d, m, &one, As, m, &zero,
                                           H, d));
                blocked_cholesky(h, H, d, flags + 1, s);
                // dx = -H^{-1} g by two triangular solves.
                NAJA_ROCBLAS(rocblas_dcopy(h, d, g, 1, dx, 1));
                NAJA_ROCBLAS(rocblas_dscal(h, d, &minus_one, dx, 1));
                NAJA_ROCBLAS(rocblas_dtrsv(h, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_non_unit, d,
                                           H, d, dx, 1));
                NAJA_ROCBLAS(rocblas_dtrsv(h, rocblas_fill_lower, rocblas_operation_transpose,
                                           rocblas_diagonal_non_unit, d, H, d, dx, 1));
                hipLaunchKernelGGL(newton_update_kernel, dim3(1), dim3(256), 0, s, g, dx, d, x, xf, lam2);
                NAJA_HIP(hipGetLastError());
                NAJA_HIP(hipMemcpyAsync(&l2, lam2, sizeof(double), hipMemcpyDeviceToHost, s));
                NAJA_HIP(hipMemcpyAsync(f, flags, sizeof(f), hipMemcpyDeviceToHost, s));
                NAJA_HIP(hipStreamSynchronize(s));
                NAJA_REQUIRE(f[0] == 0, NAJA_ERR_INVALID_ARGUMENT,
                             "analytic centre: the start (or an iterate) is not interior");
                NAJA_REQUIRE(f[1] == 0, NAJA_ERR_RUNTIME, "analytic centre: the barrier Hessian is not positive definite");
                if (l2 < tol) {
                    ++it;
                    break;
                }
            }
            NAJA_HIP(hipMemcpyAsync(x_out, xf, sizeof(float) * d, hipMemcpyDeviceToDevice, s));
            scratch_free(mem, s);
            mem = nullptr;
            NAJA_HIP(hipStreamSynchronize(s));
            if (iters_out) *iters_out = it;
            NAJA_REQUIRE(l2 < tol, NAJA_ERR_RUNTIME, "analytic centre: Newton decrement %.3e after %d steps", l2, it);
        } catch (...) {
            if (mem) scratch_free(mem, s);
            (void)hipStreamSynchronize(s);
            throw;
        }
    });
}
