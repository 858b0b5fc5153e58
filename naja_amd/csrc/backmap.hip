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

#include <algorithm>
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

// ---------------------------------------------------------------------------
// Small operands (n <= 128, d <= 128, d % 4 == 0): the rounding backmap of
// batched variants, y = c + P z with n = d = 100 for the learned-rounded
// batch. A persistent workgroup takes a run of 64-row sample tiles of one
// batch item and keeps T in LDS for all of them (about 70 KB of LDS for
// n = d = 100, so two workgroups share a CU and one's tile loads and
// epilogue overlap the other's MFMAs). Each wave owns 16 rows x 16 NB
// columns as NB blocks of v_mfma_f32_16x16x4f32; K runs in groups
// of 16 where lane group g = lane / 16 takes k = kk + 4 g .. kk + 4 g + 3
// over four consecutive MFMAs, so each operand is one 16-byte LDS read
// (A and B use the same k order, so the sums are unchanged). The LDS row
// stride S >= d has S % 8 == 4: the 8 lanes of a 16-byte read phase then
// hit distinct bank groups. The next tile's rows are loaded into registers
// while the current tile's MFMAs run; results go back through LDS so every
// tile's output block V[s0 .. s0 + 64) is written as one contiguous run. A K
// tail (d % 16 != 0) takes one MFMA per 4 k instead of padding K to 16.
// ---------------------------------------------------------------------------
constexpr int kSmRows = 64;
constexpr int kSmIB = kSmRows / 64;  // 16-row blocks per wave (four waves per tile)
constexpr int kSmMaxN = 128;
constexpr int kSmMaxD = 128;
constexpr int kSmPer = kSmRows * kSmMaxD / 4 / 256;  // float4 loads per thread per tile (<= 16)

__host__ __device__ inline int small_stride(int d)
{
    const int s = (d + 3) / 4 * 4;
    return (s % 8 == 4) ? s : s + 4;
}

typedef float floatx4 __attribute__((ext_vector_type(4)));

// Y and V may be the same array (in place, n == d): no __restrict__ on them.
// ser (optional): the same outputs also in the series layout of the
// diagnostics, ser[((z nchains + c) n + j) spcp + s] = V_z[s nchains + c][j]
// (row s nchains + c of a batch item is sample s of chain c), written from
// the LDS output tile, so no separate transpose has to read V again.
template <int NB>
__global__ __launch_bounds__(256) void backmap_small_kernel(const float* __restrict__ T,
                                                           const float* __restrict__ shift,
                                                           const float* Y, float* V, int n,
                                                           int d, int64_t ns, int tiles_per_wg, int64_t sT,
                                                           int64_t sShift, int64_t sY, int64_t sV,
                                                           float* __restrict__ ser, int nchains, int spcp)
{
    extern __shared__ float lds[];
    const int S = small_stride(d);
    float* Ts = lds;                              // 16 NB rows x S (rows >= n zero)
    float* Ys = lds + 16 * NB * S;                // kSmRows x S, then the output tile (kSmRows x n)
    T += sT * blockIdx.z;
    shift += sShift * blockIdx.z;
    Y += sY * blockIdx.z;
    V += sV * blockIdx.z;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t n_tiles = (ns + kSmRows - 1) / kSmRows;
    const int64_t t_begin = (int64_t)blockIdx.x * tiles_per_wg;
    const int64_t t_end = t_begin + tiles_per_wg < n_tiles ? t_begin + tiles_per_wg : n_tiles;
    if (t_begin >= t_end) return;  // uniform

    // T into LDS (rows >= n and columns >= d zero).
    for (int e = tid; e < 16 * NB * S; e += 256) {
        const int j = e / S, k = e % S;
        Ts[e] = (j < n && k < d) ? T[(int64_t)j * d + k] : 0.0f;
    }
    float sh[NB];
#pragma unroll
    for (int jb = 0; jb < NB; ++jb) {
        const int col = 16 * jb + (lane & 15);
        sh[jb] = col < n ? shift[col] : 0.0f;
    }

    // A tile of rows [s0, s0 + rows) is the contiguous float run
    // Y[s0 d .. (s0 + rows) d): thread tid loads float4 q = tid + 256 i.
    const int q_tile = kSmRows * d / 4;
    auto load_tile = [&](int64_t t, float4 (&r)[kSmPer]) {
        const int64_t s0 = t * kSmRows;
        const int64_t rows = ns - s0 < kSmRows ? ns - s0 : kSmRows;
        const int q_valid = (int)(rows * d / 4);
        const float4* src = reinterpret_cast<const float4*>(Y + s0 * d);
#pragma unroll
        for (int i = 0; i < kSmPer; ++i) {
            const int q = tid + 256 * i;
            r[i] = (q < q_valid) ? src[q] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    };
    float4 next[kSmPer];
    load_tile(t_begin, next);

    for (int64_t t = t_begin; t < t_end; ++t) {
        __syncthreads();  // the previous output tile has been written out
#pragma unroll
        for (int i = 0; i < kSmPer; ++i) {
            const int q = tid + 256 * i;
            if (q < q_tile) {
                const int e = 4 * q;
                const int li = (S == d) ? e : (e / d) * S + e % d;  // d % 4 == 0: a float4 stays in one row
                *reinterpret_cast<float4*>(Ys + li) = next[i];
            }
        }
        __syncthreads();
        if (t + 1 < t_end) load_tile(t + 1, next);  // in flight during the MFMAs

        floatx4 acc[kSmIB][NB];
#pragma unroll
        for (int ib = 0; ib < kSmIB; ++ib)
#pragma unroll
            for (int jb = 0; jb < NB; ++jb) acc[ib][jb] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        const int g = lane >> 4, r16 = lane & 15;
        const float* ya = Ys + (16 * kSmIB * wave + r16) * S + 4 * g;
        const float* tb = Ts + r16 * S + 4 * g;
        int kk = 0;
        for (; kk + 16 <= d; kk += 16) {
            float4 a[kSmIB], b[NB];
#pragma unroll
            for (int ib = 0; ib < kSmIB; ++ib) a[ib] = *reinterpret_cast<const float4*>(ya + ib * 16 * S + kk);
#pragma unroll
            for (int jb = 0; jb < NB; ++jb) b[jb] = *reinterpret_cast<const float4*>(tb + jb * 16 * S + kk);
#pragma unroll
            for (int ib = 0; ib < kSmIB; ++ib)
#pragma unroll
                for (int jb = 0; jb < NB; ++jb) {
                    acc[ib][jb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ib].x, b[jb].x, acc[ib][jb], 0, 0, 0);
                    acc[ib][jb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ib].y, b[jb].y, acc[ib][jb], 0, 0, 0);
                    acc[ib][jb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ib].z, b[jb].z, acc[ib][jb], 0, 0, 0);
                    acc[ib][jb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ib].w, b[jb].w, acc[ib][jb], 0, 0, 0);
                }
        }
        // K tail (d % 16 = 4, 8 or 12; d % 4 = 0): one MFMA per 4 k, lane
        // group g taking k = kk + g, so no MFMA multiplies padding.
        for (; kk < d; kk += 4) {
            float a[kSmIB], b[NB];
#pragma unroll
            for (int ib = 0; ib < kSmIB; ++ib) a[ib] = ya[ib * 16 * S + kk - 3 * g];
#pragma unroll
            for (int jb = 0; jb < NB; ++jb) b[jb] = tb[jb * 16 * S + kk - 3 * g];
#pragma unroll
            for (int ib = 0; ib < kSmIB; ++ib)
#pragma unroll
                for (int jb = 0; jb < NB; ++jb)
                    acc[ib][jb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ib], b[jb], acc[ib][jb], 0, 0, 0);
        }
        __syncthreads();  // every wave is done reading Ys
        // D layout of a 16 x 16 block: lane holds column lane % 16, register
        // e holds row 4 (lane / 16) + e. Output tile kSmRows x n, row-major.
        float* Os = Ys;
#pragma unroll
        for (int ib = 0; ib < kSmIB; ++ib)
#pragma unroll
            for (int jb = 0; jb < NB; ++jb) {
                const int col = 16 * jb + r16;
                if (col < n) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int row = 16 * kSmIB * wave + 16 * ib + 4 * g + e;
                        Os[row * n + col] = acc[ib][jb][e] + sh[jb];
                    }
                }
            }
        __syncthreads();
        const int64_t s0 = t * kSmRows;
        const int64_t rows = ns - s0 < kSmRows ? ns - s0 : kSmRows;
        const int len = (int)(rows * n);
        float* dst = V + s0 * n;
        if ((n & 3) == 0) {
            for (int q = tid; q < len / 4; q += 256)
                reinterpret_cast<float4*>(dst)[q] = reinterpret_cast<const float4*>(Os)[q];
        } else {
            for (int e = tid; e < len; e += 256) dst[e] = Os[e];
        }
        if (ser) {
            float* sz = ser + (int64_t)blockIdx.z * nchains * n * spcp;
            if (rows == kSmRows && (nchains & (nchains - 1)) == 0 && nchains <= kSmRows / 4) {
                // A whole tile, a power-of-two chain count up to 16: spt = 64 /
                // nchains samples of every chain, in spt / 4 runs of four. Thread
                // tid takes run g of series (c, j0) and then the same run of
                // columns j0 + 16, j0 + 32, ... (256 threads cover 64 series,
                // a multiple of nchains, so c stays): the four samples are LDS
                // rows nchains apart, stored as one 16-byte streaming write
                // (rows of ser are 256-byte aligned, sb a multiple of 4). The
                // addresses advance by constants, so the loop forms no index.
                const int lc = __builtin_ctz((unsigned)nchains);
                const int lg = __builtin_ctz((unsigned)(kSmRows / nchains)) - 2;  // log2(runs per series)
                const int g = tid & ((1 << lg) - 1);
                const int cs = tid >> lg;
                const int c = cs & (nchains - 1);
                const int jstep = 256 >> (lg + lc);
                const int rs = nchains * n;  // LDS floats between consecutive samples of a series
                const int64_t sb = s0 >> lc;
                int j = cs >> lc;
                const float* src = Os + (4 * g * nchains + c) * n + j;
                float* dst = sz + ((int64_t)c * n + j) * spcp + sb + 4 * g;
                const int64_t dstep = (int64_t)jstep * spcp;
#pragma unroll 2
                for (; j < n; j += jstep, src += jstep, dst += dstep) {
                    const floatx4 w = {src[0], src[rs], src[2 * rs], src[3 * rs]};
                    // Streaming stores: 5 GB of series do not stay in L2 for the diagnostics anyway.
                    __builtin_nontemporal_store(w, reinterpret_cast<floatx4*>(dst));
                }
            } else {
                for (int e = tid; e < (int)rows * n; e += 256) {
                    const int r = e % (int)rows, j = e / (int)rows;
                    const int64_t g = s0 + r;
                    sz[((g % nchains) * n + j) * spcp + g / nchains] = Os[r * n + j];
                }
            }
        }
    }
}

size_t small_lds_bytes(int n, int d)
{
    const int NB = (n + 15) / 16;
    const size_t S = (size_t)small_stride(d);
    const size_t ys = std::max((size_t)kSmRows * S, (size_t)kSmRows * n);
    // + 16 floats so the masked 16-byte reads past the last T row stay in
    // the allocation.
    return sizeof(float) * (16 * (size_t)NB * S + ys + 16);
}

template <int NB>
void launch_small(const float* T, const float* shift, const float* Y, float* V, int n, int d, int64_t ns, int batch,
                  int64_t sT, int64_t sShift, int64_t sY, int64_t sV, hipStream_t stream, float* ser, int nchains,
                  int spcp)
{
    const int64_t n_tiles = (ns + kSmRows - 1) / kSmRows;
    // About 512 workgroups in all (two per CU at this LDS size), each
    // taking a contiguous run of tiles of one batch item.
    const int64_t wpb = std::max<int64_t>(1, std::min<int64_t>(n_tiles, (512 + batch - 1) / batch));
    const int tpw = (int)((n_tiles + wpb - 1) / wpb);
    const int64_t gx = (n_tiles + tpw - 1) / tpw;
    hipLaunchKernelGGL(backmap_small_kernel<NB>, dim3((unsigned)gx, 1, (unsigned)batch), dim3(256),
                       small_lds_bytes(n, d), stream, T, shift, Y, V, n, d, ns, tpw, sT, sShift, sY, sV, ser, nchains,
                       spcp);
    NAJA_HIP(hipGetLastError());
}

}  // namespace

// The small-operand kernel needs 16-byte rows and bases for its float4 tile
// loads and stores.
bool small_ok(int n, int d, const float* Y, const float* V, int64_t sY, int64_t sV, int batch)
{
    return n <= kSmMaxN && d <= kSmMaxD && d % 4 == 0 && n >= 1 && reinterpret_cast<uintptr_t>(Y) % 16 == 0 &&
           reinterpret_cast<uintptr_t>(V) % 16 == 0 && sY % 4 == 0 && sV % 4 == 0 && batch <= 65535;
}

bool backmap_inplace_ok(int d, const float* Y, int64_t sY, int batch) { return small_ok(d, d, Y, Y, sY, sY, batch); }

bool backmap_series_ok(int n, int d, const float* Y, const float* V, int64_t sY, int64_t sV, int batch)
{
    return small_ok(n, d, Y, V, sY, sV, batch);
}

void backmap_launch_batched(const float* T, const float* shift, const float* Y, float* V, int n, int d, int64_t ns,
                            int batch, int64_t sT, int64_t sShift, int64_t sY, int64_t sV, hipStream_t stream,
                            float* ser, int nchains, int spcp)
{
    if (ns == 0 || batch == 0) return;
    const bool small = small_ok(n, d, Y, V, sY, sV, batch);
    NAJA_REQUIRE(!ser || (small && nchains > 0 && ns % nchains == 0 && spcp >= ns / nchains), NAJA_ERR_INVALID_ARGUMENT,
                 "backmap: a series-layout copy needs the small-operand kernel and whole chains");
    if (small) {
        switch ((n + 15) / 16) {
        case 1: launch_small<1>(T, shift, Y, V, n, d, ns, batch, sT, sShift, sY, sV, stream, ser, nchains, spcp); return;
        case 2: launch_small<2>(T, shift, Y, V, n, d, ns, batch, sT, sShift, sY, sV, stream, ser, nchains, spcp); return;
        case 3: launch_small<3>(T, shift, Y, V, n, d, ns, batch, sT, sShift, sY, sV, stream, ser, nchains, spcp); return;
        case 4: launch_small<4>(T, shift, Y, V, n, d, ns, batch, sT, sShift, sY, sV, stream, ser, nchains, spcp); return;
        case 5: launch_small<5>(T, shift, Y, V, n, d, ns, batch, sT, sShift, sY, sV, stream, ser, nchains, spcp); return;
        case 6: launch_small<6>(T, shift, Y, V, n, d, ns, batch, sT, sShift, sY, sV, stream, ser, nchains, spcp); return;
        case 7: launch_small<7>(T, shift, Y, V, n, d, ns, batch, sT, sShift, sY, sV, stream, ser, nchains, spcp); return;
        default: launch_small<8>(T, shift, Y, V, n, d, ns, batch, sT, sShift, sY, sV, stream, ser, nchains, spcp); return;
        }
    }
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
