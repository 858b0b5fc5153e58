// The extern "C" boundary of libnaja_amd.so (declared in include/naja_amd.h).
//
// Argument checks mirror the reference's throws: shape mismatches are
// NAJA_ERR_INVALID_ARGUMENT (backmap.cu:27-35 throws std::invalid_argument),
// an unusable workgroup request is NAJA_ERR_RUNTIME (chr.cu:549 throws
// std::runtime_error("Invalid threads per block")). HIP failures are
// reported as NAJA_ERR_HIP instead of being printed and ignored
// (helper.h:12-27).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/naja_amd.h"
#include "backmap.h"
#include "chr_kernel.h"
#include "common.h"
#include "diag.h"
#include "philox.h"

namespace naja {

namespace {
thread_local char g_error[1024] = {0};
}

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

void clear_error() { g_error[0] = 0; }

namespace {

void check_params(const naja_chr_params* p, int d)
{
    NAJA_REQUIRE(p != nullptr, NAJA_ERR_INVALID_ARGUMENT, "params is NULL");
    NAJA_REQUIRE(p->nspc > 0, NAJA_ERR_INVALID_ARGUMENT, "invalid nspc %d", p->nspc);
    NAJA_REQUIRE(p->thinning > 0, NAJA_ERR_INVALID_ARGUMENT, "invalid thinning %d", p->thinning);
    NAJA_REQUIRE(p->nchains > 0, NAJA_ERR_INVALID_ARGUMENT, "invalid nchains %d", p->nchains);
    NAJA_REQUIRE(p->pair_prob >= 0.0 && p->pair_prob <= 1.0, NAJA_ERR_INVALID_ARGUMENT, "PAIR_PROB must be in [0,1]");
    NAJA_REQUIRE(p->ksparse_prob >= 0.0 && p->ksparse_prob <= 1.0, NAJA_ERR_INVALID_ARGUMENT,
                 "invalid ksparse_prob (must be in [0,1])");
    // chr.cu:529-551 accepts threads-per-block 32..1024 (powers of two) and
    // throws runtime_error otherwise; <= 0 lets the library choose.
    NAJA_REQUIRE(p->tpb <= 0 || p->tpb == 32 || p->tpb == 64 || p->tpb == 128 || p->tpb == 256 || p->tpb == 512 ||
                     p->tpb == 1024,
                 NAJA_ERR_RUNTIME, "Invalid threads per block (%d)", p->tpb);
    NAJA_REQUIRE(p->chains_per_group == 0 || p->chains_per_group == 1 || p->chains_per_group == 4,
                 NAJA_ERR_UNSUPPORTED, "chains_per_group must be 1 or 4 (got %d)", p->chains_per_group);
    // Pair mode 2 without a schedule falls back to coordinate moves, as the
    // reference kernel does (chr.cu:324-326); a given schedule must index
    // inside the polytope.
    if (p->pair_mode == 2 && p->n_pairs > 0 && p->pair_i && p->pair_j && p->pair_c && p->pair_s) {
        std::vector<int> pi((size_t)p->n_pairs), pj((size_t)p->n_pairs);
        NAJA_HIP(hipMemcpy(pi.data(), p->pair_i, sizeof(int) * pi.size(), hipMemcpyDeviceToHost));
        NAJA_HIP(hipMemcpy(pj.data(), p->pair_j, sizeof(int) * pj.size(), hipMemcpyDeviceToHost));
        for (int q = 0; q < p->n_pairs; ++q) {
            NAJA_REQUIRE(pi[(size_t)q] >= 0 && pi[(size_t)q] < d && pj[(size_t)q] >= 0 && pj[(size_t)q] < d,
                         NAJA_ERR_INVALID_ARGUMENT, "pair schedule entry %d indexes outside [0, %d)", q, d);
        }
    }
}

}  // namespace

}  // namespace naja

extern "C" {

const char* naja_last_error(void) { return naja::g_error; }

int naja_version(void) { return 100; }

int naja_chr_sample_batched(const float* A, const double* b, int m0, int d, int n_variants, const int* extra_rows,
                            const float* A_extra, const double* b_extra, const float* y0, const naja_chr_params* p,
                            float* samples, float* y_final, void* stream_)
{
    using namespace naja;
    return guarded([&] {
        hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
        NAJA_REQUIRE(A && b && y0 && samples, NAJA_ERR_INVALID_ARGUMENT, "NULL A, b, y0 or samples");
        NAJA_REQUIRE(m0 > 0 && d > 0, NAJA_ERR_INVALID_ARGUMENT, "invalid polytope shape %d x %d", m0, d);
        NAJA_REQUIRE(n_variants > 0, NAJA_ERR_INVALID_ARGUMENT, "n_variants must be > 0");
        check_params(p, d);
        const int G = p->chains_per_group > 0 ? p->chains_per_group : 1;

        int m_max = 0;
        for (int v = 0; v < n_variants; ++v) {
            const int k = extra_rows ? extra_rows[v] : 0;
            NAJA_REQUIRE(k >= 0, NAJA_ERR_INVALID_ARGUMENT, "variant %d has a negative extra row count", v);
            NAJA_REQUIRE(k == 0 || (A_extra && b_extra), NAJA_ERR_INVALID_ARGUMENT,
                         "extra rows given without A_extra / b_extra");
            m_max = std::max(m_max, m0 + k);
        }
        const ChrShape sh = chr_choose_shape(m_max, G, p->tpb);
        const int lda = sh.rows();

        std::vector<ChrVariant> var((size_t)n_variants);
        std::vector<int64_t> eoff((size_t)n_variants);
        std::vector<int> ek((size_t)n_variants);
        int64_t a_off = 0, b_off = 0, e_off = 0;
        const int64_t per_variant_samples = (int64_t)p->nchains * p->nspc * d;
        for (int v = 0; v < n_variants; ++v) {
            const int k = extra_rows ? extra_rows[v] : 0;
            ChrVariant& V = var[(size_t)v];
            V.m = m0 + k;
            V.lda = lda;
            V.a_off = a_off;
            V.b_off = b_off;
            V.y0_col = (int64_t)v * p->nchains;
            V.sample_off = (int64_t)v * per_variant_samples;
            eoff[(size_t)v] = e_off;
            ek[(size_t)v] = k;
            a_off += (int64_t)lda * d;
            b_off += lda;
            e_off += k;
        }

        std::vector<ChrItem> items;
        for (int v = 0; v < n_variants; ++v)
            for (int c0 = 0; c0 < p->nchains; c0 += G) items.push_back(ChrItem{v, c0});

        ChrVariant* var_d = nullptr;
        ChrItem* items_d = nullptr;
        int64_t* eoff_d = nullptr;
        int* ek_d = nullptr;
        float* A_pack = nullptr;
        float* Ainv_pack = nullptr;
        double* b_pack = nullptr;
        NAJA_HIP(hipMallocAsync((void**)&var_d, sizeof(ChrVariant) * var.size(), stream));
        NAJA_HIP(hipMallocAsync((void**)&items_d, sizeof(ChrItem) * items.size(), stream));
        NAJA_HIP(hipMallocAsync((void**)&eoff_d, sizeof(int64_t) * eoff.size(), stream));
        NAJA_HIP(hipMallocAsync((void**)&ek_d, sizeof(int) * ek.size(), stream));
        NAJA_HIP(hipMallocAsync((void**)&A_pack, sizeof(float) * (size_t)a_off, stream));
        NAJA_HIP(hipMallocAsync((void**)&Ainv_pack, sizeof(float) * (size_t)a_off, stream));
        NAJA_HIP(hipMallocAsync((void**)&b_pack, sizeof(double) * (size_t)b_off, stream));
        NAJA_HIP(hipMemcpyAsync(var_d, var.data(), sizeof(ChrVariant) * var.size(), hipMemcpyHostToDevice, stream));
        NAJA_HIP(hipMemcpyAsync(items_d, items.data(), sizeof(ChrItem) * items.size(), hipMemcpyHostToDevice, stream));
        NAJA_HIP(hipMemcpyAsync(eoff_d, eoff.data(), sizeof(int64_t) * eoff.size(), hipMemcpyHostToDevice, stream));
        NAJA_HIP(hipMemcpyAsync(ek_d, ek.data(), sizeof(int) * ek.size(), hipMemcpyHostToDevice, stream));

        chr_pack(A, b, m0, d, A_extra, b_extra, var_d, eoff_d, ek_d, n_variants, A_pack, Ainv_pack, b_pack, lda, stream);

        ChrArgs a;
        std::memset(&a, 0, sizeof(a));
        a.A = A_pack;
        a.Ainv = Ainv_pack;
        a.b = b_pack;
        a.d = d;
        a.variants = var_d;
        a.items = items_d;
        a.y0 = y0;
        a.samples = samples;
        a.y_final = y_final;
        a.nchains = p->nchains;
        a.thin = p->thinning;
        a.nsteps = (int64_t)p->nspc * p->thinning;
        a.seed_lo = (uint32_t)p->seed;
        a.seed_hi = (uint32_t)(p->seed >> 32);
        a.thr_k = rng_threshold(p->ksparse_prob);
        a.thr_kp = rng_threshold(p->ksparse_prob + p->pair_prob);
        a.has_pair = p->pair_prob > 0.0 ? 1 : 0;
        a.pair_mode = p->pair_mode;
        a.pair_i = p->pair_i;
        a.pair_j = p->pair_j;
        a.pair_c = p->pair_c;
        a.pair_s = p->pair_s;
        const bool sched = p->pair_mode == 2 && p->n_pairs > 0 && p->pair_i && p->pair_j && p->pair_c && p->pair_s;
        a.n_pairs = sched ? p->n_pairs : 0;
        int k = p->ksparse_k;
        if (k < 1) k = 1;
        if (k > 32) k = 32;
        if (k > d) k = d;
        a.ks_k = k;
        a.ks_mag = (float)(1.0 / std::sqrt((double)k));
        a.resync = chr_effective_resync(p->resync_interval, d);
        a.move_log = p->move_log;
        a.nlog = p->move_log ? p->nlog : 0;
        if (a.pair_mode == 2 && a.n_pairs == 0) a.has_pair = 0;

        chr_launch(a, (int)items.size(), sh, stream);

        NAJA_HIP(hipFreeAsync(var_d, stream));
        NAJA_HIP(hipFreeAsync(items_d, stream));
        NAJA_HIP(hipFreeAsync(eoff_d, stream));
        NAJA_HIP(hipFreeAsync(ek_d, stream));
        NAJA_HIP(hipFreeAsync(A_pack, stream));
        NAJA_HIP(hipFreeAsync(Ainv_pack, stream));
        NAJA_HIP(hipFreeAsync(b_pack, stream));
    });
}

int naja_chr_sample(const float* A, const double* b, int m, int d, const float* y0, const naja_chr_params* p,
                    float* samples, float* y_final, void* stream)
{
    return naja_chr_sample_batched(A, b, m, d, 1, nullptr, nullptr, nullptr, y0, p, samples, y_final, stream);
}

int naja_backmap(const float* T, const float* shift, const float* Y, float* V, int n, int d, int64_t ns, void* stream)
{
    using namespace naja;
    return guarded([&] {
        NAJA_REQUIRE(T && shift && ((Y && V) || ns == 0), NAJA_ERR_INVALID_ARGUMENT, "backmap_into: NULL operand");
        NAJA_REQUIRE(n > 0 && d > 0 && ns >= 0, NAJA_ERR_INVALID_ARGUMENT,
                     "backmap_into: transformation shape mismatch with samples");
        backmap_launch(T, shift, Y, V, n, d, ns, reinterpret_cast<hipStream_t>(stream));
    });
}

int naja_diag(const float* samples, int dim, int n_total, int nchains, int layout, float* ess, float* rhat,
              float* ess_chain, void* stream)
{
    using namespace naja;
    return guarded([&] {
        NAJA_REQUIRE(samples && ess && rhat, NAJA_ERR_INVALID_ARGUMENT, "NULL samples, ess or rhat");
        NAJA_REQUIRE(dim > 0 && nchains > 0 && n_total > 0, NAJA_ERR_INVALID_ARGUMENT, "invalid diagnostics shape");
        NAJA_REQUIRE(layout == NAJA_LAYOUT_INTERLEAVED || layout == NAJA_LAYOUT_BLOCKS, NAJA_ERR_INVALID_ARGUMENT,
                     "invalid layout %d", layout);
        NAJA_REQUIRE(n_total % nchains == 0, NAJA_ERR_INVALID_ARGUMENT, "n_total (%d) not divisible by n_chains (%d)",
                     n_total, nchains);
        const int spc = n_total / nchains;
        NAJA_REQUIRE(spc >= 4, NAJA_ERR_INVALID_ARGUMENT, "compute_ess: fewer than 4 samples per chain");
        diag_launch(samples, dim, n_total, nchains, layout, ess, rhat, ess_chain, reinterpret_cast<hipStream_t>(stream));
    });
}

}  // extern "C"
