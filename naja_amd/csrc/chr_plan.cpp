// Sampling plans: the polytope blocks of a batch are packed once (and, with a
// rounding transform, multiplied into A' = A P_v), then launched any number of
// times with no allocation, so a launch can be captured in a hipGraph. The
// plan also owns the measuring autotuner over launch shapes.
//
// Replaces the per-call setup of naja::gpu::CoordinateHitAndRun
// (chr.cu:13-70: slack DGEMM, cuRAND state allocation, kernel launch) and, for
// batches, naja's one-model-per-job loop (job_bulk.cpp:115-189).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/naja_amd.h"
#include "backmap.h"
#include "chr_kernel.h"
#include "common.h"
#include "philox.h"

struct naja_chr_plan {
    int m0 = 0, d = 0, n_variants = 0, G = 1, m_max = 0;
    // Paired-row layout: every block whose second half negates its first is
    // packed as its first half with a second right-hand side (bn); rows_max
    // is the largest variant's packed row count in the layout in use.
    bool paired = false;
    int rows_max = 0;
    // Chains per workgroup of the current shape's launches (kChrLdsChains for
    // the LDS-column kernel, whose items come in groups of one variant) and
    // the dynamic LDS that kernel needs.
    int nc = 1;
    size_t lds_cols = 0;
    std::vector<int> pair_flags;  // [base block, extra block of each variant]
    // The packs are in the scaled-pair form (ChrArgs::scaled): the paired
    // LDS-column kernel with every block paired.
    bool sc = false;
    naja_chr_params p{};
    hipStream_t stream = nullptr;  // creation stream: allocations and frees are ordered on it
    const float* A = nullptr;
    const double* b = nullptr;
    const float* A_extra = nullptr;
    const double* b_extra = nullptr;
    const float* P = nullptr;  // optional n_variants x d x d, row-major: y = c + P z
    const float* c = nullptr;  // optional n_variants x d
    std::vector<int> extra_rows;
    naja::ChrShape shape{1, 1, 1};
    bool tuned = false;

    // host copies of the device tables (kept alive for async uploads)
    std::vector<naja::ChrVariant> var;
    std::vector<int64_t> eoff;
    std::vector<int> ek;
    std::vector<naja::ChrItem> items;

    naja::ChrVariant* var_d = nullptr;
    naja::ChrItem* items_d = nullptr;
    int64_t* eoff_d = nullptr;
    int* ek_d = nullptr;
    int* pair_d = nullptr;
    float* A_pack = nullptr;
    float* Ainv_pack = nullptr;
    double* b_pack = nullptr;
    double* bn_pack = nullptr;  // paired layout only
    double* bn_raw = nullptr;   // paired transform plans only
    float* A_raw = nullptr;   // untransformed pack (transform plans only)
    double* b_raw = nullptr;
    float* zbuf = nullptr;    // z-space samples (transform plans only)
    float* inplace = nullptr; // transform plan sampling z into this output and backmapping it in place
    int64_t pack_floats = 0;  // capacity of the pack buffers (floats)
    int64_t pack_rows = 0;    // capacity of b buffers (rows)
    // The tables and the pack buffers are each carved out of one scratch
    // allocation (every pool call costs tens of microseconds of host time).
    char* tables_mem = nullptr;
    char* pack_mem = nullptr;

    // Ordering between the creation stream (packs, allocations, frees) and
    // launches on other streams: `packed` is recorded on the creation stream
    // after every repack (pack_seq counts them), and a launch stream waits on
    // it once per repack; `launched` is recorded after a launch on another
    // stream, and the creation stream waits on it before it rewrites or
    // frees the packs.
    hipEvent_t packed = nullptr;
    hipEvent_t launched = nullptr;
    uint64_t pack_seq = 0;
    std::vector<std::pair<hipStream_t, uint64_t>> waited;  // launch stream -> pack_seq it has waited for
    bool launch_pending = false;
};

namespace naja {

namespace {

using TuneKey = std::tuple<int, int, int, int, int, int>;  // rows64, d, G, n_items, tpb, paired
std::mutex g_tune_mu;
std::map<TuneKey, ChrShape> g_tuned;

TuneKey tune_key(const naja_chr_plan& pl)
{
    return TuneKey{(pl.rows_max + 63) / 64 * 64, pl.d, pl.G, (int)pl.items.size(), pl.p.tpb > 0 ? pl.p.tpb : 0,
                   pl.paired ? 1 : 0};
}

// Packed rows of variant v in the layout in use.
int packed_rows(const naja_chr_plan& pl, int v)
{
    const int k = pl.ek[(size_t)v];
    if (!pl.paired) return pl.m0 + k;
    return (pl.pair_flags[0] ? pl.m0 / 2 : pl.m0) + (pl.pair_flags[(size_t)v + 1] ? k / 2 : k);
}

void free_async(void* ptr, hipStream_t s) { scratch_free(ptr, s); }

// Byte offsets of consecutive sub-buffers of one allocation, each 256-byte
// aligned; returns the total size.
size_t carve(std::initializer_list<size_t> bytes, size_t* off)
{
    size_t total = 0;
    int q = 0;
    for (size_t b : bytes) {
        off[q++] = total;
        total += (b + 255) / 256 * 256;
    }
    return total;
}

template <typename T>
void alloc_async(T** ptr, size_t count, hipStream_t s)
{
    scratch_alloc(ptr, count, s);
}

// Workgroup b of a launch is dispatched to XCD b % 8 (round robin), and each
// XCD has its own L2. Workgroup b therefore runs logical item
// (b % 8) * per + b / 8, per = n / 8: every XCD gets one contiguous range of
// items, so the chains of one variant share an L2 and each L2 holds the
// columns of an eighth of the variants instead of reading from all of them.
// Items past 8 * per keep their place. Results do not depend on the order:
// an item's samples depend only on its variant and chains.
// With nc chains per workgroup the same order applies to the groups of nc
// consecutive items (nc chains of one variant).
std::vector<ChrItem> xcd_order(const std::vector<ChrItem>& logical, int nc = 1)
{
    constexpr int kXcds = 8;
    const int n = (int)logical.size() / nc, per = n / kXcds;
    std::vector<ChrItem> out(logical);
    for (int b = 0; b < per * kXcds; ++b)
        for (int w = 0; w < nc; ++w)
            out[(size_t)b * nc + w] = logical[(size_t)((b % kXcds) * per + b / kXcds) * nc + w];
    return out;
}

std::vector<ChrItem> logical_items(const naja_chr_plan& pl)
{
    std::vector<ChrItem> logical;
    for (int v = 0; v < pl.n_variants; ++v)
        for (int c0 = 0; c0 < pl.p.nchains; c0 += pl.G) logical.push_back(ChrItem{v, c0});
    return logical;
}

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

// (Re)packs the blocks for `sh` (lda = 64 W R), growing the buffers if needed.
void repack(naja_chr_plan& pl, const ChrShape& sh)
{
    const hipStream_t s = pl.stream;
    if (pl.launch_pending) {  // a launch on another stream may still read the packs
        NAJA_HIP(hipStreamWaitEvent(s, pl.launched, 0));
        pl.launch_pending = false;
    }
    const int lda = sh.rows();
    int64_t a_off = 0, b_off = 0;
    const int64_t per_variant_samples = (int64_t)pl.p.nchains * pl.p.nspc * pl.d;
    for (int v = 0; v < pl.n_variants; ++v) {
        ChrVariant& V = pl.var[(size_t)v];
        V.m = packed_rows(pl, v);
        V.lda = lda;
        V.a_off = a_off;
        V.b_off = b_off;
        V.y0_col = (int64_t)v * pl.p.nchains;
        V.sample_off = (int64_t)v * per_variant_samples;
        a_off += (int64_t)lda * pl.d;
        b_off += lda;
    }
    // The LDS-column kernel: one wavefront per chain and four chains of one
    // variant per workgroup, when the launch's moves, d, chain count, kernels
    // and the columns (d x 64 W R packed rows) allow it.
    {
        const naja_chr_params* p = &pl.p;
        const bool has_pair = p->pair_prob > 0.0 && !(p->pair_mode == 2 && !(p->n_pairs > 0 && p->pair_i &&
                                                                               p->pair_j && p->pair_c && p->pair_s));
        const bool coord_only = rng_threshold(p->ksparse_prob) == 0 && !has_pair;
        int nc = 1;
        size_t lds = 0;
        if (sh.W == 1 && sh.G == 1 && coord_only && pl.d <= kChrRegStateMaxD && p->nchains % kChrLdsChains == 0 &&
            chr_has_lds_kernel(sh, pl.paired)) {
            lds = chr_lds_columns_bytes(sh, pl.d);
            if (lds <= 156u * 1024u) nc = kChrLdsChains;
        }
        if (nc != pl.nc) {
            pl.items = xcd_order(logical_items(pl), nc);
            NAJA_HIP(hipMemcpyAsync(pl.items_d, pl.items.data(), sizeof(ChrItem) * pl.items.size(),
                                    hipMemcpyHostToDevice, s));
        }
        pl.nc = nc;
        pl.lds_cols = nc > 1 ? lds : 0;
        // Scaled pairs save the clamps' instructions, which bind launches
        // of few chains (latency), and read the padding rows' +inf
        // reciprocals, which costs a launch of a workgroup per CU or more
        // (column traffic) unless the LDS holds the columns.
        const bool fills = nc == 1 && pl.items.size() >= 256;
        pl.sc = !fills && pl.paired && chr_has_scaled_kernel(sh, coord_only, pl.d <= kChrRegStateMaxD, nc) &&
                std::all_of(pl.pair_flags.begin(), pl.pair_flags.end(), [](int f) { return f != 0; });
    }
    if (a_off > pl.pack_floats || b_off > pl.pack_rows) {
        free_async(pl.pack_mem, s);
        pl.pack_mem = nullptr;
        const size_t af = sizeof(float) * (size_t)a_off, bd = sizeof(double) * (size_t)b_off;
        const size_t bnd = pl.paired ? bd : 0;
        size_t off[7];
        const size_t total = carve({af, af, bd, pl.P ? af : 0, pl.P ? bd : 0, bnd, pl.P ? bnd : 0}, off);
        alloc_async(&pl.pack_mem, total, s);
        pl.A_pack = reinterpret_cast<float*>(pl.pack_mem + off[0]);
        pl.Ainv_pack = reinterpret_cast<float*>(pl.pack_mem + off[1]);
        pl.b_pack = reinterpret_cast<double*>(pl.pack_mem + off[2]);
        pl.A_raw = pl.P ? reinterpret_cast<float*>(pl.pack_mem + off[3]) : nullptr;
        pl.b_raw = pl.P ? reinterpret_cast<double*>(pl.pack_mem + off[4]) : nullptr;
        pl.bn_pack = pl.paired ? reinterpret_cast<double*>(pl.pack_mem + off[5]) : nullptr;
        pl.bn_raw = pl.paired && pl.P ? reinterpret_cast<double*>(pl.pack_mem + off[6]) : nullptr;
        pl.pack_floats = a_off;
        pl.pack_rows = b_off;
    }
    NAJA_HIP(hipMemcpyAsync(pl.var_d, pl.var.data(), sizeof(ChrVariant) * pl.var.size(), hipMemcpyHostToDevice, s));
    const int* pairs = pl.paired ? pl.pair_d : nullptr;
    if (pl.P) {
        chr_pack(pl.A, pl.b, pl.m0, pl.d, pl.A_extra, pl.b_extra, pl.var_d, pl.eoff_d, pl.ek_d, pl.n_variants, pairs,
                 pl.A_raw, pl.Ainv_pack, pl.b_raw, pl.bn_raw, lda, s);
        chr_transform(pl.A_raw, pl.b_raw, pl.bn_raw, pl.var_d, pl.n_variants, pl.d, lda, pl.P, pl.c, pl.A_pack,
                      pl.Ainv_pack, pl.b_pack, pl.bn_pack, s, pl.sc);
    } else {
        chr_pack(pl.A, pl.b, pl.m0, pl.d, pl.A_extra, pl.b_extra, pl.var_d, pl.eoff_d, pl.ek_d, pl.n_variants, pairs,
                 pl.A_pack, pl.Ainv_pack, pl.b_pack, pl.bn_pack, lda, s, pl.sc);
    }
    pl.shape = sh;
    NAJA_HIP(hipEventRecord(pl.packed, s));
    ++pl.pack_seq;
}

// A new rounding (P, c) of a transform plan: the packed untransformed rows
// stay, only A' = A P and b' = b - A c are formed again, on the creation
// stream, after any launch that may still read the packs.
void retransform(naja_chr_plan& pl)
{
    const hipStream_t s = pl.stream;
    if (pl.launch_pending) {
        NAJA_HIP(hipStreamWaitEvent(s, pl.launched, 0));
        pl.launch_pending = false;
    }
    chr_transform(pl.A_raw, pl.b_raw, pl.bn_raw, pl.var_d, pl.n_variants, pl.d, pl.shape.rows(), pl.P, pl.c,
                  pl.A_pack, pl.Ainv_pack, pl.b_pack, pl.bn_pack, s, pl.sc);
    // Under capture (a graph replays the transform and its launches in
    // order on one stream) there is no event to record for other streams.
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    NAJA_HIP(hipStreamIsCapturing(s, &cap));
    if (cap != hipStreamCaptureStatusNone) return;
    NAJA_HIP(hipEventRecord(pl.packed, s));
    ++pl.pack_seq;
}

// Before a launch on stream s: s waits for the latest repack on the
// creation stream (once per repack and stream). A captured launch records no
// wait (an event recorded outside the capture cannot be waited on inside
// it): the packing must be complete before capture begins.
void order_after_pack(naja_chr_plan& pl, hipStream_t s)
{
    if (s == pl.stream) return;
    auto it = std::find_if(pl.waited.begin(), pl.waited.end(), [&](const auto& e) { return e.first == s; });
    if (it != pl.waited.end() && it->second == pl.pack_seq) return;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    NAJA_HIP(hipStreamIsCapturing(s, &cap));
    if (cap != hipStreamCaptureStatusNone) return;
    NAJA_HIP(hipStreamWaitEvent(s, pl.packed, 0));
    if (it != pl.waited.end())
        it->second = pl.pack_seq;
    else
        pl.waited.emplace_back(s, pl.pack_seq);
}

// After a launch on stream s: the creation stream will wait for it before
// repacking or freeing (not for captured launches: replays are ordered by
// the caller).
void note_launch(naja_chr_plan& pl, hipStream_t s)
{
    if (s == pl.stream) return;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    NAJA_HIP(hipStreamIsCapturing(s, &cap));
    if (cap != hipStreamCaptureStatusNone) return;
    NAJA_HIP(hipEventRecord(pl.launched, s));
    pl.launch_pending = true;
}

ChrArgs make_args(const naja_chr_plan& pl, const float* y0, float* samples, float* y_final, int64_t nsteps,
                  int64_t t0 = 0)
{
    const naja_chr_params* p = &pl.p;
    ChrArgs a;
    std::memset(&a, 0, sizeof(a));
    a.A = pl.A_pack;
    a.Ainv = pl.Ainv_pack;
    a.b = pl.b_pack;
    a.bn = pl.paired ? pl.bn_pack : nullptr;
    a.d = pl.d;
    a.variants = pl.var_d;
    a.items = pl.items_d;
    a.y0 = y0;
    a.samples = samples;
    a.y_final = y_final;
    a.nchains = p->nchains;
    a.thin = p->thinning;
    a.nsteps = nsteps;
    a.seed_lo = (uint32_t)p->seed;
    a.seed_hi = (uint32_t)(p->seed >> 32);
    a.thr_k = rng_threshold(p->ksparse_prob);
    a.thr_kp = rng_threshold(p->ksparse_prob + p->pair_prob);
    a.has_pair = p->pair_prob > 0.0 ? 1 : 0;
    a.pair_mode = p->pair_mode;
    const bool sched = p->pair_mode == 2 && p->n_pairs > 0 && p->pair_i && p->pair_j && p->pair_c && p->pair_s;
    a.pair_i = p->pair_i;
    a.pair_j = p->pair_j;
    a.pair_c = p->pair_c;
    a.pair_s = p->pair_s;
    a.n_pairs = sched ? p->n_pairs : 0;
    if (a.pair_mode == 2 && a.n_pairs == 0) a.has_pair = 0;
    int k = p->ksparse_k;
    if (k < 1) k = 1;
    if (k > 32) k = 32;
    if (k > pl.d) k = pl.d;
    a.ks_k = k;
    a.ks_mag = (float)(1.0 / std::sqrt((double)k));
    a.resync = chr_effective_resync(p->resync_interval, pl.d);
    a.move_log = p->move_log;
    a.nlog = p->move_log ? p->nlog : 0;
    a.t0 = t0;
    a.scaled = pl.sc ? 1 : 0;
    return a;
}

void launch(naja_chr_plan& pl, const float* y0, float* samples, float* y_final, int64_t nsteps, hipStream_t s,
            int64_t t0 = 0, void (*after_chr)(hipStream_t, void*) = nullptr, void* user = nullptr,
            float* ser = nullptr, int spcp = 0)
{
    if (pl.inplace)
        NAJA_REQUIRE(samples == pl.inplace, NAJA_ERR_INVALID_ARGUMENT,
                     "this plan samples in place into the output it was created for");
    float* zs = pl.inplace ? samples : pl.zbuf;
    float* out = pl.P ? zs : samples;
    const ChrArgs a = make_args(pl, y0, out, y_final, nsteps, t0);
    order_after_pack(pl, s);
    chr_launch(a, (int)pl.items.size(), pl.shape, s, pl.nc, pl.lds_cols);
    if (after_chr) after_chr(s, user);
    if (pl.P) {
        // y = c + P z for every saved sample of every variant (batched MFMA backmap).
        const int64_t per = (int64_t)pl.p.nchains * pl.p.nspc;
        backmap_launch_batched(pl.P, pl.c, zs, samples, pl.d, pl.d, per, pl.n_variants, (int64_t)pl.d * pl.d,
                               pl.d, per * pl.d, per * pl.d, s, ser, pl.p.nchains, spcp);
    }
    note_launch(pl, s);
}

}  // namespace

}  // namespace naja

namespace naja {
namespace {

// Plan creation behind naja_chr_plan_create; inplace_out (may be null) is the
// output a one-shot rounded call will launch into.
int plan_create(const float* A, const double* b, int m0, int d, int n_variants, const int* extra_rows,
                const float* A_extra, const double* b_extra, const float* P, const float* c, const naja_chr_params* p,
                void* stream, naja_chr_plan** out, float* inplace_out)
{
    using namespace naja;
    if (out) *out = nullptr;
    naja_chr_plan* pl = nullptr;
    const int rc = guarded([&] {
        NAJA_REQUIRE(out != nullptr, NAJA_ERR_INVALID_ARGUMENT, "NULL plan output");
        NAJA_REQUIRE(A && b, NAJA_ERR_INVALID_ARGUMENT, "NULL A or b");
        NAJA_REQUIRE(m0 > 0 && d > 0, NAJA_ERR_INVALID_ARGUMENT, "invalid polytope shape %d x %d", m0, d);
        NAJA_REQUIRE(n_variants > 0, NAJA_ERR_INVALID_ARGUMENT, "n_variants must be > 0");
        NAJA_REQUIRE((P == nullptr) == (c == nullptr), NAJA_ERR_INVALID_ARGUMENT,
                     "a rounding transform needs both P and c");
        check_params(p, d);
        pl = new naja_chr_plan();
        pl->m0 = m0;
        pl->d = d;
        pl->n_variants = n_variants;
        pl->p = *p;
        pl->G = p->chains_per_group > 0 ? p->chains_per_group : 1;
        pl->stream = reinterpret_cast<hipStream_t>(stream);
        pl->A = A;
        pl->b = b;
        pl->A_extra = A_extra;
        pl->b_extra = b_extra;
        pl->P = P;
        pl->c = c;
        pl->var.resize((size_t)n_variants);
        pl->eoff.resize((size_t)n_variants);
        pl->ek.resize((size_t)n_variants);
        int64_t e_off = 0;
        for (int v = 0; v < n_variants; ++v) {
            const int k = extra_rows ? extra_rows[v] : 0;
            NAJA_REQUIRE(k >= 0, NAJA_ERR_INVALID_ARGUMENT, "variant %d has a negative extra row count", v);
            NAJA_REQUIRE(k == 0 || (A_extra && b_extra), NAJA_ERR_INVALID_ARGUMENT,
                         "extra rows given without A_extra / b_extra");
            pl->ek[(size_t)v] = k;
            pl->eoff[(size_t)v] = e_off;
            e_off += k;
            pl->m_max = std::max(pl->m_max, m0 + k);
        }
        pl->items = xcd_order(logical_items(*pl));
        const hipStream_t s = pl->stream;
        NAJA_HIP(hipEventCreateWithFlags(&pl->packed, hipEventDisableTiming));
        NAJA_HIP(hipEventCreateWithFlags(&pl->launched, hipEventDisableTiming));
        size_t off[5];
        const size_t total = carve({sizeof(ChrVariant) * pl->var.size(), sizeof(ChrItem) * pl->items.size(),
                                    sizeof(int64_t) * pl->eoff.size(), sizeof(int) * pl->ek.size(),
                                    sizeof(int) * ((size_t)n_variants + 1)},
                                   off);
        alloc_async(&pl->tables_mem, total, s);
        pl->var_d = reinterpret_cast<ChrVariant*>(pl->tables_mem + off[0]);
        pl->items_d = reinterpret_cast<ChrItem*>(pl->tables_mem + off[1]);
        pl->eoff_d = reinterpret_cast<int64_t*>(pl->tables_mem + off[2]);
        pl->ek_d = reinterpret_cast<int*>(pl->tables_mem + off[3]);
        pl->pair_d = reinterpret_cast<int*>(pl->tables_mem + off[4]);
        NAJA_HIP(hipMemcpyAsync(pl->items_d, pl->items.data(), sizeof(ChrItem) * pl->items.size(),
                                hipMemcpyHostToDevice, s));
        NAJA_HIP(hipMemcpyAsync(pl->eoff_d, pl->eoff.data(), sizeof(int64_t) * pl->eoff.size(),
                                hipMemcpyHostToDevice, s));
        NAJA_HIP(hipMemcpyAsync(pl->ek_d, pl->ek.data(), sizeof(int) * pl->ek.size(), hipMemcpyHostToDevice, s));
        // A one-shot call may sample z straight into its output and backmap it
        // there (the small-operand backmap reads each tile whole before
        // writing it back); otherwise z goes to a buffer of the plan's own.
        const int64_t per = (int64_t)p->nchains * p->nspc;
        if (P && inplace_out && backmap_inplace_ok(d, inplace_out, per * d, n_variants))
            pl->inplace = inplace_out;
        else if (P)
            alloc_async(&pl->zbuf, (size_t)n_variants * per * d, s);
        // Row pairing: one check kernel and one read-back. The paired layout
        // is used when it packs at least a quarter fewer rows and a paired
        // kernel holds the largest variant.
        pl->pair_flags.assign((size_t)n_variants + 1, 1);
        NAJA_HIP(hipMemcpyAsync(pl->pair_d, pl->pair_flags.data(), sizeof(int) * pl->pair_flags.size(),
                                hipMemcpyHostToDevice, s));
        chr_pair_check(A, m0, d, A_extra, pl->eoff_d, pl->ek_d, n_variants, pl->pair_d, s);
        NAJA_HIP(hipMemcpyAsync(pl->pair_flags.data(), pl->pair_d, sizeof(int) * pl->pair_flags.size(),
                                hipMemcpyDeviceToHost, s));
        NAJA_HIP(hipStreamSynchronize(s));
        int64_t rows_all = 0, rows_paired = 0;
        int paired_max = 0;
        pl->paired = true;
        for (int v = 0; v < n_variants; ++v) {
            const int r = packed_rows(*pl, v);
            rows_paired += r;
            rows_all += m0 + pl->ek[(size_t)v];
            paired_max = std::max(paired_max, r);
        }
        ChrShape sh{0, pl->G, 0};
        if (rows_paired * 4 <= rows_all * 3)
            sh = chr_choose_shape(paired_max, pl->G, p->tpb, (int)pl->items.size(), true);
        pl->paired = sh.W > 0;
        pl->rows_max = pl->paired ? paired_max : pl->m_max;
        if (!pl->paired) sh = chr_choose_shape(pl->m_max, pl->G, p->tpb, (int)pl->items.size());
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tuned.find(tune_key(*pl));
            if (it != g_tuned.end()) {
                sh = it->second;
                pl->tuned = true;
            }
        }
        repack(*pl, sh);
    });
    if (rc != NAJA_OK) {
        if (pl) naja_chr_plan_destroy(pl);
        return rc;
    }
    *out = pl;
    return NAJA_OK;
}

}  // namespace
}  // namespace naja

namespace naja {
// A plan launch with a callback between the CHR kernel and the backmap (for
// the conditioned run, which places its timing events there), optionally
// writing the backmapped samples also in the diagnostics' series layout.
void chr_plan_launch_hooked(naja_chr_plan* plan, const float* y0, float* samples, hipStream_t s,
                            void (*after_chr)(hipStream_t, void*), void* user, float* ser, int spcp)
{
    NAJA_REQUIRE(plan && y0 && samples, NAJA_ERR_INVALID_ARGUMENT, "NULL plan, y0 or samples");
    NAJA_REQUIRE(!ser || plan->P, NAJA_ERR_INVALID_ARGUMENT, "a series copy comes from the rounding backmap");
    launch(*plan, y0, samples, nullptr, (int64_t)plan->p.nspc * plan->p.thinning, s, 0, after_chr, user, ser, spcp);
}

// Whether a launch of this rounded plan into `samples` can also write the
// diagnostics' series layout (its backmap runs the small-operand kernel).
bool chr_plan_series_ok(const naja_chr_plan* plan, const float* samples)
{
    if (!plan || !plan->P) return false;
    const int64_t per = (int64_t)plan->p.nchains * plan->p.nspc;
    const float* zs = plan->inplace ? samples : plan->zbuf;
    return backmap_series_ok(plan->d, plan->d, zs, samples, per * plan->d, per * plan->d, plan->n_variants);
}

// Packs a rounded plan's untransformed rows again from the caller's A, b,
// A_extra, b_extra (same counts, same shape and row pairing as at creation),
// on stream s, for a pipeline whose inputs change in place between steps; a
// naja_chr_plan_set_rounding after it forms the transformed rows.
void chr_plan_repack_rows(naja_chr_plan* plan, hipStream_t s)
{
    NAJA_REQUIRE(plan && plan->P, NAJA_ERR_INVALID_ARGUMENT, "a rounded plan is needed");
    naja_chr_plan& pl = *plan;
    const int* pairs = pl.paired ? pl.pair_d : nullptr;
    chr_pack(pl.A, pl.b, pl.m0, pl.d, pl.A_extra, pl.b_extra, pl.var_d, pl.eoff_d, pl.ek_d, pl.n_variants, pairs,
             pl.A_raw, pl.Ainv_pack, pl.b_raw, pl.bn_raw, pl.shape.rows(), s);
}
}  // namespace naja

extern "C" {

int naja_chr_plan_create(const float* A, const double* b, int m0, int d, int n_variants, const int* extra_rows,
                         const float* A_extra, const double* b_extra, const float* P, const float* c,
                         const naja_chr_params* p, void* stream, naja_chr_plan** out)
{
    return naja::plan_create(A, b, m0, d, n_variants, extra_rows, A_extra, b_extra, P, c, p, stream, out, nullptr);
}


int naja_chr_plan_launch(naja_chr_plan* plan, const float* y0, float* samples, float* y_final, void* stream)
{
    using namespace naja;
    return guarded([&] {
        NAJA_REQUIRE(plan && y0 && samples, NAJA_ERR_INVALID_ARGUMENT, "NULL plan, y0 or samples");
        launch(*plan, y0, samples, y_final, (int64_t)plan->p.nspc * plan->p.thinning,
               reinterpret_cast<hipStream_t>(stream));
    });
}

int naja_chr_plan_set_rounding(naja_chr_plan* plan, const float* P, const float* c)
{
    using namespace naja;
    return guarded([&] {
        NAJA_REQUIRE(plan && P && c, NAJA_ERR_INVALID_ARGUMENT, "NULL plan, P or c");
        NAJA_REQUIRE(plan->P != nullptr, NAJA_ERR_INVALID_ARGUMENT,
                     "the plan was created without a rounding transform");
        NAJA_REQUIRE(plan->inplace == nullptr, NAJA_ERR_INVALID_ARGUMENT,
                     "a one-shot plan sampling in place cannot change its rounding");
        plan->P = P;
        plan->c = c;
        retransform(*plan);
    });
}

int naja_chr_plan_autotune(naja_chr_plan* plan, const float* y0, void* stream, int* shape_out)
{
    using namespace naja;
    return guarded([&] {
        NAJA_REQUIRE(plan && y0, NAJA_ERR_INVALID_ARGUMENT, "NULL plan or y0");
        naja_chr_plan& pl = *plan;
        const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        const int rows = (pl.rows_max + 63) / 64 * 64;
        std::vector<ChrShape> cand;
        for (int W : {1, 4, 8, 16}) {
            const int R = chr_smallest_r(W, pl.G, rows, pl.paired);
            if (R > 0) cand.push_back(ChrShape{W, pl.G, R});
        }
        NAJA_REQUIRE(!cand.empty(), NAJA_ERR_UNSUPPORTED, "no compiled shape for %d rows", pl.rows_max);
        // Probe: about 1024 steps of the real run (at least one saved sample).
        const int nspc_probe = std::max(1, std::min(pl.p.nspc, 1024 / pl.p.thinning));
        const int64_t nsteps = (int64_t)nspc_probe * pl.p.thinning;
        float* probe = nullptr;
        alloc_async(&probe, (size_t)pl.n_variants * pl.p.nchains * pl.p.nspc * pl.d, s);
        hipEvent_t e0, e1;
        NAJA_HIP(hipEventCreate(&e0));
        NAJA_HIP(hipEventCreate(&e1));
        ChrShape best = pl.shape;
        float best_ms = 1e30f;
        for (const ChrShape& sh : cand) {
            repack(pl, sh);
            const ChrArgs a = make_args(pl, y0, (pl.P && pl.zbuf) ? pl.zbuf : probe, nullptr, nsteps);
            order_after_pack(pl, s);
            chr_launch(a, (int)pl.items.size(), sh, s, pl.nc, pl.lds_cols);  // warm-up
            NAJA_HIP(hipEventRecord(e0, s));
            chr_launch(a, (int)pl.items.size(), sh, s, pl.nc, pl.lds_cols);
            NAJA_HIP(hipEventRecord(e1, s));
            note_launch(pl, s);
            NAJA_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            NAJA_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best_ms) {
                best_ms = ms;
                best = sh;
            }
        }
        NAJA_HIP(hipEventDestroy(e0));
        NAJA_HIP(hipEventDestroy(e1));
        free_async(probe, s);
        repack(pl, best);
        pl.tuned = true;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tuned[tune_key(pl)] = best;
        }
        if (shape_out) {
            shape_out[0] = best.W;
            shape_out[1] = best.G;
            shape_out[2] = best.R;
        }
    });
}

int naja_chr_plan_shape(const naja_chr_plan* plan, int* shape_out)
{
    using namespace naja;
    return guarded([&] {
        NAJA_REQUIRE(plan && shape_out, NAJA_ERR_INVALID_ARGUMENT, "NULL plan or output");
        shape_out[0] = plan->shape.W;
        shape_out[1] = plan->shape.G;
        shape_out[2] = plan->shape.R;
    });
}

int naja_chr_plan_layout(const naja_chr_plan* plan, int* layout_out)
{
    using namespace naja;
    return guarded([&] {
        NAJA_REQUIRE(plan && layout_out, NAJA_ERR_INVALID_ARGUMENT, "NULL plan or output");
        layout_out[0] = plan->paired ? 1 : 0;
        layout_out[1] = plan->nc;
        layout_out[2] = plan->sc ? 1 : 0;
    });
}

int naja_chr_plan_destroy(naja_chr_plan* plan)
{
    using namespace naja;
    if (!plan) return NAJA_OK;
    const int rc = guarded([&] {
        const hipStream_t s = plan->stream;
        if (plan->launch_pending) NAJA_HIP(hipStreamWaitEvent(s, plan->launched, 0));
        void* bufs[] = {plan->tables_mem, plan->pack_mem, plan->zbuf};
        for (void* ptr : bufs) free_async(ptr, s);
    });
    if (plan->packed) (void)hipEventDestroy(plan->packed);
    if (plan->launched) (void)hipEventDestroy(plan->launched);
    delete plan;
    return rc;
}

int naja_chr_sample_batched_rounded(const float* A, const double* b, int m0, int d, int n_variants,
                                    const int* extra_rows, const float* A_extra, const double* b_extra,
                                    const float* P, const float* c, const float* y0, const naja_chr_params* p,
                                    float* samples, float* y_final, void* stream)
{
    using namespace naja;
    naja_chr_plan* pl = nullptr;
    int rc = guarded([&] {
        NAJA_REQUIRE(y0 && samples, NAJA_ERR_INVALID_ARGUMENT, "NULL y0 or samples");
    });
    if (rc) return rc;
    rc = plan_create(A, b, m0, d, n_variants, extra_rows, A_extra, b_extra, P, c, p, stream, &pl, samples);
    if (rc) return rc;
    rc = naja_chr_plan_launch(pl, y0, samples, y_final, stream);
    const int rc2 = naja_chr_plan_destroy(pl);
    return rc ? rc : rc2;
}

int naja_chr_sample_batched(const float* A, const double* b, int m0, int d, int n_variants, const int* extra_rows,
                            const float* A_extra, const double* b_extra, const float* y0, const naja_chr_params* p,
                            float* samples, float* y_final, void* stream)
{
    return naja_chr_sample_batched_rounded(A, b, m0, d, n_variants, extra_rows, A_extra, b_extra, nullptr, nullptr,
                                           y0, p, samples, y_final, stream);
}

int naja_chr_sample(const float* A, const double* b, int m, int d, const float* y0, const naja_chr_params* p,
                    float* samples, float* y_final, void* stream)
{
    return naja_chr_sample_batched(A, b, m, d, 1, nullptr, nullptr, nullptr, y0, p, samples, y_final, stream);
}

int naja_chr_sample_streamed(const float* A, const double* b, int m, int d, float* Y, const naja_chr_params* p,
                             int chunk_nspc, const float* T, const float* shift, int n, naja_sample_sink sink,
                             void* user, void* stream)
{
    using namespace naja;
    naja_chr_plan* pl = nullptr;
    float* dbuf[2] = {nullptr, nullptr};
    float* vbuf[2] = {nullptr, nullptr};
    float* hbuf[2] = {nullptr, nullptr};
    hipStream_t copy = nullptr;
    hipEvent_t computed[2] = {nullptr, nullptr}, copied[2] = {nullptr, nullptr};
    const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int rc = guarded([&] {
        NAJA_REQUIRE(A && b && Y && p && sink, NAJA_ERR_INVALID_ARGUMENT, "NULL A, b, Y, params or sink");
        NAJA_REQUIRE(chunk_nspc > 0, NAJA_ERR_INVALID_ARGUMENT, "chunk_nspc must be > 0");
        NAJA_REQUIRE((T == nullptr) == (shift == nullptr) && (T == nullptr || n > 0), NAJA_ERR_INVALID_ARGUMENT,
                     "a backmap needs T, shift and n > 0");
        check_params(p, d);
    });
    if (rc) return rc;
    naja_chr_params pc = *p;
    pc.nspc = std::min(chunk_nspc, p->nspc);
    rc = naja_chr_plan_create(A, b, m, d, 1, nullptr, nullptr, nullptr, nullptr, nullptr, &pc, stream, &pl);
    if (rc) return rc;
    rc = guarded([&] {
        const int nch = p->nchains;
        const int cols = T ? n : d;
        const size_t chunk_rows = (size_t)pc.nspc * nch;
        for (int k = 0; k < 2; ++k) {
            NAJA_HIP(hipMalloc(&dbuf[k], sizeof(float) * chunk_rows * d));
            if (T) NAJA_HIP(hipMalloc(&vbuf[k], sizeof(float) * chunk_rows * n));
            NAJA_HIP(hipHostMalloc((void**)&hbuf[k], sizeof(float) * chunk_rows * cols, hipHostMallocDefault));
            NAJA_HIP(hipEventCreateWithFlags(&computed[k], hipEventDisableTiming));
            NAJA_HIP(hipEventCreateWithFlags(&copied[k], hipEventDisableTiming));
        }
        NAJA_HIP(hipStreamCreateWithFlags(&copy, hipStreamNonBlocking));
        const int64_t nchunks = ((int64_t)p->nspc + pc.nspc - 1) / pc.nspc;
        std::vector<int64_t> rows(2, 0), first(2, 0);
        auto deliver = [&](int slot) {
            NAJA_HIP(hipEventSynchronize(copied[slot]));
            const int src = sink(hbuf[slot], rows[slot], cols, first[slot], user);
            NAJA_REQUIRE(src == 0, NAJA_ERR_RUNTIME, "sample sink returned %d", src);
        };
        for (int64_t k = 0; k < nchunks; ++k) {
            const int slot = (int)(k & 1);
            const int64_t s0 = k * pc.nspc;
            const int ns = (int)std::min<int64_t>(pc.nspc, p->nspc - s0);
            // The copy of chunk k - 2 must have drained this slot.
            if (k >= 2) NAJA_HIP(hipStreamWaitEvent(s, copied[slot], 0));
            // Chunk k continues the chains from Y (state persisted in place)
            // at global step s0 * thin, so the Philox stream runs on.
            launch(*pl, Y, dbuf[slot], Y, (int64_t)ns * p->thinning, s, s0 * p->thinning);
            const float* src = dbuf[slot];
            if (T) {
                backmap_launch(T, shift, dbuf[slot], vbuf[slot], n, d, (int64_t)ns * nch, s);
                src = vbuf[slot];
            }
            NAJA_HIP(hipEventRecord(computed[slot], s));
            NAJA_HIP(hipStreamWaitEvent(copy, computed[slot], 0));
            NAJA_HIP(hipMemcpyAsync(hbuf[slot], src, sizeof(float) * (size_t)ns * nch * cols, hipMemcpyDeviceToHost,
                                    copy));
            NAJA_HIP(hipEventRecord(copied[slot], copy));
            rows[slot] = (int64_t)ns * nch;
            first[slot] = s0 * nch;
            // Hand chunk k - 1 to the sink while chunk k computes.
            if (k >= 1) deliver(slot ^ 1);
        }
        deliver((int)((nchunks - 1) & 1));
        NAJA_HIP(hipStreamSynchronize(s));
    });
    for (int k = 0; k < 2; ++k) {
        if (dbuf[k]) (void)hipFree(dbuf[k]);
        if (vbuf[k]) (void)hipFree(vbuf[k]);
        if (hbuf[k]) (void)hipHostFree(hbuf[k]);
        if (computed[k]) (void)hipEventDestroy(computed[k]);
        if (copied[k]) (void)hipEventDestroy(copied[k]);
    }
    if (copy) (void)hipStreamDestroy(copy);
    const int rc2 = naja_chr_plan_destroy(pl);
    return rc ? rc : rc2;
}

}  // extern "C"
