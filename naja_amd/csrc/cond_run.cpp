// A conditioned-variant run as one hipGraph: the facet-aligned rounding of
// every variant (completed from the base polytope's facets, maximum-volume
// exchange), the variants' rows packed again from the caller's buffers (so
// bounds or rows changed in place between steps are sampled), the new
// rounding set on a sampling plan (A' = A P, b' = b - A c),
// the batched CHR launch with its backmap to y (which, for the small-operand
// backmap, also writes the samples in the diagnostics' series layout, so the
// diagnostics skip their transpose), and the batched ESS / split-R-hat,
// captured once on the creation stream and replayed by each step. Everything the graph touches is allocated at creation: the facet
// rounding's tables and status words, the plan's packs, the diagnostics
// workspace with its twiddle tables. A step replays the graph, copies the
// per-variant rounding status back to pinned host memory and synchronises
// (the caller reads the per-variant ESS and R-hat anyway). With `timed`,
// external event nodes around the CHR launch give its duration per replay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/naja_amd.h"
#include "common.h"
#include "diag.h"
#include "facet.h"

namespace naja {
// chr_plan.cpp: a plan launch with a callback between the CHR kernel and the backmap.
void chr_plan_launch_hooked(naja_chr_plan* plan, const float* y0, float* samples, hipStream_t s,
                            void (*after_chr)(hipStream_t, void*), void* user, float* ser = nullptr, int spcp = 0);
bool chr_plan_series_ok(const naja_chr_plan* plan, const float* samples);
// chr_plan.cpp: the untransformed rows packed again from the caller's inputs.
void chr_plan_repack_rows(naja_chr_plan* plan, hipStream_t s);
}  // namespace naja

struct naja_cond_run {
    int m0 = 0, d = 0, nv = 0, nchains = 0, nspc = 0;
    const float* A0 = nullptr;
    const double* b0 = nullptr;
    const float* Ae = nullptr;
    const double* be = nullptr;
    const float* y0 = nullptr;
    const float* c = nullptr;
    float* P = nullptr;
    float* samples = nullptr;
    float* ess = nullptr;
    float* rhat = nullptr;
    float* ess_min = nullptr;
    float* rhat_max = nullptr;
    float exchange_tol = 0.0f;
    char* dev_mem = nullptr;  // tables, theta, status, diagnostics workspace
    int64_t* eoff = nullptr;
    int* ek = nullptr;
    float* theta = nullptr;
    int* status = nullptr;
    void* diag_ws = nullptr;
    int* status_h = nullptr;  // pinned
    naja_chr_plan* plan = nullptr;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool timed = false;
    bool fused_series = false;  // the backmap writes the diagnostics' series layout too
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = -1.0f;
};

namespace naja {
namespace {

size_t r256(size_t b) { return (b + 255) / 256 * 256; }

// A nested C-ABI call's failure (its message stays in the error state).
void check_plan(int rc)
{
    if (rc != NAJA_OK) throw Error(rc);
}

// An event-record node appended to the capture in progress on s, in stream
// order: the next captured work depends on it.
void capture_event(hipStream_t s, hipEvent_t ev)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    unsigned long long id = 0;
    hipGraph_t g = nullptr;
    const hipGraphNode_t* deps = nullptr;
    size_t n = 0;
    NAJA_HIP(hipStreamGetCaptureInfo_v2(s, &st, &id, &g, &deps, &n));
    NAJA_REQUIRE(st == hipStreamCaptureStatusActive && g, NAJA_ERR_RUNTIME, "no capture in progress");
    hipGraphNode_t node = nullptr;
    NAJA_HIP(hipGraphAddEventRecordNode(&node, g, deps, n, ev));
    NAJA_HIP(hipStreamUpdateCaptureDependencies(s, &node, 1, hipStreamSetCaptureDependencies));
}

void event_after_chr(hipStream_t s, void* user) { capture_event(s, static_cast<naja_cond_run*>(user)->ev1); }

// The work of one step, in stream order on s (eagerly, or under capture with
// the timing events as nodes around the CHR launch).
void enqueue_step(naja_cond_run& r, hipStream_t s, bool capturing)
{
    facet_launch(r.A0, r.b0, r.m0, r.d, r.nv, r.Ae, r.be, r.eoff, r.ek, nullptr, nullptr, r.theta, 0,
                 NAJA_FACET_COMPLETE_BASE_FACETS, r.exchange_tol, nullptr, r.P, r.status, s);
    chr_plan_repack_rows(r.plan, s);
    check_plan(naja_chr_plan_set_rounding(r.plan, r.P, r.c));
    const bool ev = capturing && r.timed;
    if (ev) capture_event(s, r.ev0);
    const int n_total = r.nchains * r.nspc;
    chr_plan_launch_hooked(r.plan, r.y0, r.samples, s, ev ? event_after_chr : nullptr, &r,
                           r.fused_series ? diag_workspace_series(r.diag_ws) : nullptr,
                           diag_series_stride(n_total, r.nchains));
    diag_launch_ws(r.diag_ws, r.samples, r.d, n_total, r.nchains, NAJA_LAYOUT_INTERLEAVED, r.ess, r.rhat, s, r.nv,
                   r.ess_min, r.rhat_max, r.fused_series);
}

// Copies the rounding status back and raises the first variant's error.
void check_status(naja_cond_run& r, hipStream_t s)
{
    NAJA_HIP(hipMemcpyAsync(r.status_h, r.status, sizeof(int) * r.nv, hipMemcpyDeviceToHost, s));
    NAJA_HIP(hipStreamSynchronize(s));
    for (int v = 0; v < r.nv; ++v) {
        NAJA_REQUIRE(r.status_h[v] != NAJA_ERR_INVALID_ARGUMENT, NAJA_ERR_INVALID_ARGUMENT,
                     "conditioned run: variant %d's start is not interior", v);
        NAJA_REQUIRE(r.status_h[v] == NAJA_OK, NAJA_ERR_RUNTIME,
                     "conditioned run: variant %d has no independent facet completion", v);
    }
}

void release(naja_cond_run* r)
{
    if (!r) return;
    if (r->exec) (void)hipGraphExecDestroy(r->exec);
    if (r->graph) (void)hipGraphDestroy(r->graph);
    if (r->plan) (void)naja_chr_plan_destroy(r->plan);
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    if (r->ev0) (void)hipEventDestroy(r->ev0);
    if (r->ev1) (void)hipEventDestroy(r->ev1);
    if (r->status_h) (void)hipHostFree(r->status_h);
    if (r->dev_mem) (void)hipFree(r->dev_mem);
    delete r;
}

}  // namespace
}  // namespace naja

extern "C" int naja_cond_run_create(const float* A0, const double* b0, int m0, int d, int n_variants,
                                    const int* extra_rows, const float* extra_A, const double* extra_b, float theta,
                                    float exchange_tol, const naja_chr_params* p, const float* y0, const float* c,
                                    float* P, float* samples, float* ess, float* rhat, float* ess_min,
                                    float* rhat_max, int timed, void* stream, naja_cond_run** out)
{
    using namespace naja;
    naja_cond_run* r = nullptr;
    const int rc = guarded([&] {
        NAJA_REQUIRE(out, NAJA_ERR_INVALID_ARGUMENT, "NULL output");
        *out = nullptr;
        NAJA_REQUIRE(A0 && b0 && extra_rows && p && y0 && c && P && samples && ess && rhat && ess_min && rhat_max,
                     NAJA_ERR_INVALID_ARGUMENT, "naja_cond_run_create: NULL argument");
        NAJA_REQUIRE(m0 > 0 && d > 0 && n_variants > 0, NAJA_ERR_INVALID_ARGUMENT, "invalid shape");
        NAJA_REQUIRE(exchange_tol == 0.0f || exchange_tol >= 1.0f, NAJA_ERR_INVALID_ARGUMENT,
                     "exchange tolerance %g (0 = off, else >= 1)", (double)exchange_tol);
        NAJA_REQUIRE(p->nchains > 0 && p->nspc > 0, NAJA_ERR_INVALID_ARGUMENT, "nchains and nspc must be > 0");
        NAJA_REQUIRE((int64_t)p->nchains * p->nspc < ((int64_t)1 << 31), NAJA_ERR_UNSUPPORTED,
                     "too many samples per variant");
        r = new naja_cond_run();
        r->m0 = m0;
        r->d = d;
        r->nv = n_variants;
        r->nchains = p->nchains;
        r->nspc = p->nspc;
        r->A0 = A0;
        r->b0 = b0;
        r->Ae = extra_A;
        r->be = extra_b;
        r->y0 = y0;
        r->c = c;
        r->P = P;
        r->samples = samples;
        r->ess = ess;
        r->rhat = rhat;
        r->ess_min = ess_min;
        r->rhat_max = rhat_max;
        r->exchange_tol = exchange_tol;
        r->timed = timed != 0;
        r->stream = reinterpret_cast<hipStream_t>(stream);
        std::vector<int64_t> off((size_t)n_variants);
        int64_t tot = 0;
        int kmax = 0;
        for (int v = 0; v < n_variants; ++v) {
            NAJA_REQUIRE(extra_rows[v] >= 0, NAJA_ERR_INVALID_ARGUMENT, "variant %d has %d extra rows", v,
                         extra_rows[v]);
            off[(size_t)v] = tot;
            tot += extra_rows[v];
            kmax = std::max(kmax, extra_rows[v]);
        }
        NAJA_REQUIRE(tot == 0 || (extra_A && extra_b), NAJA_ERR_INVALID_ARGUMENT, "extra rows without extra_A/extra_b");
        facet_check_shape(m0, d, kmax, 0, NAJA_FACET_COMPLETE_BASE_FACETS);
        const int n_total = p->nchains * p->nspc;
        NAJA_REQUIRE(p->nspc <= diag_max_chain_length(), NAJA_ERR_UNSUPPORTED,
                     "conditioned run: chains of %d samples exceed the LDS transform (%d)", p->nspc,
                     diag_max_chain_length());
        const size_t b_off = r256(sizeof(int64_t) * n_variants), b_k = r256(sizeof(int) * n_variants),
                     b_th = r256(sizeof(float) * n_variants);
        const size_t b_ws = diag_workspace_bytes(d, n_total, p->nchains, n_variants);
        NAJA_HIP(hipMalloc(&r->dev_mem, b_off + 2 * b_k + b_th + b_ws));
        r->eoff = reinterpret_cast<int64_t*>(r->dev_mem);
        r->ek = reinterpret_cast<int*>(r->dev_mem + b_off);
        r->status = reinterpret_cast<int*>(r->dev_mem + b_off + b_k);
        r->theta = reinterpret_cast<float*>(r->dev_mem + b_off + 2 * b_k);
        r->diag_ws = r->dev_mem + b_off + 2 * b_k + b_th;
        NAJA_HIP(hipHostMalloc(reinterpret_cast<void**>(&r->status_h), sizeof(int) * n_variants, hipHostMallocDefault));
        const std::vector<float> th((size_t)n_variants, theta);
        NAJA_HIP(hipMemcpy(r->eoff, off.data(), sizeof(int64_t) * n_variants, hipMemcpyHostToDevice));
        NAJA_HIP(hipMemcpy(r->ek, extra_rows, sizeof(int) * n_variants, hipMemcpyHostToDevice));
        NAJA_HIP(hipMemcpy(r->theta, th.data(), sizeof(float) * n_variants, hipMemcpyHostToDevice));
        diag_workspace_init(r->diag_ws, d, n_total, p->nchains, n_variants);
        if (r->timed) {
            NAJA_HIP(hipEventCreate(&r->ev0));
            NAJA_HIP(hipEventCreate(&r->ev1));
        }
        const hipStream_t s = r->stream;
        // First step eagerly: the roundings, then the plan packed and
        // transformed with them (creation synchronises the stream), the
        // launch and the diagnostics; every status checked.
        facet_launch(A0, b0, m0, d, n_variants, extra_A, extra_b, r->eoff, r->ek, nullptr, nullptr, r->theta, 0,
                     NAJA_FACET_COMPLETE_BASE_FACETS, exchange_tol, nullptr, P, r->status, s);
        check_status(*r, s);
        check_plan(naja_chr_plan_create(A0, b0, m0, d, n_variants, extra_rows, extra_A, extra_b, P, c, p, stream,
                                        &r->plan));
        r->fused_series = chr_plan_series_ok(r->plan, samples);
        check_plan(naja_chr_plan_launch(r->plan, y0, samples, nullptr, stream));
        diag_launch_ws(r->diag_ws, samples, d, n_total, p->nchains, NAJA_LAYOUT_INTERLEAVED, ess, rhat, s, n_variants,
                       ess_min, rhat_max);
        NAJA_HIP(hipStreamSynchronize(s));
        // Capture: the same work, rounding included.
        NAJA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_step(*r, s, true);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        NAJA_HIP(hipStreamEndCapture(s, &r->graph));
        NAJA_HIP(hipGraphInstantiate(&r->exec, r->graph, nullptr, nullptr, 0));
        *out = r;
    });
    if (rc != NAJA_OK) {  // the clean-up's own calls reset the error state: keep the failure's message
        const std::string msg = naja_last_error();
        naja::release(r);
        naja::set_error("%s", msg.c_str());
    }
    return rc;
}

extern "C" int naja_cond_run_step(naja_cond_run* run, void* stream)
{
    using namespace naja;
    return guarded([&] {
        NAJA_REQUIRE(run && run->exec, NAJA_ERR_INVALID_ARGUMENT, "NULL run");
        const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        NAJA_HIP(hipGraphLaunch(run->exec, s));
        check_status(*run, s);
        if (run->timed) NAJA_HIP(hipEventElapsedTime(&run->last_ms, run->ev0, run->ev1));
    });
}

extern "C" int naja_cond_run_sampling_ms(const naja_cond_run* run, float* ms)
{
    using namespace naja;
    return guarded([&] {
        NAJA_REQUIRE(run && ms, NAJA_ERR_INVALID_ARGUMENT, "NULL run or output");
        NAJA_REQUIRE(run->timed, NAJA_ERR_INVALID_ARGUMENT, "the run was created without timing events");
        *ms = run->last_ms;
    });
}

extern "C" int naja_cond_run_destroy(naja_cond_run* run)
{
    naja::release(run);
    return NAJA_OK;
}
