/*
 * ORACLE (test infrastructure only). Scalar CPU restatement of naja's
 * coordinate hit-and-run step, used by tests/, the smoke check and the CPU
 * baseline leg of bench.py as the checker. Nothing in the product library
 * links or calls this file.
 *
 * Step rule follows the reference kernel chrKernel, chr.cu:281-482
 * (also restated by the reference in tests/test_chr_pair_steps.cpp:17-101):
 *   1. draw the move (chr.cu:285-374): k-sparse / pair / coordinate by one
 *      uniform, then the coordinate i, then the partner j or pair index;
 *   2. optional resync s = b - A x every resync_interval steps (chr.cu:386-395);
 *   3. per row ae = (A u)_r and the chord ends from the rows' inverse
 *      distances ae/s (chr.cu:397-434);
 *   4. alpha uniform on [lo, hi], x += alpha u (chr.cu:435-453);
 *   5. save x when t % thin == 0 into column (t/thin)*nchains + chain
 *      (chr.cu:283-284, chr.cu:458-461);
 *   6. s -= alpha ae (chr.cu:464-482).
 * Deliberate, documented differences from the reference (shared with the
 * HIP kernel so the two agree bit for bit):
 *   - Philox4x32-10 counter stream (philox.h) replaces cuRAND MRG32k3a;
 *     chains of one group (chains_per_group G) share the move draw;
 *   - FP32 state (x, slack, A) with explicit fmaf and IEEE division; the
 *     slack is computed in FP64 as s = b - A x at t = 0 and on every resync
 *     (the reference computes the initial slack by DGEMM, chr.cu:34-40);
 *     resync_interval 0 selects an automatic window (orc_effective_resync);
 *   - the chord is computed in the direct form q_r = s_r / ae_r (IEEE
 *     division; for a coordinate move q_r = RN(s_r * RN(1/A[r,i])), the
 *     reciprocal being a fixed property of A): hi = the least q with the
 *     sign bit clear, lo = the
 *     greatest q with the sign bit set, found as an unsigned and a signed
 *     integer minimum over the float bit patterns; alpha =
 *     fmaf(U, hi - lo, lo). In exact arithmetic hi = 1/max(ae/s) and
 *     lo = 1/min(ae/s), the reference's ends; here each end is one rounding
 *     of the exact value instead of two. ae = 0 gives +-inf and 0/0 NaN,
 *     neither of which can win a minimum against a real facet;
 *   - a slack that FP32 rounding leaves <= 0, after an update or a resync,
 *     is stored as +0: a point just outside a facet is treated as on it
 *     (q = +-0 pins that end of the chord at the current point), so the
 *     chain cannot run through the violated facet. The FP64 reference has
 *     no clamp.
 * The FP64 variant (orc_chr_run_f64) keeps every quantity in double on the
 * same random stream; it measures FP32 drift.
 */
#include <math.h>
#include <float.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "philox.h"

#define ORC_MOVE_COORD 0
#define ORC_MOVE_PAIR 1
#define ORC_MOVE_KSPARSE 2
#define ORC_KMAX 32

typedef struct {
    int type;
    int i, j;
    float ci, cj;
    double ci64, cj64;
    int k;
    int idx[ORC_KMAX];
    float coef[ORC_KMAX];
    double coef64[ORC_KMAX];
} orc_move;

typedef struct {
    int m, d, nchains, nspc, thin;
    uint64_t seed;
    uint32_t variant;
    double pair_prob;
    int pair_mode;
    const int* pair_i;
    const int* pair_j;
    const float* pair_c;
    const float* pair_s;
    int n_pairs;
    double ksparse_prob;
    int ksparse_k;
    int resync_interval;
    int chains_per_group;
} orc_params;

/* Move draw, chr.cu:285-374 restated on the Philox layout of philox.h. */
static void orc_draw_move(const orc_params* P, uint64_t t, uint32_t group, orc_move* mv)
{
    const uint32_t k0 = (uint32_t)P->seed;
    const uint32_t k1 = (uint32_t)(P->seed >> 32) ^ P->variant;
    uint32_t ctr[4] = {(uint32_t)t, (uint32_t)(t >> 32), group, 0u};
    uint32_t w[4];
    orc_philox4x32_10(ctr, k0, k1, w);

    const uint64_t thr_k = orc_threshold(P->ksparse_prob);
    const uint64_t thr_kp = orc_threshold(P->ksparse_prob + P->pair_prob);
    const int want_ks = (P->ksparse_prob > 0.0 && (uint64_t)w[0] < thr_k);
    const int want_pair = (!want_ks && P->pair_prob > 0.0 && (uint64_t)w[0] < thr_kp);

    mv->type = ORC_MOVE_COORD;
    mv->i = (int)orc_index(w[1], (uint32_t)P->d);
    mv->j = -1;
    mv->ci = 1.0f; mv->cj = 0.0f;
    mv->ci64 = 1.0; mv->cj64 = 0.0;
    int k = P->ksparse_k;
    if (k < 1) k = 1;
    if (k > ORC_KMAX) k = ORC_KMAX;
    if (k > P->d) k = P->d;
    mv->k = k;

    if (want_ks) {
        mv->type = ORC_MOVE_KSPARSE;
        const double mag64 = 1.0 / sqrt((double)k);
        const float mag = (float)mag64;
        for (int q = 0; q < k; ++q) {
            uint32_t c2[4] = {(uint32_t)t, (uint32_t)(t >> 32), group, 1u + (uint32_t)(q >> 2)};
            uint32_t v[4];
            orc_philox4x32_10(c2, k0, k1, v);
            mv->idx[q] = (int)orc_index(v[q & 3], (uint32_t)P->d);
            const int neg = (int)((w[3] >> q) & 1u);
            mv->coef[q] = neg ? -mag : mag;
            mv->coef64[q] = neg ? -mag64 : mag64;
        }
    } else if (want_pair && P->d >= 2) {
        if (P->pair_mode == 1) {
            const int r = (int)orc_index(w[2], (uint32_t)(P->d - 1));
            mv->type = ORC_MOVE_PAIR;
            mv->j = (r >= mv->i) ? r + 1 : r;
            mv->ci = 0.70710678118654752440f;
            mv->cj = -0.70710678118654752440f;
            mv->ci64 = 0.70710678118654752440;
            mv->cj64 = -0.70710678118654752440;
        } else if (P->pair_mode == 2 && P->n_pairs > 0 && P->pair_i && P->pair_j && P->pair_c && P->pair_s) {
            const int p = (int)orc_index(w[2], (uint32_t)P->n_pairs);
            const float c = P->pair_c[p], s = P->pair_s[p];
            mv->type = ORC_MOVE_PAIR;
            mv->i = P->pair_i[p];
            mv->j = P->pair_j[p];
            if ((w[3] & 1u) == 0u) { mv->ci = c; mv->cj = s; }
            else { mv->ci = -s; mv->cj = c; }
            if ((w[3] >> 1) & 1u) { mv->ci = -mv->ci; mv->cj = -mv->cj; }
            mv->ci64 = (double)mv->ci;
            mv->cj64 = (double)mv->cj;
        }
    }
}

static float orc_alpha_uniform(const orc_params* P, uint64_t t, uint32_t chain)
{
    const uint32_t k0 = (uint32_t)P->seed;
    const uint32_t k1 = (uint32_t)(P->seed >> 32) ^ P->variant;
    uint32_t ctr[4] = {(uint32_t)t, (uint32_t)(t >> 32), chain >> 2, ORC_ALPHA_BLOCK};
    uint32_t w[4];
    orc_philox4x32_10(ctr, k0, k1, w);
    return orc_uniform(w[chain & 3u]);
}

/* ---------------------------------------------------------------- FP32 */

typedef struct {
    const orc_params* P;
    const float* A;      /* column-major m x d */
    const double* b;
    const float* y0;     /* d x nchains */
    float* samples;      /* d x (nchains*nspc) */
    float* y_final;      /* d x nchains, optional */
    float* s_final;      /* m x nchains, optional */
    int32_t* move_log;   /* nlog x 3 (type,i,j) of this chain, optional */
    int64_t nlog;
    int c_begin, c_end;
} orc_job32;

/* s <= 0 (and -0) -> +0 as a signed integer max over the bits; a NaN with the
   sign clear passes and never wins a minimum. Applied whenever a slack is
   formed (resync and update), so stored slacks are never negative. */
static float orc_clamp_slack(float v)
{
    int32_t bits;
    memcpy(&bits, &v, 4);
    if (bits < 0) bits = 0;
    float out;
    memcpy(&out, &bits, 4);
    return out;
}

static void orc_resync32(const float* A, const double* b, const float* x, float* s, int m, int d)
{
    for (int r = 0; r < m; ++r) {
        double acc = 0.0;
        for (int c = 0; c < d; ++c) acc = fma((double)A[(size_t)c * m + r], (double)x[c], acc);
        s[r] = orc_clamp_slack((float)(b[r] - acc));
    }
}

static void* orc_chain_worker32(void* arg)
{
    orc_job32* J = (orc_job32*)arg;
    const orc_params* P = J->P;
    const int m = P->m, d = P->d;
    const int G = P->chains_per_group > 0 ? P->chains_per_group : 1;
    float* x = (float*)malloc(sizeof(float) * d);
    float* s = (float*)malloc(sizeof(float) * m);
    float* ae = (float*)malloc(sizeof(float) * m);
    orc_move mv;
    const int64_t nsteps = (int64_t)P->nspc * P->thin;
    for (int c = J->c_begin; c < J->c_end; ++c) {
        memcpy(x, J->y0 + (size_t)c * d, sizeof(float) * d);
        const uint32_t group = (uint32_t)(c / G);
        for (int64_t t = 0; t < nsteps; ++t) {
            orc_draw_move(P, (uint64_t)t, group, &mv);
            if (J->move_log && c == 0 && t < J->nlog) {
                J->move_log[t * 3 + 0] = mv.type;
                J->move_log[t * 3 + 1] = mv.i;
                J->move_log[t * 3 + 2] = mv.j;
            }
            if (t == 0 || (P->resync_interval > 0 && t % P->resync_interval == 0))
                orc_resync32(J->A, J->b, x, s, m, d);
            uint32_t hi_bits = 0xFFFFFFFFu;   /* unsigned min over the bits of q */
            int32_t lo_bits = 0x7FFFFFFF;      /* signed min over the bits of q */
            const float* Ai = J->A + (size_t)mv.i * m;
            const float* Aj = (mv.type == ORC_MOVE_PAIR) ? J->A + (size_t)mv.j * m : NULL;
            for (int r = 0; r < m; ++r) {
                float a;
                if (mv.type == ORC_MOVE_COORD) {
                    a = Ai[r];
                } else if (mv.type == ORC_MOVE_PAIR) {
                    a = fmaf(Ai[r], mv.ci, Aj[r] * mv.cj);
                } else {
                    a = 0.0f;
                    for (int q = 0; q < mv.k; ++q) a = fmaf(J->A[(size_t)mv.idx[q] * m + r], mv.coef[q], a);
                }
                ae[r] = a;
                const float sp = s[r]; /* stored clamped */
                /* coordinate moves: q = RN(s * RN(1/a)); other moves: RN(s / ae) */
                const float q = (mv.type == ORC_MOVE_COORD) ? sp * (1.0f / a) : sp / a;
                uint32_t qb;
                memcpy(&qb, &q, 4);
                if (qb < hi_bits) hi_bits = qb;
                if ((int32_t)qb < lo_bits) lo_bits = (int32_t)qb;
            }
            float lo, hi;
            memcpy(&hi, &hi_bits, 4);
            memcpy(&lo, &lo_bits, 4);
            const float U = orc_alpha_uniform(P, (uint64_t)t, (uint32_t)c);
            const float alpha = fmaf(U, hi - lo, lo);
            if (mv.type == ORC_MOVE_COORD) {
                x[mv.i] = fmaf(alpha, mv.ci, x[mv.i]);
            } else if (mv.type == ORC_MOVE_PAIR) {
                x[mv.i] = fmaf(alpha, mv.ci, x[mv.i]);
                x[mv.j] = fmaf(alpha, mv.cj, x[mv.j]);
            } else {
                for (int q = 0; q < mv.k; ++q) x[mv.idx[q]] = fmaf(alpha, mv.coef[q], x[mv.idx[q]]);
            }
            if (t % P->thin == 0) {
                const int64_t col = (t / P->thin) * P->nchains + c;
                memcpy(J->samples + (size_t)col * d, x, sizeof(float) * d);
            }
            const float na = -alpha;
            for (int r = 0; r < m; ++r) s[r] = orc_clamp_slack(fmaf(na, ae[r], s[r]));
        }
        if (J->y_final) memcpy(J->y_final + (size_t)c * d, x, sizeof(float) * d);
        if (J->s_final) memcpy(J->s_final + (size_t)c * m, s, sizeof(float) * m);
    }
    free(x); free(s); free(ae);
    return NULL;
}

/* ---------------------------------------------------------------- FP64 */

typedef struct {
    const orc_params* P;
    const double* A;
    const double* b;
    const double* y0;
    double* samples;
    int c_begin, c_end;
} orc_job64;

static void* orc_chain_worker64(void* arg)
{
    orc_job64* J = (orc_job64*)arg;
    const orc_params* P = J->P;
    const int m = P->m, d = P->d;
    const int G = P->chains_per_group > 0 ? P->chains_per_group : 1;
    double* x = (double*)malloc(sizeof(double) * d);
    double* s = (double*)malloc(sizeof(double) * m);
    double* ae = (double*)malloc(sizeof(double) * m);
    orc_move mv;
    const int64_t nsteps = (int64_t)P->nspc * P->thin;
    for (int c = J->c_begin; c < J->c_end; ++c) {
        memcpy(x, J->y0 + (size_t)c * d, sizeof(double) * d);
        const uint32_t group = (uint32_t)(c / G);
        for (int64_t t = 0; t < nsteps; ++t) {
            orc_draw_move(P, (uint64_t)t, group, &mv);
            if (t == 0 || (P->resync_interval > 0 && t % P->resync_interval == 0)) {
                for (int r = 0; r < m; ++r) {
                    double acc = 0.0;
                    for (int cc = 0; cc < d; ++cc) acc = fma(J->A[(size_t)cc * m + r], x[cc], acc);
                    s[r] = J->b[r] - acc;
                }
            }
            double hi = INFINITY, lo = -INFINITY;
            for (int r = 0; r < m; ++r) {
                double a;
                if (mv.type == ORC_MOVE_COORD) {
                    a = J->A[(size_t)mv.i * m + r];
                } else if (mv.type == ORC_MOVE_PAIR) {
                    a = fma(J->A[(size_t)mv.i * m + r], mv.ci64, J->A[(size_t)mv.j * m + r] * mv.cj64);
                } else {
                    a = 0.0;
                    for (int q = 0; q < mv.k; ++q) a = fma(J->A[(size_t)mv.idx[q] * m + r], mv.coef64[q], a);
                }
                ae[r] = a;
                const double q = ((s[r] > 0.0) ? s[r] : 0.0) / a;
                if (!signbit(q)) { if (q < hi) hi = q; }
                else if (q > lo) lo = q;
            }
            const double U = (double)orc_alpha_uniform(P, (uint64_t)t, (uint32_t)c);
            const double alpha = fma(U, hi - lo, lo);
            if (mv.type == ORC_MOVE_COORD) {
                x[mv.i] = fma(alpha, mv.ci64, x[mv.i]);
            } else if (mv.type == ORC_MOVE_PAIR) {
                x[mv.i] = fma(alpha, mv.ci64, x[mv.i]);
                x[mv.j] = fma(alpha, mv.cj64, x[mv.j]);
            } else {
                for (int q = 0; q < mv.k; ++q) x[mv.idx[q]] = fma(alpha, mv.coef64[q], x[mv.idx[q]]);
            }
            if (t % P->thin == 0) {
                const int64_t col = (t / P->thin) * P->nchains + c;
                memcpy(J->samples + (size_t)col * d, x, sizeof(double) * d);
            }
            for (int r = 0; r < m; ++r) s[r] = fma(-alpha, ae[r], s[r]);
        }
    }
    free(x); free(s); free(ae);
    return NULL;
}

static double orc_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/*
 * Effective FP64 resync interval. The reference keeps its slack in FP64 and
 * resyncs only on request (chr.cu:386, default 0). With FP32 slack a request
 * of 0 means an automatic window of max(1024, 16 d) steps, a negative value
 * means never (only the initial s = b - A y).
 */
int orc_effective_resync(int requested, int d)
{
    if (requested > 0) return requested;
    if (requested < 0) return 0;
    return d * 16 > 1024 ? d * 16 : 1024;
}

static int orc_check(const orc_params* P)
{
    if (P->m <= 0 || P->d <= 0 || P->nchains <= 0 || P->nspc <= 0 || P->thin <= 0) return -1;
    if (P->pair_mode == 2 && P->n_pairs > 0) {
        for (int p = 0; p < P->n_pairs; ++p)
            if (P->pair_i[p] < 0 || P->pair_i[p] >= P->d || P->pair_j[p] < 0 || P->pair_j[p] >= P->d) return -2;
    }
    return 0;
}

/*
 * Run nchains chains of nspc*thin steps. Threads split the chains. Returns
 * the wall time of the sampling loop (seconds, steady clock) or a negative
 * error code.
 */
double orc_chr_run_f32(const float* A, const double* b, int m, int d,
                       const float* y0, int nchains, int nspc, int thin,
                       uint64_t seed, uint32_t variant,
                       double pair_prob, int pair_mode,
                       const int* pair_i, const int* pair_j, const float* pair_c, const float* pair_s, int n_pairs,
                       double ksparse_prob, int ksparse_k, int resync_interval, int chains_per_group,
                       float* samples, float* y_final, float* s_final,
                       int32_t* move_log, int64_t nlog, int nthreads)
{
    orc_params P = {m, d, nchains, nspc, thin, seed, variant, pair_prob, pair_mode,
                    pair_i, pair_j, pair_c, pair_s, n_pairs, ksparse_prob, ksparse_k,
                    orc_effective_resync(resync_interval, d), chains_per_group};
    int rc = orc_check(&P);
    if (rc) return (double)rc;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > nchains) nthreads = nchains;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    orc_job32* jobs = (orc_job32*)malloc(sizeof(orc_job32) * nthreads);
    const double t0 = orc_now();
    for (int k = 0; k < nthreads; ++k) {
        jobs[k].P = &P; jobs[k].A = A; jobs[k].b = b; jobs[k].y0 = y0;
        jobs[k].samples = samples; jobs[k].y_final = y_final; jobs[k].s_final = s_final;
        jobs[k].move_log = move_log; jobs[k].nlog = nlog;
        jobs[k].c_begin = (int)((int64_t)nchains * k / nthreads);
        jobs[k].c_end = (int)((int64_t)nchains * (k + 1) / nthreads);
        pthread_create(&th[k], NULL, orc_chain_worker32, &jobs[k]);
    }
    for (int k = 0; k < nthreads; ++k) pthread_join(th[k], NULL);
    const double dt = orc_now() - t0;
    free(th); free(jobs);
    return dt;
}

double orc_chr_run_f64(const double* A, const double* b, int m, int d,
                       const double* y0, int nchains, int nspc, int thin,
                       uint64_t seed, uint32_t variant,
                       double pair_prob, int pair_mode,
                       const int* pair_i, const int* pair_j, const float* pair_c, const float* pair_s, int n_pairs,
                       double ksparse_prob, int ksparse_k, int resync_interval, int chains_per_group,
                       double* samples, int nthreads)
{
    orc_params P = {m, d, nchains, nspc, thin, seed, variant, pair_prob, pair_mode,
                    pair_i, pair_j, pair_c, pair_s, n_pairs, ksparse_prob, ksparse_k,
                    orc_effective_resync(resync_interval, d), chains_per_group};
    int rc = orc_check(&P);
    if (rc) return (double)rc;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > nchains) nthreads = nchains;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
    orc_job64* jobs = (orc_job64*)malloc(sizeof(orc_job64) * nthreads);
    const double t0 = orc_now();
    for (int k = 0; k < nthreads; ++k) {
        jobs[k].P = &P; jobs[k].A = A; jobs[k].b = b; jobs[k].y0 = y0; jobs[k].samples = samples;
        jobs[k].c_begin = (int)((int64_t)nchains * k / nthreads);
        jobs[k].c_end = (int)((int64_t)nchains * (k + 1) / nthreads);
        pthread_create(&th[k], NULL, orc_chain_worker64, &jobs[k]);
    }
    for (int k = 0; k < nthreads; ++k) pthread_join(th[k], NULL);
    const double dt = orc_now() - t0;
    free(th); free(jobs);
    return dt;
}

/* Raw Philox block and the index / uniform maps, exported for the KAT tests. */
void orc_philox_block(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4])
{
    orc_philox4x32_10(ctr, k0, k1, out);
}

/* Move stream of one group for steps [0, n): (type, i, j) triples. */
void orc_move_stream(int d, uint64_t seed, uint32_t variant, uint32_t group,
                     double pair_prob, int pair_mode,
                     const int* pair_i, const int* pair_j, const float* pair_c, const float* pair_s, int n_pairs,
                     double ksparse_prob, int ksparse_k, int64_t n, int32_t* out)
{
    orc_params P;
    memset(&P, 0, sizeof(P));
    P.d = d; P.seed = seed; P.variant = variant; P.pair_prob = pair_prob; P.pair_mode = pair_mode;
    P.pair_i = pair_i; P.pair_j = pair_j; P.pair_c = pair_c; P.pair_s = pair_s; P.n_pairs = n_pairs;
    P.ksparse_prob = ksparse_prob; P.ksparse_k = ksparse_k;
    orc_move mv;
    for (int64_t t = 0; t < n; ++t) {
        orc_draw_move(&P, (uint64_t)t, group, &mv);
        out[t * 3 + 0] = mv.type; out[t * 3 + 1] = mv.i; out[t * 3 + 2] = mv.j;
    }
}

/*
 * Rounding transform restated for the tests: A' = A P (FP32, fmaf over k in
 * order), b' = b - A c (FP64 fma chain), A column-major m x d, P row-major
 * d x d. Matches the GPU's transform of the packed blocks element for element.
 */
void orc_transform(const float* A, const double* b, int m, int d, const float* P, const float* c,
                   float* Aout, double* bout)
{
    for (int j = 0; j < d; ++j)
        for (int r = 0; r < m; ++r) {
            float acc = 0.0f;
            for (int k = 0; k < d; ++k) acc = fmaf(A[(size_t)k * m + r], P[(size_t)k * d + j], acc);
            Aout[(size_t)j * m + r] = acc;
        }
    for (int r = 0; r < m; ++r) {
        double acc = 0.0;
        for (int k = 0; k < d; ++k) acc = fma((double)A[(size_t)k * m + r], (double)c[k], acc);
        bout[r] = b[r] - acc;
    }
}
