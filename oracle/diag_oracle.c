/*
 * ORACLE (test infrastructure only). CPU restatement of naja's MCMC
 * diagnostics, src/validate/mcmc.cpp:17-194, used by tests/ and bench.py's
 * CPU baseline as the checker; the product never links it.
 *
 *  - ESS (mcmc.cpp:60-76, :91-124): per (dim, chain) series of length n;
 *    n < 4 -> n; mean summed in double, cast to float (mcmc.cpp:25-27);
 *    centred in float; biased autocovariance acov[k] = (1/n) sum x_t x_{t+k}
 *    (mcmc.cpp:17-57, which obtains it from a float kissfft of size 2n);
 *    acov[0] < 1e-30 -> 1; tau = 1 + 2 sum_{k>=1, k<n/2} (rho_{2k-1}+rho_{2k})
 *    stopping before the first negative pair, in float; tau >= 1; ESS = n/tau;
 *    per dim the minimum over chains.
 *    Here the autocovariance is computed in double (radix-2 FFT of the next
 *    power of two >= 2n, or directly for small n) and rounded to float;
 *    oracle/ref_ess_kiss.c runs the same rule on the reference's own kissfft.
 *  - split-R-hat (mcmc.cpp:126-194): sequential FP64 moments per half chain,
 *    Bessel-corrected, B, W, var_hat, R = sqrt(var_hat/W) (1 if W <= 1e-30).
 *  - percentiles (mcmc.cpp:78-87): linear interpolation on sorted values.
 * Chain grouping: layout 0 reads chain c as rows s*nchains + c (how
 * samples.npy is written, npy.h:96-141 with chr.cu:460); layout 1 reads rows
 * [c*spc, (c+1)*spc) as the reference validator does (mcmc.cpp:103-110).
 */
#include <math.h>
#include <float.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static void orc_fft_c2c(double* re, double* im, int n, int inverse)
{
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const double ang = (inverse ? 2.0 : -2.0) * M_PI / (double)len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < len / 2; ++k) {
                const double wr = cos(ang * k), wi = sin(ang * k);
                const int a = i + k, b = i + k + len / 2;
                const double xr = re[b] * wr - im[b] * wi;
                const double xi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - xr; im[b] = im[a] - xi;
                re[a] += xr; im[a] += xi;
            }
        }
    }
}

/* acov[0..n-1] of the centred float series xc (double accumulate, float out). */
void orc_autocov(const float* xc, int n, float* acov, int direct)
{
    if (direct) {
        for (int k = 0; k < n; ++k) {
            double acc = 0.0;
            for (int t = 0; t + k < n; ++t) acc += (double)xc[t] * (double)xc[t + k];
            acov[k] = (float)(acc / n);
        }
        return;
    }
    int N = 1;
    while (N < 2 * n) N <<= 1;
    double* re = (double*)calloc((size_t)N, sizeof(double));
    double* im = (double*)calloc((size_t)N, sizeof(double));
    for (int i = 0; i < n; ++i) re[i] = (double)xc[i];
    orc_fft_c2c(re, im, N, 0);
    for (int i = 0; i < N; ++i) { re[i] = re[i] * re[i] + im[i] * im[i]; im[i] = 0.0; }
    orc_fft_c2c(re, im, N, 1);
    for (int k = 0; k < n; ++k) acov[k] = (float)(re[k] / ((double)N * (double)n));
    free(re); free(im);
}

/* mcmc.cpp:25-30: mean in double, cast to float, centre in float. */
void orc_center(const float* x, int n, float* xc)
{
    double sum = 0.0;
    for (int i = 0; i < n; ++i) sum += x[i];
    const float mean = (float)(sum / n);
    for (int i = 0; i < n; ++i) xc[i] = x[i] - mean;
}

/* mcmc.cpp:60-76: truncated pair-sum estimator from a float autocovariance. */
float orc_ess_from_acov(const float* acov, int n)
{
    if (n < 4) return (float)n;
    if (acov[0] < 1e-30f) return 1.0f;
    const float var = acov[0];
    float tau = 1.0f;
    for (int k = 1; k < n / 2; ++k) {
        const float pair_sum = acov[2 * k - 1] / var + acov[2 * k] / var;
        if (pair_sum < 0.0f) break;
        tau += 2.0f * pair_sum;
    }
    if (tau < 1.0f) tau = 1.0f;
    return (float)n / tau;
}

float orc_ess_one_chain(const float* x, int n, int direct)
{
    if (n < 4) return (float)n;
    float* xc = (float*)malloc(sizeof(float) * n);
    float* acov = (float*)malloc(sizeof(float) * n);
    orc_center(x, n, xc);
    orc_autocov(xc, n, acov, direct);
    const float e = orc_ess_from_acov(acov, n);
    free(xc); free(acov);
    return e;
}

static int64_t orc_row(int layout, int s, int c, int nchains, int spc)
{
    return layout == 0 ? (int64_t)s * nchains + c : (int64_t)c * spc + s;
}

/*
 * samples: n_total rows of dim floats (row-major, as in samples.npy).
 * ess_chain (optional): dim x nchains per-chain ESS. Returns 0 or -1.
 */
int orc_ess(const float* samples, int dim, int n_total, int nchains, int layout,
            float* ess, float* ess_chain, int direct)
{
    const int spc = n_total / nchains;
    if (spc < 4) return -1;
    float* x = (float*)malloc(sizeof(float) * spc);
    for (int d = 0; d < dim; ++d) {
        float mn = FLT_MAX;
        for (int c = 0; c < nchains; ++c) {
            for (int s = 0; s < spc; ++s) x[s] = samples[orc_row(layout, s, c, nchains, spc) * dim + d];
            const float e = orc_ess_one_chain(x, spc, direct);
            if (ess_chain) ess_chain[(int64_t)d * nchains + c] = e;
            if (e < mn) mn = e;
        }
        ess[d] = mn;
    }
    free(x);
    return 0;
}

int orc_split_rhat(const float* samples, int dim, int n_total, int nchains, int layout,
                   float* rhat, int* above)
{
    const int spc = n_total / nchains;
    const int half = spc / 2;
    if (half < 2) return -1;
    const int mh = 2 * nchains;
    double* means = (double*)malloc(sizeof(double) * mh);
    double* vars = (double*)malloc(sizeof(double) * mh);
    above[0] = above[1] = 0;
    for (int d = 0; d < dim; ++d) {
        double grand = 0.0;
        for (int c = 0; c < nchains; ++c) {
            for (int h = 0; h < 2; ++h) {
                double s = 0.0, s2 = 0.0;
                for (int i = 0; i < half; ++i) {
                    const double v = (double)samples[orc_row(layout, h * half + i, c, nchains, spc) * dim + d];
                    s += v;
                    s2 += v * v;
                }
                const double mean = s / half;
                double var = (s2 / half) - mean * mean;
                var = var * half / (half - 1);
                means[2 * c + h] = mean;
                vars[2 * c + h] = var;
                grand += mean;
            }
        }
        const double gm = grand / mh;
        double B = 0.0;
        for (int j = 0; j < mh; ++j) { const double df = means[j] - gm; B += df * df; }
        B *= (double)half / (mh - 1);
        double W = 0.0;
        for (int j = 0; j < mh; ++j) W += vars[j];
        W /= mh;
        const double var_hat = ((half - 1.0) / half) * W + B / half;
        const double r = (W > 1e-30) ? sqrt(var_hat / W) : 1.0;
        rhat[d] = (float)r;
        if (r > 1.1) ++above[0];
        if (r > 1.2) ++above[1];
    }
    free(means); free(vars);
    return 0;
}

static int orc_cmpf(const void* a, const void* b)
{
    const float x = *(const float*)a, y = *(const float*)b;
    return (x > y) - (x < y);
}

/* mcmc.cpp:78-87 */
float orc_percentile(const float* v, int n, float q)
{
    if (n <= 0) return 0.0f;
    float* w = (float*)malloc(sizeof(float) * n);
    memcpy(w, v, sizeof(float) * n);
    qsort(w, (size_t)n, sizeof(float), orc_cmpf);
    const float pos = q * (float)(n - 1);
    const int lo = (int)floorf(pos), hi = (int)ceilf(pos);
    float r;
    if (lo == hi) r = w[lo];
    else { const float t = pos - (float)lo; r = w[lo] * (1.0f - t) + w[hi] * t; }
    free(w);
    return r;
}
