/*
 * ORACLE cross-check (test infrastructure only). Linked, by oracle/Makefile's
 * `ref` target, against the reference's own vendored kissfft sources
 * (extern/kissfft/kiss_fft.c, extern/kissfft/kiss_fftr.c, compiled in place
 * under /root/reference) into oracle/_ref/libref_kiss.so.
 *
 * Restates the reference's autocovariance exactly as mcmc.cpp:17-57 performs
 * it: float kiss real FFT of size 2n, |F|^2 in float, inverse real FFT,
 * times 1/(2n * n) in float. Tests compare the oracle's double-precision
 * autocovariance (diag_oracle.c) and its ESS against this.
 */
#include <stdlib.h>
#include <string.h>

#include "kiss_fftr.h"

int ref_kiss_autocov(const float* data, int n, float* acov)
{
    const int nfft = 2 * n;
    kiss_fftr_cfg fwd = kiss_fftr_alloc(nfft, 0, NULL, NULL);
    kiss_fftr_cfg inv = kiss_fftr_alloc(nfft, 1, NULL, NULL);
    if (!fwd || !inv) return -1;
    double sum = 0.0;
    for (int i = 0; i < n; ++i) sum += data[i];
    const float mean = (float)(sum / n);
    float* buf = (float*)calloc((size_t)nfft, sizeof(float));
    for (int i = 0; i < n; ++i) buf[i] = data[i] - mean;
    kiss_fft_cpx* freq = (kiss_fft_cpx*)malloc(sizeof(kiss_fft_cpx) * (size_t)(nfft / 2 + 1));
    kiss_fftr(fwd, buf, freq);
    for (int k = 0; k <= nfft / 2; ++k) {
        const float re = freq[k].r, im = freq[k].i;
        freq[k].r = re * re + im * im;
        freq[k].i = 0.0f;
    }
    float* raw = (float*)malloc(sizeof(float) * (size_t)nfft);
    kiss_fftri(inv, freq, raw);
    const float norm = 1.0f / ((float)nfft * (float)n);
    for (int k = 0; k < n; ++k) acov[k] = raw[k] * norm;
    free(buf); free(freq); free(raw);
    kiss_fftr_free(fwd);
    kiss_fftr_free(inv);
    return 0;
}
