/*
 * ORACLE (test infrastructure only). FP64 restatement of naja's backmap,
 * src/gpu/backmap.cu:21-48: V = T * Y + shift * 1^T (cublasDgemm followed by
 * add_shift_cols_kernel). Layouts follow the product boundary:
 *   T     : n x d row-major (row r = reaction r, as the CSV stores it)
 *   Y     : ns x d row-major (sample rows, as samples.npy stores them)
 *   V     : ns x n row-major (backmapped sample rows)
 * Shape errors return -1, mirroring backmap.cu:27-35's invalid_argument.
 */
#include <stdint.h>

int orc_backmap(const double* T, const double* shift, const float* Y, double* V,
                int n, int d, int64_t ns)
{
    if (n <= 0 || d <= 0 || ns < 0) return -1;
    for (int64_t s = 0; s < ns; ++s) {
        const float* y = Y + s * d;
        for (int r = 0; r < n; ++r) {
            const double* t = T + (int64_t)r * d;
            double acc = 0.0;
            for (int k = 0; k < d; ++k) acc += t[k] * (double)y[k];
            V[s * n + r] = acc + shift[r];
        }
    }
    return 0;
}
