// Affine-hull reduction of a polytope at its start point
// (job_sampling.cu:196-235).
//
// Rows whose slack at x0 is at most `tol` form the equality set A_eq. A basis
// B (n x k) of ker(A_eq) is taken from a full-pivot LU with a relative pivot
// threshold (the reference uses Eigen's FullPivLU with setThreshold(1e-10)
// and its kernel()). Sampling then runs on z in R^k with
//   A' = A B,  b' = b - A x0,  z0 = 0,   and  x = x0 + B z.
//
// LU rule (Eigen FullPivLU::computeInPlace semantics, restated):
//   step k: pivot = entry of largest |.| in the trailing block; ties go to the
//   first entry in column-major order (smallest column, then smallest row);
//   stop when that largest value is exactly 0. Swap rows k/p and columns k/q,
//   divide the column below the pivot by it, rank-1 update of the trailing
//   block. maxpivot = largest |pivot| seen.
// Rank: pivots with |u_kk| > threshold * maxpivot (strict).
// Kernel basis: one column per non-pivot (column-permuted) index f, in
//   increasing order: x[q[f]] = 1, other free entries 0, and the pivot
//   entries solve the upper-triangular system U_pp z = -U_pf by back
//   substitution (column-oriented: divide, then axpy into the rows above).
// All host FP64; nothing here touches the GPU.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"

namespace {

struct FullPivLU {
    int rows = 0, cols = 0, nonzero = 0;
    double maxpivot = 0.0;
    std::vector<double> lu;  // row-major rows x cols: strict lower = L, upper = U
    std::vector<int> q;      // permuted column j holds original column q[j]

    double& at(int i, int j) { return lu[(size_t)i * cols + j]; }
    double at(int i, int j) const { return lu[(size_t)i * cols + j]; }

    FullPivLU(const double* a, int r, int c) : rows(r), cols(c), lu(a, a + (size_t)r * c), q(c)
    {
        for (int j = 0; j < c; ++j) q[j] = j;
        const int size = std::min(r, c);
        nonzero = size;
        for (int k = 0; k < size; ++k) {
            double best = -1.0;
            int bi = k, bj = k;
            for (int i = k; i < rows; ++i)
                for (int j = k; j < cols; ++j) {
                    const double v = std::fabs(at(i, j));
                    if (v > best || (v == best && (j < bj || (j == bj && i < bi)))) {
                        best = v;
                        bi = i;
                        bj = j;
                    }
                }
            if (best == 0.0) {
                nonzero = k;
                break;
            }
            if (best > maxpivot) maxpivot = best;
            if (bi != k)
                for (int j = 0; j < cols; ++j) std::swap(at(k, j), at(bi, j));
            if (bj != k) {
                for (int i = 0; i < rows; ++i) std::swap(at(i, k), at(i, bj));
                std::swap(q[k], q[bj]);
            }
            const double piv = at(k, k);
            for (int i = k + 1; i < rows; ++i) at(i, k) /= piv;
            for (int i = k + 1; i < rows; ++i) {
                const double l = at(i, k);
                double* ri = &at(i, 0);
                const double* rk = &at(k, 0);
                for (int j = k + 1; j < cols; ++j) ri[j] -= l * rk[j];
            }
        }
    }

    // Row indices k whose pivot passes the threshold, increasing.
    std::vector<int> pivots(double threshold) const
    {
        std::vector<int> p;
        const double thr = maxpivot * threshold;
        for (int k = 0; k < nonzero; ++k)
            if (std::fabs(at(k, k)) > thr) p.push_back(k);
        return p;
    }

    // Kernel basis, row-major cols x dimker.
    std::vector<double> kernel(double threshold, int* dimker_out) const
    {
        const std::vector<int> piv = pivots(threshold);
        const int r = (int)piv.size(), dimker = cols - r;
        std::vector<char> is_piv(cols, 0);
        for (int k : piv) is_piv[k] = 1;
        std::vector<int> fr;
        for (int j = 0; j < cols; ++j)
            if (!is_piv[j]) fr.push_back(j);
        std::vector<double> B((size_t)cols * dimker, 0.0), z(r);
        for (int f = 0; f < dimker; ++f) {
            const int col = fr[f];
            // rhs_t = -U[piv[t], col]; the U part of row piv[t] starts at column piv[t]
            for (int t = 0; t < r; ++t) z[t] = col >= piv[t] ? -at(piv[t], col) : 0.0;
            for (int t = r - 1; t >= 0; --t) {
                z[t] /= at(piv[t], piv[t]);
                for (int u = 0; u < t; ++u) z[u] -= at(piv[u], piv[t]) * z[t];
            }
            for (int t = 0; t < r; ++t) B[(size_t)q[piv[t]] * dimker + f] = z[t];
            B[(size_t)q[col] * dimker + f] = 1.0;
        }
        *dimker_out = dimker;
        return B;
    }
};

}  // namespace

using namespace naja;

extern "C" int naja_kernel_basis(const double* Aeq, int r, int n, double threshold, double* basis, int* dim_out)
{
    return guarded([&]() {
        NAJA_REQUIRE(r >= 0 && n >= 1 && dim_out && basis && (r == 0 || Aeq) && threshold >= 0.0,
                     NAJA_ERR_INVALID_ARGUMENT, "naja_kernel_basis: need r >= 0, n >= 1, threshold >= 0");
        if (r == 0) {
            std::fill(basis, basis + (size_t)n * n, 0.0);
            for (int i = 0; i < n; ++i) basis[(size_t)i * n + i] = 1.0;
            *dim_out = n;
            return;
        }
        FullPivLU lu(Aeq, r, n);
        int k = 0;
        const std::vector<double> B = lu.kernel(threshold, &k);
        std::copy(B.begin(), B.end(), basis);
        *dim_out = k;
    });
}

extern "C" int naja_affine_hull(const double* A, const double* b, const double* x0, int m, int n, double tol,
                                double threshold, double* basis, int* dim_out, int* n_tight_out)
{
    return guarded([&]() {
        NAJA_REQUIRE(m >= 1 && n >= 1 && A && b && x0 && basis && dim_out && tol >= 0.0 && threshold >= 0.0,
                     NAJA_ERR_INVALID_ARGUMENT, "naja_affine_hull: need m, n >= 1, tol >= 0, threshold >= 0");
        std::vector<double> Aeq;
        int r = 0;
        for (int i = 0; i < m; ++i) {
            const double* a = A + (size_t)i * n;
            double ax = 0.0;
            for (int j = 0; j < n; ++j) ax += a[j] * x0[j];
            if (b[i] - ax <= tol) {
                Aeq.insert(Aeq.end(), a, a + n);
                ++r;
            }
        }
        if (n_tight_out) *n_tight_out = r;
        const int rc = naja_kernel_basis(Aeq.data(), r, n, threshold, basis, dim_out);
        if (rc != NAJA_OK) throw Error(rc);
    });
}
