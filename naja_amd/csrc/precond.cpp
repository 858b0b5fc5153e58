// FP64 rounding preconditioners, on the host.
//
// * Dikin rounding: dikin_precondition (dikin_precondition.cpp:36-183),
//   restated with this library's own symmetric eigensolver (Householder
//   tridiagonalisation + implicit QL) instead of Eigen's
//   SelfAdjointEigenSolver.
// * Dikin-ellipsoid whitening: the full log-barrier Hessian at the centre,
//   H = sum_r a_r a_r^T / s_r^2 over every row, factored as H = L L^T; the
//   map y = c + L^{-T} z sends the Dikin ellipsoid {y : (y-c)^T H (y-c) <= 1}
//   to the unit ball. The pilot chains of the learned preconditioner run in
//   this frame.
// * Low-rank whitening: P = (I + V V^T)^{-1/2} (or the expansion
//   (I + V V^T)^{1/2}) from a d x r factor V, through the r x r eigenproblem
//   of V^T V (the learned preconditioner's output).
//
// All matrices are row-major FP64 in host memory. The sampler's rounding
// transform takes P with y = c + P z, so every routine here returns that P.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "../../include/naja_amd.h"
#include "common.h"

namespace naja {
namespace {

// Symmetric eigendecomposition. a: n x n row-major (read), evals ascending,
// evecs: n x n row-major with eigenvector j in column j.
void sym_eig(int n, const double* a, double* evals, double* evecs)
{
    std::vector<double> A(a, a + (size_t)n * n);
    std::vector<double> Q((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
    std::vector<double> diag(n), off(n, 0.0), v(n), p(n);
    auto at = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };

    // Householder: column k below the subdiagonal is zeroed by
    // H = I - 2 v v^T acting on rows/columns k+1..n-1.
    for (int k = 0; k + 2 < n; ++k) {
        const int len = n - k - 1;
        double nrm2 = 0.0;
        for (int i = 0; i < len; ++i) nrm2 += at(k + 1 + i, k) * at(k + 1 + i, k);
        const double nrm = std::sqrt(nrm2);
        const double x0 = at(k + 1, k);
        if (nrm == 0.0) continue;
        const double alpha = (x0 > 0.0) ? -nrm : nrm;
        for (int i = 0; i < len; ++i) v[i] = at(k + 1 + i, k);
        v[0] -= alpha;
        double vn2 = 0.0;
        for (int i = 0; i < len; ++i) vn2 += v[i] * v[i];
        if (vn2 == 0.0) continue;
        const double inv = 1.0 / std::sqrt(vn2);
        for (int i = 0; i < len; ++i) v[i] *= inv;
        // B <- H B H on the trailing block: p = B v, K = v.p, q = p - K v,
        // B -= 2 (v q^T + q v^T).
        double K = 0.0;
        for (int i = 0; i < len; ++i) {
            double s = 0.0;
            const double* row = &A[(size_t)(k + 1 + i) * n + k + 1];
            for (int j = 0; j < len; ++j) s += row[j] * v[j];
            p[i] = s;
            K += v[i] * s;
        }
        for (int i = 0; i < len; ++i) p[i] -= K * v[i];
        for (int i = 0; i < len; ++i) {
            double* row = &A[(size_t)(k + 1 + i) * n + k + 1];
            for (int j = 0; j < len; ++j) row[j] -= 2.0 * (v[i] * p[j] + p[i] * v[j]);
        }
        at(k + 1, k) = alpha;
        at(k, k + 1) = alpha;
        for (int i = 1; i < len; ++i) at(k + 1 + i, k) = at(k, k + 1 + i) = 0.0;
        // Q <- Q H on columns k+1..n-1.
        for (int r = 0; r < n; ++r) {
            double* row = &Q[(size_t)r * n + k + 1];
            double s = 0.0;
            for (int j = 0; j < len; ++j) s += row[j] * v[j];
            s *= 2.0;
            for (int j = 0; j < len; ++j) row[j] -= s * v[j];
        }
    }
    for (int i = 0; i < n; ++i) diag[i] = at(i, i);
    for (int i = 0; i + 1 < n; ++i) off[i] = at(i + 1, i);
    off[n - 1] = 0.0;

    // Implicit QL with Wilkinson-type shifts on the tridiagonal (diag, off),
    // off[i] coupling i and i+1; plane rotations accumulate into Q's columns.
    // An off-diagonal entry counts as zero when it is negligible against the
    // matrix's scale (not only against its neighbours: near-zero eigenvalues
    // of a rank-deficient matrix would otherwise never deflate).
    const double eps = std::numeric_limits<double>::epsilon();
    double tnorm = 0.0;
    for (int i = 0; i < n; ++i)
        tnorm = std::max(tnorm, std::fabs(diag[i]) + std::fabs(off[i]) + (i > 0 ? std::fabs(off[i - 1]) : 0.0));
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m + 1 < n; ++m) {
                const double dd = std::fabs(diag[m]) + std::fabs(diag[m + 1]);
                if (std::fabs(off[m]) <= eps * dd || std::fabs(off[m]) <= eps * tnorm) break;
            }
            if (m == l) break;
            NAJA_REQUIRE(++iter <= 64, NAJA_ERR_RUNTIME, "symmetric eigensolver did not converge (n=%d)", n);
            double g = (diag[l + 1] - diag[l]) / (2.0 * off[l]);
            double r = std::hypot(g, 1.0);
            g = diag[m] - diag[l] + off[l] / (g + std::copysign(r, g));
            double s = 1.0, c = 1.0, pp = 0.0;
            int i = m - 1;
            bool deflated = false;
            for (; i >= l; --i) {
                const double f = s * off[i];
                const double bb = c * off[i];
                r = std::hypot(f, g);
                off[i + 1] = r;
                if (r == 0.0) {
                    diag[i + 1] -= pp;
                    off[m] = 0.0;
                    deflated = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = diag[i + 1] - pp;
                r = (diag[i] - g) * s + 2.0 * c * bb;
                pp = s * r;
                diag[i + 1] = g + pp;
                g = c * r - bb;
                for (int k = 0; k < n; ++k) {
                    double* qk = &Q[(size_t)k * n];
                    const double fq = qk[i + 1];
                    qk[i + 1] = s * qk[i] + c * fq;
                    qk[i] = c * qk[i] - s * fq;
                }
            }
            if (deflated) continue;
            diag[l] -= pp;
            off[l] = g;
            off[m] = 0.0;
        }
    }
    // Sort ascending, permuting eigenvector columns.
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return diag[x] < diag[y]; });
    for (int j = 0; j < n; ++j) {
        evals[j] = diag[order[j]];
        for (int i = 0; i < n; ++i) evecs[(size_t)i * n + j] = Q[(size_t)i * n + order[j]];
    }
}

// Chord length through x_c along v (dikin_precondition.cpp:13-32): rows
// with |A v| < 1e-15 are skipped; an empty interval gives 0.
double chord_along(int m, int d, const double* A, const double* slack, const double* v)
{
    double tmin = -std::numeric_limits<double>::infinity();
    double tmax = std::numeric_limits<double>::infinity();
    for (int i = 0; i < m; ++i) {
        double av = 0.0;
        const double* row = A + (size_t)i * d;
        for (int j = 0; j < d; ++j) av += row[j] * v[j];
        if (std::fabs(av) < 1e-15) continue;
        const double t = slack[i] / av;
        if (av > 0.0)
            tmax = std::min(tmax, t);
        else
            tmin = std::max(tmin, t);
    }
    if (tmax <= tmin) return 0.0;
    return tmax - tmin;
}

struct DikinOut {
    double* P;      // d x d (nullable): the reference's P (stretches pinched directions)
    double* P_inv;  // d x d (nullable): y = x_c + P_inv z
    double* b_new;  // m (nullable): b - A x_c
    double* A_new;  // m x d (nullable): A P_inv
    int n_corrected;
};

// dikin_precondition (dikin_precondition.cpp:36-183).
void dikin_one(int m, int d, const double* A, const double* b, const double* x_c, int k_directions,
               double tight_factor, double target_chord, double alpha_max, DikinOut& out)
{
    NAJA_REQUIRE(m > 0 && d > 0, NAJA_ERR_INVALID_ARGUMENT, "dikin_precondition: empty polytope");
    std::vector<double> slack(m), delta(m), N((size_t)m * d, 0.0);
    for (int i = 0; i < m; ++i) {  // :49-65
        const double* row = A + (size_t)i * d;
        double ax = 0.0, n2 = 0.0;
        for (int j = 0; j < d; ++j) {
            ax += row[j] * (x_c ? x_c[j] : 0.0);
            n2 += row[j] * row[j];
        }
        slack[i] = b[i] - ax;
        const double nrm = std::sqrt(n2);
        if (nrm < 1e-15) {
            delta[i] = std::numeric_limits<double>::infinity();
        } else {
            delta[i] = slack[i] / nrm;
            for (int j = 0; j < d; ++j) N[(size_t)i * d + j] = row[j] / nrm;
        }
    }
    double min_delta = std::numeric_limits<double>::infinity();  // :67-74
    for (int i = 0; i < m; ++i)
        if (delta[i] > 0.0 && std::isfinite(delta[i])) min_delta = std::min(min_delta, delta[i]);
    const double threshold = tight_factor * min_delta;
    std::vector<int> tight;
    for (int i = 0; i < m; ++i)
        if (delta[i] > 0.0 && delta[i] <= threshold) tight.push_back(i);

    auto emit = [&](const std::vector<double>& P, const std::vector<double>& Pinv, int kk) {
        if (out.P) std::memcpy(out.P, P.data(), sizeof(double) * d * d);
        if (out.P_inv) std::memcpy(out.P_inv, Pinv.data(), sizeof(double) * d * d);
        if (out.b_new) std::memcpy(out.b_new, slack.data(), sizeof(double) * m);
        if (out.A_new) {
            for (int i = 0; i < m; ++i)
                for (int j = 0; j < d; ++j) {
                    double s = 0.0;
                    for (int k = 0; k < d; ++k) s += A[(size_t)i * d + k] * Pinv[(size_t)k * d + j];
                    out.A_new[(size_t)i * d + j] = s;
                }
        }
        out.n_corrected = kk;
    };
    std::vector<double> I((size_t)d * d, 0.0);
    for (int j = 0; j < d; ++j) I[(size_t)j * d + j] = 1.0;
    if (tight.empty()) {  // :84-93 identity transform, shifted so x_c maps to 0
        emit(I, I, 0);
        return;
    }
    std::vector<double> H((size_t)d * d, 0.0);  // :95-100
    for (int i : tight) {
        const double w = 1.0 / (delta[i] * delta[i]);
        const double* n_i = &N[(size_t)i * d];
        for (int r = 0; r < d; ++r) {
            const double wr = w * n_i[r];
            double* hr = &H[(size_t)r * d];
            for (int c = 0; c < d; ++c) hr[c] += wr * n_i[c];
        }
    }
    std::vector<double> ev(d), V((size_t)d * d);  // :102-124
    sym_eig(d, H.data(), ev.data(), V.data());
    const int k = std::min(k_directions, d);
    const double median = ev[d / 2];
    int kk = 0;
    for (int j = d - 1; j >= 0 && kk < k; --j) {
        if (ev[j] > 10.0 * std::max(median, 1e-10))
            ++kk;
        else
            break;
    }
    if (kk == 0) kk = 1;
    std::vector<double> Vk((size_t)kk * d);  // Vk[j*d + r] = eigenvector (d-1-j), component r
    for (int j = 0; j < kk; ++j)
        for (int r = 0; r < d; ++r) Vk[(size_t)j * d + r] = V[(size_t)r * d + (d - 1 - j)];

    if (target_chord <= 0.0) {  // :126-142 median coordinate chord over <= 50 axes
        std::vector<double> chords, e(d, 0.0);
        for (int j = 0; j < std::min(d, 50); ++j) {
            e[j] = 1.0;
            const double c = chord_along(m, d, A, slack.data(), e.data());
            e[j] = 0.0;
            if (c > 0.0) chords.push_back(c);
        }
        if (!chords.empty()) {
            std::sort(chords.begin(), chords.end());
            target_chord = chords[chords.size() / 2];
        } else {
            target_chord = 1.0;
        }
    }
    std::vector<double> alphas(kk);  // :144-154
    for (int j = 0; j < kk; ++j) {
        const double cj = chord_along(m, d, A, slack.data(), &Vk[(size_t)j * d]);
        if (cj < 1e-15)
            alphas[j] = std::min(alpha_max, 10.0);
        else
            alphas[j] = std::min(std::max(target_chord / cj, 1.0), alpha_max);
    }
    std::vector<double> P(I), Pinv(I);  // :156-167
    for (int j = 0; j < kk; ++j) {
        const double* vj = &Vk[(size_t)j * d];
        const double a = alphas[j] - 1.0, ai = 1.0 / alphas[j] - 1.0;
        for (int r = 0; r < d; ++r)
            for (int c = 0; c < d; ++c) {
                P[(size_t)r * d + c] += a * vj[r] * vj[c];
                Pinv[(size_t)r * d + c] += ai * vj[r] * vj[c];
            }
    }
    emit(P, Pinv, kk);
}

// Full Dikin-ellipsoid whitening at x_c: H = A^T diag(1/s^2) A = L L^T,
// P = L^{-T} (upper triangular), so P P^T = H^{-1}.
void barrier_one(int m, int d, const double* A, const double* b, const double* x_c, double* P)
{
    std::vector<double> s(m), H((size_t)d * d, 0.0);
    for (int i = 0; i < m; ++i) {
        const double* row = A + (size_t)i * d;
        double ax = 0.0;
        for (int j = 0; j < d; ++j) ax += row[j] * (x_c ? x_c[j] : 0.0);
        s[i] = b[i] - ax;
        NAJA_REQUIRE(s[i] > 0.0, NAJA_ERR_INVALID_ARGUMENT,
                     "Dikin ellipsoid: the centre is not interior (row %d slack %g)", i, s[i]);
        const double w = 1.0 / (s[i] * s[i]);
        for (int r = 0; r < d; ++r) {
            const double wr = w * row[r];
            if (wr == 0.0) continue;
            double* hr = &H[(size_t)r * d];
            for (int c = 0; c <= r; ++c) hr[c] += wr * row[c];
        }
    }
    // Cholesky, lower triangle in place.
    for (int j = 0; j < d; ++j) {
        double sum = H[(size_t)j * d + j];
        for (int k = 0; k < j; ++k) sum -= H[(size_t)j * d + k] * H[(size_t)j * d + k];
        NAJA_REQUIRE(sum > 0.0, NAJA_ERR_INVALID_ARGUMENT,
                     "Dikin ellipsoid: the barrier Hessian is singular (the polytope is unbounded)");
        const double ljj = std::sqrt(sum);
        H[(size_t)j * d + j] = ljj;
        for (int i = j + 1; i < d; ++i) {
            double t = H[(size_t)i * d + j];
            for (int k = 0; k < j; ++k) t -= H[(size_t)i * d + k] * H[(size_t)j * d + k];
            H[(size_t)i * d + j] = t / ljj;
        }
    }
    // M = L^{-1} (lower), then P = M^T.
    std::vector<double> M((size_t)d * d, 0.0);
    for (int c = 0; c < d; ++c) {
        for (int i = c; i < d; ++i) {
            double t = (i == c) ? 1.0 : 0.0;
            for (int k = c; k < i; ++k) t -= H[(size_t)i * d + k] * M[(size_t)k * d + c];
            M[(size_t)i * d + c] = t / H[(size_t)i * d + i];
        }
    }
    for (int r = 0; r < d; ++r)
        for (int c = 0; c < d; ++c) P[(size_t)r * d + c] = M[(size_t)c * d + r];
}

// P = (I + V V^T)^{p} for V (d x r row-major), p = -1/2 or +1/2:
// V^T V = R diag(lam) R^T, Q = V R diag(lam)^{-1/2},
// P = I + Q diag((1 + lam)^{p} - 1) Q^T.
void lowrank_one(int d, int r, const double* V, double* P, bool expand)
{
    std::vector<double> G((size_t)r * r, 0.0), lam(r), R((size_t)r * r);
    for (int i = 0; i < r; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = 0.0;
            for (int k = 0; k < d; ++k) s += V[(size_t)k * r + i] * V[(size_t)k * r + j];
            G[(size_t)i * r + j] = G[(size_t)j * r + i] = s;
        }
    sym_eig(r, G.data(), lam.data(), R.data());
    std::memset(P, 0, sizeof(double) * d * d);
    for (int j = 0; j < d; ++j) P[(size_t)j * d + j] = 1.0;
    const double lmax = std::max(lam[r - 1], 0.0);
    std::vector<double> q(d);
    for (int j = 0; j < r; ++j) {
        if (!(lam[j] > 1e-12 * lmax) || lam[j] <= 0.0) continue;
        const double inv = 1.0 / std::sqrt(lam[j]);
        for (int k = 0; k < d; ++k) {
            double s = 0.0;
            for (int t = 0; t < r; ++t) s += V[(size_t)k * r + t] * R[(size_t)t * r + j];
            q[k] = s * inv;
        }
        const double coef = (expand ? std::sqrt(1.0 + lam[j]) : 1.0 / std::sqrt(1.0 + lam[j])) - 1.0;
        for (int a = 0; a < d; ++a) {
            const double ca = coef * q[a];
            for (int b = 0; b < d; ++b) P[(size_t)a * d + b] += ca * q[b];
        }
    }
}

// Stacks [A0; E_v] and [b0; f_v] for variant v of a ragged batch.
void stack_variant(int m0, int d, const double* A0, const double* b0, const int* ks, const int64_t* koff,
                   const double* Ae, const double* be, int v, std::vector<double>& A, std::vector<double>& b)
{
    const int k = ks[v];
    A.resize((size_t)(m0 + k) * d);
    b.resize(m0 + k);
    std::memcpy(A.data(), A0, sizeof(double) * m0 * d);
    std::memcpy(b.data(), b0, sizeof(double) * m0);
    if (k > 0) {
        std::memcpy(A.data() + (size_t)m0 * d, Ae + (size_t)koff[v] * d, sizeof(double) * k * d);
        std::memcpy(b.data() + m0, be + koff[v], sizeof(double) * k);
    }
}

std::vector<int64_t> ragged_offsets(int nv, const int* ks)
{
    std::vector<int64_t> off(nv + 1, 0);
    for (int v = 0; v < nv; ++v) {
        NAJA_REQUIRE(ks[v] >= 0, NAJA_ERR_INVALID_ARGUMENT, "variant %d has %d extra rows", v, ks[v]);
        off[v + 1] = off[v] + ks[v];
    }
    return off;
}

}  // namespace
}  // namespace naja

using namespace naja;

extern "C" int naja_symmetric_eigen(const double* a, int n, double* evals, double* evecs)
{
    return guarded([&]() {
        NAJA_REQUIRE(a && evals && evecs && n > 0, NAJA_ERR_INVALID_ARGUMENT, "naja_symmetric_eigen: bad arguments");
        sym_eig(n, a, evals, evecs);
    });
}

extern "C" int naja_precond_dikin(const double* A, const double* b, const double* x_c, int m, int d,
                                  int k_directions, double tight_factor, double target_chord, double alpha_max,
                                  double* P, double* P_inv, double* A_new, double* b_new, int* n_corrected)
{
    return guarded([&]() {
        NAJA_REQUIRE(A && b && m > 0 && d > 0, NAJA_ERR_INVALID_ARGUMENT, "dikin_precondition: empty polytope");
        DikinOut out{P, P_inv, b_new, A_new, 0};
        dikin_one(m, d, A, b, x_c, k_directions, tight_factor, target_chord, alpha_max, out);
        if (n_corrected) *n_corrected = out.n_corrected;
    });
}

extern "C" int naja_precond_dikin_batched(const double* A0, const double* b0, int m0, int d, int n_variants,
                                          const int* extra_rows, const double* extra_A, const double* extra_b,
                                          const double* x_c, int k_directions, double tight_factor,
                                          double target_chord, double alpha_max, double* P_inv,
                                          int* n_corrected, int nthreads)
{
    return guarded([&]() {
        NAJA_REQUIRE(A0 && b0 && m0 > 0 && d > 0 && n_variants > 0 && extra_rows && P_inv,
                     NAJA_ERR_INVALID_ARGUMENT, "naja_precond_dikin_batched: bad arguments");
        const auto koff = ragged_offsets(n_variants, extra_rows);
        NAJA_REQUIRE(koff[n_variants] == 0 || (extra_A && extra_b), NAJA_ERR_INVALID_ARGUMENT,
                     "naja_precond_dikin_batched: extra rows without extra_A/extra_b");
        parallel_for(n_variants, nthreads, [&](int v) {
            std::vector<double> A, b;
            stack_variant(m0, d, A0, b0, extra_rows, koff.data(), extra_A, extra_b, v, A, b);
            DikinOut out{nullptr, P_inv + (size_t)v * d * d, nullptr, nullptr, 0};
            dikin_one(m0 + extra_rows[v], d, A.data(), b.data(), x_c ? x_c + (size_t)v * d : nullptr,
                      k_directions, tight_factor, target_chord, alpha_max, out);
            if (n_corrected) n_corrected[v] = out.n_corrected;
        });
    });
}

extern "C" int naja_precond_ellipsoid_batched(const double* A0, const double* b0, int m0, int d, int n_variants,
                                              const int* extra_rows, const double* extra_A,
                                              const double* extra_b, const double* x_c, double* P, int nthreads)
{
    return guarded([&]() {
        NAJA_REQUIRE(A0 && b0 && m0 > 0 && d > 0 && n_variants > 0 && extra_rows && P, NAJA_ERR_INVALID_ARGUMENT,
                     "naja_precond_ellipsoid_batched: bad arguments");
        const auto koff = ragged_offsets(n_variants, extra_rows);
        NAJA_REQUIRE(koff[n_variants] == 0 || (extra_A && extra_b), NAJA_ERR_INVALID_ARGUMENT,
                     "naja_precond_ellipsoid_batched: extra rows without extra_A/extra_b");
        parallel_for(n_variants, nthreads, [&](int v) {
            std::vector<double> A, b;
            stack_variant(m0, d, A0, b0, extra_rows, koff.data(), extra_A, extra_b, v, A, b);
            barrier_one(m0 + extra_rows[v], d, A.data(), b.data(), x_c ? x_c + (size_t)v * d : nullptr,
                        P + (size_t)v * d * d);
        });
    });
}

extern "C" int naja_precond_lowrank(const double* V, int n_variants, int d, int rank, int expand, double* P,
                                    int nthreads)
{
    return guarded([&]() {
        NAJA_REQUIRE(V && P && n_variants > 0 && d > 0 && rank > 0 && rank <= d, NAJA_ERR_INVALID_ARGUMENT,
                     "naja_precond_lowrank: bad arguments");
        parallel_for(n_variants, nthreads, [&](int v) {
            lowrank_one(d, rank, V + (size_t)v * d * rank, P + (size_t)v * d * d, expand != 0);
        });
    });
}
