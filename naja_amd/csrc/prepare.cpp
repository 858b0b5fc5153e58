// Producers of conditioned variants (host, FP64):
//
//  * naja_cube_center: the start-point LP of axis_aligned_cube_center_lp_gurobi
//    (feasible_start_lp.cpp:33-87): maximise r subject to
//    A x + ||a_i||_1 r <= b, r >= 0, x free -- the centre of the largest
//    axis-aligned cube inside the polytope. The reference calls Gurobi; this
//    is an own Mehrotra predictor-corrector interior-point method on the
//    inequality form, with rows scaled by their l1 norms (the same LP).
//  * naja_extra_from_bounds: build_extra_from_bounds (cmd_prepare.cpp:111-156),
//    tightened flux bounds as rows +-T[i,:] over the rounded variables.
//
// LP method: variables y = (x, r); constraints G y <= h with rows
// (a_i / w_i, 1) | b_i / w_i for every row of l1 norm w_i > 0 and (0, -1) | 0;
// objective min c.y with c = (0, -1). Each iteration solves the normal
// equations (G^T D G) dy = rhs, D = z / s, by a dense Cholesky factorisation
// (tiny pivots are replaced by a huge one, which zeroes directions no row
// constrains), takes an affine step to get the centring parameter
// sigma = (mu_aff / mu)^3, then the corrected step with a 0.995 fraction to
// the boundary. The iterate with the largest exact radius
// r(x) = min_i (b_i - a_i.x) / w_i is returned, so the cube at the returned
// centre lies inside the polytope by construction.
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "common.h"

namespace {

using namespace naja;

struct CubeLp {
    int n = 0, mk = 0;        // variables, kept rows
    std::vector<double> at;   // n x mk: scaled rows, transposed (at[j*mk + i] = a_ij / w_i)
    std::vector<double> h;    // mk + 1
    int nthreads = 1;

    // g = G y (mk + 1)
    void G(const double* y, double* g) const
    {
        std::fill(g, g + mk, y[n]);
        for (int j = 0; j < n; ++j) {
            const double xj = y[j];
            if (xj == 0.0) continue;
            const double* col = &at[(size_t)j * mk];
            for (int i = 0; i < mk; ++i) g[i] += col[i] * xj;
        }
        g[mk] = -y[n];
    }

    // out = G^T v (n + 1)
    void GT(const double* v, double* out) const
    {
        for (int j = 0; j < n; ++j) {
            const double* col = &at[(size_t)j * mk];
            double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
            int i = 0;
            for (; i + 3 < mk; i += 4) {
                s0 += col[i] * v[i];
                s1 += col[i + 1] * v[i + 1];
                s2 += col[i + 2] * v[i + 2];
                s3 += col[i + 3] * v[i + 3];
            }
            for (; i < mk; ++i) s0 += col[i] * v[i];
            out[j] = (s0 + s1) + (s2 + s3);
        }
        double s = 0.0;
        for (int i = 0; i < mk; ++i) s += v[i];
        out[n] = s - v[mk];
    }

    // Exact cube radius at x: min_i (h_i - a~_i.x).
    double radius(const double* y, std::vector<double>& g) const
    {
        std::vector<double> y0(y, y + n + 1);
        y0[n] = 0.0;
        G(y0.data(), g.data());
        double r = std::numeric_limits<double>::infinity();
        for (int i = 0; i < mk; ++i) {
            const double v = h[i] - g[i];
            if (std::isnan(v)) return -std::numeric_limits<double>::infinity();
            r = std::min(r, v);
        }
        return r;
    }

    // N = G^T diag(d) G, (n+1) x (n+1) row-major, lower triangle filled.
    void normal(const std::vector<double>& d, std::vector<double>& N, std::vector<double>& ad) const
    {
        const int p = n + 1;
        std::fill(N.begin(), N.end(), 0.0);
        parallel_for(n, nthreads, [&](int j) {
            const double* col = &at[(size_t)j * mk];
            double* o = &ad[(size_t)j * mk];
            for (int i = 0; i < mk; ++i) o[i] = col[i] * d[i];
        });
        const int chunk = 512;
        parallel_for(nthreads, nthreads, [&](int t) {
            for (int c0 = 0; c0 < mk; c0 += chunk) {
                const int c1 = std::min(mk, c0 + chunk);
                for (int j = t; j < n; j += nthreads) {
                    const double* aj = &ad[(size_t)j * mk];
                    double* Nj = &N[(size_t)j * p];
                    for (int k = 0; k <= j; ++k) {
                        const double* ak = &at[(size_t)k * mk];
                        double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
                        int i = c0;
                        for (; i + 3 < c1; i += 4) {
                            s0 += aj[i] * ak[i];
                            s1 += aj[i + 1] * ak[i + 1];
                            s2 += aj[i + 2] * ak[i + 2];
                            s3 += aj[i + 3] * ak[i + 3];
                        }
                        for (; i < c1; ++i) s0 += aj[i] * ak[i];
                        Nj[k] += (s0 + s1) + (s2 + s3);
                    }
                }
            }
        });
        double dsum = 0.0;
        for (int i = 0; i < mk; ++i) dsum += d[i];
        for (int j = 0; j < n; ++j) {
            const double* aj = &ad[(size_t)j * mk];
            double s = 0.0;
            for (int i = 0; i < mk; ++i) s += aj[i];
            N[(size_t)n * p + j] = s;
        }
        N[(size_t)n * p + n] = dsum + d[mk];
    }
};

// In-place lower Cholesky of a row-major p x p matrix (lower triangle used).
void cholesky(int p, std::vector<double>& N, int nthreads)
{
    double dmax = 0.0;
    for (int k = 0; k < p; ++k) dmax = std::max(dmax, N[(size_t)k * p + k]);
    const double tiny = 1e-30 * std::max(dmax, 1e-300);
    for (int j = 0; j < p; ++j) {
        double* Lj = &N[(size_t)j * p];
        double s = Lj[j];
        for (int k = 0; k < j; ++k) s -= Lj[k] * Lj[k];
        const double ljj = s > tiny ? std::sqrt(s) : 1e64;
        Lj[j] = ljj;
        auto row = [&](int i) {
            double* Li = &N[(size_t)i * p];
            double t = Li[j];
            for (int k = 0; k < j; ++k) t -= Li[k] * Lj[k];
            Li[j] = t / ljj;
        };
        const int rows = p - j - 1;
        if (rows > 256 && nthreads > 1) {
            parallel_for(nthreads, nthreads, [&](int t) {
                for (int i = j + 1 + t; i < p; i += nthreads) row(i);
            });
        } else {
            for (int i = j + 1; i < p; ++i) row(i);
        }
    }
}

void chol_solve(int p, const std::vector<double>& L, double* x)
{
    for (int i = 0; i < p; ++i) {
        double s = x[i];
        const double* Li = &L[(size_t)i * p];
        for (int k = 0; k < i; ++k) s -= Li[k] * x[k];
        x[i] = s / Li[i];
    }
    for (int i = p - 1; i >= 0; --i) {
        double s = x[i];
        for (int k = i + 1; k < p; ++k) s -= L[(size_t)k * p + i] * x[k];
        x[i] = s / L[(size_t)i * p + i];
    }
}

double step_to_boundary(const std::vector<double>& v, const std::vector<double>& dv)
{
    double a = 1.0;
    for (size_t i = 0; i < v.size(); ++i)
        if (dv[i] < 0.0) a = std::min(a, -v[i] / dv[i]);
    return a;
}

void cube_center(const double* A, const double* b, int m, int n, const double* x_hint, double* x_out,
                 double* r_out, int nthreads)
{
    NAJA_REQUIRE(m > 0 && n > 0, NAJA_ERR_INVALID_ARGUMENT, "axis_aligned_cube_center: empty A");
    CubeLp lp;
    lp.n = n;
    lp.nthreads = nthreads <= 0 ? (int)std::max(1u, std::thread::hardware_concurrency()) : nthreads;
    std::vector<int> keep;
    std::vector<double> w(m);
    for (int i = 0; i < m; ++i) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += std::fabs(A[(size_t)i * n + j]);
        NAJA_REQUIRE(std::isfinite(s), NAJA_ERR_INVALID_ARGUMENT, "axis_aligned_cube_center: non-finite row l1 norm");
        NAJA_REQUIRE(std::isfinite(b[i]), NAJA_ERR_INVALID_ARGUMENT, "axis_aligned_cube_center: non-finite b");
        w[i] = s;
        if (s > 0.0)
            keep.push_back(i);
        else
            NAJA_REQUIRE(b[i] >= 0.0, NAJA_ERR_RUNTIME,
                         "axis_aligned_cube_center: LP infeasible (row %d is 0 <= %g)", i, b[i]);
    }
    NAJA_REQUIRE(!keep.empty(), NAJA_ERR_RUNTIME, "axis_aligned_cube_center: LP unbounded (no nonzero row)");
    const int mk = (int)keep.size(), M = mk + 1, p = n + 1;
    lp.mk = mk;
    lp.at.resize((size_t)n * mk);
    lp.h.resize(M);
    for (int t = 0; t < mk; ++t) {
        const int i = keep[t];
        for (int j = 0; j < n; ++j) lp.at[(size_t)j * mk + t] = A[(size_t)i * n + j] / w[i];
        lp.h[t] = b[i] / w[i];
    }
    lp.h[mk] = 0.0;

    std::vector<double> y(p, 0.0), s(M), z(M, 1.0 / mk), g(M), gt(p), rd(p), rp(M);
    if (x_hint)
        for (int j = 0; j < n; ++j) y[j] = x_hint[j];
    double hmax = 0.0, habs = 0.0;
    for (double v : lp.h) {
        hmax = std::max(hmax, std::fabs(v));
        habs += std::fabs(v);
    }
    const double theta = 0.1 * (habs / M + 1e-12);
    lp.G(y.data(), g.data());
    for (int i = 0; i < M; ++i) s[i] = std::max(lp.h[i] - g[i], theta);

    std::vector<double> N((size_t)p * p), ad((size_t)n * mk), d(M), dy(p), ds(M), dz(M), tmp(M), rhs(p);
    double best_r = -std::numeric_limits<double>::infinity(), best_rd = std::numeric_limits<double>::infinity();
    std::vector<double> best_x(y.begin(), y.begin() + n);
    bool converged = false;
    const double tol = 1e-9;

    auto keep_best = [&]() {
        const double r = lp.radius(y.data(), g);
        if (r > best_r) {
            best_r = r;
            std::copy(y.begin(), y.begin() + n, best_x.begin());
        }
    };
    // Newton direction for complementarity target rc (s o z form).
    auto direction = [&](const std::vector<double>& rc) {
        for (int i = 0; i < M; ++i) tmp[i] = (-rc[i] + z[i] * rp[i]) / s[i];
        lp.GT(tmp.data(), gt.data());
        for (int j = 0; j < p; ++j) dy[j] = -rd[j] - gt[j];
        chol_solve(p, N, dy.data());
        lp.G(dy.data(), g.data());
        for (int i = 0; i < M; ++i) {
            ds[i] = -rp[i] - g[i];
            dz[i] = (-rc[i] - z[i] * ds[i]) / s[i];
        }
    };

    std::vector<double> rc(M);
    for (int it = 0; it < 80; ++it) {
        keep_best();
        lp.GT(z.data(), rd.data());
        rd[n] -= 1.0;  // + c
        lp.G(y.data(), g.data());
        double rp_max = 0.0, rd_max = 0.0, mu = 0.0, hz = 0.0;
        for (int i = 0; i < M; ++i) {
            rp[i] = g[i] + s[i] - lp.h[i];
            rp_max = std::max(rp_max, std::fabs(rp[i]));
            mu += s[i] * z[i];
            hz += lp.h[i] * z[i];
        }
        mu /= M;
        for (int j = 0; j < p; ++j) rd_max = std::max(rd_max, std::fabs(rd[j]));
        if (!std::isfinite(rp_max) || !std::isfinite(rd_max) || !std::isfinite(mu)) break;
        const double gap = std::fabs(-y[n] + hz) / std::max(std::fabs(hz), 1e-300);
        if (rp_max / (1.0 + hmax) < tol && rd_max < tol && gap < tol) {
            converged = true;
            break;
        }
        best_rd = std::min(best_rd, rd_max);
        if (best_rd < 1e-6 && rd_max > 100.0 * best_rd) break;  // numerical breakdown: keep the best iterate

        for (int i = 0; i < M; ++i) d[i] = z[i] / s[i];
        lp.normal(d, N, ad);
        cholesky(p, N, lp.nthreads);

        for (int i = 0; i < M; ++i) rc[i] = s[i] * z[i];
        direction(rc);
        const double ap = step_to_boundary(s, ds), adu = step_to_boundary(z, dz);
        double mu_aff = 0.0;
        for (int i = 0; i < M; ++i) mu_aff += (s[i] + ap * ds[i]) * (z[i] + adu * dz[i]);
        mu_aff /= M;
        const double sigma = std::pow(mu_aff / mu, 3.0);
        for (int i = 0; i < M; ++i) rc[i] = s[i] * z[i] + ds[i] * dz[i] - sigma * mu;
        direction(rc);
        const double ap2 = std::min(1.0, 0.995 * step_to_boundary(s, ds));
        const double ad2 = std::min(1.0, 0.995 * step_to_boundary(z, dz));
        for (int j = 0; j < p; ++j) y[j] += ap2 * dy[j];
        for (int i = 0; i < M; ++i) {
            s[i] += ap2 * ds[i];
            z[i] += ad2 * dz[i];
        }
    }
    keep_best();
    NAJA_REQUIRE(converged || best_rd < 1e-7, NAJA_ERR_RUNTIME,
                 "axis_aligned_cube_center: LP did not converge (infeasible or unbounded)");
    NAJA_REQUIRE(best_r >= -1e-12 * (1.0 + hmax), NAJA_ERR_RUNTIME,
                 "axis_aligned_cube_center: LP infeasible (best centre violates a row by %g)", -best_r);
    std::copy(best_x.begin(), best_x.end(), x_out);
    // the radius in the model's row scaling (same rows, same l1 weights)
    double r = std::numeric_limits<double>::infinity();
    for (int i : keep) {
        double ax = 0.0;
        for (int j = 0; j < n; ++j) ax += A[(size_t)i * n + j] * x_out[j];
        r = std::min(r, (b[i] - ax) / w[i]);
    }
    *r_out = std::max(r, 0.0);
}

}  // namespace

extern "C" int naja_cube_center(const double* A, const double* b, int m, int n, const double* x_hint,
                                double* x_out, double* r_out, int nthreads)
{
    return guarded([&]() {
        NAJA_REQUIRE(A && b && x_out && r_out, NAJA_ERR_INVALID_ARGUMENT, "naja_cube_center: null argument");
        cube_center(A, b, m, n, x_hint, x_out, r_out, nthreads);
    });
}

extern "C" int naja_extra_from_bounds(const double* T, const double* shift, int n, int d, const double* lb_base,
                                      const double* ub_base, const double* lb_new, const double* ub_new, double eps,
                                      double* extra_A, double* extra_b, int* k_out)
{
    return guarded([&]() {
        NAJA_REQUIRE(T && shift && lb_base && ub_base && lb_new && ub_new && extra_A && extra_b && k_out && n > 0 &&
                         d > 0 && eps >= 0.0,
                     NAJA_ERR_INVALID_ARGUMENT, "naja_extra_from_bounds: bad arguments");
        // validate_bounds (cmd_prepare.cpp:80-93) on both bound sets
        for (int pass = 0; pass < 2; ++pass) {
            const double* lb = pass ? lb_base : lb_new;
            const double* ub = pass ? ub_base : ub_new;
            for (int i = 0; i < n; ++i) {
                NAJA_REQUIRE(std::isfinite(lb[i]) && std::isfinite(ub[i]), NAJA_ERR_INVALID_ARGUMENT,
                             "non-finite bound at index %d", i);
                NAJA_REQUIRE(lb[i] <= ub[i], NAJA_ERR_INVALID_ARGUMENT, "l_bounds > u_bounds at index %d", i);
            }
        }
        int k = 0;
        for (int i = 0; i < n; ++i) {
            const double* Ti = T + (size_t)i * d;
            if (ub_new[i] < ub_base[i] - eps) {
                std::copy(Ti, Ti + d, extra_A + (size_t)k * d);
                extra_b[k++] = ub_new[i] - shift[i];
            }
            if (lb_new[i] > lb_base[i] + eps) {
                for (int j = 0; j < d; ++j) extra_A[(size_t)k * d + j] = -Ti[j];
                extra_b[k++] = -lb_new[i] + shift[i];
            }
        }
        *k_out = k;
    });
}
