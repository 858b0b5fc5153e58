"""Test polytopes shared by the oracle and GPU tests (all written explicitly
or generated with numpy's seeded PCG64, so fixtures are portable)."""
import numpy as np


def cube(d):
    """[-1, 1]^d as A = [I; -I], b = 1 (test_gpu_sampling_cube_moments.cu:22-60)."""
    A = np.vstack([np.eye(d), -np.eye(d)])
    b = np.ones(2 * d)
    return A, b, np.zeros(d)


def simplex(n):
    """{x >= 0, sum x <= 1} in R^n, start 1/(n+1) (test_gpu_sampling_simplex_moments.cu:25-60)."""
    A = np.vstack([-np.eye(n), np.ones((1, n))])
    b = np.concatenate([np.zeros(n), [1.0]])
    return A, b, np.full(n, 1.0 / (n + 1))


def random_polytope(d, m, seed):
    """m random N(0,1) rows with b = 1 around the origin
    (test_gpu_sampling_feasibility.cu:120-193), plus the cube [-3, 3]^d so the
    polytope is bounded."""
    rng = np.random.default_rng(seed)
    A = np.vstack([rng.standard_normal((m, d)), np.eye(d), -np.eye(d)])
    b = np.concatenate([np.ones(m), 3.0 * np.ones(2 * d)])
    return A, b, np.zeros(d)


def thin_slab(d, eps):
    """[-1,1]^d intersected with |x0 + x1 + x2| <= eps (test_gpu_thin_polytope.cu:48-90)."""
    A, b, x0 = cube(d)
    row = np.zeros(d)
    row[:3] = 1.0
    A = np.vstack([A, row, -row])
    b = np.concatenate([b, [eps, eps]])
    return A, b, x0


def genome_like(n, d, seed, nnz=(8, 40)):
    """The genome-scale reduced polytope recipe: T (n x d) with 8-40 N(0,1)
    nonzeros per row, densified by a random orthogonal rotation; shift ~ N(0,1);
    margins log-uniform on [1e-3, 1e2]; A = [T; -T], b = [m_hi; m_lo], y0 = 0."""
    rng = np.random.default_rng(seed)
    T0 = np.zeros((n, d))
    for r in range(n):
        k = rng.integers(nnz[0], nnz[1] + 1)
        cols = rng.choice(d, size=k, replace=False)
        T0[r, cols] = rng.standard_normal(k)
    Q, R = np.linalg.qr(rng.standard_normal((d, d)))
    Q = Q * np.sign(np.diag(R))
    T = T0 @ Q
    shift = rng.standard_normal(n)
    m_lo = 10.0 ** rng.uniform(-3, 2, n)
    m_hi = 10.0 ** rng.uniform(-3, 2, n)
    A = np.vstack([T, -T])
    b = np.concatenate([m_hi, m_lo])
    lb = shift - m_lo
    ub = shift + m_hi
    return dict(A=A, b=b, y0=np.zeros(d), T=T, shift=shift, lb=lb, ub=ub)
