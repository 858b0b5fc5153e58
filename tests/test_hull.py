"""Affine-hull reduction (job_sampling.cu:196-235): library vs the oracle's
restatement, pinned by closed forms, scipy's SVD null space and A_eq B = 0."""
import numpy as np
import pytest
import scipy.linalg

import oracle.hull as ohull
from naja_amd import NajaInvalidArgument
from naja_amd import hull


def test_closed_forms():
    # x1 + x2 = c: first maximum in column order is (0, 0); free column 1
    np.testing.assert_array_equal(hull.kernel_basis([[1.0, 1.0]]), [[-1.0], [1.0]])
    # 2 x2 = c: pivot column 1, free columns 0 and 2 in permuted order (q = [1, 0, 2])
    np.testing.assert_array_equal(hull.kernel_basis([[0.0, 2.0, 0.0]]), [[1.0, 0.0], [0.0, 0.0], [0.0, 1.0]])
    # x1 - x3 = 0, x2 + x3 = 0: kernel spanned by (1, -1, 1)
    B = hull.kernel_basis([[1.0, 0.0, -1.0], [0.0, 1.0, 1.0]])
    assert B.shape == (3, 1)
    np.testing.assert_allclose(B[:, 0] / B[2, 0], [1.0, -1.0, 1.0], rtol=0, atol=1e-15)
    # no rows: identity; full rank: empty basis
    np.testing.assert_array_equal(hull.kernel_basis(np.zeros((0, 3))), np.eye(3))
    assert hull.kernel_basis(np.eye(3)).shape == (3, 0)
    # all-zero rows: nothing is a pivot
    np.testing.assert_array_equal(hull.kernel_basis(np.zeros((2, 3))), np.eye(3))


def _rank_deficient(rng, r, n, rank, dup=0):
    M = rng.standard_normal((rank, n)) @ np.diag(10 ** rng.uniform(-2, 2, n))
    C = rng.standard_normal((r, rank))
    A = C @ M
    if dup:
        A = np.vstack([A, -A[:dup]])  # equality pairs as opposite inequalities
    return A


@pytest.mark.parametrize("r,n,rank,dup,seed", [(3, 7, 3, 0, 1), (12, 9, 5, 4, 2), (40, 30, 22, 10, 3),
                                               (6, 6, 6, 0, 4), (25, 60, 17, 25, 5)])
def test_kernel_basis_matches_oracle_and_svd(r, n, rank, dup, seed):
    rng = np.random.default_rng(seed)
    A = _rank_deficient(rng, r, n, rank, dup)
    B = hull.kernel_basis(A)
    np.testing.assert_array_equal(B, ohull.kernel_basis(A))  # same FP64 operation order
    N = scipy.linalg.null_space(A)
    assert B.shape == (n, N.shape[1]) == (n, n - rank)
    if B.shape[1]:
        # same span: the projection onto the SVD null space leaves B unchanged
        np.testing.assert_allclose(N @ (N.T @ B), B, rtol=0, atol=1e-9 * np.abs(B).max())
        assert np.linalg.matrix_rank(B) == B.shape[1]
        scale = np.abs(A).max() * np.abs(B).max() * n
        assert np.abs(A @ B).max() <= 1e-12 * scale
        # one unit entry per column on distinct rows, zero at the other free rows
        ones = [np.flatnonzero(B[:, f] == 1.0) for f in range(B.shape[1])]
        rows = [o[0] for o in ones]
        assert len(set(rows)) == B.shape[1]
        np.testing.assert_array_equal(B[rows], np.eye(B.shape[1]))


def test_planted_mistake_is_caught():
    """A sign error in the back substitution breaks A_eq B = 0."""
    rng = np.random.default_rng(9)
    A = _rank_deficient(rng, 5, 8, 4)
    B = ohull.kernel_basis(A)
    Bad = B.copy()
    nz = np.flatnonzero(np.abs(B[:, 0]) != 1.0)[0]
    Bad[nz, 0] = -Bad[nz, 0]
    assert np.abs(A @ Bad).max() > 1e-6


def test_affine_hull_reduction():
    # box [-1, 1]^3 and the equality x0 + x1 = 0.2 as two opposite rows
    A = np.vstack([np.eye(3), -np.eye(3), [[1.0, 1.0, 0.0], [-1.0, -1.0, 0.0]]])
    b = np.concatenate([np.ones(6), [0.2, -0.2]])
    x0 = np.array([0.1, 0.1, 0.0])
    h = hull.affine_hull(A, b, x0, tol=1e-9)
    rows, Bo = ohull.affine_hull(A, b, x0, 1e-9)
    assert h["enabled"] and h["n_tight"] == len(rows) == 2 and rows == [6, 7]
    np.testing.assert_array_equal(h["basis"], Bo)
    B = h["basis"]
    assert B.shape == (3, 2)
    np.testing.assert_allclose(A[6] @ B, 0.0, atol=1e-15)
    np.testing.assert_allclose(h["A"], A @ B)
    np.testing.assert_allclose(h["b"], b - A @ x0)
    np.testing.assert_array_equal(h["x0"], np.zeros(2))
    # a transform of the unreduced variables composes
    T = np.arange(12.0).reshape(4, 3)
    s = np.ones(4)
    T2, s2 = hull.compose_transform(T, s, h)
    z = np.array([0.3, -0.4])
    np.testing.assert_allclose(T2 @ z + s2, T @ (x0 + B @ z) + s, rtol=0, atol=1e-13)


def test_affine_hull_no_tight_rows_or_full_rank():
    A = np.vstack([np.eye(2), -np.eye(2)])
    h = hull.affine_hull(A, np.ones(4), np.zeros(2), tol=1e-9)
    assert not h["enabled"] and h["n_tight"] == 0 and h["basis"].shape == (2, 2)
    # a vertex: tight rows span R^2, kernel dimension 0 -> no reduction
    h = hull.affine_hull(A, np.ones(4), np.ones(2), tol=1e-9)
    assert not h["enabled"] and h["n_tight"] == 2 and h["basis"].shape == (2, 0)


def test_affine_hull_bad_arguments():
    with pytest.raises(NajaInvalidArgument):
        hull.affine_hull(np.eye(2), np.ones(2), np.zeros(2), tol=-1.0)
