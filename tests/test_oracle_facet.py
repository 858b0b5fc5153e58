"""The oracle's facet-aligned rounding (oracle/precond.py facet, the rule of
naja_precond_facet_device) pinned by what the mathematics fixes: a box keeps
its own axes, the completion rows are an orthonormal basis of the subspace
the chosen facets leave free, every independent slab is chosen, and after the
exchange no single swap of a facet into the basis grows |det N| by more than
the tolerance (checked by determinants, not by the exchange's own
arithmetic)."""
import itertools

import numpy as np
import pytest

from oracle import precond as op


def slabs(d, nslab, seed, w_lo=1e-3, w_hi=1e-1):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((nslab, d))
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    w = 10 ** rng.uniform(np.log10(w_lo), np.log10(w_hi), nslab)
    E = np.vstack([a, -a]) if nslab else np.zeros((0, d))
    f = np.concatenate([w, w]) if nslab else np.zeros(0)
    return E, f


def cube(d):
    return np.vstack([np.eye(d), -np.eye(d)]), np.ones(2 * d)


def test_box_keeps_its_axes():
    """No extra rows, a box [l, u] around an off-centre interior point: the
    completion from the base facets is the box's own axes, so P is a scaled
    signed permutation (here, with the greedy order over equal residuals, the
    identity up to the slack scaling)."""
    rng = np.random.default_rng(1)
    d = 9
    lo, hi = -rng.uniform(0.5, 2, d), rng.uniform(0.5, 2, d)
    A0 = np.vstack([np.eye(d), -np.eye(d)])
    b0 = np.concatenate([hi, -lo])
    xc = 0.1 * (lo + hi)
    P, sel, N = op.facet(A0, b0, np.zeros((0, d)), np.zeros(0), None, 0.02, x_c=xc,
                         completion=op.COMPLETE_BASE_FACETS)
    assert (sel >= 0).all()
    Pn = P / np.abs(P).max(axis=0, keepdims=True)
    assert np.allclose(np.abs(Pn), np.eye(d)[:, np.argmax(np.abs(Pn), axis=0)], atol=1e-12)
    assert sorted(np.argmax(np.abs(Pn), axis=0)) == list(range(d))


@pytest.mark.parametrize("nslab", [1, 7, 20, 33])
def test_completion_is_an_orthonormal_basis_of_the_free_subspace(nslab):
    d = 40
    A0, b0 = cube(d)
    E, f = slabs(d, nslab, seed=nslab)
    P, sel, N = op.facet(A0, b0, E, f, None, 0.02, completion=op.COMPLETE_BASE_FACETS)
    nf = E.shape[0]
    fac = sel < nf
    assert fac.sum() == nslab  # every independent slab, one side each
    assert len(set(np.asarray(sel)[fac] % nslab)) == nslab
    F, S = N[~fac], N[fac]
    np.testing.assert_allclose(F @ F.T, np.eye(d - nslab), atol=1e-12)
    assert np.abs(F @ S.T).max() <= 1e-12 * np.abs(S).max()
    # the chosen facet rows are the slabs' dual rows a / w exactly
    U = E / f[:, None]
    np.testing.assert_array_equal(S, U[np.asarray(sel)[fac]])
    np.testing.assert_allclose(N @ P, np.eye(d), atol=1e-9)


@pytest.mark.parametrize("d,nslab,seed", [(3, 6, 0), (4, 9, 1), (5, 12, 2), (6, 14, 3)])
def test_exchange_leaves_no_swap_that_grows_the_volume(d, nslab, seed):
    """After the exchange, for every facet candidate i and basis row k, the
    determinant with row k replaced by u_i is at most tol times |det N|
    (Cramer: that ratio is C_ik), by brute-force determinants; and the
    exchange never lowers |det N|."""
    tol = 1.02
    A0, b0 = cube(d)
    E, f = slabs(d, nslab, seed=seed, w_lo=1e-2, w_hi=1.0)
    U = E / f[:, None]
    _, sel0, N0 = op.facet(A0, b0, E, f, None, 0.02, completion=op.COMPLETE_BASE_FACETS)
    _, sel, N = op.facet(A0, b0, E, f, None, 0.02, completion=op.COMPLETE_BASE_FACETS, exchange_tol=tol)
    assert (sel < E.shape[0]).all()
    det = abs(np.linalg.det(N))
    assert det >= abs(np.linalg.det(N0)) * (1 - 1e-12)
    for i, k in itertools.product(range(U.shape[0]), range(d)):
        M = N.copy()
        M[k] = U[i]
        assert abs(np.linalg.det(M)) <= tol * det * (1 + 1e-4)


def test_exchange_reaches_the_largest_volume_on_a_tiny_case():
    """d = 2, five slabs: the maximum-volume pair of slab normals found by
    enumerating every pair equals the exchange's (a local maximum is global
    here: every pair is one swap from every other)."""
    rng = np.random.default_rng(4)
    d = 2
    A0, b0 = cube(d)
    ang = rng.uniform(0, np.pi, 5)
    a = np.stack([np.cos(ang), np.sin(ang)], 1)
    w = rng.uniform(0.05, 0.5, 5)
    E, f = np.vstack([a, -a]), np.concatenate([w, w])
    U = a / w[:, None]
    best = max(abs(np.linalg.det(U[list(p)])) for p in itertools.combinations(range(5), 2))
    _, sel, N = op.facet(A0, b0, E, f, None, 0.02, completion=op.COMPLETE_BASE_FACETS, exchange_tol=1.0 + 1e-6)
    assert abs(np.linalg.det(N)) >= best * (1 - 1e-9)


def test_base_facets_completion_rejects_base_rows_as_facets():
    A0, b0 = cube(4)
    with pytest.raises(ValueError):
        op.facet(A0, b0, np.zeros((0, 4)), np.zeros(0), None, 0.02, include_base=True,
                 completion=op.COMPLETE_BASE_FACETS)
