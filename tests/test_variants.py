"""Conditioned-variant producers: the start-point LP (feasible_start_lp.cpp:33-87)
against closed forms and scipy's HiGHS, and build_extra_from_bounds
(cmd_prepare.cpp:111-156) against the oracle's restatement."""
import numpy as np
import pytest

import oracle.lp as olp
from naja_amd import NajaError, NajaInvalidArgument
from naja_amd import variants as V


def gem_like(rng, d, n):
    T = rng.standard_normal((n, d)) * (rng.random((n, d)) < 0.15)
    Q, _ = np.linalg.qr(rng.standard_normal((d, d)))
    T = T @ Q
    shift = rng.standard_normal(n)
    m_lo, m_hi = 10 ** rng.uniform(-3, 1, n), 10 ** rng.uniform(-3, 1, n)
    return T, shift, m_lo, m_hi


def test_box_centre_as_the_reference_test():
    """tests/test_feasible_start_lp.cpp:8-34: box [-1, 1]^3 -> x = 0, r = 1."""
    n = 3
    A = np.vstack([np.eye(n), -np.eye(n)])
    b = np.ones(2 * n)
    x, r = V.cube_center(A, b)
    assert abs(r - 1.0) <= 1e-6 and np.abs(x).max() <= 1e-7
    slack = b - A @ x
    assert slack.min() >= -1e-9 and slack.min() >= r - 1e-6


@pytest.mark.parametrize("n", [1, 2, 5, 40])
def test_simplex_closed_form(n):
    # x >= 0, sum x <= 1: x_i >= r and n r + n r <= 1 -> r = 1 / (2n), x_i = r (unique)
    A = np.vstack([-np.eye(n), np.ones((1, n))])
    b = np.r_[np.zeros(n), 1.0]
    x, r = V.cube_center(A, b)
    assert abs(r - 1 / (2 * n)) <= 1e-10
    np.testing.assert_allclose(x, 1 / (2 * n), rtol=0, atol=1e-9)
    xo, ro = olp.cube_center(A, b)
    assert abs(ro - 1 / (2 * n)) <= 1e-9  # the oracle meets the closed form too


def test_non_unique_centre_slab():
    # |x0| <= 1, x1 free: r = 1 and any x1 is optimal
    A = np.array([[1.0, 0.0], [-1.0, 0.0]])
    x, r = V.cube_center(A, np.ones(2), x_hint=np.array([0.3, 7.0]))
    assert abs(r - 1.0) <= 1e-9 and abs(x[0]) <= 1e-8


@pytest.mark.parametrize("d,n,seed", [(20, 40, 1), (100, 300, 2), (256, 700, 3)])
def test_gem_polytopes_match_highs(d, n, seed):
    rng = np.random.default_rng(seed)
    T, shift, m_lo, m_hi = gem_like(rng, d, n)
    A = np.vstack([T, -T])
    b = np.r_[m_hi, m_lo]
    x, r = V.cube_center(A, b)
    _, ro = olp.cube_center(A, b)
    assert abs(r - ro) <= 1e-7 * ro
    w = np.abs(A).sum(1)
    assert (b - A @ x - w * r).min() >= -1e-12 * np.abs(b).max()


def test_conditioned_variant_matches_highs():
    rng = np.random.default_rng(4)
    d = 60
    A0 = np.vstack([np.eye(d), -np.eye(d)])
    U = rng.standard_normal((150, d))
    U /= np.linalg.norm(U, axis=1)[:, None]
    A = np.vstack([A0, U])
    b = np.r_[np.ones(2 * d), 10 ** rng.uniform(-2, 0, 150)]
    x, r = V.cube_center(A, b)
    _, ro = olp.cube_center(A, b)
    assert abs(r - ro) <= 1e-7 * ro and V.feasible(A, b, x)


def test_lp_errors():
    with pytest.raises(NajaError):  # infeasible: x <= -1 and x >= 1
        V.cube_center(np.array([[1.0], [-1.0]]), np.array([-1.0, -1.0]))
    with pytest.raises(NajaError):  # unbounded radius
        V.cube_center(np.array([[1.0, 0.0]]), np.array([1.0]))
    with pytest.raises(NajaError):  # zero row with negative rhs
        V.cube_center(np.array([[0.0], [1.0], [-1.0]]), np.array([-1.0, 1.0, 1.0]))
    with pytest.raises(NajaInvalidArgument):
        V.cube_center(np.array([[np.inf]]), np.array([1.0]))


def test_extra_from_bounds_matches_oracle():
    rng = np.random.default_rng(7)
    d, n = 12, 30
    T, shift, m_lo, m_hi = gem_like(rng, d, n)
    lb, ub = shift - m_lo, shift + m_hi
    lbn, ubn = lb.copy(), ub.copy()
    ubn[::3] -= 0.5 * m_hi[::3]
    lbn[1::4] += 0.25 * m_lo[1::4]
    ubn[5] = ub[5] - 0.5e-9   # within eps: no row
    lbn[7] = lb[7] + 2e-9     # beyond eps: a row
    EA, Eb = V.extra_from_bounds(T, shift, lb, ub, lbn, ubn)
    OA, Ob = olp.extra_from_bounds(T, shift, lb, ub, lbn, ubn)
    np.testing.assert_array_equal(EA, OA)
    np.testing.assert_array_equal(Eb, Ob)
    assert len(Eb) == len(range(0, n, 3)) - (5 % 3 == 0) + len(range(1, n, 4)) + (7 % 4 != 1)
    # the rows encode exactly the new bounds: x = T y + shift <= ub_new at equality
    i = 0
    y = np.linalg.lstsq(T[i:i + 1], [ubn[i] - shift[i]], rcond=None)[0]
    assert abs(EA[0] @ y - Eb[0]) <= 1e-12
    # unchanged bounds -> no rows
    EA0, Eb0 = V.extra_from_bounds(T, shift, lb, ub, lb, ub)
    assert EA0.shape == (0, d) and Eb0.shape == (0,)


def test_extra_from_bounds_validation():
    T = np.eye(2)
    with pytest.raises(NajaInvalidArgument):
        V.extra_from_bounds(T, np.zeros(2), [0, 0], [1, 1], [0.5, 0], [0.4, 1])  # l > u
    with pytest.raises(NajaInvalidArgument):
        V.extra_from_bounds(T, np.zeros(2), [0, 0], [1, 1], [0, 0], [np.nan, 1])
    with pytest.raises(ValueError):
        V.extra_from_bounds(T, np.zeros(2), [0, 0], [1, 1], [0, 0, 0], [1, 1, 1])


def test_conditioned_variants_starts_are_feasible():
    rng = np.random.default_rng(11)
    d, n, nv = 10, 25, 6
    T, shift, m_lo, m_hi = gem_like(rng, d, n)
    A0, b0 = np.vstack([T, -T]), np.r_[m_hi, m_lo]
    lb, ub = shift - m_lo, shift + m_hi
    LB = np.tile(lb, (nv, 1))
    UB = np.tile(ub, (nv, 1))
    for v in range(nv):
        j = rng.choice(n, 4, replace=False)
        if v % 2:
            # exclude the base start y = 0 but keep the polytope non-empty: lb_new
            # halfway to the flux at an interior point y1 on a random ray
            u = rng.standard_normal(d)
            Au = A0 @ u
            y1 = 0.5 * (b0[Au > 0] / Au[Au > 0]).min() * u
            j0 = int(np.argmax(T @ y1))
            LB[v, j0] = shift[j0] + 0.5 * (T[j0] @ y1)
        else:
            UB[v, j] = shift[j] + 0.5 * m_hi[j]
    out = V.conditioned_variants(A0, b0, np.zeros(d), T, shift, lb, ub, LB, UB)
    assert list(out["recentred"]) == [False, True] * (nv // 2)
    for v in range(nv):
        A = np.vstack([A0, out["extra_A"][v]])
        b = np.r_[b0, out["extra_b"][v]]
        assert V.feasible(A, b, out["starts"][v])
        # the sampled fluxes respect the conditioned bounds at the start
        flux = T @ out["starts"][v] + shift
        assert (flux <= UB[v] + 1e-9).all() and (flux >= LB[v] - 1e-9).all()


def test_infeasible_condition_is_reported_like_highs():
    """Conditions that empty the variant (flux 1 duplicates flux 0, which must
    lie above 0.6 m while flux 1 stays below 0.4 m) make the LP infeasible:
    HiGHS says so and the library raises (the reference throws on Gurobi's
    LP status)."""
    rng = np.random.default_rng(11)
    d, n = 10, 25
    T, shift, m_lo, m_hi = gem_like(rng, d, n)
    T[0] = rng.standard_normal(d)
    T[1], shift[1], m_hi[1], m_lo[1] = T[0], shift[0], m_hi[0], m_lo[0]
    A0, b0 = np.vstack([T, -T]), np.r_[m_hi, m_lo]
    lb, ub = shift - m_lo, shift + m_hi
    LB, UB = lb.copy(), ub.copy()
    LB[0] = shift[0] + 0.6 * m_hi[0]
    UB[1] = shift[1] + 0.4 * m_hi[1]
    ea, eb = V.extra_from_bounds(T, shift, lb, ub, LB, UB)
    A, b = np.vstack([A0, ea]), np.r_[b0, eb]
    with pytest.raises(RuntimeError):
        olp.cube_center(A, b)
    with pytest.raises(NajaError):
        V.cube_center(A, b)
