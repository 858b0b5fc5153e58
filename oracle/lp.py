"""ORACLE — test infrastructure only: the start-point LP of
axis_aligned_cube_center_lp_gurobi (feasible_start_lp.cpp:33-87) written out
as the reference formulates it (maximise r; rows a_i.x + ||a_i||_1 r <= b_i;
x free, r >= 0) and solved by scipy's HiGHS, an LP solver independent of the
library's interior-point method. Its optimal value pins the library's radius.
"""
from __future__ import annotations

import numpy as np
from scipy.optimize import linprog


def cube_center(A, b):
    A = np.asarray(A, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64).ravel()
    m, n = A.shape
    w = np.abs(A).sum(1)
    res = linprog(np.r_[np.zeros(n), -1.0], A_ub=np.c_[A, w], b_ub=b,
                  bounds=[(None, None)] * n + [(0, None)], method="highs")
    if res.status != 0:
        raise RuntimeError(f"axis_aligned_cube_center: LP status {res.status}")
    return res.x[:n], float(res.x[n])


def extra_from_bounds(T, shift, lb_base, ub_base, lb_new, ub_new, eps=1e-9):
    """cmd_prepare.cpp:127-139, one flux at a time."""
    rows, rhs = [], []
    for i in range(len(lb_base)):
        if ub_new[i] < ub_base[i] - eps:
            rows.append(np.asarray(T[i], dtype=np.float64))
            rhs.append(ub_new[i] - shift[i])
        if lb_new[i] > lb_base[i] + eps:
            rows.append(-np.asarray(T[i], dtype=np.float64))
            rhs.append(-lb_new[i] + shift[i])
    d = np.asarray(T).shape[1]
    return (np.array(rows).reshape(-1, d), np.array(rhs, dtype=np.float64))
