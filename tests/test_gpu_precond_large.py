"""The Dikin-ellipsoid rounding at any size on the GPU
(naja_precond_ellipsoid_large: FP64 Hessian by rocBLAS, blocked Cholesky in
64 x 64 LDS blocks) against the oracle's LAPACK factor: the same P to FP32
rounding and P^T H P = I, up to the large-scale dimension d = 2400 on a
reduced row count."""
import numpy as np
import pytest

from oracle import precond as op
from tests.polytopes import genome_like

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,d,off_centre", [(1000, 300, False), (700, 130, True), (3000, 2400, False)])
def test_large_ellipsoid_matches_oracle(gpu, n, d, off_centre):
    import naja_amd

    t = gpu
    g = genome_like(n, d, seed=n)
    A32 = g["A"].astype(np.float32)
    A = A32.astype(np.float64)
    xc = None
    if off_centre:  # a point inside: a tenth of the way to the tightest row
        xc = (1e-4 * np.random.default_rng(d).standard_normal(d)).astype(np.float32)
        assert (g["b"] - A @ xc > 0).all()
    print(f"[large ellipsoid] m={A.shape[0]} d={d} off_centre={off_centre}", flush=True)
    P = naja_amd.ellipsoid_rounding_large(t.from_numpy(A32).cuda(), t.from_numpy(g["b"]).cuda(),
                                          x_c=t.from_numpy(xc).cuda() if xc is not None else None)
    t.cuda.synchronize()
    P = P.cpu().numpy().astype(np.float64)
    Pref = op.ellipsoid(A, g["b"], None if xc is None else xc.astype(np.float64))
    s = g["b"] - (A @ xc if xc is not None else 0.0)
    H = (A.T / s ** 2) @ A
    err = np.abs(P - Pref).max() / np.abs(Pref).max()
    ident = np.abs(P.T @ H @ P - np.eye(d)).max()
    print(f"[large ellipsoid] rel err {err:.2e}, |P^T H P - I| {ident:.2e}", flush=True)
    assert np.allclose(np.tril(P, -1), 0.0)  # upper triangular
    assert err <= 1e-5 and ident <= 1e-4


def test_large_ellipsoid_rejects_exterior_centre(gpu):
    import naja_amd

    t = gpu
    A = np.vstack([np.eye(5), -np.eye(5)]).astype(np.float32)
    with pytest.raises(naja_amd.NajaInvalidArgument):
        naja_amd.ellipsoid_rounding_large(t.from_numpy(A).cuda(), t.ones(10, dtype=t.float64, device="cuda"),
                                          x_c=t.full((5,), 2.0, device="cuda"))


@pytest.mark.parametrize("shape", ["genome_like_300", "genome_scale_512"])
def test_large_facet_matches_oracle(gpu, shape):
    """naja_precond_facet_large against the oracle's FP64 restatement of the
    ordered greedy rule (tightest dual rows first, kept while 2 % of the norm
    survives the projection): the same P = N^-1 to FP32 rounding, N P = I for
    the oracle's accepted rows, and a rounded chain that stays inside."""
    import naja_amd
    from naja_amd import workloads

    t = gpu
    if shape == "genome_like_300":
        g = genome_like(1000, 300, seed=1000)
    else:  # the genome-scale configuration: n = 2583 reactions, d = 512
        g = workloads.genome_scale(n=2583, d=512, seed=0)
    A32 = g["A"].astype(np.float32)
    A = A32.astype(np.float64)
    d = A.shape[1]
    At, bt = t.from_numpy(A32).cuda(), t.from_numpy(g["b"]).cuda()
    P = naja_amd.facet_rounding_large(At, bt, theta=0.02)
    t.cuda.synchronize()
    Pg = P.cpu().numpy().astype(np.float64)
    Pref, acc = op.facet_ordered(A, g["b"], 0.02)
    N = (A / g["b"][:, None])[acc]
    err = np.abs(Pg - Pref).max() / np.abs(Pref).max()
    ident = np.abs(N @ Pg - np.eye(d)).max()
    print(f"[large facet {shape}] rel err {err:.2e}, |N P - I| {ident:.2e}", flush=True)
    assert err <= 1e-5 and ident <= 1e-4
    X0 = t.zeros((1, 3, d), device="cuda")
    batch = naja_amd.VariantBatch(A32, g["b"], [np.zeros((0, d))], [np.zeros(0)])
    Y = naja_amd.coordinate_hit_and_run_batched(batch, None, None, None, X0, 200, 5, 3, seed=4, pair_prob=0.3,
                                                P=P[None].contiguous(), c=t.zeros((1, d), device="cuda"))
    t.cuda.synchronize()
    y = Y[0].double()
    viol = float((y @ At.double().T - bt[None, :]).max())
    assert viol <= 1e-5 * max(1.0, float(bt.abs().max()))


def test_large_facet_needs_enough_facets(gpu):
    import naja_amd

    t = gpu
    A = np.vstack([np.eye(6), -np.eye(6)])[:, :6].astype(np.float32)
    A[:, 5] = 0.0  # no facet bounds coordinate 5
    A = np.vstack([A, np.zeros((1, 6), np.float32)])
    b = np.ones(A.shape[0])
    with pytest.raises(naja_amd.NajaError):
        naja_amd.facet_rounding_large(t.from_numpy(A).cuda(), t.from_numpy(b).cuda())


@pytest.mark.parametrize("n,d", [(1000, 300), (2583, 512)])
def test_analytic_center_matches_oracle(gpu, n, d):
    """naja_analytic_center (damped Newton, FP64 on the GPU) equals the
    oracle's numpy Newton iteration on a genome-like polytope whose start is
    1e-3 from its closest facet, and the barrier gradient vanishes there."""
    import naja_amd
    from naja_amd import workloads

    t = gpu
    g = workloads.genome_scale(n=n, d=d, seed=0) if n == 2583 else genome_like(n, d, seed=n)
    A32 = g["A"].astype(np.float32)
    A = A32.astype(np.float64)
    x, steps = naja_amd.analytic_center(t.from_numpy(A32).cuda(), t.from_numpy(g["b"]).cuda())
    x = x.cpu().numpy().astype(np.float64)
    xr, steps_r = op.analytic_center(A, g["b"])
    s = g["b"] - A @ x
    grad = A.T @ (1.0 / s)
    scale = np.abs(A.T) @ (1.0 / s)
    print(f"[analytic centre n={n} d={d}] steps {steps} (oracle {steps_r}), max |x - x_ref| "
          f"{np.abs(x - xr).max():.2e}, min slack {s.min():.2e}, max |grad| / scale {np.max(np.abs(grad) / scale):.2e}",
          flush=True)
    assert (s > 0).all()
    # x is the FP32 rounding of the FP64 iterate: compare in the metric of the
    # slacks it determines.
    np.testing.assert_allclose(g["b"] - A @ xr, s, rtol=1e-3, atol=1e-9)
    assert np.max(np.abs(grad) / scale) <= 1e-3
