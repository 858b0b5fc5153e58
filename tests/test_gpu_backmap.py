"""GPU MFMA backmap vs the FP64 oracle (backmap.cu:21-48 semantics)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

SHAPES = [(8, 5, 1000), (37, 7, 129), (2583, 512, 4000), (130, 64, 257), (1, 1, 1), (300, 33, 0),
          (100, 100, 5003), (128, 128, 333), (64, 127, 100), (128, 4, 64), (12, 8, 1000), (17, 36, 300),
          (100, 100, 128), (100, 100, 129), (129, 100, 300)]


@pytest.mark.parametrize("n,d,ns", SHAPES)
def test_backmap_matches_fp64_oracle(gpu, n, d, ns):
    import naja_amd

    t = gpu
    rng = np.random.default_rng(n * 1000 + d)
    T = rng.standard_normal((n, d))
    shift = 0.5 * rng.standard_normal(n)
    Y = rng.standard_normal((ns, d)).astype(np.float32)
    V = naja_amd.backmap(t.from_numpy(T.astype(np.float32)), t.from_numpy(shift.astype(np.float32)),
                         t.from_numpy(Y))
    t.cuda.synchronize()
    V = V.cpu().numpy().astype(np.float64)
    assert V.shape == (ns, n)
    if ns == 0:
        return
    ref = oracle.backmap(T.astype(np.float32).astype(np.float64), shift.astype(np.float32).astype(np.float64), Y)
    # FP32 accumulation of d products: error bound ~ d * eps32 * sum|T||Y|
    scale = np.abs(Y.astype(np.float64)) @ np.abs(T.T) + np.abs(shift)
    err = np.abs(V - ref) / scale
    print(f"backmap n={n} d={d} ns={ns}: max scaled err {err.max():.3e}", flush=True)
    assert err.max() <= 4 * d * 2 ** -24 + 1e-7


def test_backmap_reconstruction(gpu):
    """test_gpu_backmap_correctness.cu:84-112: the reduced preimage of V through
    T's pseudo-inverse reproduces Y."""
    import naja_amd

    t = gpu
    rng = np.random.default_rng(2024)
    T = rng.standard_normal((8, 5))
    shift = 0.5 * rng.standard_normal(8)
    Y = rng.uniform(-1, 1, (400, 5)).astype(np.float32)
    V = naja_amd.backmap(t.from_numpy(T.astype(np.float32)), t.from_numpy(shift.astype(np.float32)),
                         t.from_numpy(Y)).cpu().numpy().astype(np.float64)
    Y_rec = (V - shift.astype(np.float32)) @ np.linalg.pinv(T.astype(np.float32).astype(np.float64)).T
    assert np.abs(Y_rec - Y).max() <= 1e-5


def test_backmap_shape_errors(gpu):
    import naja_amd

    t = gpu
    with pytest.raises(naja_amd.NajaInvalidArgument):
        naja_amd.backmap(t.zeros((4, 3)), t.zeros(4), t.zeros((5, 2)))
    with pytest.raises(naja_amd.NajaInvalidArgument):
        naja_amd.backmap(t.zeros((4, 3)), t.zeros(3), t.zeros((5, 3)))
