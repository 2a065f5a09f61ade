"""GPU: lane layout of v_mfma_f64_16x16x4f64 and the 64x64 MFMA tile GEMM
used by every assembly step, against numpy."""
import numpy as np
import pytest

from espreso_amd import core

pytestmark = pytest.mark.gpu


def test_mfma_f64_lane_layout():
    rng = np.random.default_rng(0)
    a = rng.standard_normal(64)
    b = rng.standard_normal(64)
    out = core.debug_mfma(a, b)
    A = np.zeros((16, 4))
    B = np.zeros((4, 16))
    for l in range(64):
        A[l % 16, l // 16] = a[l]     # lane l holds A[l%16][l/16]
        B[l // 16, l % 16] = b[l]     # lane l holds B[l/16][l%16]
    D = A @ B
    got = np.zeros((16, 16))
    for l in range(64):
        for v in range(4):
            got[l // 16 + 4 * v, l % 16] = out[l, v]  # value v of lane l: row l/16 + 4v
    print("max layout error", np.abs(got - D).max(), flush=True)
    assert np.allclose(got, D, rtol=0, atol=1e-13)


@pytest.mark.parametrize("ta", [False, True])
@pytest.mark.parametrize("tb", [False, True])
@pytest.mark.parametrize("shape", [(64, 64, 64), (37, 50, 29), (64, 64, 200), (5, 64, 3)])
def test_tile_gemm(ta, tb, shape):
    M, N, K = shape
    rng = np.random.default_rng(M * 1000 + N * 10 + K)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    C = rng.standard_normal((M, N))
    ref = -1.5 * ((A.T if ta else A) @ (B.T if tb else B)) + 0.5 * C
    got = core.debug_tile_gemm(A, B, C, ta=ta, tb=tb, alpha=-1.5, beta=0.5)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("shape", [0, 1])
def test_mfma_rate_loop(shape):
    """The FP64 MFMA issue-rate diagnostic runs and gives a plausible rate
    (published FP64 matrix peak of the part: 78.6 TFLOP/s)."""
    t = core.debug_mfma_rate(shape, 8)
    print("mfma rate shape", shape, t, "TFLOP/s", flush=True)
    assert 5.0 < t < 120.0

