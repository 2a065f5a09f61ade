"""GPU: lane layout of v_mfma_f64_16x16x4f64 and the 64x64 MFMA tile GEMM
used by every assembly step, against numpy."""
import itertools

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


def _mfma4_layouts():
    """Candidate lane maps of v_mfma_f64_4x4x4f64 (4 blocks of 4x4x4): the
    lane's three 2-bit fields (bits 0-1, 2-3, 4-5) assigned to the block and
    the two matrix indices of A(i,k), B(k,j) and D(i,j), in every order."""
    orders = list(itertools.permutations(range(3)))
    for oa in orders:
        for ob in orders:
            for od in orders:
                yield oa, ob, od


def _lane(order, blk, x, y):
    """Lane of (block, first index x, second index y): order[f] is the bit
    field (0: bits 0-1, 1: bits 2-3, 2: bits 4-5) of value f (0 block, 1 x, 2 y)."""
    vals = (blk, x, y)
    return sum(vals[f] << (2 * order[f]) for f in range(3))


def _mfma4_model(a, b, layout):
    oa, ob, od = layout
    out = np.zeros(64)
    for blk in range(4):
        A = np.array([[a[_lane(oa, blk, i, k)] for k in range(4)] for i in range(4)])
        B = np.array([[b[_lane(ob, blk, k, j)] for j in range(4)] for k in range(4)])
        D = A @ B
        for i in range(4):
            for j in range(4):
                out[_lane(od, blk, i, j)] = D[i, j]
    return out


# block in bits 2-3; A(i,k): i in bits 0-1, k in 4-5; B(k,j): j in 0-1, k in 4-5;
# D(i,j): j in 0-1, i in 4-5
MFMA4_LAYOUT = ((1, 0, 2), (1, 2, 0), (1, 2, 0))


def test_mfma4_model_identifies_layouts():
    """CPU: the random-operand test below tells every candidate layout apart."""
    rng = np.random.default_rng(1)
    a = rng.standard_normal(64)
    b = rng.standard_normal(64)
    outs = {lay: _mfma4_model(a, b, lay) for lay in _mfma4_layouts()}
    ref = outs[MFMA4_LAYOUT]
    same = [lay for lay, o in outs.items() if np.allclose(o, ref, rtol=0, atol=1e-13)]
    assert same == [MFMA4_LAYOUT]


def test_mfma_f64_4x4x4_lane_layout():
    """Lane map of the 4-block 4x4x4 FP64 MFMA: block b in lane bits 2-3;
    A(i,k) in lane i + 4b + 16k, B(k,j) in lane j + 4b + 16k, D(i,j) in lane
    j + 4b + 16i (the 16x16x4 map restricted to its four diagonal blocks)."""
    rng = np.random.default_rng(1)
    a = rng.standard_normal(64)
    b = rng.standard_normal(64)
    got = core.debug_mfma4(a, b)
    match = [lay for lay in _mfma4_layouts() if np.allclose(_mfma4_model(a, b, lay), got, rtol=0, atol=1e-13)]
    print("4x4x4 layouts matching the device:", match, flush=True)
    assert match == [MFMA4_LAYOUT]


@pytest.mark.parametrize("ta", [False, True])
@pytest.mark.parametrize("tb", [False, True])
@pytest.mark.parametrize("shape", [(64, 64, 64), (37, 50, 29), (64, 64, 200), (5, 64, 3), (64, 7, 17), (1, 1, 1),
                                   (64, 64, 48), (33, 64, 111)])
def test_tile_gemm(ta, tb, shape):
    M, N, K = shape
    rng = np.random.default_rng(M * 1000 + N * 10 + K)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    C = rng.standard_normal((M, N))
    ref = -1.5 * ((A.T if ta else A) @ (B.T if tb else B)) + 0.5 * C
    got = core.debug_tile_gemm(A, B, C, ta=ta, tb=tb, alpha=-1.5, beta=0.5)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    # beta 0 must not read C (a NaN there stays out of the result)
    got0 = core.debug_tile_gemm(A, B, np.full((M, N), np.nan), ta=ta, tb=tb, alpha=2.0, beta=0.0)
    ref0 = 2.0 * ((A.T if ta else A) @ (B.T if tb else B))
    assert np.abs(got0 - ref0).max() <= 1e-12 * np.abs(ref0).max()


@pytest.mark.parametrize("shape,chains,wpc", [(0, 8, 0), (1, 8, 0), (2, 8, 1), (3, 2, 2)])
def test_mfma_rate_loop(shape, chains, wpc):
    """The FP64 MFMA issue-rate diagnostic runs and gives a plausible rate
    (published FP64 matrix peak of the part: 78.6 TFLOP/s): one operand pair
    (shapes 0, 1) and the operand grid of a register-blocked tile (2, 3), the
    latter with the occupancy capped."""
    t = core.debug_mfma_rate(shape, chains, wpc)
    print("mfma rate shape", shape, "chains", chains, "workgroups/CU", wpc, t, "TFLOP/s", flush=True)
    assert 5.0 < t < 120.0

