"""The allocation-free projector espreso_feti_project_device against
espreso_feti_project, through the C ABI.

Both run the same kernels (the allocation-free one with its coarse vectors in
the library's static device scratch instead of the instance's buffers), so
the results must be bitwise equal.  Two calls in a row on different
instances, the second with a smaller problem (fewer multipliers and a smaller
coarse space) than the first, so that the shared scratch holds the first
call's values when the second starts; then a hybrid instance, a call on the
caller's own stream, and the state check.
"""
import sys

import numpy as np
import pytest

from espreso_amd import core
from espreso_amd import problems as PB

pytestmark = pytest.mark.gpu

BIG = PB.scaled(PB.HTFETI_512x10, clusters=(2, 2, 2), elements=(4, 4, 4))   # 64 subdomains
SMALL = PB.scaled(PB.GRAVITY, traction_seed=9, traction_scale=1e5)         # 16 subdomains


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check(spec, method, seed, use_stream=False):
    import torch
    g = core.GridFETI(spec, method=method, preconditioner="dirichlet")
    g.setup()
    n = g.n_lambda
    say(f"case: {spec['name']} {method}, {n} multipliers, coarse {g.sizes()[2]}, stream {use_stream}")
    x = np.random.default_rng(seed).standard_normal(n)
    want = g.project(x)
    dx = torch.tensor(x, dtype=torch.float64, device="cuda")
    dy = torch.full_like(dx, float("nan"))
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream().cuda_stream if use_stream else None
    g.project_device(dx.data_ptr(), dy.data_ptr(), stream)
    torch.cuda.synchronize()
    got = dy.cpu().numpy()
    g.close()
    assert np.all(np.isfinite(got))
    assert np.array_equal(got, want), np.abs(got - want).max()
    return n


def test_two_calls_second_smaller():
    n1 = check(BIG, "total", 1)
    n2 = check(SMALL, "total", 2)
    assert n2 < n1


def test_hybrid_and_caller_stream():
    check(BIG, "hybrid", 3)
    check(SMALL, "total", 4, use_stream=True)


def test_refused_before_setup():
    g = core.GridFETI(SMALL, preconditioner="dirichlet")
    with pytest.raises(RuntimeError, match="setup"):
        g.project_device(0, 0)
    g.close()
