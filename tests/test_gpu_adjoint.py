"""GPU adjoint: re-solves on the assembled operators and the compliance
gradient with respect to per-element Young's moduli (torch.autograd shim).

Bars: re-solve with new loads equal to a fresh solve (1e-10); gradient equal
to the oracle's reference -u_e^T K_e0 u_e from the undecomposed direct solve
(1e-8 relative to max |g|, FP64 path; 1e-6 with FP32-stored F_i and
refinement); Euler's relation sum_e E_e dc/dE_e = -c (1e-8); central finite
differences on the 40^3 workload (h = 1e-6 relative): a random direction over
all elements (1e-5) and 20 single elements, whose tolerance is set by the
size of the change of c against the accuracy of c itself: the compliances of
the differences are solved to a relative dual residual of 1e-13, and c is
taken to be good to 1e-13 c (printed with each case).
"""
import sys

import numpy as np
import pytest
import torch

import oracle_lib as O
from espreso_amd import autograd as AG
from espreso_amd import core
from espreso_amd import problems as PB

pytestmark = pytest.mark.gpu

SMALL = PB.scaled(PB.ADJOINT_64x10, domains=(2, 2, 2), elements=(4, 4, 4))
OPTS = dict(preconditioner="dirichlet", precision=1e-12, max_iterations=2000)
FD_OPTS = dict(OPTS, precision=1e-13, max_iterations=4000)  # compliances of the finite differences


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def young(spec, seed=11):
    n = core.n_elements(spec)
    rng = np.random.default_rng(seed)
    return spec["young"] * (0.1 + 0.9 * rng.random(n))


def grad_of(spec, E, **opts):
    t = torch.tensor(E, dtype=torch.float64, requires_grad=True)
    c = AG.compliance(t, spec, **dict(OPTS, **opts))
    c.backward()
    return float(c), t.grad.numpy().copy()


def compliance(spec, E, **opts):
    g = core.GridFETI(dict(spec, material_elem=E), **dict(OPTS, **opts))
    g.setup()
    g.solve()
    c = g.load_work()
    g.close()
    return c


def test_resolve_matches_fresh_solve():
    say("case: resolve with new loads")
    spec = PB.scaled(PB.GRAVITY, traction_seed=4, traction_scale=1e5)
    g = core.GridFETI(spec, **OPTS)
    g.setup()
    g.solve()
    u1 = g.nodal_solution()
    f = g.nodal_load()
    g.set_nodal_rhs(2.0 * f)
    g.resolve()
    u2 = g.nodal_solution()
    assert np.abs(u2 - 2.0 * u1).max() <= 1e-9 * np.abs(u1).max()
    rng = np.random.default_rng(2)
    f3 = rng.standard_normal(f.shape) * np.abs(f).max()
    f3[np.abs(f) == 0] = 0.0  # keep the loads off the clamped face
    g.set_nodal_rhs(f3)
    st = g.resolve()
    u3 = g.nodal_solution()
    g.close()
    # fresh solve of the same loads: superposition of the original problem's pieces
    g2 = core.GridFETI(spec, **OPTS)
    g2.setup()
    g2.set_nodal_rhs(f3)
    g2.resolve()
    u4 = g2.nodal_solution()
    g2.close()
    assert np.abs(u3 - u4).max() <= 1e-9 * np.abs(u4).max()
    assert st["iterations"] > 0


@pytest.mark.parametrize("method", ["total", "hybrid"])
def test_gradient_vs_oracle(method):
    say("case: gradient vs oracle", method)
    E = young(SMALL)
    spec = dict(SMALL, material_elem=E)
    spec.pop("young_seed")
    c, g = grad_of(spec, E, method=method)
    gref, _ = O.compliance_gradient(spec)
    err = np.abs(g - gref).max() / np.abs(gref).max()
    say("gradient rel err", err, "euler", abs(E @ g + c) / c)
    assert err <= 1e-8
    assert abs(E @ g + c) <= 1e-8 * c


def test_gradient_fp32_schur():
    say("case: gradient with FP32 F_i")
    E = young(SMALL)
    spec = dict(SMALL, material_elem=E)
    _, g64 = grad_of(spec, E)
    _, g32 = grad_of(spec, E, sc_precision="fp32")
    err = np.abs(g32 - g64).max() / np.abs(g64).max()
    say("fp32 vs fp64 gradient", err)
    assert err <= 1e-6


def test_gradient_finite_differences_40cube():
    spec = PB.ADJOINT_64x10
    E = young(spec, seed=int(spec["young_seed"] % 1000))
    spec = dict(spec, material_elem=E)
    spec.pop("young_seed")
    say("case: 40^3 gradient vs central differences")
    c, g = grad_of(spec, E)
    say("compliance", c, "euler", abs(E @ g + c) / c)
    assert abs(E @ g + c) <= 1e-8 * c
    # directional derivative along a random direction over all elements
    rng = np.random.default_rng(7)
    d = rng.choice([-1.0, 1.0], size=E.size) * E
    h = 1e-6
    fd = (compliance(spec, E + h * d, **FD_OPTS) - compliance(spec, E - h * d, **FD_OPTS)) / (2 * h)
    ad = g @ d
    say("directional: fd", fd, "adjoint", ad, "rel", abs(fd - ad) / abs(ad))
    assert abs(fd - ad) <= 1e-5 * abs(ad)
    # single elements: h = 1e-6 E_e changes c by ~2e-6 E_e |g_e|; the solves
    # give c to ~1e-13 c, which bounds what the difference can resolve
    errs = []
    for e in rng.choice(E.size, 20, replace=False):
        he = 1e-6 * E[e]
        Ep, Em = E.copy(), E.copy()
        Ep[e] += he
        Em[e] -= he
        fd = (compliance(spec, Ep, **FD_OPTS) - compliance(spec, Em, **FD_OPTS)) / (2 * he)
        noise = 1e-13 * c / he
        errs.append(abs(fd - g[e]) / abs(g[e]))
        say(f"element {e}: fd {fd:.9e} adjoint {g[e]:.9e} rel {errs[-1]:.2e} (resolvable {noise / abs(g[e]):.1e})")
        assert abs(fd - g[e]) <= 1e-5 * abs(g[e]) + 3 * noise
    say("single-element relative errors: median", float(np.median(errs)), "max", float(np.max(errs)))
