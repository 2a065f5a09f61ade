"""Host-side pins of the adjoint (compliance gradient) path.

* the product's closed-form unit element equals the oracle's Gauss element
* per-element Young's modulus input reproduces the seeded field, product and
  oracle agree on K_i with it
* the oracle's reference gradient dc/dE_e = -u_e^T K_e0 u_e: Euler's relation
  sum_e E_e dc/dE_e = -c (c is homogeneous of degree -1 in E) and central
  finite differences of the undecomposed direct solve on a tiny mesh
"""
import numpy as np
import pytest

import oracle_lib as O
from espreso_amd import core
from espreso_amd import problems as PB

TINY = PB.scaled(PB.ADJOINT_64x10, domains=(2, 2, 2), elements=(2, 2, 2))


def _young(spec):
    n = core.n_elements(spec)
    rng = np.random.default_rng(11)
    return spec["young"] * (0.1 + 0.9 * rng.random(n))


def test_unit_element_matches_gauss():
    for spec in (PB.ADJOINT_64x10, PB.HEAT_SIMPLE, PB.scaled(PB.GRAVITY, length=(1.0, 2.0, 3.0))):
        Kp = core.unit_element(spec)
        Ko = O.unit_element(spec)
        assert np.abs(Kp - Ko).max() <= 1e-12 * np.abs(Ko).max()


def test_material_elem_input_matches_oracle():
    spec = dict(TINY, material_elem=_young(TINY))
    spec.pop("young_seed")
    o = O.Oracle(spec, preconditioner=0)
    for s in (0, 5):
        e = core.grid_export(spec, s)
        K = np.zeros((e["n"], e["n"]))
        for r in range(e["n"]):
            K[r, e["colidx"][e["rowptr"][r]:e["rowptr"][r + 1]]] = e["vals"][e["rowptr"][r]:e["rowptr"][r + 1]]
        Ko = o.dense(s, 0)
        assert np.abs(K - Ko).max() <= 1e-10 * np.abs(Ko).max()
    o.close()


def _compliance(spec):
    u = O.global_solve(spec).reshape(-1)
    o = O.Oracle(spec, preconditioner=0)
    f = np.zeros_like(u)
    nd = o.ndof
    for s in range(o.n_sub):
        dofs = (o.l2g(s)[:, None] * nd + np.arange(nd)).reshape(-1)
        np.add.at(f, dofs, o.rhs(s))
    o.close()
    return f @ u


def test_reference_gradient_euler_relation():
    spec = dict(TINY, material_elem=_young(TINY))
    g, u = O.compliance_gradient(spec)
    c = _compliance(spec)
    assert c > 0
    assert abs(spec["material_elem"] @ g + c) <= 1e-10 * c


def test_reference_gradient_finite_differences():
    E = _young(TINY)
    spec = dict(TINY, material_elem=E)
    g, _ = O.compliance_gradient(spec)
    rng = np.random.default_rng(5)
    for e in rng.choice(len(E), 6, replace=False):
        h = 1e-4 * E[e]
        Ep, Em = E.copy(), E.copy()
        Ep[e] += h
        Em[e] -= h
        fd = (_compliance(dict(TINY, material_elem=Ep)) - _compliance(dict(TINY, material_elem=Em))) / (2 * h)
        assert abs(fd - g[e]) <= 1e-6 * abs(g[e]), (e, fd, g[e])
