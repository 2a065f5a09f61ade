"""GPU parity for uneven subdomains and multiple load cases (resolve).

Bars: primal within 1e-8 relative of the undecomposed direct solve of the same
global mesh (FP64), for the workload's own loads and for seeded U(-1,1) nodal
load cases solved one after another on the assembled operators.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from espreso_amd import core
from espreso_amd import problems as PB

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


HEAT = dict(PB.HEAT_UNEVEN_2000, domains=(3, 3, 4), slab_elements=PB.uneven_slabs((3, 3, 4), 5, 3, 7),
            name="heat_uneven_36")
ELAST = dict(PB.ELASTICITY_UNEVEN_512, domains=(3, 2, 3), slab_elements=PB.uneven_slabs((3, 2, 3), 6, 3, 6),
             name="elasticity_uneven_18")


def with_material(spec, seed=3):
    """Per-subdomain conductivity (log-uniform [0.1, 10]) as an explicit element field."""
    if spec["physics"] != 1:
        return spec
    sl = core.slabs(spec)
    off = [np.concatenate([[0], np.cumsum(v)]) for v in sl]
    NE = [o[-1] for o in off]
    rng = np.random.default_rng(seed)
    ks = 10.0 ** rng.uniform(-1, 1, size=[len(v) for v in sl])
    ez, ey, ex = np.meshgrid(np.arange(NE[2]), np.arange(NE[1]), np.arange(NE[0]), indexing="ij")
    sx = np.searchsorted(off[0], ex.ravel(), side="right") - 1
    sy = np.searchsorted(off[1], ey.ravel(), side="right") - 1
    sz = np.searchsorted(off[2], ez.ravel(), side="right") - 1
    s = dict(spec, material_elem=ks[sx, sy, sz])
    s.pop("conductivity_seed", None)
    return s


def test_uneven_seeded_conductivity():
    """The builder's seeded per-subdomain conductivity on uneven subdomains,
    against the direct solve with the same values as an element field."""
    say("case: uneven heat, seeded conductivity")
    g = core.GridFETI(HEAT, preconditioner="dirichlet", precision=1e-11, max_iterations=3000)
    g.setup()
    g.solve()
    u = g.nodal_solution()
    ref = O.undecomposed(HEAT)
    err = rel(u, O.global_solve(ref))
    res = O.global_residual(ref, u)
    say("uneven seeded err", err, "residual", res)
    assert err < 1e-8
    assert res < 1e-6
    g.close()


@pytest.mark.parametrize("name", ["heat", "elasticity"])
def test_uneven_solve_and_load_cases(name):
    spec = with_material(HEAT if name == "heat" else ELAST)
    say("case: uneven", spec["name"])
    g = core.GridFETI(spec, preconditioner="dirichlet", precision=1e-11, max_iterations=3000)
    g.setup()
    st = g.solve()
    ug = O.global_solve(PB.uniform_equivalent(spec))
    err = rel(g.nodal_solution(), ug)
    say("uneven", name, "iterations", st["iterations"], "err", err)
    assert err < 1e-8
    loads = PB.nodal_load_cases(spec, 3)
    for q in range(loads.shape[0]):
        g.set_nodal_rhs(loads[q])
        st = g.resolve()
        uq = O.global_solve_load(PB.uniform_equivalent(spec), loads[q])
        err = rel(g.nodal_solution(), uq)
        say("load case", q, "iterations", st["iterations"], "err", err)
        assert err < 1e-8
    g.close()
