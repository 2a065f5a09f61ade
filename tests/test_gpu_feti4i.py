"""GPU: the FETI4I drop-in (element matrices -> subdomains -> HIP core)
against the oracle's undecomposed direct solve."""
import numpy as np
import pytest

import oracle_lib as O
from espreso_amd import core
from espreso_amd import problems as PB
from feti4i_inputs import element_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("subdomains,prec", [(6, 3), (8, 1), (3, 0)])
def test_feti4i_solve_matches_global(subdomains, prec):
    spec = PB.scaled(PB.GRAVITY, clusters=(1, 1, 1), domains=(6, 4, 4), elements=(1, 1, 1), traction_seed=4,
                     traction_scale=1e5)
    print("case: FETI4I gravity", subdomains, prec, flush=True)
    elements, size, rhs, dirichlet, o = element_problem(spec)
    f4 = core.FETI4I(elements, size, rhs, dirichlet, subdomains=subdomains, preconditioner=prec, precision=1e-10)
    u = f4.solve()
    ug = O.global_solve(spec).ravel()
    assert np.abs(u - ug).max() / np.abs(ug).max() < 1e-8
    st = f4.stats()
    assert st[0] > 0 and st[1] <= st[2]
    # new right-hand side on the same instance
    f4.update_rhs(2.0 * rhs)
    u2 = f4.solve()
    assert np.abs(u2 - 2.0 * ug).max() / np.abs(ug).max() < 1e-8
    f4.close()


def test_feti4i_heat():
    spec = PB.scaled(PB.HEAT_SIMPLE, clusters=(1, 1, 1), domains=(8, 6, 4), elements=(1, 1, 1))
    print("case: FETI4I heat", flush=True)
    elements, size, rhs, dirichlet, o = element_problem(spec)
    f4 = core.FETI4I(elements, size, rhs, dirichlet, subdomains=5, precision=1e-10)
    u = f4.solve()
    assert abs(u.min() - 100) < 1e-7 and abs(u.max() - 300) < 1e-7 and abs(u.mean() - 200) < 1e-7
    assert f4.info()["max_kdim"] == 1
    f4.close()


@pytest.mark.parametrize("subdomains,prec", [(6, 3), (8, 1), (2, 3)])
def test_feti4i_hybrid_matches_global(subdomains, prec):
    """HYBRID_FETI through FETI4I: the instance is one cluster glued by corners;
    same solution as the undecomposed direct solve, and as TOTAL_FETI."""
    spec = PB.scaled(PB.GRAVITY, clusters=(1, 1, 1), domains=(6, 4, 4), elements=(1, 1, 1), traction_seed=4,
                     traction_scale=1e5)
    print("case: FETI4I hybrid gravity", subdomains, prec, flush=True)
    elements, size, rhs, dirichlet, o = element_problem(spec)
    f4 = core.FETI4I(elements, size, rhs, dirichlet, subdomains=subdomains, preconditioner=prec, precision=1e-10,
                     method="hybrid")
    u = f4.solve()
    ug = O.global_solve(spec).ravel()
    info = f4.info()
    print("hybrid: iterations", info["iterations"], "coarse", info["n_coarse"], "corner rows", info["n_corner_rows"],
          flush=True)
    assert np.abs(u - ug).max() / np.abs(ug).max() < 1e-8
    # the cluster's rigid body modes, plus relative rotations of subdomain
    # pairs that a single corner node glues only at one point
    assert 6 <= info["n_coarse"] <= 6 * subdomains
    f4.update_rhs(2.0 * rhs)
    u2 = f4.solve()
    assert np.abs(u2 - 2.0 * ug).max() / np.abs(ug).max() < 1e-8
    f4.close()


def test_feti4i_hybrid_heat():
    spec = PB.scaled(PB.HEAT_SIMPLE, clusters=(1, 1, 1), domains=(8, 6, 4), elements=(1, 1, 1))
    print("case: FETI4I hybrid heat", flush=True)
    elements, size, rhs, dirichlet, o = element_problem(spec)
    f4 = core.FETI4I(elements, size, rhs, dirichlet, subdomains=5, precision=1e-10, method="hybrid")
    u = f4.solve()
    assert abs(u.min() - 100) < 1e-7 and abs(u.max() - 300) < 1e-7 and abs(u.mean() - 200) < 1e-7
    assert f4.info()["n_coarse"] == 1  # one corner fixes the constant mode of a pair
    f4.close()


@pytest.mark.parametrize("method", ["total", "hybrid"])
def test_feti4i_one_subdomain(method):
    """FETI4I with a single subdomain (HYBRID_FETI: one cluster without corner
    rows) against the direct solve."""
    spec = PB.scaled(PB.GRAVITY, clusters=(1, 1, 1), domains=(3, 2, 2), elements=(1, 1, 1), traction_seed=4,
                     traction_scale=1e5)
    print("case: FETI4I one subdomain", method, flush=True)
    elements, size, rhs, dirichlet, o = element_problem(spec)
    f4 = core.FETI4I(elements, size, rhs, dirichlet, subdomains=1, preconditioner=3, precision=1e-10, method=method)
    u = f4.solve()
    ug = O.global_solve(spec).ravel()
    err = np.abs(u - ug).max() / np.abs(ug).max()
    print("FETI4I one subdomain", method, "err", err, flush=True)
    assert err < 1e-8
    f4.close()
