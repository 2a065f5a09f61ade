"""FETI4I front end, host part (no device): decomposition of element
matrices into subdomains, algebraic kernels, B1 rows."""
import numpy as np
import pytest

import oracle_lib as O
from espreso_amd import core
from espreso_amd import problems as PB
from feti4i_inputs import element_problem

SPEC = PB.scaled(PB.GRAVITY, clusters=(1, 1, 1), domains=(6, 4, 4), elements=(1, 1, 1), traction_seed=4,
                 traction_scale=1e5)


@pytest.fixture(scope="module")
def inputs():
    return element_problem(SPEC)


def test_decomposition_reassembles_and_kernels(inputs):
    elements, size, rhs, dirichlet, o = inputs
    f4 = core.FETI4I(elements, size, rhs, dirichlet, subdomains=6)
    info = f4.info()
    assert info["subdomains"] == 6 and info["max_kdim"] == 6
    K = np.zeros((size, size))
    f = np.zeros(size)
    for s in range(6):
        d = f4.subdomain(s)
        n = len(d["gdofs"])
        Ks = np.zeros((n, n))
        for r in range(n):
            sl = slice(d["rowptr"][r], d["rowptr"][r + 1])
            Ks[r, d["colidx"][sl]] = d["vals"][sl]
        g = d["gdofs"]
        K[np.ix_(g, g)] += Ks
        f[g] += d["f"]
        R = d["R"]
        assert R.shape[1] == 6
        assert np.linalg.norm(Ks @ R) <= 1e-9 * np.linalg.norm(Ks)
        assert np.allclose(R.T @ R, np.eye(6), atol=1e-10)
    Kref = np.zeros((size, size))
    for nodes, dofs, Ke in elements:
        Kref[np.ix_(dofs, dofs)] += Ke
    assert np.abs(K - Kref).max() <= 1e-13 * np.abs(Kref).max()
    assert np.allclose(f, rhs, rtol=0, atol=1e-12 * np.abs(rhs).max())
    f4.close()


def test_b1_rows_hold_for_the_solution(inputs):
    elements, size, rhs, dirichlet, o = inputs
    f4 = core.FETI4I(elements, size, rhs, dirichlet, subdomains=5)
    u = O.global_solve(SPEC).ravel()
    nl = f4.info()["n_lambda"]
    Bu = np.zeros(nl)
    cnt = np.zeros(nl, int)
    dmap = dict(dirichlet)
    cval = np.zeros(nl)
    for s in range(5):
        d = f4.subdomain(s)
        us = u[d["gdofs"]]
        np.add.at(Bu, d["lam"], d["sign"] * us[d["dof"]])
        np.add.at(cnt, d["lam"], 1)
        for l, dof in zip(d["lam"], d["dof"]):
            g = d["gdofs"][dof]
            if g in dmap:
                cval[l] = dmap[g]
    assert set(np.unique(cnt)) <= {1, 2}
    assert np.allclose(Bu, cval, atol=1e-12)
    f4.close()


def test_invalid_options_are_reported(inputs):
    elements, size, rhs, dirichlet, o = inputs
    with pytest.raises(RuntimeError, match="SUBDOMAINS"):
        core.FETI4I(elements, size, rhs, dirichlet, subdomains=0)
    with pytest.raises(RuntimeError, match="PRECONDITIONER"):
        core.FETI4I(elements, size, rhs, dirichlet, subdomains=2, preconditioner=7)


@pytest.mark.parametrize("subdomains", [2, 6])
def test_hybrid_corners(inputs, subdomains):
    """HYBRID_FETI corners: every B0 row holds +1 and -1 on two subdomain
    copies of the same instance DOF; every group of interface nodes with the
    same owner set gets one corner node (all DOFs of the node) per
    consecutive owner pair; rows are unique."""
    elements, size, rhs, dirichlet, o = inputs
    f4 = core.FETI4I(elements, size, rhs, dirichlet, subdomains=subdomains, method="hybrid")
    info = f4.info()
    assert info["method"] == 1 and info["n_corner_rows"] > 0
    nd = info["ndof_node"]
    rows = {}
    owners = {}
    for s in range(subdomains):
        d = f4.subdomain(s)
        for g in d["gdofs"]:
            owners.setdefault(int(g) // nd, set()).add(s)
        for r, dof, sg in zip(*f4.corners(s)):
            rows.setdefault(int(r), []).append((s, int(d["gdofs"][dof]), float(sg)))
    assert sorted(rows) == list(range(info["n_corner_rows"]))
    for r, ent in rows.items():
        assert len(ent) == 2
        (s1, g1, a), (s2, g2, b) = ent
        assert s1 != s2 and g1 == g2 and {a, b} == {1.0, -1.0}
    # groups by owner set; one middle node per group, glued along consecutive owners
    groups = {}
    for node, ow in owners.items():
        if len(ow) > 1:
            groups.setdefault(tuple(sorted(ow)), []).append(node)
    expected = sum((len(k) - 1) * nd for k in groups)
    assert info["n_corner_rows"] == expected
    corner_nodes = {g // nd for ent in rows.values() for _, g, _ in ent}
    for key, nodes in groups.items():
        nodes.sort()
        assert nodes[len(nodes) // 2] in corner_nodes
    f4.close()


def test_hybrid_option_is_accepted_and_others_rejected(inputs):
    elements, size, rhs, dirichlet, o = inputs
    with pytest.raises(RuntimeError, match="FETI_METHOD"):
        core.FETI4I(elements, size, rhs, dirichlet, subdomains=2, method=2)
    f4 = core.FETI4I(elements, size, rhs, dirichlet, subdomains=2, method=0)
    assert f4.info()["n_corner_rows"] == 0 and f4.info()["method"] == 0
    f4.close()
