"""Element-level inputs for the FETI4I tests, taken from the oracle: with one
element per oracle subdomain, each subdomain stiffness matrix is exactly one
HEXA8 element matrix and its load vector the element load."""
import numpy as np

import oracle_lib as O


def element_problem(spec_elems):
    """spec with elements (1,1,1): returns (elements, size, rhs, dirichlet, oracle)."""
    o = O.Oracle(spec_elems, preconditioner=0)
    nd = o.ndof
    size = o.n_nodes * nd
    elements = []
    rhs = np.zeros(size)
    for s in range(o.n_sub):
        nodes = o.l2g(s)
        dofs = (np.repeat(nodes * nd, nd) + np.tile(np.arange(nd), len(nodes))).astype(np.int32)
        K = o.dense(s, 0)
        elements.append((nodes, dofs, K))
        rhs[dofs] += o.rhs(s)
    # Dirichlet DOFs: every multiplier row with a single entry
    lamcount = np.zeros(o.n_lambda, int)
    owner = {}
    for s in range(o.n_sub):
        lam, dof, sign, scale = o.lambdas(s)
        nodes = o.l2g(s)
        for l, d in zip(lam, dof):
            lamcount[l] += 1
            owner[l] = nodes[d // nd] * nd + d % nd
    c = o.vec_c()
    dirichlet = sorted({(owner[l], c[l]) for l in range(o.n_lambda) if lamcount[l] == 1})
    return elements, size, rhs, dirichlet, o
