"""Pins the oracle's later additions before the GPU tests lean on them.

  * global direct solve with a nodal load  -> numpy dense solve of the same
    assembled K (a random load, all three components)
  * uneven subdomain slabs                 -> the undecomposed mesh is the same
    as the uniform one (identical global solutions), and the TFETI solve of
    the uneven decomposition equals it (<= 1e-8)
  * loads replaced on a factorised oracle  -> the global solve with the same
    nodal load (<= 1e-8), twice in a row
  * blocked, threaded coarse factorisation -> P^2 = P and G P = 0 with a coarse
    space of several blocks (and above the threading threshold), and P against
    numpy brute force on a two-block coarse space
  * SUPER_DIRICHLET block                  -> equals the Dirichlet Schur block
    when K_ii is diagonal (one interior DOF: heat on 2^3-element boxes, the
    reference's own remark that the two coincide when K_rr is not
    diagonalised), the FETI solve with it equals the global solve, and a
    wrong diagonal would break the elasticity check against the sparse
    formula (brute force)
  * KKT of the reference's normalisation   -> ||B u - c|| with ||c|| < 1 is
    normalised by ||c|| + 1
"""
import numpy as np
import pytest

import oracle_lib as O
from espreso_amd import core
from espreso_amd import problems as PB

UNEVEN = [
    dict(PB.scaled(PB.HEAT_SIMPLE, clusters=(1, 1, 1), domains=(3, 2, 2), elements=(4, 4, 4),
                   slab_elements=((2, 5, 3), (4, 2), (3, 3)), conductivity_seed=0), name="heat_uneven_small"),
    dict(PB.scaled(PB.GRAVITY, clusters=(1, 2, 1), domains=(2, 1, 2), elements=(4, 4, 4),
                   slab_elements=((3, 2), (2, 4), (4, 2)), length=(1.0, 1.0, 1.0)), name="elast_uneven_small"),
]


def _dense_global(spec):
    o = O.Oracle(spec, preconditioner=0)
    nd = o.ndof
    n = o.n_nodes * nd
    K = np.zeros((n, n))
    for s in range(o.n_sub):
        g = np.repeat(o.l2g(s) * nd, nd) + np.tile(np.arange(nd), len(o.l2g(s)))
        K[np.ix_(g, g)] += o.dense(s, 0)
    return o, K


def test_global_solve_load_vs_numpy_dense():
    spec = PB.scaled(PB.GRAVITY, clusters=(1, 1, 1), domains=(2, 1, 2), elements=(2, 3, 2))
    o, K = _dense_global(spec)
    n = K.shape[0]
    nn = [5, 4, 5]
    fixed = np.zeros(n, bool)
    fixed[:3 * nn[0] * nn[1]] = True  # z = 0 face (node-major, 3 components)
    free = ~fixed
    g = np.random.default_rng(17).uniform(-1, 1, size=(o.n_nodes, 3)) * 1e4
    u = np.zeros(n)
    u[free] = np.linalg.solve(K[np.ix_(free, free)], g.ravel()[free])
    got = O.global_solve_load(spec, g).ravel()
    assert np.linalg.norm(got - u) / np.linalg.norm(u) < 1e-12
    # a dropped component or a transposed load would fail: the load is not symmetric in x, y, z
    assert np.linalg.norm(O.global_solve_load(spec, g[:, ::-1]).ravel() - u) > 1e-3 * np.linalg.norm(u)


@pytest.mark.parametrize("spec", UNEVEN, ids=lambda s: s["name"])
def test_uneven_oracle(spec):
    ug = O.global_solve(PB.uniform_equivalent(spec))
    assert np.allclose(O.global_solve(spec), ug, rtol=0, atol=1e-12 * np.abs(ug).max())
    o = O.Oracle(spec, preconditioner=3, precision=1e-11)
    u, _, st = o.solve()
    assert np.linalg.norm(u - ug) / np.linalg.norm(ug) < 1e-8
    # the same decomposition as the product's builder: same multiplier rows
    assert o.n_lambda == core.grid_export(spec, 0)["c"].size
    assert st["kkt_primal"] < 1e-9


def test_set_nodal_load_resolves():
    spec = PB.scaled(PB.GRAVITY, traction_seed=5, traction_scale=1e5)
    o = O.Oracle(spec, preconditioner=1, precision=1e-11)
    rng = np.random.default_rng(3)
    for _ in range(2):
        g = rng.uniform(-1, 1, size=(o.n_nodes, o.ndof)) * 1e5
        o.set_nodal_load(g)
        u, _, _ = o.solve()
        ug = O.global_solve_load(spec, g)
        assert np.linalg.norm(u - ug) / np.linalg.norm(ug) < 1e-8


def test_coarse_blocks_bruteforce():
    # 18 subdomains: 108 coarse rows = two blocks of the coarse factorisation
    spec = PB.scaled(PB.GRAVITY, clusters=(3, 3, 2), domains=(1, 1, 1), elements=(2, 2, 2))
    o = O.Oracle(spec, preconditioner=0)
    assert o.n_coarse == 108
    nl = o.n_lambda
    Gm = np.zeros((o.n_coarse, nl))
    off = 0
    for s in range(o.n_sub):
        lam, dof, sign, _ = o.lambdas(s)
        R = o.dense(s, 2)
        B = np.zeros((nl, R.shape[0]))
        B[lam, dof] = sign
        Gm[off:off + 6] = -(B @ R).T
        off += 6
    x = np.random.default_rng(1).standard_normal(nl)
    P = np.eye(nl) - Gm.T @ np.linalg.solve(Gm @ Gm.T, Gm)
    assert np.allclose(o.project(x), P @ x, rtol=0, atol=1e-10 * np.abs(x).max())


@pytest.mark.slow
def test_coarse_threaded_invariants():
    # 4800 coarse rows: the threaded triangular sweeps (> 4096 rows) and 50 blocks
    spec = PB.scaled(PB.GRAVITY, clusters=(10, 10, 8), domains=(1, 1, 1), elements=(1, 1, 1),
                     length=(10.0, 10.0, 8.0))
    o = O.Oracle(spec, preconditioner=0)
    assert o.n_coarse == 4800
    x = np.random.default_rng(2).standard_normal(o.n_lambda)
    px = o.project(x)
    assert np.abs(o.apply_G(px)).max() < 1e-9 * np.abs(o.apply_G(x)).max()
    assert np.allclose(o.project(px), px, rtol=0, atol=1e-10 * np.abs(px).max())


def test_super_dirichlet_block():
    # heat, 2^3-element subdomains with every outer face prescribed: each box
    # has one interior node, K_ii is 1 x 1 and the simplified block equals the
    # Schur complement
    heat = PB.scaled(PB.HEAT_SIMPLE, clusters=(2, 1, 1), domains=(1, 2, 2), elements=(2, 2, 2),
                     fixed=(1, 1, 1, 1, 1, 1), fixed_value=(100.0, 300.0, 0.0, 0.0, 0.0, 0.0))
    o = O.Oracle(heat, preconditioner=4)
    for s in range(o.n_sub):
        assert o.info(s)["n"] - o.info(s)["m"] == 1
        S = o.dense(s, 3)
        assert np.allclose(o.dense(s, 5), S, rtol=0, atol=1e-12 * np.abs(S).max())
    # elasticity: K_ii is not diagonal; the block is K_bb - K_bi diag(K_ii)^-1 K_ib
    el = PB.scaled(PB.GRAVITY, elements=(3, 2, 2))
    o = O.Oracle(el, preconditioner=4)
    for s in (0, o.n_sub - 1):
        K = o.dense(s, 0)
        b = o.boundary(s)
        i = np.setdiff1d(np.arange(K.shape[0]), b)
        ref = K[np.ix_(b, b)] - K[np.ix_(b, i)] @ (K[np.ix_(i, b)] / np.diag(K)[i][:, None])
        got = o.dense(s, 5)
        assert np.allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
        assert np.abs(got - o.dense(s, 3)).max() > 1e-6 * np.abs(ref).max()  # differs from the full Schur
    # the preconditioner applies the block from the sparse K: equal to the dense blocks
    nl = o.n_lambda
    x = np.random.default_rng(6).standard_normal(nl)
    y = np.zeros(nl)
    for s in range(o.n_sub):
        lam, dof, sign, scale = o.lambdas(s)
        b = o.boundary(s)
        pos = np.searchsorted(b, dof)
        v = np.zeros(len(b))
        np.add.at(v, pos, sign * scale * x[lam])
        w = o.dense(s, 5) @ v
        np.add.at(y, lam, sign * scale * w[pos])
    assert np.allclose(o.apply_prec(x), y, rtol=0, atol=1e-11 * np.abs(y).max())


@pytest.mark.parametrize("spec", [PB.GRAVITY, PB.HEAT_SIMPLE], ids=["gravity", "heat"])
def test_super_dirichlet_solve(spec):
    ug = O.global_solve(spec)
    o = O.Oracle(spec, preconditioner=4, precision=1e-10)
    u, _, st = o.solve()
    assert np.linalg.norm(u - ug) / np.linalg.norm(ug) < 1e-8
    it4 = st["iterations"]
    it1 = O.Oracle(spec, preconditioner=1, precision=1e-10).solve()[2]["iterations"]
    assert it4 <= it1 + 2  # at least as good as lumped on these meshes


def test_kkt_dual_normalisation():
    # heat_simple prescribes 100 and 300: ||c|| >= 1, plain relative residual;
    # gravity prescribes zeros: ||c|| = 0 -> normalised by 1 (itersolver.cpp:330-332)
    for spec in (PB.HEAT_SIMPLE, PB.GRAVITY):
        o = O.Oracle(spec, preconditioner=3, precision=1e-10)
        u, lam, st = o.solve()
        c = o.vec_c()
        Bu = np.zeros(o.n_lambda)
        for s in range(o.n_sub):
            l, dof, sign, _ = o.lambdas(s)
            Bu[l] += sign * o.primal(s)[dof]
        nc = np.linalg.norm(c)
        want = np.linalg.norm(Bu - c) / (nc + 1.0 if nc < 1 else nc)
        assert abs(st["kkt_dual"] - want) <= 1e-6 * want + 1e-300
