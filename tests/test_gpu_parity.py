"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bars: assembled matrices and dual operators (same generalised inverse as the
oracle's kplus=1 mode) within 1e-10 relative (FP64 reassociation only);
primal solutions within 1e-8 relative of the undecomposed direct solve
(FP64, stop at precision 1e-10) and of the golden values.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from espreso_amd import core
from espreso_amd import problems as PB

pytestmark = pytest.mark.gpu

SMALL = PB.scaled(PB.GRAVITY, traction_seed=9, traction_scale=1e5)          # 16 sub of 5^3
MULTI = PB.scaled(PB.GRAVITY, clusters=(1, 1, 1), domains=(2, 2, 2), elements=(8, 8, 8),
                  length=(1.0, 1.0, 1.0), young=2.1e11, traction_seed=3, traction_scale=1e3)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def small_pair():
    say("case: SMALL dirichlet fp64")
    g = core.GridFETI(SMALL, preconditioner="dirichlet", precision=1e-10)
    g.setup()
    o = O.Oracle(SMALL, preconditioner=3, precision=1e-10, kplus=1)
    yield g, o
    g.close()
    o.close()


@pytest.mark.parametrize("spec_name", ["SMALL", "MULTI"])
def test_assembled_blocks(spec_name, small_pair):
    spec = SMALL if spec_name == "SMALL" else MULTI
    say("case: assembled blocks", spec_name)
    if spec_name == "SMALL":
        g, o = small_pair
    else:
        g = core.GridFETI(spec, preconditioner="dirichlet")
        g.setup()
        o = O.Oracle(spec, preconditioner=3, kplus=1)
    for s in range(o.n_sub):
        assert np.array_equal(g.boundary(s), o.boundary(s))
        SD = g.dense(s, 1)
        assert rel(SD, o.dense(s, 3)) < 1e-10, s
        F = g.dense(s, 0)
        Fo = o.dense(s, 4)
        assert rel(F, Fo) < 1e-9, s
    if spec_name != "SMALL":
        g.close()


def test_dual_operators(small_pair):
    g, o = small_pair
    say("case: dual operators")
    rng = np.random.default_rng(1)
    x = rng.standard_normal(o.n_lambda)
    assert g.n_lambda == o.n_lambda
    assert rel(g.apply_F(x), o.apply_F(x)) < 1e-10
    assert rel(g.apply_prec(x), o.apply_prec(x)) < 1e-10
    assert rel(g.project(x), o.project(x)) < 1e-10
    assert rel(g.vec_b(), o.vec_b()) < 1e-10


@pytest.mark.parametrize("prec,oprec", [("lumped", 1), ("weight", 2), ("none", 0)])
def test_other_preconditioners(prec, oprec):
    say("case: preconditioner", prec)
    g = core.GridFETI(SMALL, preconditioner=prec)
    g.setup()
    o = O.Oracle(SMALL, preconditioner=oprec, kplus=1)
    x = np.random.default_rng(2).standard_normal(o.n_lambda)
    assert rel(g.apply_prec(x), o.apply_prec(x)) < 1e-12
    g.close()


def test_solve_small_vs_global(small_pair):
    g, o = small_pair
    say("case: solve SMALL")
    st = g.solve()
    u = g.nodal_solution()
    ug = O.global_solve(SMALL)
    assert rel(u, ug) < 1e-8
    uo, _, sto = o.solve()
    assert abs(st["iterations"] - sto["iterations"]) <= 1
    for s in range(o.n_sub):
        assert rel(g.primal(s, o.info(s)["n"]), o.primal(s)) < 1e-8


def test_pressure_patch_golden():
    say("case: pressure patch")
    g = core.GridFETI(PB.PRESSURE_PATCH, preconditioner="lumped", precision=1e-8)
    g.setup()
    g.solve()
    u = g.nodal_solution()
    nn = (11, 21, 11)
    z1 = np.linalg.norm(u, axis=1).reshape(nn[2], nn[1], nn[0])[-1]
    assert np.allclose(z1, 1.0, rtol=0, atol=1e-7)
    g.close()


def test_gravity_emr_golden():
    say("case: gravity")
    g = core.GridFETI(PB.GRAVITY, preconditioner="dirichlet", precision=1e-8)
    g.setup()
    g.solve()
    mag = np.linalg.norm(g.nodal_solution(), axis=1)
    gold = PB.GRAVITY_EMR
    assert abs(mag.max() - gold["max"]) / gold["max"] < 1e-3
    assert abs(mag.mean() - gold["avg"]) / gold["avg"] < 1e-3
    g.close()


def test_heat_golden():
    say("case: heat")
    g = core.GridFETI(PB.HEAT_SIMPLE, preconditioner="dirichlet", precision=1e-10)
    g.setup()
    g.solve()
    u = g.nodal_solution()[:, 0]
    assert abs(u.min() - 100) < 1e-6 and abs(u.max() - 300) < 1e-6 and abs(u.mean() - 200) < 1e-6
    g.close()


@pytest.mark.parametrize("prec", ["dirichlet", "lumped"])
def test_tfeti_64x5_vs_global(prec):
    say("case: 64 subdomains of 5^3", prec)
    spec = PB.TFETI_64x5
    g = core.GridFETI(spec, preconditioner=prec, precision=1e-10)
    g.setup()
    st = g.solve()
    ug = O.global_solve(spec)
    assert rel(g.nodal_solution(), ug) < 1e-8
    assert st["iterations"] < 200
    g.close()


@pytest.mark.parametrize("sc", ["fp32", "bf16"])
def test_mixed_precision_refinement(sc):
    say("case: mixed precision", sc)
    g = core.GridFETI(MULTI, preconditioner="dirichlet", precision=1e-10, sc_precision=sc)
    g.setup()
    st = g.solve()
    ug = O.global_solve(MULTI)
    err = rel(g.nodal_solution(), ug)
    say("mixed", sc, "err", err, "iters", st["iterations"], "refine", st["refinement_steps"])
    assert err < 1e-8
    # stored F_i is the FP64 F_i rounded to the storage precision
    F64 = g.dense(0, 0)
    Flp = g.dense(0, 2)
    tol = 2 ** -23 if sc == "fp32" else 2 ** -8
    assert np.abs(Flp - F64).max() <= 2 * tol * np.abs(F64).max()
    g.close()


def test_invalid_input_is_reported():
    say("case: invalid input")
    s = core.FETISolver(1, preconditioner="dirichlet")
    with pytest.raises(RuntimeError):
        s.setup()  # no subdomain data
    rp = np.array([0, 1], np.int32)
    with pytest.raises(RuntimeError):
        s.set_subdomain(0, (rp, np.array([0], np.int32), np.array([np.nan])), np.zeros(1), np.ones((1, 1)),
                        [0], [0], [1.0], [1.0])
    s.close()


@pytest.mark.parametrize("sc,sd", [("fp64", "bf16"), ("fp32", "bf16"), ("bf16", "fp32"), ("fp64", "fp32")])
def test_mixed_storage_precisions(sc, sd):
    """F_i and S_D stored in different precisions: each keeps its own strip
    geometry; the solution stays within 1e-8 of the direct solve and the stored
    S_D within its rounding of the oracle's."""
    say("case: F_i", sc, "S_D", sd)
    g = core.GridFETI(MULTI, preconditioner="dirichlet", precision=1e-10, sc_precision=sc, sd_precision=sd)
    g.setup()
    st = g.solve()
    err = rel(g.nodal_solution(), O.global_solve(MULTI))
    items = g.symv_items()
    say("F_i", sc, "S_D", sd, "err", err, "iters", st["iterations"], "items", items)
    assert err < 1e-8
    used = {0, core.SC[sc], core.SC[sd]}
    assert all((items[p] > 0) == (p in used) for p in range(3))
    o = O.Oracle(MULTI, preconditioner=3, kplus=1)
    tol = {"fp64": 1e-12, "fp32": 2 ** -23, "bf16": 2 ** -8}
    for s in (0, o.n_sub - 1):
        Sd = o.dense(s, 3)
        assert np.abs(g.dense(s, 1) - Sd).max() <= 2 * tol[sd] * np.abs(Sd).max()
        F = g.dense(s, 0)  # FP64 F_i; which=2 is F_i as stored
        assert np.abs(g.dense(s, 2) - F).max() <= 2 * tol[sc] * np.abs(F).max()
    o.close()
    g.close()


@pytest.mark.parametrize("sc", ["fp32", "bf16"])
def test_low_precision_dirichlet(sc):
    say("case: S_D stored in", sc)
    g = core.GridFETI(MULTI, preconditioner="dirichlet", precision=1e-10, sc_precision=sc, sd_precision=sc)
    g.setup()
    st = g.solve()
    err = rel(g.nodal_solution(), O.global_solve(MULTI))
    say("low-precision S_D", sc, "err", err, "iters", st["iterations"])
    assert err < 1e-8
    o = O.Oracle(MULTI, preconditioner=3, kplus=1)
    tol = 2 ** -23 if sc == "fp32" else 2 ** -8
    for s in (0, o.n_sub - 1):
        Sd = o.dense(s, 3)
        assert np.abs(g.dense(s, 1) - Sd).max() <= 2 * tol * np.abs(Sd).max()
    o.close()
    g.close()


@pytest.mark.parametrize("method", ["total", "hybrid"])
def test_many_subdomains_paired_substitution(method):
    """512 subdomains of 4^3 elements: enough column tiles that the assembly's
    forward substitution takes two tiles per workgroup (the smaller problems
    above take one).  Checked through size-independent properties: the dual
    residual reaches 1e-10 and the primal solution's residual in the oracle's
    FP64 global assembly is below 1e-7."""
    spec = PB.scaled(PB.HTFETI_512x10, elements=(4, 4, 4))
    say("case: 512 subdomains of 4^3,", method)
    g = core.GridFETI(spec, preconditioner="dirichlet", precision=1e-10, method=method, max_iterations=2000)
    g.setup()
    st = g.solve()
    u = g.nodal_solution()
    res = O.global_residual(O.undecomposed(spec), u)
    say("paired substitution", method, "iterations", st["iterations"], "residual", res)
    assert st["final_norm"] < st["tol"]
    assert res < 1e-7
    g.close()


SINGLE = {
    "elasticity": PB.scaled(PB.GRAVITY, clusters=(1, 1, 1), domains=(1, 1, 1), elements=(4, 4, 4), traction_seed=5,
                            traction_scale=1e5),
    "heat": PB.scaled(PB.HEAT_SIMPLE, clusters=(1, 1, 1), domains=(1, 1, 1), elements=(5, 4, 3)),
    # subdomains of one element: every DOF on the interface, no interior DOFs
    "one_element_subdomains": PB.scaled(PB.GRAVITY, clusters=(1, 1, 1), domains=(3, 2, 2), elements=(1, 1, 1),
                                        traction_seed=7, traction_scale=1e5),
    # eight clusters of one subdomain each: no corner rows anywhere, clusters glued by B1 only
    "clusters_of_one": PB.scaled(PB.GRAVITY, clusters=(2, 2, 2), domains=(1, 1, 1), elements=(3, 3, 3),
                                 traction_seed=6, traction_scale=1e5),
}


@pytest.mark.parametrize("method", ["total", "hybrid"])
@pytest.mark.parametrize("name", list(SINGLE))
def test_single_subdomain(method, name):
    """Extreme decompositions: the multipliers of a lone subdomain are only the
    Dirichlet rows of B1; with HYBRID_FETI a cluster of one subdomain has no
    corner rows (n0 = 0); a one-element subdomain has no interior DOFs."""
    spec = SINGLE[name]
    say("case: single subdomain,", name, method)
    g = core.GridFETI(spec, preconditioner="dirichlet", precision=1e-10, method=method)
    g.setup()
    st = g.solve()
    err = rel(g.nodal_solution(), O.global_solve(spec))
    say("single subdomain", name, method, "iterations", st["iterations"], "err", err)
    assert err < 1e-8
    g.close()


@pytest.mark.parametrize("sc", ["fp64", "fp32"])
def test_large_subdomains(sc):
    """Subdomains of 16^3 elements (about 4600 interface DOFs each): the SYMV's
    LDS window (2 m + R values) exceeds 64 KB.  Checked through the dual
    residual and the primal residual in the oracle's FP64 global assembly."""
    spec = PB.scaled(PB.GRAVITY, clusters=(1, 1, 1), domains=(2, 2, 2), elements=(16, 16, 16), traction_seed=8,
                     traction_scale=1e5)
    say("case: 8 subdomains of 16^3,", sc)
    g = core.GridFETI(spec, preconditioner="dirichlet", precision=1e-10, sc_precision=sc, max_iterations=2000)
    g.setup()
    st = g.solve()
    res = O.global_residual(O.undecomposed(spec), g.nodal_solution())
    say("large subdomains", sc, "iterations", st["iterations"], "residual", res)
    assert st["final_norm"] < st["tol"]
    assert res < 1e-7
    g.close()
