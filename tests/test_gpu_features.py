"""GPU checks of the solver-layer features around the dual loop, through the C ABI.

  * SUPER_DIRICHLET preconditioner: M x against the oracle's dense
    K_bb - K_bi diag(K_ii)^-1 K_ib (1e-10), and TFETI/HTFETI solves with it
    against the undecomposed direct solve (1e-8)
  * KKT report: the device's kkt_primal / kkt_dual equal the same residuals
    computed on the host from the returned u, lambda, K_i and c (1e-6
    relative), and are small for a converged solve (hybrid included)
  * opt-in output truncation equals trunc(u D)/D of the untruncated solution
  * K updates: update_values / update_material then setup + solve equal a
    fresh instance on the new values (1e-10)
  * set_dual after setup: a different size is refused, new Dirichlet values
    of the same rows re-solve to the direct solution of the changed problem
  * status: an unconverged solve returns ESPRESO_NOT_CONVERGED with the stats
  * the SYMV's LDS guard refuses an interface that does not fit (no launch)
  * K values handed over as lower triangles (sorted rows) solve like the full
    values (unsorted rows), and a structurally asymmetric K is refused
  * low-precision block re-solves leave the instance's own solution in place,
    and the mixed-precision stats report the outer ||b||
"""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle_lib as O
from espreso_amd import core
from espreso_amd import problems as PB

pytestmark = pytest.mark.gpu

SMALL = PB.scaled(PB.GRAVITY, traction_seed=9, traction_scale=1e5)          # 16 subdomains of 5^3
HYB = PB.scaled(PB.HTFETI_512x10, clusters=(2, 2, 2), elements=(4, 4, 4))    # 64 subdomains, 8 clusters


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_super_dirichlet_operator():
    say("case: super dirichlet operator")
    g = core.GridFETI(SMALL, preconditioner="super_dirichlet")
    g.setup()
    o = O.Oracle(SMALL, preconditioner=4, kplus=1)
    x = np.random.default_rng(4).standard_normal(o.n_lambda)
    assert rel(g.apply_prec(x), o.apply_prec(x)) < 1e-10
    g.close()


@pytest.mark.parametrize("method,spec", [("total", SMALL), ("hybrid", HYB)], ids=["tfeti", "htfeti"])
def test_super_dirichlet_solve(method, spec):
    say("case: super dirichlet solve", method)
    g = core.GridFETI(spec, preconditioner="super_dirichlet", method=method, precision=1e-10, max_iterations=3000)
    g.setup()
    st = g.solve()
    err = rel(g.nodal_solution(), O.global_solve(spec))
    say("super dirichlet", method, st["iterations"], "it, err", err)
    assert err < 1e-8 and st["converged"] == 1
    g.close()


def _host_kkt(g, spec):
    """kkt_primal / kkt_dual of the reference's formulas (TFETI, B1 only) from
    the returned per-subdomain u, lambda, K_i and c."""
    lam = g.dual()
    num = den = 0.0
    Bu = np.zeros(g.n_lambda)
    c = None
    for s in range(g.n_sub):
        e = core.grid_export(spec, s)
        n = e["n"]
        K = np.zeros((n, n))
        for r in range(n):
            K[r, e["colidx"][e["rowptr"][r]:e["rowptr"][r + 1]]] = e["vals"][e["rowptr"][r]:e["rowptr"][r + 1]]
        u = g.primal(s, n)
        btl = np.zeros(n)
        np.add.at(btl, e["dof"], e["sign"] * lam[e["lam"]])
        t = e["f"] - btl
        num += np.sum((K @ u - t) ** 2)
        den += np.sum(t ** 2)
        np.add.at(Bu, e["lam"], e["sign"] * u[e["dof"]])
        c = e["c"]
    nc = np.linalg.norm(c)
    return np.sqrt(num / den), np.linalg.norm(Bu - c) / (nc + 1.0 if nc < 1 else nc)


def test_kkt_report():
    say("case: kkt tfeti")
    g = core.GridFETI(SMALL, preconditioner="dirichlet", precision=1e-10)
    g.setup()
    st = g.solve()
    kp, kd = _host_kkt(g, SMALL)
    say("kkt", st["kkt_primal"], kp, st["kkt_dual"], kd)
    assert abs(st["kkt_primal"] - kp) <= 1e-6 * kp + 1e-15
    assert abs(st["kkt_dual"] - kd) <= 1e-6 * kd + 1e-15
    assert st["kkt_primal"] < 1e-7 and st["kkt_dual"] < 1e-7
    g.close()
    say("case: kkt htfeti")
    h = core.GridFETI(HYB, method="hybrid", preconditioner="dirichlet", precision=1e-10, max_iterations=3000)
    h.setup()
    st = h.solve()
    say("kkt hybrid", st["kkt_primal"], st["kkt_dual"])
    # the corner forces B0^T lambda0 enter the residual: without them it would be O(1)
    assert st["kkt_primal"] < 1e-7 and st["kkt_dual"] < 1e-7
    h.close()


def test_truncation_opt_in():
    say("case: truncation")
    a = core.GridFETI(SMALL, preconditioner="dirichlet", precision=1e-6)
    a.setup()
    a.solve()
    b = core.GridFETI(SMALL, preconditioner="dirichlet", precision=1e-6, truncate_output=True)
    b.setup()
    b.solve()
    for s in (0, a.n_sub - 1):
        n = core.grid_export(SMALL, s)["n"]
        ua = a.primal(s, n)
        ub = b.primal(s, n)
        umax = max(np.abs(a.primal(q, core.grid_export(SMALL, q)["n"])).max() for q in range(a.n_sub))
        D = 1.0 / 1e-6 / max(1.0, 10.0 ** np.ceil(np.log10(umax)))
        want = np.trunc(D * ua) / D
        assert np.abs(ub - want).max() <= 2.0 / D  # one truncation step (solves may differ in the last digits)
        assert np.abs(ua - want).max() > 0        # the default output is not truncated
    a.close()
    b.close()


def test_update_values_equals_fresh():
    say("case: update values")
    E1 = 2.1e11 * (0.1 + 0.9 * np.random.default_rng(1).random(core.n_elements(SMALL)))
    E2 = 2.1e11 * (0.1 + 0.9 * np.random.default_rng(2).random(core.n_elements(SMALL)))
    s1, s2 = dict(SMALL, material_elem=E1), dict(SMALL, material_elem=E2)
    g = core.GridFETI(s1, preconditioner="lumped", precision=1e-11)
    g.setup()
    g.solve()
    for s in range(g.n_sub):
        g.update_values(s, core.grid_export(s2, s)["vals"])
    with pytest.raises(RuntimeError, match="setup"):
        g.solve()  # the operators are stale until setup
    g.setup()
    st = g.solve()
    u_upd = g.nodal_solution()
    f = core.GridFETI(s2, preconditioner="lumped", precision=1e-11)
    f.setup()
    stf = f.solve()
    say("update_values", st["iterations"], stf["iterations"], rel(u_upd, f.nodal_solution()))
    assert rel(u_upd, f.nodal_solution()) < 1e-10
    assert rel(u_upd, O.global_solve(s2)) < 1e-8
    # the grid path: new element moduli recomputed on the device
    g.update_material(E1)
    g.setup()
    g.solve()
    ref = core.GridFETI(s1, preconditioner="lumped", precision=1e-11)
    ref.setup()
    ref.solve()
    assert rel(g.nodal_solution(), ref.nodal_solution()) < 1e-10
    for h in (g, f, ref):
        h.close()


def test_set_dual_after_setup():
    say("case: set_dual after setup")
    g = core.GridFETI(PB.HEAT_SIMPLE, preconditioner="dirichlet", precision=1e-11)
    g.setup()
    g.solve()
    with pytest.raises(RuntimeError, match="n_lambda differs"):
        g.set_dual(np.zeros(g.n_lambda + 1))
    other = dict(PB.HEAT_SIMPLE, fixed_value=(50.0, 250.0, 0, 0, 0, 0))
    g.set_dual(core.grid_export(other, 0)["c"])
    g.resolve()
    u = g.nodal_solution()
    assert rel(u, O.global_solve(other)) < 1e-8
    g.close()


def test_not_converged_status():
    say("case: not converged status")
    g = core.GridFETI(SMALL, preconditioner="dirichlet", precision=1e-30, max_iterations=5)
    g.setup()
    st = core.Stats()
    rc = core.lib().espreso_feti_solve(g.h, C.byref(st))
    assert rc == core.NOT_CONVERGED
    assert st.converged == 0 and st.iterations == 5 and st.final_norm > st.tol
    g.close()


def test_lds_guard_refuses_wide_interface():
    """A chain of n DOFs with every DOF on a Dirichlet row: the interface is the
    whole subdomain (n > the LDS the SYMV workgroup can hold); set-up must
    fail with an argument error before any kernel runs."""
    say("case: LDS guard")
    n = 11000
    rows, cols, vals = [], [], []
    for i in range(n):
        for j, v in ((i - 1, -1.0), (i, 2.0 if 0 < i < n - 1 else 1.0), (i + 1, -1.0)):
            if 0 <= j < n:
                rows.append(i)
                cols.append(j)
                vals.append(v)
    rp = np.searchsorted(np.array(rows), np.arange(n + 1)).astype(np.int32)
    s = core.FETISolver(1, preconditioner="lumped")
    R = np.full((n, 1), 1.0 / np.sqrt(n))
    lam = np.arange(n, dtype=np.int32)
    s.set_subdomain(0, (rp, np.array(cols, np.int32), np.array(vals)), np.zeros(n), R, lam, lam,
                    np.ones(n), np.ones(n))
    s.set_dual(np.zeros(n))
    with pytest.raises(RuntimeError, match="LDS"):
        s.setup()
    s.close()


def _chain(n, drop=None, reverse=False):
    """CSR of the chain Laplacian (2 on the diagonal, -1 beside it), 1 at
    both ends; drop = (i, j) leaves that entry out, reverse orders every row's
    columns downward."""
    rp, cols, vals = [0], [], []
    for i in range(n):
        row = [(j, v) for j, v in ((i - 1, -1.0), (i, 2.0 if 0 < i < n - 1 else 1.0), (i + 1, -1.0))
               if 0 <= j < n and (i, j) != drop]
        if reverse:
            row = row[::-1]
        cols += [j for j, _ in row]
        vals += [v for _, v in row]
        rp.append(len(cols))
    return np.array(rp, np.int32), np.array(cols, np.int32), np.array(vals)


def _chain_solver(K, n, c):
    s = core.FETISolver(1, preconditioner="dirichlet", precision=1e-12)
    R = np.full((n, 1), 1.0 / np.sqrt(n))
    dof = np.arange(0, n, 2, dtype=np.int32)  # every other DOF prescribed
    lam = np.arange(len(dof), dtype=np.int32)
    f = np.sin(np.arange(n, dtype=np.float64))
    s.set_subdomain(0, K, f, R, lam, dof, np.ones(len(dof)), np.ones(len(dof)))
    s.set_dual(c[:len(dof)])
    return s


def test_lower_triangle_values_and_asymmetric_structure():
    """Sorted rows: the K values cross as the lower triangle and the device
    mirrors the rest; reversed rows: the full values cross.  Both solve the
    same chain problem to the same primal (1e-12 relative).  Dropping one
    lower entry (its mirror, an upper entry of another row, stays) makes the
    structure asymmetric: set-up refuses it."""
    say("case: lower-triangle K values, asymmetric structure")
    n = 60
    c = np.cos(np.arange(n, dtype=np.float64))
    us = []
    for reverse in (False, True):
        s = _chain_solver(_chain(n, reverse=reverse), n, c)
        s.setup()
        st = s.solve()
        assert st["converged"] == 1
        us.append(s.primal(0, n))
        s.close()
    err = np.abs(us[0] - us[1]).max() / np.abs(us[1]).max()
    say(f"  sorted vs reversed rows: {err:.2e}")
    assert err <= 1e-12
    s = _chain_solver(_chain(n, drop=(11, 10)), n, c)
    with pytest.raises(RuntimeError, match="structurally symmetric"):
        s.setup()
    s.close()


def test_low_precision_block_keeps_solution():
    say("case: low-precision block keeps the instance solution")
    g = core.GridFETI(SMALL, preconditioner="dirichlet", precision=1e-10, sc_precision="fp32")
    g.setup()
    st = g.solve()
    u0 = g.nodal_solution()
    lam0 = g.dual()
    loads = PB.nodal_load_cases(SMALL, 2)
    g.resolve_block_nodal(loads)
    assert np.array_equal(g.nodal_solution(), u0)
    assert np.array_equal(g.dual(), lam0)
    # the mixed-precision stats give the outer ||b||, as the FP64 path does
    f = core.GridFETI(SMALL, preconditioner="dirichlet", precision=1e-10)
    f.setup()
    st64 = f.solve()
    assert abs(st["norm_b"] - st64["norm_b"]) <= 1e-10 * st64["norm_b"]
    g.close()
    f.close()
