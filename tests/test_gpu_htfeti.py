"""GPU parity of the hybrid FETI path (options.method = 1) through the C ABI.

B0 of both types: corner rows and kernel rows (the reference's default).
Against the numpy HTFETI restatement (oracle/htfeti.py): the projected dual
operator P F~ P and the initial projected residual, which do not depend on
the choice of generalised inverse (the GPU uses the boundary-kernel F_i, the
restatement the fix-point K_reg^-1), within 1e-9 relative; PCG iteration
counts within one at the same stopping threshold.  Against the undecomposed
direct solve and the golden HYBRID_FETI .emr values: primal within 1e-8
relative.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from espreso_amd import core
from espreso_amd import problems as PB

pytestmark = pytest.mark.gpu

CUBE = PB.scaled(PB.GRAVITY, clusters=(2, 2, 1), domains=(2, 2, 2), elements=(3, 3, 3), traction_seed=5,
                 traction_scale=1e4, length=(1.0, 1.0, 1.0), young=2.1e11)
SPECS = {"pressure": PB.PRESSURE_PATCH, "heat": PB.HEAT_SIMPLE, "cube": CUBE}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module", params=[(n, b0) for n in sorted(SPECS) for b0 in ("corners", "kernels")],
                ids=lambda p: f"{p[0]}-{p[1]}")
def pair(request):
    name, b0 = request.param
    spec = dict(SPECS[name], b0=b0)
    say("case: hybrid", name, "B0", b0)
    g = core.GridFETI(spec, method="hybrid", preconditioner="dirichlet", precision=1e-10)
    g.setup()
    H = O.hybrid(spec, dirichlet=True, b0=b0)
    yield name, spec, g, H
    g.close()


def test_coarse_dimension(pair):
    name, spec, g, H = pair
    assert g.sizes()[2] == H.ncoarse == len(H.subs) * (6 if H.ndof == 3 else 1)


def test_projected_operators(pair):
    name, spec, g, H = pair
    say("case: hybrid projected operators", name)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(H.nlam)
    assert rel(g.project(x), H.project(x)) < 1e-10
    PFP = g.project(g.apply_F(g.project(x)))
    assert rel(PFP, H.project(H.apply_F(H.project(x)))) < 1e-9
    # d = B K~+ f - c depends on the generalised inverse, the initial projected
    # residual P (d - F~ x0) with x0 = G^T (G G^T)^-1 e does not
    x0 = H.project(H.e, 1)
    r0 = g.project(g.vec_b() - g.apply_F(x0))
    assert rel(r0, H.project(H.vec_d() - H.apply_F(x0))) < 1e-9
    assert rel(g.apply_prec(x), H.apply_prec(x)) < 1e-10


def test_solve_vs_global(pair):
    name, spec, g, H = pair
    say("case: hybrid solve", name)
    st = g.solve()
    u = g.nodal_solution()
    ug = O.global_solve(spec)
    assert rel(u, ug) < 1e-8
    # same stopping threshold: the reference's rule takes min(||P r0||, ||d||)
    # and ||d|| depends on the generalised inverse
    _, _, it = H.solve(preconditioner=1, tol=st["tol"])
    say("hybrid iterations gpu", st["iterations"], "oracle", it)
    assert abs(st["iterations"] - it) <= 1
    if name == "pressure":
        nn = [spec["clusters"][d] * spec["domains"][d] * spec["elements"][d] + 1 for d in range(3)]
        z1 = np.linalg.norm(u, axis=1).reshape(nn[2], nn[1], nn[0])[-1]
        # golden HEXA8.HYBRID_FETI.emr: Z1 displacement min = max = 1.000000e+00
        assert np.allclose(z1, 1.0, rtol=0, atol=1e-7)
    if name == "heat":
        T = u[:, 0]
        # golden HEXA8.HYBRID_FETI.emr: max(TOP) 100, min(BOTTOM) 300, average 200
        assert abs(T.min() - 100) < 1e-6 and abs(T.max() - 300) < 1e-6 and abs(T.mean() - 200) < 1e-6


@pytest.mark.parametrize("sc,sd", [("fp64", "fp64"), ("fp32", "fp32"), ("bf16", "bf16"), ("fp32", "bf16")])
def test_clusters_of_eight_vs_global(sc, sd):
    spec = PB.scaled(PB.HTFETI_512x10, clusters=(2, 2, 2), elements=(5, 5, 5))
    say("case: hybrid 8 clusters of 2x2x2 subdomains of 5^3", sc, sd)
    g = core.GridFETI(spec, method="hybrid", preconditioner="dirichlet", precision=1e-10, sc_precision=sc,
                      sd_precision=sd)
    g.setup()
    st = g.solve()
    ug = O.global_solve(spec)
    err = rel(g.nodal_solution(), ug)
    say("hybrid 8x8", sc, sd, "err", err, "iterations", st["iterations"], "coarse", st["n_coarse"])
    assert err < 1e-8
    assert st["n_coarse"] == 8 * 6
    # FP64 takes 50 iterations; a corner correction that does not cancel the
    # corner directions (BF16-rounded on both sides) took 1602
    assert st["iterations"] < 300
    g.close()


@pytest.mark.parametrize("sc,sd", [("fp64", "fp64"), ("fp32", "bf16")])
def test_clusters_of_eight_kernel_rows(sc, sd):
    """B0 KERNELS on 8 clusters of 2x2x2 subdomains: 12 face pairs x 6 rows per
    cluster, each row a face's worth of weighted DOFs."""
    spec = PB.scaled(PB.HTFETI_512x10, clusters=(2, 2, 2), elements=(5, 5, 5), b0="kernels")
    say("case: hybrid 8 clusters, B0 kernels", sc, sd)
    g = core.GridFETI(spec, method="hybrid", preconditioner="dirichlet", precision=1e-10, sc_precision=sc,
                      sd_precision=sd)
    g.setup()
    st = g.solve()
    err = rel(g.nodal_solution(), O.global_solve(spec))
    say("hybrid kernels", sc, sd, "err", err, "iterations", st["iterations"], "coarse", st["n_coarse"])
    assert err < 1e-8
    assert st["n_coarse"] == 8 * 6
    assert st["kkt_primal"] < 1e-7
    g.close()


@pytest.mark.parametrize("sc", ["fp64", "fp32"])
def test_hybrid_operator_symmetric(sc):
    """The hybrid F as applied in the iterations: with FP32 storage the
    cluster correction takes F_i B0^T in FP32, i.e. the columns of the stored
    F_i, so the operator stays symmetric up to FP32 accumulation:
    |x'Fy - y'Fx| relative 1e-5 (FP64: 1e-12), on projected x, y."""
    spec = PB.scaled(PB.HTFETI_512x10, clusters=(2, 2, 2), elements=(5, 5, 5))
    say("case: hybrid operator symmetry", sc)
    g = core.GridFETI(spec, method="hybrid", preconditioner="dirichlet", sc_precision=sc)
    g.setup()
    rng = np.random.default_rng(7)
    x = g.project(rng.standard_normal(g.n_lambda))
    y = g.project(rng.standard_normal(g.n_lambda))
    Fx, Fy = g.apply_F(x), g.apply_F(y)
    asym = abs(y @ Fx - x @ Fy) / (np.linalg.norm(x) * np.linalg.norm(Fy))
    say("hybrid", sc, "asymmetry", asym)
    assert asym < (1e-12 if sc == "fp64" else 1e-5)
    g.close()


@pytest.mark.parametrize("prec", ["none", "lumped", "weight", "dirichlet"])
def test_clusters_of_eight_preconditioners(prec):
    """HTFETI with every preconditioner reaches the direct solve."""
    spec = PB.scaled(PB.HTFETI_512x10, clusters=(2, 2, 2), elements=(4, 4, 4))
    say("case: hybrid 8 clusters, preconditioner", prec)
    g = core.GridFETI(spec, method="hybrid", preconditioner=prec, precision=1e-10, max_iterations=3000)
    g.setup()
    st = g.solve()
    err = rel(g.nodal_solution(), O.global_solve(spec))
    say("hybrid preconditioner", prec, "iterations", st["iterations"], "err", err)
    assert err < 1e-8
    g.close()


def test_iteration_limit():
    """An unreachable tolerance stops at max_iterations and reports it (the
    stop test runs one iteration ahead; the last iterate is the limit's)."""
    spec = PB.scaled(PB.HTFETI_512x10, clusters=(2, 2, 2), elements=(4, 4, 4))
    say("case: iteration limit")
    g = core.GridFETI(spec, method="hybrid", preconditioner="dirichlet", precision=1e-30, max_iterations=7)
    g.setup()
    with pytest.warns(RuntimeWarning, match="stop test"):
        st = g.solve()  # ESPRESO_NOT_CONVERGED: the iterate and the stats come back
    say("iteration limit", st["iterations"], st["final_norm"], st["tol"])
    assert st["iterations"] == 7
    assert st["final_norm"] > st["tol"]
    assert st["converged"] == 0
    assert np.isfinite(g.nodal_solution()).all()
    # the same 7 iterations without the limit being the cause: a looser run that needs more
    g2 = core.GridFETI(spec, method="hybrid", preconditioner="dirichlet", precision=1e-10, max_iterations=3000)
    g2.setup()
    st2 = g2.solve()
    assert st2["iterations"] > 7
    g.close()
    g2.close()
