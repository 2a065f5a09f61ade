"""GPU parity at the benchmark sizes, against the oracle's full-size solutions.

The oracle (the CPU restatement of the reference's TFETI solve,
itersolver.cpp:1406-1724) solved every workload below at its real size to a
relative dual residual of 1e-11; tests/golden/make_fullsize.py stored its
nodal solution at 8192 seeded nodes (2048 per case for the multi-case
fixtures), its 2-norm and a seeded checksum over every DOF.  The HIP path
solves the same workload here and each of those four differences must stay
within 1e-8 relative (the primal-error bar of the FP64 path; the mixed
storages reach it through the FP64 iterative refinement):

  * 512 subdomains of 10^3 elements, HYBRID_FETI, Dirichlet, FP64, with B0
    corner rows and with B0 kernel rows
  * 1000 subdomains of 12^3 elements, F_i FP32 and S_D BF16, TOTAL_FETI and
    HYBRID_FETI; and the FP64 TOTAL_FETI path
  * 2000 uneven subdomains (4..14 elements per direction), heat with seeded
    per-subdomain conductivity and elasticity, the workload's own load and 32
    seeded U(-1,1) nodal load cases solved together (block PCG)
"""
import sys

import numpy as np
import pytest

from espreso_amd import core
from espreso_amd import problems as PB

sys.path.insert(0, __file__.rsplit("/", 1)[0] + "/golden")
import fullsize  # noqa: E402

pytestmark = pytest.mark.gpu

BAR = 1e-8


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fixture(workload):
    fx = fullsize.load(workload)
    assert fx is not None, f"tests/golden/fullsize_{workload}.npz is missing (tests/golden/make_fullsize.py writes it)"
    return fx


def check(fx, u, case=0, what=""):
    d = fullsize.compare(fx, u, case)
    say(f"  {what} case {case}: " + ", ".join(f"{k} {v:.2e}" for k, v in d.items()))
    for k, v in d.items():
        assert v <= BAR, f"{what} case {case}: {k} = {v:.3e} > {BAR}"
    return d


def solve(spec, **opts):
    g = core.GridFETI(spec, **opts)
    g.setup()
    st = g.solve()
    say(f"  {spec['name']} {opts}: {st['iterations']} iterations, ||Pr|| {st['final_norm']:.2e} tol {st['tol']:.2e}, "
        f"refinement {st['refinement_steps']}, kkt {st['kkt_primal']:.1e} {st['kkt_dual']:.1e}")
    assert st["converged"] == 1
    return g, st


def test_elasticity_512sub_10e3_htfeti_fp64():
    say("case: elasticity_512sub_10e3 HYBRID_FETI dirichlet fp64, precision 1e-10")
    fx = fixture("elasticity_512sub_10e3")
    g, st = solve(PB.HTFETI_512x10, method="hybrid", preconditioner="dirichlet", precision=1e-10,
                  max_iterations=2000)
    check(fx, g.nodal_solution(), what="512sub hybrid fp64")
    g.close()


def test_elasticity_512sub_10e3_htfeti_kernel_rows():
    say("case: elasticity_512sub_10e3 HYBRID_FETI with B0 kernel rows, dirichlet fp64, precision 1e-10")
    fx = fixture("elasticity_512sub_10e3")
    g, st = solve(dict(PB.HTFETI_512x10, b0="kernels"), method="hybrid", preconditioner="dirichlet",
                  precision=1e-10, max_iterations=2000)
    check(fx, g.nodal_solution(), what="512sub hybrid kernels fp64")
    g.close()


@pytest.mark.parametrize("method,sc,sd", [("total", "fp32", "bf16"), ("hybrid", "fp32", "bf16"),
                                          ("total", "fp64", "fp64")],
                         ids=["tfeti-fp32-bf16", "htfeti-fp32-bf16", "tfeti-fp64"])
def test_elasticity_1000sub_12e3(method, sc, sd):
    say(f"case: elasticity_1000sub_12e3 {method} F_i {sc} S_D {sd}, precision 1e-10")
    fx = fixture("elasticity_1000sub_12e3")
    g, st = solve(PB.MIXED_1000x12, method=method, preconditioner="dirichlet", precision=1e-10,
                  max_iterations=3000, sc_precision=sc, sd_precision=sd)
    check(fx, g.nodal_solution(), what=f"1000sub {method} {sc}/{sd}")
    g.close()


@pytest.mark.parametrize("spec", [PB.HEAT_UNEVEN_2000, PB.ELASTICITY_UNEVEN_2000], ids=["heat", "elasticity"])
def test_uneven_2000sub_32_load_cases(spec):
    say(f"case: {spec['name']} own load + 32 nodal load cases as a block, precision 1e-10")
    fx = fixture(spec["name"])
    assert len(fx["norm2"]) == 33
    g, st = solve(spec, preconditioner="dirichlet", precision=1e-10, max_iterations=3000)
    check(fx, g.nodal_solution(), 0, spec["name"])
    loads = PB.nodal_load_cases(spec, 32)
    u, its, norms = g.resolve_block_nodal(loads)
    say(f"  block: iterations {int(its.min())}..{int(its.max())}")
    worst = 0.0
    for c in range(32):
        worst = max(worst, max(check(fx, u[c], c + 1, spec["name"]).values()))
    say(f"  worst over the 32 cases {worst:.2e}")
    g.close()
