"""GPU: the SYMV work-item autotuner.  The first set-up times the candidate
sizes on the assembled operators and keeps the fastest; the choice is one of
the candidates, ESPRESO_AUTOTUNE=0 keeps the default, and the solution does not
depend on it (1e-9 relative between the two runs, 1e-8 to the direct solve)."""
import os

import numpy as np
import pytest

import oracle_lib as O
from espreso_amd import core
from espreso_amd import problems as PB

pytestmark = pytest.mark.gpu

CANDIDATES = {96 * 1024, 192 * 1024, 320 * 1024, 640 * 1024, 1280 * 1024}
SPEC = PB.scaled(PB.GRAVITY, traction_seed=4, traction_scale=1e5)


def solve(tune, **opts):
    if not tune:
        os.environ["ESPRESO_AUTOTUNE"] = "0"
    try:
        g = core.GridFETI(SPEC, preconditioner="dirichlet", precision=1e-11, max_iterations=3000, **opts)
        g.setup()
        g.solve()
        items = g.symv_items()
        u = g.nodal_solution()
        g.setup()  # a second set-up keeps the choice
        assert g.symv_items() == items
        g.close()
        return items, u
    finally:
        os.environ.pop("ESPRESO_AUTOTUNE", None)


@pytest.mark.parametrize("sc", ["fp64", "fp32"])
def test_autotuned_items(sc):
    print("case: autotune", sc, flush=True)
    tuned, ut = solve(True, sc_precision=sc)
    fixed, uf = solve(False, sc_precision=sc)
    print("tuned", tuned, "fixed", fixed, flush=True)
    p = core.SC[sc]
    assert tuned[0] in CANDIDATES
    assert tuned[p] in CANDIDATES
    assert all(tuned[q] == 0 for q in (1, 2) if q != p)
    assert fixed == tuple(320 * 1024 if q in (0, p) else 0 for q in range(3))
    ref = O.global_solve(SPEC)
    assert np.abs(ut - uf).max() <= 1e-9 * np.abs(uf).max()
    assert np.abs(ut - ref).max() <= 1e-8 * np.abs(ref).max()
