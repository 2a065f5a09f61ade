"""GPU: structured-grid K_i assembled on the device from the unit element
matrix and the element materials.

Bars: the device values equal the host builder's bit for bit (same elements,
same order, separately rounded products and sums); with identical inputs the
solutions of the two runs differ only by the order of the SYMV's LDS atomics,
i.e. within the solve tolerance: 1e-9 relative to each other and 1e-8 to the
undecomposed direct solve, for TFETI and HYBRID_FETI, Dirichlet and lumped
preconditioners (the lumped K_bb is gathered from the device values).
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from espreso_amd import core
from espreso_amd import problems as PB
from test_gpu_uneven import with_material

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


CASES = {
    "gravity_hybrid": PB.scaled(PB.GRAVITY, traction_seed=4, traction_scale=1e5, method="hybrid"),
    "random_E": PB.scaled(PB.ADJOINT_64x10, domains=(2, 2, 2), elements=(4, 3, 5)),
    "heat_uneven": with_material(dict(PB.HEAT_UNEVEN_2000, domains=(3, 2, 2),
                                      slab_elements=PB.uneven_slabs((3, 2, 2), 5, 2, 6), name="heat_uneven_12")),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("prec", ["dirichlet", "lumped"])
def test_device_assembly(name, prec):
    spec = CASES[name]
    method = spec.get("method", "total")
    say("case: device assembly", name, prec)
    opts = dict(preconditioner=prec, precision=1e-11, max_iterations=3000, method=method)
    gd = core.GridFETI(dict(spec, device_assembly=True), **opts)
    gd.setup()
    for i in range(gd.n_sub):
        host = core.grid_export(spec, i)
        dev = gd.values(i, len(host["vals"]))
        assert np.array_equal(dev, host["vals"]), (i, np.abs(dev - host["vals"]).max())
    std = gd.solve()
    ud = gd.nodal_solution()
    gd.close()
    gh = core.GridFETI(spec, **opts)
    gh.setup()
    sth = gh.solve()
    uh = gh.nodal_solution()
    gh.close()
    ug = O.global_solve(PB.uniform_equivalent(spec) if "slab_elements" in spec else spec)
    say(f"iterations device {std['iterations']} host {sth['iterations']}, device vs host {rel(ud, uh):.2e}, "
        f"vs direct {rel(ud, ug):.2e}")
    assert rel(ud, uh) < 1e-9
    assert rel(ud, ug) < 1e-8


def test_device_assembly_hybrid_mixed_storage():
    """K_i assembled on the device, HYBRID_FETI, F_i in FP32 and S_D in BF16."""
    spec = dict(PB.scaled(PB.HTFETI_512x10, clusters=(2, 2, 1), elements=(4, 4, 4)), device_assembly=True)
    print("case: device assembly, hybrid, fp32/bf16", flush=True)
    g = core.GridFETI(spec, method="hybrid", preconditioner="dirichlet", precision=1e-10, sc_precision="fp32",
                      sd_precision="bf16")
    g.setup()
    g.solve()
    ref = dict(spec)
    ref.pop("device_assembly")
    u = g.nodal_solution()
    ug = O.global_solve(ref)
    err = np.abs(u - ug).max() / np.abs(ug).max()
    print("device assembly mixed err", err, flush=True)
    assert err < 1e-8
    g.close()
