"""Structured grid, host part of the device assembly: the builder run for the
device (structure, loads, kernels, B1 only) gives bit for bit the B1 rows and
K-diagonal scaling weights, loads, kernels and CSR structure of the full host
build -- the scaling diagonal is summed over the same elements in the same
order as the assembled K."""
import numpy as np
import pytest

from espreso_amd import core
from espreso_amd import problems as PB

CASES = {
    "gravity": PB.scaled(PB.GRAVITY, traction_seed=4, traction_scale=1e5),
    "random_E": PB.scaled(PB.ADJOINT_64x10, domains=(2, 2, 2), elements=(3, 4, 3)),
    "heat_uneven": dict(PB.HEAT_UNEVEN_2000, domains=(3, 2, 2), slab_elements=PB.uneven_slabs((3, 2, 2), 5, 2, 5)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_structure_only_build_matches_full_build(name):
    spec = CASES[name]
    n_sub = core.grid_export(spec, 0)["n_sub"]
    for i in range(n_sub):
        full = core.grid_export(spec, i)
        dev = core.grid_export(dict(spec, device_assembly=True), i)
        assert not np.any(dev["vals"])
        assert np.any(full["vals"])
        for k in ("rowptr", "colidx", "f", "R", "lam", "dof", "sign", "scale", "c"):
            assert np.array_equal(np.asarray(full[k]), np.asarray(dev[k])), (name, i, k)
