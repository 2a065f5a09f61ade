"""CPU checks of the product library: it loads, exports every entry point its
headers declare, and its host-side problem builder reproduces the oracle's
decomposed problem (no device calls)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_lib as O
from espreso_amd import core
from espreso_amd import problems as PB

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", text, flags=re.M)))


@pytest.mark.parametrize("header", ["espreso_amd.h", "feti4i.h"])
def test_library_exports_header(header):
    if not os.path.exists(os.path.join(ROOT, "include", header)):
        pytest.skip(f"{header} not present")
    L = C.CDLL(core.LIB_PATH)
    names = _declared(header)
    assert names, "no declarations parsed"
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    assert set(_declared("espreso_amd.h")) == set(core.SIGNATURES)


def _dense(g):
    n = g["n"]
    K = np.zeros((n, n))
    for r in range(n):
        sl = slice(g["rowptr"][r], g["rowptr"][r + 1])
        K[r, g["colidx"][sl]] = g["vals"][sl]
    return K


@pytest.mark.parametrize("spec,scaling", [
    (PB.scaled(PB.GRAVITY, traction_seed=5, traction_scale=1e5), 1),
    (PB.PRESSURE_PATCH, 0),
    (PB.HEAT_SIMPLE, 1),
    (PB.scaled(PB.HEAT_SIMPLE, conductivity_seed=11), 1),
    (PB.scaled(PB.GRAVITY, elements=(3, 2, 2), young_seed=3), 1),
])
def test_grid_builder_matches_oracle(spec, scaling):
    o = O.Oracle(spec, preconditioner=0, scaling=scaling)
    for s in range(o.n_sub):
        g = core.grid_export(spec, s, scaling=scaling)
        assert g["n_sub"] == o.n_sub
        Ko = o.dense(s, 0)
        K = _dense(g)
        assert np.abs(K - Ko).max() <= 1e-13 * np.abs(Ko).max()
        assert np.allclose(K, K.T, rtol=0, atol=1e-14 * np.abs(K).max())
        fo = o.rhs(s)
        assert np.abs(g["f"] - fo).max() <= 1e-13 * max(np.abs(fo).max(), 1e-300)
        lam, dof, sign, scale = o.lambdas(s)
        assert np.array_equal(g["lam"], lam) and np.array_equal(g["dof"], dof)
        assert np.array_equal(g["sign"], sign)
        assert np.abs(g["scale"] - scale).max() < 1e-14
        Ro = o.dense(s, 2)
        assert np.abs(g["R"] @ g["R"].T - Ro @ Ro.T).max() < 1e-12
        assert np.linalg.norm(K @ g["R"]) <= 1e-9 * np.linalg.norm(K)
    assert np.array_equal(g["c"], o.vec_c())


def test_grid_export_rejects_bad_input():
    spec = PB.scaled(PB.GRAVITY)
    with pytest.raises(RuntimeError):
        core.grid_export(spec, 10_000)
    bad = PB.scaled(PB.GRAVITY, elements=(0, 5, 5))
    with pytest.raises(RuntimeError):
        core.grid_export(bad, 0)
