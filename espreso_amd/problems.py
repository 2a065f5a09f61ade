"""Named structured-box problems (the reference's GRID generator inputs).

Each problem is a plain dict that maps one-to-one onto ``espreso_grid_problem``
in include/espreso_amd.h: a box [0,Lx]x[0,Ly]x[0,Lz] split into
clusters x domains x elements (reference: src/input/meshgenerator/
composition/gridgenerator.cpp:44-69), HEXA8 elements, material, loads and
Dirichlet faces.  Faces are ordered x0, x1, y0, y1, z0, z1; ``fixed`` holds a
bitmask of the fixed components per face.
"""

E_STEEL = 2.1e11

# --- workloads of the benchmark suite -------------------------------------

# 64 subdomains of 5^3 HEXA8 elasticity elements, gravity -z, z=0 clamped.
TFETI_64x5 = dict(
    name="elasticity_64sub_5e3",
    physics=0, clusters=(1, 1, 1), domains=(4, 4, 4), elements=(5, 5, 5),
    length=(1.0, 1.0, 1.0), young=E_STEEL, poisson=0.3, density=7850.0,
    accel=(0.0, 0.0, -9.81), fixed=(0, 0, 0, 0, 7, 0),
)

# 512 subdomains of 10^3 elements (3993 DOFs each), clusters of 2x2x2,
# gravity plus a seeded U(-1,1) nodal traction on the x=1 face.
HTFETI_512x10 = dict(
    name="elasticity_512sub_10e3",
    physics=0, clusters=(4, 4, 4), domains=(2, 2, 2), elements=(10, 10, 10),
    length=(1.0, 1.0, 1.0), young=E_STEEL, poisson=0.3, density=7850.0,
    accel=(0.0, 0.0, -9.81), fixed=(0, 0, 0, 0, 7, 0),
    traction_seed=20261202, traction_scale=1.0, method="hybrid",
)

# 1000 subdomains of 12^3 elements (6591 DOFs each), TFETI: the coarse space
# holds the 6 rigid-body modes of every floating subdomain (G G^T 6000 x 6000).
MIXED_1000x12 = dict(
    name="elasticity_1000sub_12e3",
    physics=0, clusters=(5, 5, 5), domains=(2, 2, 2), elements=(12, 12, 12),
    length=(1.0, 1.0, 1.0), young=E_STEEL, poisson=0.3, density=7850.0,
    accel=(0.0, 0.0, -9.81), fixed=(0, 0, 0, 0, 7, 0), method="total",
)

# 40^3 cube in 4^3 subdomains, per-element Young's modulus, downward traction
# on z=1 (unit pressure pulling down = negative pressure on z1).
ADJOINT_64x10 = dict(
    name="elasticity_64sub_10e3_randomE",
    physics=0, clusters=(1, 1, 1), domains=(4, 4, 4), elements=(10, 10, 10),
    length=(1.0, 1.0, 1.0), young=E_STEEL, poisson=0.3, density=0.0,
    fixed=(0, 0, 0, 0, 7, 0), pressure=(0, 0, 0, 0, 0, 1.0),
    young_seed=20261205,
)

# Uneven subdomains (boxes of 4..14 elements per direction, seeded per slab so
# that neighbouring boxes stay conforming), 2000 subdomains in one cluster,
# heat with conductivity log-uniform in [0.1, 10] per subdomain; the load
# cases are seeded U(-1,1) nodal sources applied with resolve().
def uneven_slabs(counts, seed, lo=4, hi=14):
    """Seeded element counts per subdomain slab for (NSx, NSy, NSz) slabs."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return tuple(tuple(int(v) for v in rng.integers(lo, hi + 1, size=n)) for n in counts)


HEAT_UNEVEN_2000 = dict(
    name="heat_uneven_2000sub",
    physics=1, clusters=(1, 1, 1), domains=(10, 10, 20), elements=(9, 9, 9),
    slab_elements=uneven_slabs((10, 10, 20), 20261203),
    length=(1.0, 1.0, 2.0), conductivity=1.0, conductivity_seed=20261204,
    fixed=(0, 0, 0, 0, 1, 0), heat_source=1.0,
)

# the elasticity variant of the uneven workload: 2000 subdomains of 4..14
# elements per direction (seeded slabs), gravity, z=0 clamped; its load cases
# are seeded U(-1,1) nodal forces (nodal_load_cases)
ELASTICITY_UNEVEN_2000 = dict(
    name="elasticity_uneven_2000sub",
    physics=0, clusters=(1, 1, 1), domains=(10, 10, 20), elements=(9, 9, 9),
    slab_elements=uneven_slabs((10, 10, 20), 20261208),
    length=(1.0, 1.0, 2.0), young=E_STEEL, poisson=0.3, density=7850.0,
    accel=(0.0, 0.0, -9.81), fixed=(0, 0, 0, 0, 7, 0),
)

ELASTICITY_UNEVEN_512 = dict(
    name="elasticity_uneven_512sub",
    physics=0, clusters=(1, 1, 1), domains=(8, 8, 8), elements=(9, 9, 9),
    slab_elements=uneven_slabs((8, 8, 8), 20261206),
    length=(1.0, 1.0, 1.0), young=E_STEEL, poisson=0.3, density=7850.0,
    accel=(0.0, 0.0, -9.81), fixed=(0, 0, 0, 0, 7, 0),
)


def nodal_load_cases(spec, n_cases, seed=20261207):
    """Seeded U(-1,1) nodal loads, (n_cases, n_nodes, ndof)."""
    import numpy as np
    from .core import slabs
    nn = [sum(v) + 1 for v in slabs(spec)]
    nd = 3 if spec.get("physics", 0) == 0 else 1
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, size=(n_cases, nn[0] * nn[1] * nn[2], nd))


def uniform_equivalent(spec):
    """The same global mesh as one subdomain (for the undecomposed direct solve)."""
    from .core import slabs
    s = dict(spec)
    s.pop("slab_elements", None)
    s["clusters"] = (1, 1, 1)
    s["domains"] = (1, 1, 1)
    s["elements"] = tuple(sum(v) for v in slabs(spec))
    return s


# --- the reference's own benchmark inputs (golden .emr values) -----------

E_PATCH = 2.1e11
PZ = 70 * E_PATCH / 5200
PX = 3 * PZ / 7

# benchmarks/linearElasticity3D/steadystate/pressure/espreso.ecf (HEXA8 args
# 2 2 1 / 1 2 2 / 5 5 5): rollers on X0/Y0/Z0, normal pressure on X1/Y1/Z1.
# Closed form u = (0, 0, -0.01 z): |u| = 1 on Z1.
PRESSURE_PATCH = dict(
    name="pressure_patch",
    physics=0, clusters=(2, 2, 1), domains=(1, 2, 2), elements=(5, 5, 5),
    length=(100.0, 100.0, 100.0), young=E_PATCH, poisson=0.3, density=7850.0,
    fixed=(1, 0, 2, 0, 4, 0), pressure=(0.0, PX, 0.0, PX, 0.0, PZ),
)

# benchmarks/linearElasticity3D/steadystate/gravity/espreso.ecf + HEXA8.emr:
# |u| min 0, max 1.794600e-01, avg 1.113490e-01.
GRAVITY = dict(
    name="gravity",
    physics=0, clusters=(2, 2, 1), domains=(1, 2, 2), elements=(5, 5, 5),
    length=(100.0, 100.0, 100.0), young=2.1e9, poisson=0.3, density=7850.0,
    accel=(0.0, 0.0, 9.8066), fixed=(0, 0, 0, 0, 7, 0),
)
GRAVITY_EMR = dict(min=0.0, max=1.794600e-01, avg=1.113490e-01)

# benchmarks/diffusion3D/steadystate/simple/espreso.ecf (HEXA8 args
# 3 2 1 / 1 2 3 / 4 3 4): T=100 on x=0, T=300 on x=1; golden max(TOP)=100,
# min(BOTTOM)=300, avg=200.  With a diagonal conductivity the exact solution
# is linear in x.
HEAT_SIMPLE = dict(
    name="heat_simple",
    physics=1, clusters=(3, 2, 1), domains=(1, 2, 3), elements=(4, 3, 4),
    length=(1.0, 1.0, 1.0), conductivity=1.0,
    fixed=(1, 1, 0, 0, 0, 0), fixed_value=(100.0, 300.0, 0, 0, 0, 0),
)

ALL = {p["name"]: p for p in (TFETI_64x5, HTFETI_512x10, MIXED_1000x12, ADJOINT_64x10, HEAT_UNEVEN_2000,
                              ELASTICITY_UNEVEN_2000, ELASTICITY_UNEVEN_512, PRESSURE_PATCH, GRAVITY, HEAT_SIMPLE)}


def scaled(spec, **over):
    """Copy of ``spec`` with fields replaced (used for small test variants)."""
    s = dict(spec)
    s.update(over)
    return s
