"""ctypes mirror of the reference's FETISolver interface over the C ABI.

Reference interface: espreso::LinearSolver {update, solve, finalize, precision}
(src/linearsolver/linearsolver.h:9-23) implemented by FETISolver
(src/solver/generic/FETISolver.h:22-76).  ``FETISolver`` below keeps that
shape: construct with options, ``update`` with subdomain data (K_i, f_i,
R_i, B1 rows, scaling, c), ``solve`` returns the per-subdomain primal.
``GridFETI`` drives the same core from the structured-grid builder.

No CPU fallback exists: a missing library raises ``LibraryNotBuilt``.
"""
import ctypes as C
import os
import warnings

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libespreso_amd.so")

PREC = {"none": 0, "lumped": 1, "weight": 2, "dirichlet": 3, "super_dirichlet": 4}
NOT_CONVERGED = 1  # ESPRESO_NOT_CONVERGED: solved, stop test not met (stats filled)
SC = {"fp64": 0, "fp32": 1, "bf16": 2}


class LibraryNotBuilt(RuntimeError):
    pass


class GridProblem(C.Structure):
    _fields_ = [
        ("physics", C.c_int32),
        ("clusters", C.c_int32 * 3),
        ("domains", C.c_int32 * 3),
        ("elements", C.c_int32 * 3),
        ("length", C.c_double * 3),
        ("young", C.c_double),
        ("poisson", C.c_double),
        ("density", C.c_double),
        ("accel", C.c_double * 3),
        ("conductivity", C.c_double),
        ("heat_source", C.c_double),
        ("fixed", C.c_int32 * 6),
        ("fixed_value", C.c_double * 6),
        ("pressure", C.c_double * 6),
        ("traction_seed", C.c_uint64),
        ("traction_scale", C.c_double),
        ("young_seed", C.c_uint64),
        ("conductivity_seed", C.c_uint64),
        ("material_elem", C.POINTER(C.c_double)),
        ("slab_elements", C.POINTER(C.c_int32)),
        ("device_assembly", C.c_int32),
        ("b0_type", C.c_int32),
    ]


class Options(C.Structure):
    _fields_ = [
        ("method", C.c_int32),
        ("preconditioner", C.c_int32),
        ("max_iterations", C.c_int32),
        ("scaling", C.c_int32),
        ("sc_precision", C.c_int32),
        ("device", C.c_int32),
        ("precision", C.c_double),
        ("memory_budget_gb", C.c_double),
        ("sd_precision", C.c_int32),
        ("truncate_output", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("n_lambda", C.c_int32),
        ("n_coarse", C.c_int32),
        ("n_subdomains", C.c_int32),
        ("final_norm", C.c_double),
        ("tol", C.c_double),
        ("norm_b", C.c_double),
        ("time_assembly", C.c_double),
        ("time_coarse", C.c_double),
        ("time_pcg", C.c_double),
        ("time_primal", C.c_double),
        ("time_total", C.c_double),
        ("sc_bytes", C.c_double),
        ("sd_bytes", C.c_double),
        ("assembly_flops", C.c_double),
        ("refinement_steps", C.c_int32),
        ("pad", C.c_int32),
        ("time_load", C.c_double),
        ("symv_bytes", C.c_double),
        ("symv_bytes_sd", C.c_double),
        ("converged", C.c_int32),
        ("pad2", C.c_int32),
        ("kkt_primal", C.c_double),
        ("kkt_dual", C.c_double),
        ("time_build", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k not in ("pad", "pad2")}


# every entry point of include/espreso_amd.h with its ctypes signature
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)
_H = C.c_void_p
SIGNATURES = {
    "espreso_last_error": (C.c_char_p, []),
    "espreso_version": (C.c_int, []),
    "espreso_grid_sizes": (C.c_int, [C.POINTER(GridProblem), _I, _I, _I]),
    "espreso_grid_load": (C.c_int, [_H, C.POINTER(GridProblem)]),
    "espreso_grid_l2g": (C.c_int, [_H, C.c_int32, _I]),
    "espreso_grid_export": (C.c_int, [C.POINTER(GridProblem), C.c_int32, C.c_int32, _I, _I, _I, _D, _D, _D, _I, _I,
                                      _D, _D, _D]),
    "espreso_grid_export_cluster": (C.c_int, [C.POINTER(GridProblem), C.c_int32, _I, _I, _I, _I, _D]),
    "espreso_grid_nodal_solution": (C.c_int, [_H, _D]),
    "espreso_grid_unit_element": (C.c_int, [C.POINTER(GridProblem), _D]),
    "espreso_grid_set_nodal_rhs": (C.c_int, [_H, _D]),
    "espreso_grid_element_sensitivity": (C.c_int, [_H, _D, _D, _D]),
    "espreso_grid_update_material": (C.c_int, [_H, _D]),
    "espreso_feti_update_values": (C.c_int, [_H, C.c_int32, _D]),
    "espreso_feti_set_rhs": (C.c_int, [_H, C.c_int32, _D]),
    "espreso_feti_load_work": (C.c_int, [_H, _D]),
    "espreso_grid_nodal_load": (C.c_int, [_H, _D]),
    "espreso_feti_resolve": (C.c_int, [_H, C.POINTER(Stats)]),
    "espreso_feti_resolve_block": (C.c_int, [_H, C.c_int32, _D, _D, _I, _D, C.POINTER(Stats)]),
    "espreso_feti_subdomain_values": (C.c_int, [_H, C.c_int32, _D]),
    "espreso_feti_symv_items": (C.c_int, [_H, C.POINTER(C.c_int64)]),
    "espreso_grid_resolve_block": (C.c_int, [_H, C.c_int32, _D, _D, _I, _D, C.POINTER(Stats)]),
    "espreso_feti_create": (C.c_int, [C.POINTER(_H), C.c_int32, C.POINTER(Options)]),
    "espreso_feti_destroy": (C.c_int, [_H]),
    "espreso_feti_sizes": (C.c_int, [_H, _I]),
    "espreso_feti_set_subdomain": (C.c_int, [_H, C.c_int32, C.c_int32, _I, _I, _D, _D, C.c_int32, _D,
                                             C.c_int32, _I, _I, _D, _D]),
    "espreso_feti_set_dual": (C.c_int, [_H, C.c_int32, _D]),
    "espreso_feti_set_cluster": (C.c_int, [_H, C.c_int32, C.c_int32, C.c_int32, _I, _I, _D]),
    "espreso_feti_setup": (C.c_int, [_H]),
    "espreso_feti_solve": (C.c_int, [_H, C.POINTER(Stats)]),
    "espreso_feti_primal": (C.c_int, [_H, C.c_int32, _D]),
    "espreso_feti_dual": (C.c_int, [_H, _D]),
    "espreso_feti_apply_F": (C.c_int, [_H, _D, _D]),
    "espreso_feti_apply_prec": (C.c_int, [_H, _D, _D]),
    "espreso_feti_project": (C.c_int, [_H, _D, _D]),
    "espreso_feti_project_device": (C.c_int, [_H, C.c_void_p, C.c_void_p, C.c_void_p]),
    "espreso_feti_vec_b": (C.c_int, [_H, _D]),
    "espreso_feti_subdomain_dense": (C.c_int, [_H, C.c_int32, C.c_int32, _D]),
    "espreso_feti_subdomain_boundary": (C.c_int, [_H, C.c_int32, _I, _I]),
    "espreso_feti_time_apply": (C.c_int, [_H, C.c_int32, C.c_int32, _D, _D]),
    "espreso_debug_mfma": (C.c_int, [_D, _D, _D]),
    "espreso_debug_copy": (C.c_int, [C.c_int64, C.c_int32, _D]),
    "espreso_debug_stream": (C.c_int, [C.c_int64, C.c_int32, C.c_int32, _D]),
    "espreso_debug_mfma_rate": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, _D]),
    "espreso_debug_tile_gemm": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _D, C.c_int32, _D,
                                          C.c_int32, _D, C.c_int32, C.c_double, C.c_double]),
}


def debug_mfma(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.zeros(320)
    _check(lib().espreso_debug_mfma(_dp(a), _dp(b), _dp(out)))
    return out[:256].reshape(64, 4)


def debug_mfma4(a, b):
    """One v_mfma_f64_4x4x4f64 (4 blocks) with lane l holding a[l], b[l]: the
    result of each lane."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.zeros(320)
    _check(lib().espreso_debug_mfma(_dp(a), _dp(b), _dp(out)))
    return out[256:].copy()


def debug_copy(nbytes, reps=5):
    ms = C.c_double()
    _check(lib().espreso_debug_copy(int(nbytes), reps, C.byref(ms)))
    return ms.value


def debug_stream(nbytes, reps=5, workgroups=4096):
    """Read-only HBM stream: mean ms per pass over nbytes."""
    ms = C.c_double()
    _check(lib().espreso_debug_stream(int(nbytes), reps, workgroups, C.byref(ms)))
    return ms.value


def debug_mfma_rate(shape=0, chains=8, workgroups_per_cu=0):
    """FP64 MFMA issue rate in TFLOP/s (shape 0: 16x16x4, 1: 4x4x4 with 4 blocks);
    workgroups_per_cu > 0 caps the occupancy (waves per SIMD)."""
    t = C.c_double()
    _check(lib().espreso_debug_mfma_rate(shape, chains, workgroups_per_cu, C.byref(t)))
    return t.value


def debug_tile_gemm(A, B, C, ta=False, tb=False, alpha=1.0, beta=0.0):
    """C = alpha op(A) op(B) + beta C on one 64x64 MFMA tile (column-major)."""
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    Cm = np.asfortranarray(C, dtype=np.float64).copy()
    M, N = Cm.shape
    K = A.shape[0] if ta else A.shape[1]
    a = A.ravel(order="F")
    b = B.ravel(order="F")
    c = Cm.ravel(order="F")
    _check(lib().espreso_debug_tile_gemm(int(ta), int(tb), M, N, K, _dp(a), A.shape[0], _dp(b), B.shape[0], _dp(c),
                                         M, alpha, beta))
    return c.reshape(N, M).T

_lib = None


def lib():
    """Load libespreso_amd.so (raises LibraryNotBuilt when absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryNotBuilt(f"{LIB_PATH} is missing: run __graft_entry__.build() (make -C espreso_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc, solve=False):
    """Raise on a failed call.  For solve-like calls ESPRESO_NOT_CONVERGED
    returns normally (the stats say converged = 0) with a warning, as the
    reference's solver reports an unconverged PCG and still returns its
    iterate."""
    if rc == NOT_CONVERGED and solve:
        warnings.warn("FETI: stop test not met within the iteration limit", RuntimeWarning, stacklevel=3)
        return
    if rc != 0:
        raise RuntimeError(f"espreso_amd error {rc}: {lib().espreso_last_error().decode()}")


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


def grid_problem(spec):
    p = GridProblem()
    p.physics = spec.get("physics", 0)
    for k in ("clusters", "domains", "elements"):
        getattr(p, k)[:] = spec[k]
    p.length[:] = spec.get("length", (1.0, 1.0, 1.0))
    p.young = spec.get("young", 2.1e11)
    p.poisson = spec.get("poisson", 0.3)
    p.density = spec.get("density", 7850.0)
    p.accel[:] = spec.get("accel", (0.0, 0.0, 0.0))
    p.conductivity = spec.get("conductivity", 1.0)
    p.heat_source = spec.get("heat_source", 0.0)
    p.fixed[:] = spec.get("fixed", (0,) * 6)
    p.fixed_value[:] = spec.get("fixed_value", (0.0,) * 6)
    p.pressure[:] = spec.get("pressure", (0.0,) * 6)
    p.traction_seed = spec.get("traction_seed", 0)
    p.traction_scale = spec.get("traction_scale", 1.0)
    p.young_seed = spec.get("young_seed", 0)
    p.conductivity_seed = spec.get("conductivity_seed", 0)
    p.device_assembly = int(bool(spec.get("device_assembly", False)))
    p.b0_type = {"corners": 0, "kernels": 1}[spec.get("b0", "corners")]
    keep = []  # arrays the struct points to must outlive it
    ye = spec.get("material_elem")
    if ye is not None:
        ye = np.ascontiguousarray(ye, dtype=np.float64)
        keep.append(ye)
        p.material_elem = _dp(ye)
    sl = spec.get("slab_elements")
    if sl is not None:
        flat = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.int32) for v in sl]), dtype=np.int32)
        assert flat.size == sum(spec["clusters"][d] * spec["domains"][d] for d in range(3))
        keep.append(flat)
        p.slab_elements = _ip(flat)
    p._keep = keep
    return p


def slabs(spec):
    """Elements of every subdomain slab per direction (uniform unless spec
    has slab_elements)."""
    if spec.get("slab_elements") is not None:
        return [list(map(int, v)) for v in spec["slab_elements"]]
    return [[spec["elements"][d]] * (spec["clusters"][d] * spec["domains"][d]) for d in range(3)]


def subdomain_elements(spec, s):
    """(ex, ey, ez) elements of subdomain s (cluster-major ids, x fastest)."""
    dom, cl = spec["domains"], spec["clusters"]
    dpc = dom[0] * dom[1] * dom[2]
    c, d = s // dpc, s % dpc
    cc = (c % cl[0], (c // cl[0]) % cl[1], c // (cl[0] * cl[1]))
    dd = (d % dom[0], (d // dom[0]) % dom[1], d // (dom[0] * dom[1]))
    sl = slabs(spec)
    return tuple(sl[q][cc[q] * dom[q] + dd[q]] for q in range(3))


def n_elements(spec):
    return int(np.prod([sum(v) for v in slabs(spec)]))


def unit_element(spec):
    """Unit-modulus element matrix of the grid builder (closed-form integrals)."""
    p = grid_problem(spec)
    nd = 3 if spec.get("physics", 0) == 0 else 1
    Ke = np.zeros((8 * nd) ** 2)
    _check(lib().espreso_grid_unit_element(C.byref(p), _dp(Ke)))
    return Ke.reshape(8 * nd, 8 * nd)


METHOD = {"total": 0, "TOTAL_FETI": 0, "hybrid": 1, "HYBRID_FETI": 1}


def make_options(preconditioner="dirichlet", precision=1e-8, max_iterations=500, scaling=1, sc_precision="fp64",
                 device=0, memory_budget_gb=0.0, method="total", sd_precision="fp64", truncate_output=False):
    o = Options()
    o.method = METHOD[method] if isinstance(method, str) else int(method)
    o.preconditioner = PREC[preconditioner] if isinstance(preconditioner, str) else int(preconditioner)
    o.max_iterations = max_iterations
    o.scaling = scaling
    o.sc_precision = SC[sc_precision] if isinstance(sc_precision, str) else int(sc_precision)
    o.device = device
    o.precision = precision
    o.memory_budget_gb = memory_budget_gb
    o.sd_precision = SC[sd_precision] if isinstance(sd_precision, str) else int(sd_precision)
    o.truncate_output = int(bool(truncate_output))
    return o


def grid_export(spec, i, scaling=1):
    """Host-only export of subdomain i of the structured-grid builder."""
    p = grid_problem(spec)
    sz = np.zeros(6, dtype=np.int32)
    nul = None
    _check(lib().espreso_grid_export(C.byref(p), scaling, i, _ip(sz), nul, nul, nul, nul, nul, nul, nul, nul, nul,
                                     nul))
    n, nnz, k, nl, ntot = (int(v) for v in sz[:5])
    rp = np.zeros(n + 1, np.int32)
    ci = np.zeros(nnz, np.int32)
    kv = np.zeros(nnz)
    f = np.zeros(n)
    R = np.zeros(n * k)
    lam = np.zeros(nl, np.int32)
    dof = np.zeros(nl, np.int32)
    sign = np.zeros(nl)
    scale = np.zeros(nl)
    c = np.zeros(ntot)
    _check(lib().espreso_grid_export(C.byref(p), scaling, i, _ip(sz), _ip(rp), _ip(ci), _dp(kv), _dp(f), _dp(R),
                                     _ip(lam), _ip(dof), _dp(sign), _dp(scale), _dp(c)))
    return dict(n=n, rowptr=rp, colidx=ci, vals=kv, f=f, R=R.reshape(k, n).T, lam=lam, dof=dof, sign=sign,
                scale=scale, c=c, n_sub=int(sz[5]))


def grid_export_cluster(spec, i):
    """Host-only export of subdomain i's cluster id and B0 corner entries."""
    p = grid_problem(spec)
    cl = np.zeros(1, np.int32)
    n0 = np.zeros(1, np.int32)
    _check(lib().espreso_grid_export_cluster(C.byref(p), i, _ip(cl), _ip(n0), None, None, None))
    row = np.zeros(int(n0[0]), np.int32)
    dof = np.zeros(int(n0[0]), np.int32)
    sign = np.zeros(int(n0[0]))
    _check(lib().espreso_grid_export_cluster(C.byref(p), i, _ip(cl), _ip(n0), _ip(row), _ip(dof), _dp(sign)))
    return dict(cluster=int(cl[0]), row=row, dof=dof, sign=sign)


class FETISolver:
    """One cluster of subdomains on one GPU (reference: FETISolver)."""

    def __init__(self, n_subdomains, **opts):
        self.n_sub = n_subdomains
        self.opts = make_options(**opts)
        h = C.c_void_p()
        _check(lib().espreso_feti_create(C.byref(h), n_subdomains, C.byref(self.opts)))
        self.h = h
        self.n_lambda = None
        self.stats = None

    def close(self):
        if getattr(self, "h", None):
            lib().espreso_feti_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- LinearSolver::update --
    def set_subdomain(self, i, K_csr, f, R, lam, dof, sign, scale):
        rp, ci, kv = (np.ascontiguousarray(a) for a in K_csr)
        rp = rp.astype(np.int32)
        ci = ci.astype(np.int32)
        kv = kv.astype(np.float64)
        f = np.ascontiguousarray(f, dtype=np.float64)
        R = np.asfortranarray(R, dtype=np.float64)
        lam = np.ascontiguousarray(lam, dtype=np.int32)
        dof = np.ascontiguousarray(dof, dtype=np.int32)
        sign = np.ascontiguousarray(sign, dtype=np.float64)
        scale = np.ascontiguousarray(scale, dtype=np.float64)
        n = len(f)
        Rf = np.ascontiguousarray(R.T).ravel()  # column-major n x k
        _check(lib().espreso_feti_set_subdomain(self.h, i, n, _ip(rp), _ip(ci), _dp(kv), _dp(f), R.shape[1], _dp(Rf),
                                                len(lam), _ip(lam), _ip(dof), _dp(sign), _dp(scale)))

    def set_cluster(self, i, cluster, row, dof, sign):
        row = np.ascontiguousarray(row, np.int32)
        dof = np.ascontiguousarray(dof, np.int32)
        sign = np.ascontiguousarray(sign, np.float64)
        _check(lib().espreso_feti_set_cluster(self.h, i, cluster, len(row), _ip(row), _ip(dof), _dp(sign)))

    def set_dual(self, c):
        c = np.ascontiguousarray(c, dtype=np.float64)
        self.n_lambda = len(c)
        _check(lib().espreso_feti_set_dual(self.h, len(c), _dp(c)))

    def setup(self):
        _check(lib().espreso_feti_setup(self.h))

    # -- LinearSolver::solve --
    def solve(self):
        st = Stats()
        _check(lib().espreso_feti_solve(self.h, C.byref(st)), solve=True)
        self.stats = st.as_dict()
        return self.stats

    def update_values(self, i, values):
        """New K_i values on the structure given at set_subdomain; the next
        setup() re-assembles F_i and S_D (FETISolver::update with K)."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        _check(lib().espreso_feti_update_values(self.h, i, _dp(v)))

    def set_rhs(self, i, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        _check(lib().espreso_feti_set_rhs(self.h, i, _dp(f)))

    def load_work(self):
        """f^T u of the last solution (compliance)."""
        c = np.zeros(1)
        _check(lib().espreso_feti_load_work(self.h, _dp(c)))
        return float(c[0])

    def resolve(self):
        """Solve with the loads given by set_rhs, reusing the assembled operators."""
        st = Stats()
        _check(lib().espreso_feti_resolve(self.h, C.byref(st)), solve=True)
        self.stats = st.as_dict()
        return self.stats

    def resolve_block(self, f):
        """Solve several load cases together on the assembled operators.
        f: (nrhs, sum_i n_i) subdomain-local loads (subdomain order).  Returns
        (u of the same shape, iterations per case, final ||P r|| per case)."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        nrhs = f.shape[0]
        u = np.zeros_like(f)
        it = np.zeros(nrhs, np.int32)
        nr = np.zeros(nrhs)
        st = Stats()
        _check(lib().espreso_feti_resolve_block(self.h, nrhs, _dp(f), _dp(u), _ip(it), _dp(nr), C.byref(st)),
               solve=True)
        self.block_stats = st.as_dict()
        return u, it, nr

    def primal(self, i, n):
        u = np.zeros(n)
        _check(lib().espreso_feti_primal(self.h, i, _dp(u)))
        return u

    def dual(self):
        lam = np.zeros(self.n_lambda)
        _check(lib().espreso_feti_dual(self.h, _dp(lam)))
        return lam

    def _op(self, name, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n_lambda)
        _check(getattr(lib(), name)(self.h, _dp(x), _dp(y)))
        return y

    def apply_F(self, x):
        return self._op("espreso_feti_apply_F", x)

    def apply_prec(self, x):
        return self._op("espreso_feti_apply_prec", x)

    def project(self, x):
        return self._op("espreso_feti_project", x)

    def project_device(self, x_ptr, y_ptr, stream=None):
        """y = P x on device vectors (addresses of n_lambda doubles), launched
        on `stream` (hipStream_t address, None: the instance's stream) and not
        waited for; allocates nothing (espreso_feti_project_device)."""
        _check(lib().espreso_feti_project_device(self.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr),
                                                 C.c_void_p(stream) if stream else None))

    def vec_b(self):
        b = np.zeros(self.n_lambda)
        _check(lib().espreso_feti_vec_b(self.h, _dp(b)))
        return b

    def boundary(self, i):
        m = C.c_int32()
        _check(lib().espreso_feti_subdomain_boundary(self.h, i, C.byref(m), None))
        b = np.zeros(m.value, dtype=np.int32)
        _check(lib().espreso_feti_subdomain_boundary(self.h, i, C.byref(m), _ip(b)))
        return b

    def symv_items(self):
        """SYMV work-item sizes chosen by the autotuner: (FP64, FP32, BF16),
        0 for a precision nothing is stored in."""
        t = (C.c_int64 * 3)()
        _check(lib().espreso_feti_symv_items(self.h, t))
        return tuple(int(v) for v in t)

    def values(self, i, nnz):
        """K_i CSR values as the device holds them (after setup)."""
        v = np.zeros(nnz)
        _check(lib().espreso_feti_subdomain_values(self.h, i, _dp(v)))
        return v

    def dense(self, i, which):
        m = len(self.boundary(i))
        out = np.zeros(m * m)
        _check(lib().espreso_feti_subdomain_dense(self.h, i, which, _dp(out)))
        return out.reshape(m, m).T.copy()

    def sizes(self):
        s = np.zeros(3, dtype=np.int32)
        _check(lib().espreso_feti_sizes(self.h, _ip(s)))
        return [int(v) for v in s]

    def time_apply(self, op=0, reps=10):
        a = C.c_double()
        s = C.c_double()
        _check(lib().espreso_feti_time_apply(self.h, op, reps, C.byref(a), C.byref(s)))
        return a.value, s.value


class GridFETI(FETISolver):
    """FETISolver fed by the structured-grid builder."""

    def __init__(self, spec, **opts):
        self.spec = spec
        self.problem = grid_problem(spec)
        ns, nn, nd = C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib().espreso_grid_sizes(C.byref(self.problem), C.byref(ns), C.byref(nn), C.byref(nd)))
        self.n_nodes, self.ndof = nn.value, nd.value
        super().__init__(ns.value, **opts)
        _check(lib().espreso_grid_load(self.h, C.byref(self.problem)))
        self.n_lambda = self.sizes()[1]

    def nodal_solution(self):
        u = np.zeros(self.n_nodes * self.ndof)
        _check(lib().espreso_grid_nodal_solution(self.h, _dp(u)))
        return u.reshape(self.n_nodes, self.ndof)

    def nodal_load(self):
        g = np.zeros(self.n_nodes * self.ndof)
        _check(lib().espreso_grid_nodal_load(self.h, _dp(g)))
        return g.reshape(self.n_nodes, self.ndof)

    def set_nodal_rhs(self, g):
        """New nodal load (n_nodes x ndof) for resolve()."""
        g = np.ascontiguousarray(g, dtype=np.float64).reshape(-1)
        assert g.size == self.n_nodes * self.ndof
        _check(lib().espreso_grid_set_nodal_rhs(self.h, _dp(g)))

    def update_material(self, material_elem):
        """New per-element Young's modulus / conductivity (element index
        ex + NEx*(ey + NEy*ez)); K_i is recomputed on the GPU and the next
        setup() re-assembles the dual operators."""
        m = np.ascontiguousarray(material_elem, dtype=np.float64).reshape(-1)
        assert m.size == n_elements(self.spec)
        _check(lib().espreso_grid_update_material(self.h, _dp(m)))

    def element_sensitivity(self, u, w=None):
        """-w_e^T K_e0 u_e for every element (GPU); w defaults to u."""
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        w = u if w is None else np.ascontiguousarray(w, dtype=np.float64).reshape(-1)
        out = np.zeros(n_elements(self.spec))
        _check(lib().espreso_grid_element_sensitivity(self.h, _dp(u), _dp(w), _dp(out)))
        return out

    def l2g(self, i):
        ex, ey, ez = subdomain_elements(self.spec, i)
        n = (ex + 1) * (ey + 1) * (ez + 1)
        g = np.zeros(n, dtype=np.int32)
        _check(lib().espreso_grid_l2g(self.h, i, _ip(g)))
        return g

    def resolve_block_nodal(self, loads):
        """Block re-solve of nodal load cases (nrhs, n_nodes, ndof): loads are
        split over the subdomain copies of a node by multiplicity (as
        set_nodal_rhs), solutions averaged over the copies (as nodal_solution).
        Returns (u (nrhs, n_nodes, ndof), iterations, final norms)."""
        loads = np.ascontiguousarray(loads, dtype=np.float64).reshape(len(loads), self.n_nodes, self.ndof)
        nrhs = loads.shape[0]
        un = np.zeros_like(loads)
        it = np.zeros(nrhs, np.int32)
        nr = np.zeros(nrhs)
        st = Stats()
        _check(lib().espreso_grid_resolve_block(self.h, nrhs, _dp(loads), _dp(un), _ip(it), _dp(nr), C.byref(st)),
               solve=True)
        self.block_stats = st.as_dict()
        return un, it, nr


# ---------------------------------------------------------------- FETI4I ----
_F4_SIGS = {
    "FETI4ISetDefaultIntegerOptions": (None, [_I]),
    "FETI4ISetDefaultRealOptions": (None, [_D]),
    "FETI4ICreateStiffnessMatrix": (None, [C.POINTER(_H), C.c_int32, C.c_int32]),
    "FETI4IAddElement": (None, [_H, C.c_int32, C.c_int32, _I, C.c_int32, _I, _D]),
    "FETI4ICreateInstance": (None, [C.POINTER(_H), _H, C.c_int32, _D, _I, C.c_int32, _I, C.c_int32, _I, _D, _I,
                                    _D]),
    "FETI4ISolve": (None, [_H, C.c_int32, _D]),
    "FETI4IUpdateStiffnessMatrix": (None, [_H, _H]),
    "FETI4IUpdateRhs": (None, [_H, C.c_int32, _D]),
    "FETI4IUpdateDirichlet": (None, [_H, C.c_int32, _I, _D]),
    "FETI4IDestroy": (None, [_H]),
    "FETI4IGetLastStatus": (C.c_int, []),
    "FETI4IGetLastError": (C.c_char_p, []),
    "FETI4IInstanceInfo": (C.c_int, [_H, _I]),
    "FETI4IExportSubdomain": (C.c_int, [_H, C.c_int32, _I, _I, _I, _I, _D, _D, _D, _I, _I, _D, _D]),
    "FETI4IExportCorners": (C.c_int, [_H, C.c_int32, _I, _I, _I, _D]),
    "FETI4IGetStats": (C.c_int, [_H, _D]),
}
_f4_ready = False


def f4lib():
    global _f4_ready
    L = lib()
    if not _f4_ready:
        for name, (res, args) in _F4_SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _f4_ready = True
    return L


def _f4check(solve=False):
    st = f4lib().FETI4IGetLastStatus()
    if st == NOT_CONVERGED and solve:
        warnings.warn("FETI4I: stop test not met within FETI4I_MAX_ITERATIONS", RuntimeWarning, stacklevel=3)
        return
    if st != 0:
        raise RuntimeError(f"FETI4I error {st}: {f4lib().FETI4IGetLastError().decode()}")


class FETI4I:
    """Thin Python view of the FETI4I C API (include/feti4i.h)."""

    def __init__(self, elements, size, rhs, dirichlet=(), subdomains=4, preconditioner=3, precision=1e-8,
                 max_iterations=500, index_base=0, method="total"):
        L = f4lib()
        io = np.zeros(10, np.int32)
        ro = np.zeros(1)
        L.FETI4ISetDefaultIntegerOptions(_ip(io))
        L.FETI4ISetDefaultRealOptions(_dp(ro))
        io[0] = subdomains
        io[1] = max_iterations
        io[2] = METHOD[method] if isinstance(method, str) else method
        io[3] = preconditioner
        ro[0] = precision
        self.index_base = index_base
        self.m = self._matrix(elements)
        rhs = np.ascontiguousarray(rhs, np.float64)
        l2g = np.arange(size, dtype=np.int32) + index_base
        di = np.ascontiguousarray([d for d, _ in dirichlet], np.int32)
        dv = np.ascontiguousarray([v for _, v in dirichlet], np.float64)
        inst = C.c_void_p()
        L.FETI4ICreateInstance(C.byref(inst), self.m, size, _dp(rhs), _ip(l2g), 0, None, len(di), _ip(di), _dp(dv),
                               _ip(io), _dp(ro))
        _f4check()
        self.inst = inst
        self.size = size

    def _matrix(self, elements):
        L = f4lib()
        m = C.c_void_p()
        L.FETI4ICreateStiffnessMatrix(C.byref(m), 0, self.index_base)
        _f4check()
        for nodes, dofs, K in elements:
            nodes = np.ascontiguousarray(nodes, np.int32)
            dofs = np.ascontiguousarray(dofs, np.int32)
            K = np.ascontiguousarray(K, np.float64)
            L.FETI4IAddElement(m, 3, len(nodes), _ip(nodes), len(dofs), _ip(dofs), _dp(K))
        _f4check()
        return m

    def update_stiffness(self, elements):
        """FETI4IUpdateStiffnessMatrix with a new element list (same structure:
        values only are re-assembled and the decomposition is kept)."""
        m = self._matrix(elements)
        try:
            f4lib().FETI4IUpdateStiffnessMatrix(self.inst, m)
            _f4check()
        finally:
            f4lib().FETI4IDestroy(m)

    def info(self):
        v = np.zeros(8, np.int32)
        f4lib().FETI4IInstanceInfo(self.inst, _ip(v))
        _f4check()
        return dict(subdomains=int(v[0]), n_lambda=int(v[1]), n_coarse=int(v[2]), iterations=int(v[3]),
                    ndof_node=int(v[4]), max_kdim=int(v[5]), n_corner_rows=int(v[6]), method=int(v[7]))

    def corners(self, i):
        """HYBRID_FETI B0 entries of subdomain i: (row, local DOF, sign)."""
        L = f4lib()
        n = np.zeros(1, np.int32)
        L.FETI4IExportCorners(self.inst, i, _ip(n), None, None, None)
        _f4check()
        row = np.zeros(int(n[0]), np.int32)
        dof = np.zeros(int(n[0]), np.int32)
        sign = np.zeros(int(n[0]))
        L.FETI4IExportCorners(self.inst, i, _ip(n), _ip(row), _ip(dof), _dp(sign))
        _f4check()
        return row, dof, sign

    def subdomain(self, i):
        L = f4lib()
        sz = np.zeros(4, np.int32)
        nul = None
        L.FETI4IExportSubdomain(self.inst, i, _ip(sz), nul, nul, nul, nul, nul, nul, nul, nul, nul, nul)
        _f4check()
        n, nnz, k, nl = (int(x) for x in sz)
        g = np.zeros(n, np.int32)
        rp = np.zeros(n + 1, np.int32)
        ci = np.zeros(nnz, np.int32)
        kv = np.zeros(nnz)
        f = np.zeros(n)
        R = np.zeros(n * k)
        lam = np.zeros(nl, np.int32)
        dof = np.zeros(nl, np.int32)
        sign = np.zeros(nl)
        scale = np.zeros(nl)
        L.FETI4IExportSubdomain(self.inst, i, _ip(sz), _ip(g), _ip(rp), _ip(ci), _dp(kv), _dp(f), _dp(R), _ip(lam),
                                _ip(dof), _dp(sign), _dp(scale))
        _f4check()
        return dict(gdofs=g, rowptr=rp, colidx=ci, vals=kv, f=f, R=R.reshape(k, n).T, lam=lam, dof=dof, sign=sign,
                    scale=scale)

    def solve(self):
        u = np.zeros(self.size)
        f4lib().FETI4ISolve(self.inst, self.size, _dp(u))
        _f4check(solve=True)
        return u

    def stats(self):
        s = np.zeros(6)
        f4lib().FETI4IGetStats(self.inst, _dp(s))
        _f4check()
        return s

    def update_rhs(self, rhs):
        rhs = np.ascontiguousarray(rhs, np.float64)
        f4lib().FETI4IUpdateRhs(self.inst, len(rhs), _dp(rhs))
        _f4check()

    def close(self):
        L = f4lib()
        if getattr(self, "inst", None):
            L.FETI4IDestroy(self.inst)
            self.inst = None
        if getattr(self, "m", None):
            L.FETI4IDestroy(self.m)
            self.m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
