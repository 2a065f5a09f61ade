"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

Test infrastructure: the checker for the HIP product.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")


class OracleProblem(C.Structure):
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
    ]


class OracleOptions(C.Structure):
    _fields_ = [
        ("method", C.c_int32),
        ("preconditioner", C.c_int32),
        ("max_iterations", C.c_int32),
        ("scaling", C.c_int32),
        ("kplus", C.c_int32),
        ("threads", C.c_int32),
        ("precision", C.c_double),
    ]


class OracleStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("n_lambda", C.c_int32),
        ("n_coarse", C.c_int32),
        ("n_subdomains", C.c_int32),
        ("final_norm", C.c_double),
        ("tol", C.c_double),
        ("norm_b", C.c_double),
        ("kkt_primal", C.c_double),
        ("kkt_dual", C.c_double),
        ("time_setup", C.c_double),
        ("time_solve", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        _lib = C.CDLL(LIB_PATH)
        d = C.POINTER(C.c_double)
        i32 = C.POINTER(C.c_int32)
        _lib.oracle_create.restype = C.c_void_p
        _lib.oracle_create.argtypes = [C.POINTER(OracleProblem), C.POINTER(OracleOptions)]
        _lib.oracle_destroy.argtypes = [C.c_void_p]
        _lib.oracle_sizes.argtypes = [C.c_void_p, i32]
        _lib.oracle_subdomain_info.argtypes = [C.c_void_p, C.c_int32, i32]
        _lib.oracle_subdomain_boundary.argtypes = [C.c_void_p, C.c_int32, i32]
        _lib.oracle_subdomain_lambdas.argtypes = [C.c_void_p, C.c_int32, i32, i32, d, d]
        _lib.oracle_subdomain_dense.argtypes = [C.c_void_p, C.c_int32, C.c_int32, d]
        _lib.oracle_subdomain_rhs.argtypes = [C.c_void_p, C.c_int32, d]
        _lib.oracle_subdomain_l2g.argtypes = [C.c_void_p, C.c_int32, i32]
        _lib.oracle_subdomain_primal.argtypes = [C.c_void_p, C.c_int32, d]
        for name in ("oracle_vec_c", "oracle_vec_b", "oracle_vec_e"):
            getattr(_lib, name).argtypes = [C.c_void_p, d]
        for name in ("oracle_apply_F", "oracle_apply_prec", "oracle_project", "oracle_apply_G"):
            getattr(_lib, name).argtypes = [C.c_void_p, d, d]
        _lib.oracle_solve.argtypes = [C.c_void_p, d, d, C.POINTER(OracleStats)]
        _lib.oracle_set_nodal_load.argtypes = [C.c_void_p, d]
        _lib.oracle_global_solve.argtypes = [C.POINTER(OracleProblem), d]
        _lib.oracle_unit_element.argtypes = [C.POINTER(OracleProblem), d]
        _lib.oracle_global_solve_load.argtypes = [C.POINTER(OracleProblem), d, d]
        _lib.oracle_global_sizes.argtypes = [C.POINTER(OracleProblem), i32, i32]
        _lib.oracle_global_residual.argtypes = [C.POINTER(OracleProblem), d, d]
        _lib.oracle_global_residual_vector.argtypes = [C.POINTER(OracleProblem), d, d]
        _lib.oracle_seeded_conductivity.argtypes = [C.c_uint64, C.c_int32, C.c_double, d]
        _lib.oracle_seeded_conductivity.restype = None
        _lib.oracle_cpu_sample.argtypes = [C.POINTER(OracleProblem), C.POINTER(OracleOptions), C.c_int32, d]
        _lib.oracle_cpu_baseline.argtypes = [C.POINTER(OracleProblem), C.POINTER(OracleOptions), C.c_int32, d]
        _lib.oracle_last_error.restype = C.c_char_p
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def make_problem(spec):
    """Build an OracleProblem from a dict (see espreso_amd.problems)."""
    p = OracleProblem()
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
    p.fixed[:] = spec.get("fixed", (0, 0, 0, 0, 0, 0))
    p.fixed_value[:] = spec.get("fixed_value", (0.0,) * 6)
    p.pressure[:] = spec.get("pressure", (0.0,) * 6)
    p.traction_seed = spec.get("traction_seed", 0)
    p.traction_scale = spec.get("traction_scale", 1.0)
    p.young_seed = spec.get("young_seed", 0)
    p.conductivity_seed = spec.get("conductivity_seed", 0)
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


def unit_element(spec):
    """Unit-modulus HEXA8 element matrix (Gauss quadrature), row-major."""
    p = make_problem(spec)
    nd = 3 if spec.get("physics", 0) == 0 else 1
    Ke = np.zeros((8 * nd) ** 2)
    lib().oracle_unit_element(C.byref(p), _dp(Ke))
    return Ke.reshape(8 * nd, 8 * nd)


def element_dofs(spec):
    """Global DOF indices (n_elements x 8*ndof) of every element, element index
    ex + NEx*(ey + NEy*ez), corners in the reference's HEXA8 order."""
    NE = [spec["clusters"][d] * spec["domains"][d] * spec["elements"][d] for d in range(3)]
    NN = [n + 1 for n in NE]
    nd = 3 if spec.get("physics", 0) == 0 else 1
    corner = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
    ez, ey, ex = np.meshgrid(np.arange(NE[2]), np.arange(NE[1]), np.arange(NE[0]), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    nodes = np.stack([(ex + c[0]) + NN[0] * ((ey + c[1]) + NN[1] * (ez + c[2])) for c in corner], axis=1)
    return (nodes[:, :, None] * nd + np.arange(nd)[None, None, :]).reshape(len(ex), 8 * nd)


def make_options(preconditioner=3, precision=1e-8, max_iterations=500, scaling=1, kplus=0, threads=0):
    o = OracleOptions()
    o.method = 0
    o.preconditioner = preconditioner
    o.max_iterations = max_iterations
    o.scaling = scaling
    o.kplus = kplus
    o.threads = threads
    o.precision = precision
    return o


def global_solve(spec):
    p = make_problem(spec)
    nn = C.c_int32()
    nd = C.c_int32()
    lib().oracle_global_sizes(C.byref(p), C.byref(nn), C.byref(nd))
    u = np.zeros(nn.value * nd.value)
    rc = lib().oracle_global_solve(C.byref(p), _dp(u))
    if rc:
        raise RuntimeError(lib().oracle_last_error().decode())
    return u.reshape(nn.value, nd.value)


def global_solve_load(spec, f_nodal):
    """Undecomposed direct solve with a nodal load (n_nodes x ndof)."""
    p = make_problem(spec)
    nn = C.c_int32()
    nd = C.c_int32()
    lib().oracle_global_sizes(C.byref(p), C.byref(nn), C.byref(nd))
    f = np.ascontiguousarray(f_nodal, dtype=np.float64).reshape(-1)
    assert f.size == nn.value * nd.value
    u = np.zeros(nn.value * nd.value)
    rc = lib().oracle_global_solve_load(C.byref(p), _dp(f), _dp(u))
    if rc:
        raise RuntimeError(lib().oracle_last_error().decode())
    return u.reshape(nn.value, nd.value)


def global_residual(spec, u):
    p = make_problem(spec)
    u = np.ascontiguousarray(u, dtype=np.float64).ravel()
    r = C.c_double()
    lib().oracle_global_residual(C.byref(p), _dp(u), C.byref(r))
    return r.value


def undecomposed(spec):
    """The spec as one subdomain over the same global mesh, for the undecomposed
    oracle solve and residual.  A seeded per-subdomain conductivity becomes an
    explicit element field first, so that uneven subdomains keep their values."""
    from espreso_amd import core
    from espreso_amd import problems as PB
    s = dict(spec)
    if s.get("physics") == 1 and s.get("conductivity_seed") and s.get("material_elem") is None:
        cl, dom = s["clusters"], s["domains"]
        ns = [cl[d] * dom[d] for d in range(3)]
        ks = np.zeros(ns[0] * ns[1] * ns[2])
        lib().oracle_seeded_conductivity(s["conductivity_seed"], ks.size, s.get("conductivity", 1.0), _dp(ks))
        off = [np.concatenate([[0], np.cumsum(v)]) for v in core.slabs(s)]
        ez, ey, ex = np.meshgrid(*(np.arange(off[d][-1]) for d in (2, 1, 0)), indexing="ij")
        S = [np.searchsorted(off[d], e.ravel(), side="right") - 1 for d, e in ((0, ex), (1, ey), (2, ez))]
        # subdomain id: cluster-major, x fastest within the cluster and across clusters
        c = [S[d] // dom[d] for d in range(3)]
        q = [S[d] % dom[d] for d in range(3)]
        sid = (c[0] + cl[0] * (c[1] + cl[1] * c[2])) * (dom[0] * dom[1] * dom[2]) + q[0] + dom[0] * (q[1] + dom[1] * q[2])
        s["material_elem"] = ks[sid]
        s.pop("conductivity_seed")
    return PB.uniform_equivalent(s)


def global_residual_vector(spec, u):
    """K u - f per nodal DOF (0 on fixed DOFs), K and f of the undecomposed problem."""
    p = make_problem(spec)
    u = np.ascontiguousarray(u, dtype=np.float64)
    r = np.zeros(u.size)
    lib().oracle_global_residual_vector(C.byref(p), _dp(u.ravel()), _dp(r))
    return r.reshape(u.shape)


class Oracle:
    """Decomposed TFETI restatement (one handle = one assembled problem)."""

    def __init__(self, spec, **opts):
        self.spec = spec
        self._p = make_problem(spec)
        self._o = make_options(**opts)
        self.h = lib().oracle_create(C.byref(self._p), C.byref(self._o))
        if not self.h:
            raise RuntimeError(lib().oracle_last_error().decode())
        s = np.zeros(6, dtype=np.int32)
        lib().oracle_sizes(self.h, _ip(s))
        self.n_sub, self.n_lambda, self.n_coarse, self.n_nodes, self.ndof, self.kdim = (int(x) for x in s)

    def close(self):
        if self.h:
            lib().oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def info(self, i):
        v = np.zeros(3, dtype=np.int32)
        lib().oracle_subdomain_info(self.h, i, _ip(v))
        return dict(n=int(v[0]), m=int(v[1]), nlam=int(v[2]))

    def boundary(self, i):
        b = np.zeros(self.info(i)["m"], dtype=np.int32)
        lib().oracle_subdomain_boundary(self.h, i, _ip(b))
        return b

    def lambdas(self, i):
        k = self.info(i)["nlam"]
        lam = np.zeros(k, dtype=np.int32)
        dof = np.zeros(k, dtype=np.int32)
        sign = np.zeros(k)
        scale = np.zeros(k)
        lib().oracle_subdomain_lambdas(self.h, i, _ip(lam), _ip(dof), _dp(sign), _dp(scale))
        return lam, dof, sign, scale

    def dense(self, i, which):
        inf = self.info(i)
        n, m = inf["n"], inf["m"]
        shape = {0: (n, n), 1: (n, n), 2: (self.kdim, n), 3: (m, m), 4: (m, m), 5: (m, m)}[which]
        out = np.zeros(shape[0] * shape[1])
        lib().oracle_subdomain_dense(self.h, i, which, _dp(out))
        return out.reshape(shape).T if which != 2 else out.reshape(shape).T

    def rhs(self, i):
        f = np.zeros(self.info(i)["n"])
        lib().oracle_subdomain_rhs(self.h, i, _dp(f))
        return f

    def l2g(self, i):
        g = np.zeros(self.info(i)["n"] // self.ndof, dtype=np.int32)
        lib().oracle_subdomain_l2g(self.h, i, _ip(g))
        return g

    def _vec(self, name, n):
        v = np.zeros(n)
        getattr(lib(), name)(self.h, _dp(v))
        return v

    def vec_c(self):
        return self._vec("oracle_vec_c", self.n_lambda)

    def vec_b(self):
        return self._vec("oracle_vec_b", self.n_lambda)

    def vec_e(self):
        return self._vec("oracle_vec_e", self.n_coarse)

    def _op(self, name, x, n_out):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(n_out)
        getattr(lib(), name)(self.h, _dp(x), _dp(y))
        return y

    def apply_F(self, x):
        return self._op("oracle_apply_F", x, self.n_lambda)

    def apply_prec(self, x):
        return self._op("oracle_apply_prec", x, self.n_lambda)

    def project(self, x):
        return self._op("oracle_project", x, self.n_lambda)

    def apply_G(self, x):
        return self._op("oracle_apply_G", x, self.n_coarse)

    def solve(self):
        u = np.zeros(self.n_nodes * self.ndof)
        lam = np.zeros(self.n_lambda)
        st = OracleStats()
        rc = lib().oracle_solve(self.h, _dp(u), _dp(lam), C.byref(st))
        if rc:
            raise RuntimeError(lib().oracle_last_error().decode())
        stats = {k: getattr(st, k) for k, _ in OracleStats._fields_}
        return u.reshape(self.n_nodes, self.ndof), lam, stats

    def primal(self, i):
        u = np.zeros(self.info(i)["n"])
        lib().oracle_subdomain_primal(self.h, i, _dp(u))
        return u

    def set_nodal_load(self, f_nodal):
        """Replace the loads by a nodal load (n_nodes x ndof); the next solve()
        reuses the factorisations."""
        f = np.ascontiguousarray(f_nodal, dtype=np.float64).reshape(-1)
        assert f.size == self.n_nodes * self.ndof
        lib().oracle_set_nodal_load(self.h, _dp(f))


def cpu_sample(spec, sample, threads=0, preconditioner=3):
    p = make_problem(spec)
    o = make_options(preconditioner=preconditioner, threads=threads)
    t = np.zeros(3)
    rc = lib().oracle_cpu_sample(C.byref(p), C.byref(o), sample, _dp(t))
    if rc:
        raise RuntimeError(lib().oracle_last_error().decode())
    return dict(setup_per_sub=t[0], kplus_per_sub=t[1], sd_apply_per_sub=t[2])


def cpu_baseline(spec, sample, threads=0):
    """Explicit-S_i CPU restatement timed on a sample of subdomains plus the
    full-size coarse set-up and per-iteration vector work (seconds)."""
    p = make_problem(spec)
    o = make_options(preconditioner=3, threads=threads, kplus=1)
    t = np.zeros(6)
    rc = lib().oracle_cpu_baseline(C.byref(p), C.byref(o), sample, _dp(t))
    if rc:
        raise RuntimeError(lib().oracle_last_error().decode())
    return dict(setup_per_sub=t[0], f_apply_per_sub=t[1], sd_apply_per_sub=t[2], coarse_setup=t[3],
                iteration_rest=t[4], sample=int(t[5]))


def hybrid(spec, dirichlet=False, b0="corners"):
    """Numpy HTFETI restatement (oracle/htfeti.py) fed by the C oracle's
    decomposition: K^+ = K_reg^-1 (fix-point regularisation), analytic R, f,
    B1, c, and B0 from the subdomain node maps: corner rows (b0 "corners") or
    the kernel rows of the face pairs ("kernels")."""
    from oracle import htfeti
    o = Oracle(spec, preconditioner=3 if dirichlet else 0, kplus=0)
    try:
        Kp, R, f, B1, l2g, SD = [], [], [], [], [], []
        for s in range(o.n_sub):
            Kp.append(np.linalg.inv(o.dense(s, 1)))
            R.append(np.ascontiguousarray(o.dense(s, 2)))
            f.append(o.rhs(s))
            B1.append(o.lambdas(s))
            l2g.append(o.l2g(s))
            if dirichlet:
                SD.append((o.boundary(s), o.dense(s, 3)))
        c = o.vec_c()
        ndof, nsub = o.ndof, o.n_sub
    finally:
        o.close()
    dpc = spec["domains"][0] * spec["domains"][1] * spec["domains"][2]
    if b0 == "kernels":
        from espreso_amd.core import slabs
        NN = [sum(v) + 1 for v in slabs(spec)]
        B0 = htfeti.kernel_rows(l2g, R, ndof, dpc, NN)
    else:
        B0 = htfeti.corners(l2g, ndof, dpc)
    H = htfeti.HybridFETI(Kp, R, f, B1, c, B0, SD if dirichlet else None)
    H.l2g = l2g
    H.ndof = ndof
    H.B0data = B0
    return H


def compliance_gradient(spec):
    """dc/dE_e = -u_e^T K_e0 u_e of the compliance c = f^T u from the
    undecomposed direct solve and the Gauss-quadrature element matrix."""
    u = global_solve(spec).reshape(-1)
    Ke0 = unit_element(spec)
    ue = u[element_dofs(spec)]
    return -np.einsum("ei,ij,ej->e", ue, Ke0, ue), u
