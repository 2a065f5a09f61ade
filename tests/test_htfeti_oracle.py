"""Pins the numpy HTFETI restatement (oracle/htfeti.py) and checks the
product's corner (B0) generation against it.

Pins, each fixed by the mathematics rather than by a retyped formula:
  * corners      -> every B0 row joins one +1 and one -1 copy of the same
                    node and component in two subdomains of one cluster; the
                    glued cluster keeps exactly the rigid-body kernel
  * K~+          -> B0 K~+ x = 0 for every x; K~+ is a generalised inverse of
                    the B0-constrained cluster operator (u - u0 lies in the
                    cluster kernel for x = K u0, B0 u0 = 0)
  * hybrid PCG   -> primal equals the undecomposed direct solve and the
                    golden HYBRID_FETI .emr values
"""
import numpy as np
import pytest

import oracle_lib as O
from espreso_amd import core
from espreso_amd import problems as PB

CUBE = PB.scaled(PB.GRAVITY, clusters=(2, 2, 1), domains=(2, 2, 2), elements=(3, 3, 3), traction_seed=5,
                 traction_scale=1e4, length=(1.0, 1.0, 1.0), young=2.1e11)
SPECS = {"pressure": PB.PRESSURE_PATCH, "heat": PB.HEAT_SIMPLE, "cube": CUBE}


@pytest.fixture(scope="module", params=sorted(SPECS))
def hyb(request):
    spec = SPECS[request.param]
    return spec, O.hybrid(spec, dirichlet=True)


def test_corner_rows(hyb):
    spec, H = hyb
    rows = {}
    for s, (cl, r, d, v) in enumerate(H.B0data):
        for q in range(len(r)):
            node = H.l2g[s][d[q] // H.ndof]
            rows.setdefault((cl, int(r[q])), []).append((s, int(node), int(d[q] % H.ndof), v[q]))
    assert rows, "no corners"
    for (cl, r), ent in rows.items():
        assert len(ent) == 2
        (s1, n1, c1, v1), (s2, n2, c2, v2) = ent
        assert s1 != s2 and H.cluster[s1] == H.cluster[s2] == cl
        assert n1 == n2 and c1 == c2 and v1 * v2 == -1.0


def test_cluster_kernel_dimension(hyb):
    spec, H = hyb
    kfull = 6 if H.ndof == 3 else 1
    for C in H.cl:
        assert C["N"].shape[1] == kfull


def test_product_corners_match(hyb):
    spec, H = hyb
    for s in range(H.nsub):
        e = core.grid_export_cluster(spec, s)
        cl, r, d, v = H.B0data[s]
        assert e["cluster"] == cl
        assert np.array_equal(e["row"], r) and np.array_equal(e["dof"], d) and np.array_equal(e["sign"], v)


def test_kplus_constrained_ginverse(hyb):
    spec, H = hyb
    rng = np.random.default_rng(7)
    o = O.Oracle(spec, preconditioner=0, kplus=0)
    K = [o.dense(s, 0) for s in range(o.n_sub)]
    o.close()
    for cl, subs in enumerate(H.subs):
        C = H.cl[cl]
        B0 = np.hstack([C["B0"][s] for s in subs])
        # arbitrary x: constraints hold
        xs = [rng.standard_normal(H.n[s]) for s in subs]
        us = H.kplus_cluster(cl, xs)
        assert np.abs(B0 @ np.concatenate(us)).max() < 1e-9 * max(np.abs(np.concatenate(us)).max(), 1e-300)
        # x = K u0 with B0 u0 = 0: u - u0 in the cluster kernel
        u0 = rng.standard_normal(B0.shape[1])
        u0 -= B0.T @ np.linalg.solve(B0 @ B0.T, B0 @ u0)
        offs = np.cumsum([0] + [H.n[s] for s in subs])
        xs = [K[s] @ u0[offs[i]:offs[i + 1]] for i, s in enumerate(subs)]
        du = np.concatenate(H.kplus_cluster(cl, xs)) - u0
        Rc = np.vstack([H.Rt[s] for s in subs])
        resid = du - Rc @ np.linalg.lstsq(Rc, du, rcond=None)[0]
        assert np.abs(resid).max() < 1e-8 * np.abs(u0).max()


@pytest.mark.parametrize("prec", [0, 1])
def test_hybrid_solve_vs_global(hyb, prec):
    spec, H = hyb
    u, lam, it = H.solve(precision=1e-10, preconditioner=prec)
    ug = O.global_solve(spec)
    ref = np.abs(ug).max()
    for s in range(H.nsub):
        us = u[s].reshape(-1, H.ndof)
        assert np.abs(us - ug[H.l2g[s]]).max() < 1e-8 * ref, s
    if spec is PB.PRESSURE_PATCH:
        nn = [spec["clusters"][d] * spec["domains"][d] * spec["elements"][d] + 1 for d in range(3)]
        z1 = np.zeros(nn[0] * nn[1] * nn[2])
        for s in range(H.nsub):
            z1[H.l2g[s]] = np.linalg.norm(u[s].reshape(-1, 3), axis=1)
        top = z1.reshape(nn[2], nn[1], nn[0])[-1]
        # golden HEXA8.HYBRID_FETI.emr: Z1 displacement min = max = 1.000000e+00
        assert np.allclose(top, 1.0, atol=1e-7)
    if spec is PB.HEAT_SIMPLE:
        T = np.concatenate([u[s] for s in range(H.nsub)])
        # golden HEXA8.HYBRID_FETI.emr: T in [100, 300]
        assert abs(T.min() - 100) < 1e-6 and abs(T.max() - 300) < 1e-6


# ---- B0 KERNELS (equalityconstraints.cpp:425-502) ----
# Pins: each row holds entries of exactly the two subdomains of one face
# pair of a cluster, on the same nodes (the shared face) with opposite signs;
# B0 u = 0 for every continuous nodal field and the rows are independent
# (the face restriction of a kernel is full rank); the glued cluster keeps
# exactly the rigid-body kernel; the hybrid primal equals the direct solve.
@pytest.fixture(scope="module", params=sorted(SPECS))
def hybk(request):
    spec = SPECS[request.param]
    return spec, O.hybrid(spec, dirichlet=True, b0="kernels")


def test_kernel_rows_structure(hybk):
    spec, H = hybk
    rows = {}
    for s, (cl, r, d, v) in enumerate(H.B0data):
        for q in range(len(r)):
            node = int(H.l2g[s][d[q] // H.ndof])
            rows.setdefault((cl, int(r[q])), {}).setdefault(s, {})[(node, int(d[q] % H.ndof))] = v[q]
    assert rows
    kfull = 6 if H.ndof == 3 else 1
    per_pair = {}
    for (cl, r), by_sub in rows.items():
        assert len(by_sub) == 2
        (s1, e1), (s2, e2) = sorted(by_sub.items())
        assert H.cluster[s1] == H.cluster[s2] == cl
        assert set(e1) == set(e2)  # the same face nodes and components
        assert all(e1[k] == -e2[k] for k in e1)
        per_pair[(s1, s2)] = per_pair.get((s1, s2), 0) + 1
    assert set(per_pair.values()) == {kfull}


def test_kernel_rows_continuity(hybk):
    spec, H = hybk
    rng = np.random.default_rng(3)
    nn = 1 + max(int(L.max()) for L in H.l2g)
    for cl, subs in enumerate(H.subs):
        B0 = np.hstack([H.cl[cl]["B0"][s] for s in subs])
        g = rng.standard_normal(nn * H.ndof).reshape(nn, H.ndof)
        u = np.concatenate([g[H.l2g[s]].ravel() for s in subs])
        assert np.abs(B0 @ u).max() <= 1e-12 * np.abs(B0).max() * np.abs(u).max() * B0.shape[1]
        assert np.linalg.matrix_rank(B0) == B0.shape[0]


def test_kernel_rows_cluster_kernel(hybk):
    spec, H = hybk
    kfull = 6 if H.ndof == 3 else 1
    for C in H.cl:
        assert C["N"].shape[1] == kfull


@pytest.mark.parametrize("prec", [0, 1])
def test_kernel_rows_solve_vs_global(hybk, prec):
    spec, H = hybk
    u, lam, it = H.solve(precision=1e-10, preconditioner=prec)
    ug = O.global_solve(spec)
    ref = np.abs(ug).max()
    for s in range(H.nsub):
        assert np.abs(u[s].reshape(-1, H.ndof) - ug[H.l2g[s]]).max() < 1e-8 * ref, s


def test_product_kernel_rows_match(hybk):
    """The grid builder's kernel rows against the restatement's: the same
    rows (pair by pair, in the same order) on the same DOFs, spanning the
    same space (both builders may take different bases of the kernel)."""
    spec, H = hybk
    kspec = dict(spec, b0="kernels")
    for s in range(H.nsub):
        e = core.grid_export_cluster(kspec, s)
        cl, r, d, v = H.B0data[s]
        assert e["cluster"] == cl
        assert np.array_equal(np.unique(e["row"]), np.unique(r))
        for rows in np.split(np.unique(r), np.unique(r).size // (6 if H.ndof == 3 else 1)):
            P = np.zeros((rows.size, H.n[s]))
            Q = np.zeros((rows.size, H.n[s]))
            for i, row in enumerate(rows):
                mp, mq = e["row"] == row, r == row
                P[i, e["dof"][mp]] = e["sign"][mp]
                Q[i, d[mq]] = v[mq]
            assert np.array_equal(np.nonzero(P.any(axis=0))[0], np.nonzero(Q.any(axis=0))[0])
            assert np.linalg.matrix_rank(np.vstack([P, Q]), tol=1e-10 * np.abs(Q).max()) == rows.size
