"""Hybrid total FETI (HTFETI) restated in numpy -- test infrastructure only.

Only the tests, the smoke check and the benchmark's CPU-baseline leg use this
module; the product never imports it.  It works on the dense per-subdomain
matrices of the C oracle (liboracle.so: K^+ from the fix-point regularised
K_reg, analytic kernels R, loads f, B1 rows, Dirichlet blocks S_D) and
restates the reference's hybrid path:

  corners()        corner nodes   src/mesh/preprocessing/fetidata.cpp:146-187
                   B0 rows        src/assembler/constraints/equalityconstraints.cpp:504-552
  kernel_rows()    face pairs     src/mesh/preprocessing/fetidata.cpp:40-140
                   B0 KERNELS     src/assembler/constraints/equalityconstraints.cpp:425-502
  HybridFETI       G0             src/solver/specific/cluster.cpp:1226-1291
                   F0             cluster.cpp:1293-1478
                   Sa             cluster.cpp:1480-2016 (regularised on the cluster kernel)
                   K~+ x          cluster.cpp:796-916 (multKplusGlobal_Kinv)
                   G per cluster  cluster.cpp:2018-2159
                   F, d, e        itersolvercpu.cpp:13-46, cluster.cpp:2660-2769
                   PCG, primal    src/solver/generic/itersolver.cpp:1406-1724, :372-431

Pinned by tests/test_htfeti_oracle.py against what the mathematics fixes:
K~+ is a generalised inverse of the B0-constrained cluster operator, every
K~+ x satisfies B0 u = 0, the cluster kernel has the rigid-body dimension,
and the hybrid PCG primal equals the undecomposed direct solve and the
reference's golden HYBRID_FETI .emr values.
"""
import numpy as np


def corners(l2g, ndof, dpc):
    """B0 corner entries of every subdomain.

    l2g[s]: global node of each local node of subdomain s (local DOF =
    node * ndof + component); subdomains s // dpc form one cluster.
    Interface nodes of a cluster are grouped by the full set of subdomains
    owning them, in order of first appearance over ascending node ids; a
    group that also touches other clusters gives a corner every size/3 of
    its nodes (fetidata.cpp:158-171), any other group its middle node
    (:172-180).  Each corner glues consecutive pairs of the cluster's
    owners, +1 on the first and -1 on the second, one row per DOF, rows
    numbered per cluster (equalityconstraints.cpp:514-536).
    Returns a list of (cluster, rows, dofs, signs) per subdomain.
    """
    nsub = len(l2g)
    pos = [{int(g): i for i, g in enumerate(L)} for L in l2g]
    owners = {}
    for s, L in enumerate(l2g):
        for g in L:
            owners.setdefault(int(g), []).append(s)
    out = [(s // dpc, [], [], []) for s in range(nsub)]
    ncl = (nsub + dpc - 1) // dpc
    for cl in range(ncl):
        subs = range(cl * dpc, min(nsub, (cl + 1) * dpc))
        nodes = sorted({int(g) for s in subs for g in l2g[s]})
        groups = {}
        for g in nodes:
            own = tuple(sorted(owners[g]))
            if sum(1 for o in own if o // dpc == cl) < 2:
                continue
            groups.setdefault(own, []).append(g)
        row = 0
        for own, mem in groups.items():
            loc = [o for o in own if o // dpc == cl]
            if len(loc) != len(own):
                picks = mem[::max(1, len(mem) // 3)]
            else:
                picks = [mem[len(mem) // 2]]
            for g in picks:
                for a in range(len(loc) - 1):
                    s1, s2 = loc[a], loc[a + 1]
                    for c in range(ndof):
                        for s, sg in ((s1, 1.0), (s2, -1.0)):
                            out[s][1].append(row)
                            out[s][2].append(pos[s][g] * ndof + c)
                            out[s][3].append(sg)
                        row += 1
    return [(c, np.array(r, np.int64), np.array(d, np.int64), np.array(v)) for c, r, d, v in out]


def face_pairs(l2g, dpc, NN):
    """Pairs (d, nd), d < nd, of subdomains of one cluster that share a face,
    with the shared face nodes ascending (fetidata.cpp:40-140: the nodes of
    the element faces of d whose nodes are all shared with nd).  On the
    structured grid (NN nodes per direction) the nodes two boxes share form a
    full rectangle of nodes when they share a face, a line or a point when
    they share only an edge or a corner; a rectangle spans two directions."""
    sets = [set(int(g) for g in L) for L in l2g]
    out = []
    for d in range(len(l2g)):
        for nd in range(d + 1, len(l2g)):
            if d // dpc != nd // dpc:
                continue
            shared = sorted(sets[d] & sets[nd])
            if not shared:
                continue
            g = np.array(shared)
            xyz = np.stack([g % NN[0], (g // NN[0]) % NN[1], g // (NN[0] * NN[1])], axis=1)
            if int((xyz.max(axis=0) > xyz.min(axis=0)).sum()) == 2:
                out.append((d, nd, shared))
    return out


def kernel_rows(l2g, R, ndof, dpc, NN):
    """B0 entries of the KERNELS variant (equalityconstraints.cpp:425-502),
    the reference's default B0 type (src/config/ecf/solver/feti.cpp:90).

    For every face pair (d, nd) of a cluster (face_pairs), the master is the
    subdomain of the pair with more kernel columns (d on a tie); each kernel
    column c of the master gives one row: sum over the face nodes n and
    components k of R_master(n, k, c) (u_d(n, k) - u_nd(n, k)), i.e. +R on d's
    copies and -R on nd's.  Rows are numbered per cluster in the order of
    the pairs (d ascending, then nd).  R[s] is n_s x k_s, local DOF = node *
    ndof + component.  Returns a list of (cluster, rows, dofs, values) per
    subdomain, in the layout of corners()."""
    nsub = len(l2g)
    pos = [{int(g): i for i, g in enumerate(L)} for L in l2g]
    out = [(s // dpc, [], [], []) for s in range(nsub)]
    nrow = {}
    for d, nd, shared in face_pairs(l2g, dpc, NN):
        cl = d // dpc
        master = nd if R[nd].shape[1] > R[d].shape[1] else d
        r0 = nrow.get(cl, 0)
        for c in range(R[master].shape[1]):
            for s, sg in ((d, 1.0), (nd, -1.0)):
                for g in shared:
                    for k in range(ndof):
                        out[s][1].append(r0 + c)
                        out[s][2].append(pos[s][g] * ndof + k)
                        out[s][3].append(sg * R[master][pos[master][g] * ndof + k, c])
        nrow[cl] = r0 + R[master].shape[1]
    return [(c, np.array(r, np.int64), np.array(d, np.int64), np.array(v)) for c, r, d, v in out]


def null_space(A, rtol=1e-10):
    """Orthonormal basis of ker A (SVD)."""
    u, s, vt = np.linalg.svd(A)
    smax = s.max() if s.size else 0.0
    rank = int((s > rtol * smax).sum()) if smax > 0 else 0
    return vt[rank:].T.copy()


class HybridFETI:
    """Dense HTFETI of one decomposed problem (all clusters).

    Kplus[s] (n x n generalised inverse of K_s), R[s] (n x k), f[s];
    B1[s] = (lam, dof, sign, scale); c (lambda vector); B0 = corners();
    SD[s] = (boundary DOFs, S_D) for the Dirichlet preconditioner or None.
    """

    def __init__(self, Kplus, R, f, B1, c, B0, SD=None):
        self.Kplus, self.R, self.f = Kplus, R, f
        self.nsub = len(Kplus)
        self.nlam = len(c)
        self.c = np.asarray(c, float)
        self.SD = SD
        n = [K.shape[0] for K in Kplus]
        # B1 rows of each subdomain: (lambda, DOF, value) triples, values +-1 and +-scale
        self.n = n
        self.B1 = [(np.asarray(B1[s][0], np.int64), np.asarray(B1[s][1], np.int64), np.asarray(B1[s][2], float),
                    np.asarray(B1[s][2], float) * np.asarray(B1[s][3], float)) for s in range(self.nsub)]
        # clusters
        self.cluster = [int(B0[s][0]) for s in range(self.nsub)]
        ncl = max(self.cluster) + 1
        self.subs = [[s for s in range(self.nsub) if self.cluster[s] == cl] for cl in range(ncl)]
        self.cl = []
        for cl, subs in enumerate(self.subs):
            n0 = 1 + max((int(B0[s][1].max()) for s in subs if B0[s][1].size), default=-1)
            B0s = {}
            for s in subs:
                M = np.zeros((n0, n[s]))
                M[B0[s][1], B0[s][2]] = B0[s][3]
                B0s[s] = M
            # F0 = sum B0 K+ B0^T (cluster.cpp:1293-1478), G0 = -(B0 R)^T stacked (:1226-1291)
            F0 = sum(B0s[s] @ Kplus[s] @ B0s[s].T for s in subs)
            G0 = np.vstack([-(B0s[s] @ R[s]).T for s in subs])
            F0inv = np.linalg.inv(F0)
            Sa = G0 @ F0inv @ G0.T
            N = null_space(G0.T)
            rho = max(np.abs(np.diag(Sa)).max(), 1.0)
            Sa_reg = Sa + rho * N @ N.T
            kofs = np.cumsum([0] + [R[s].shape[1] for s in subs])
            self.cl.append(dict(B0=B0s, F0inv=F0inv, G0=G0, Sainv=np.linalg.inv(Sa_reg), N=N, kofs=kofs))
        # coarse space: cluster kernels R~_s = R_s N_s (cluster.cpp:2018-2159)
        self.Rt = [None] * self.nsub
        self.coff = np.cumsum([0] + [C["N"].shape[1] for C in self.cl])
        self.ncoarse = int(self.coff[-1])
        for cl, subs in enumerate(self.subs):
            C = self.cl[cl]
            for i, s in enumerate(subs):
                self.Rt[s] = R[s] @ C["N"][C["kofs"][i]:C["kofs"][i + 1]]
        self.G = np.zeros((self.ncoarse, self.nlam))
        for cl, subs in enumerate(self.subs):
            for s in subs:
                lam, dof, sg, _ = self.B1[s]
                for b in range(self.Rt[s].shape[1]):
                    self.G[self.coff[cl] + b] -= np.bincount(lam, sg * self.Rt[s][dof, b], minlength=self.nlam)
        self.GGt = self.G @ self.G.T
        self.e = np.zeros(self.ncoarse)
        for cl, subs in enumerate(self.subs):
            for s in subs:
                self.e[self.coff[cl]:self.coff[cl + 1]] -= self.Rt[s].T @ f[s]

    # ---- cluster.cpp:796-916 ----
    def kplus_cluster(self, cl, xs):
        """K~+ on the subdomains of cluster cl: list of x_s -> list of u_s."""
        C = self.cl[cl]
        subs = self.subs[cl]
        g0 = sum(C["B0"][s] @ (self.Kplus[s] @ xs[i]) for i, s in enumerate(subs))
        e0 = np.concatenate([-self.R[s].T @ xs[i] for i, s in enumerate(subs)])
        t = C["F0inv"] @ g0
        alpha = C["Sainv"] @ (C["G0"] @ t - e0)
        lam0 = C["F0inv"] @ (g0 - C["G0"].T @ alpha)
        out = []
        for i, s in enumerate(subs):
            a = alpha[C["kofs"][i]:C["kofs"][i + 1]]
            out.append(self.Kplus[s] @ (xs[i] - C["B0"][s].T @ lam0) + self.R[s] @ a)
        return out

    def kplus(self, xs):
        u = [None] * self.nsub
        for cl, subs in enumerate(self.subs):
            for s, v in zip(subs, self.kplus_cluster(cl, [xs[s] for s in subs])):
                u[s] = v
        return u

    def Bt_apply(self, s, lam, scaled=False):
        """B_s^T lam (or the scaled B~_s^T lam)."""
        l, dof, sg, ss = self.B1[s]
        return np.bincount(dof, (ss if scaled else sg) * lam[l], minlength=self.n[s])

    def B_apply(self, s, u, scaled=False):
        l, dof, sg, ss = self.B1[s]
        return np.bincount(l, (ss if scaled else sg) * u[dof], minlength=self.nlam)

    def apply_F(self, lam):
        u = self.kplus([self.Bt_apply(s, lam) for s in range(self.nsub)])
        return sum(self.B_apply(s, u[s]) for s in range(self.nsub))

    def vec_d(self):
        u = self.kplus(self.f)
        return sum(self.B_apply(s, u[s]) for s in range(self.nsub)) - self.c

    def project(self, x, mode=0):
        if mode == 1:
            d = np.linalg.solve(self.GGt, x)
            return self.G.T @ d
        d = np.linalg.solve(self.GGt, self.G @ x)
        if mode == 2:
            return d
        return x - self.G.T @ d

    def apply_prec(self, x):
        y = np.zeros_like(x)
        for s in range(self.nsub):
            bd, S = self.SD[s]
            v = self.Bt_apply(s, x, scaled=True)[bd]
            w = np.zeros(self.n[s])
            w[bd] = S @ v
            y += self.B_apply(s, w, scaled=True)
        return y

    # ---- itersolver.cpp:1406-1724 (Solve_RegCG) and :372-431 (primal) ----
    def solve(self, precision=1e-8, max_iterations=1000, preconditioner=0, tol=None):
        """Projected PCG; `tol` overrides the stopping threshold on ||P r||."""
        b = self.vec_d()
        x = self.project(self.e, 1)
        r = b - self.apply_F(x)
        u = self.project(r)
        if tol is None:
            tol = precision * min(np.linalg.norm(u), np.linalg.norm(b))
        w = np.zeros_like(b)
        y = np.zeros_like(b)
        p = None
        it = 0
        for it in range(1, max_iterations + 1):
            wp, yp = w, y
            w = self.project(r)
            y = w if preconditioner == 0 else self.project(self.apply_prec(w))
            p = y.copy() if it == 1 else y + (y @ w) / (yp @ wp) * p
            Ap = self.apply_F(p)
            a = (y @ w) / (p @ Ap)
            x = x + a * p
            r = r - a * Ap
            if np.linalg.norm(w) < tol:
                break
        amp = self.project(r, 2)
        u = self.kplus([self.f[s] - self.Bt_apply(s, x) for s in range(self.nsub)])
        for cl, subs in enumerate(self.subs):
            for s in subs:
                u[s] = u[s] + self.Rt[s] @ amp[self.coff[cl]:self.coff[cl + 1]]
        return u, x, it
