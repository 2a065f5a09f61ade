// FETI4I front end (include/feti4i.h) over the MI355X FETI core.
//
// Takes what the reference's library interface takes -- element matrices,
// a right-hand side, local-to-global DOF numbers and Dirichlet conditions
// (reference: src/app/apiwrapper/wrapper.cpp:126-262, the intended
// input::API::load of src/input/api and Precomputed physics of
// src/assembler/physics/precomputed.cpp:26-93) -- and turns it into the
// subdomain data of espreso_feti_set_subdomain:
//   1. elements are split into FETI4I_SUBDOMAINS parts by recursive
//      bisection of the element graph (BFS level order from a
//      pseudo-peripheral element);
//   2. each subdomain assembles K_i from its elements; f_i is the
//      right-hand side divided by the number of subdomains sharing the DOF
//      (precomputed.cpp:62-66);
//   3. B1: one Dirichlet row per subdomain copy of a Dirichlet DOF, redundant
//      gluing of shared DOFs with K-diagonal scaling
//      (equalityconstraints.cpp:25-423);
//   4. the kernel R_i is detected algebraically (the interface forces
//      ALGEBRAIC regularisation, wrapper.cpp:118): Cholesky of K_i in
//      reverse Cuthill-McKee order, pivots below 1e-8 of their diagonal are
//      fixed, and each fixed DOF p gives the kernel vector [-K_rr^-1 K_rp; e_p];
//   5. FETI4ISolve averages the subdomain copies of each DOF
//      (precomputed.cpp:82-93).
// The decomposition runs on the host at FETI4ICreateInstance; the GPU core is
// created and set up at the first FETI4ISolve.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/espreso_amd.h"
#include "../../include/feti4i.h"

namespace {

thread_local int g_status = 0;
thread_local std::string g_error;
std::mutex g_mu;
std::set<void *> g_matrices, g_instances;

struct F4Error : std::runtime_error {
    int code;
    F4Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define F4_CHECK(cond, code, msg)                \
    do {                                         \
        if (!(cond)) throw F4Error((code), (msg)); \
    } while (0)

struct Subdomain {
    std::vector<int> gdofs;  // instance DOFs, ascending
    std::vector<int> rowptr, colidx;
    std::vector<double> vals, f, R;
    int kdim = 0;
    std::vector<int> lam, dof;
    std::vector<double> sign, scale;
    std::vector<int> b0row, b0dof;  // HYBRID_FETI corners (cluster-local row, local DOF)
    std::vector<double> b0sign;
};

}  // namespace

struct FETI4IStructMatrix {
    int type = 0, base = 0;
    std::vector<int64_t> nptr{0}, dptr{0}, vptr{0};
    std::vector<int> nodes, dofs;
    std::vector<double> vals;
    int nelem() const { return (int)dptr.size() - 1; }
};

struct FETI4IStructInstance {
    FETI4IStructMatrix K;
    int size = 0;
    std::vector<double> rhs;
    std::vector<int> l2g;
    std::vector<int> dir_idx;
    std::vector<double> dir_val;
    int nparts = 4, max_it = 200, method = 0, prec = 3, cg = 0, verbose = 1;
    double precision = 1e-5;
    bool built = false;
    std::vector<Subdomain> sub;
    int nlam = 0, ndof_node = 1, max_kdim = 0, n0 = 0;
    std::vector<double> c;
    std::vector<std::vector<int>> elems;  // elements of every subdomain (decomposition)
    espreso_feti *h = nullptr;
    bool gpu_ready = false;
    bool needs_setup = false;             // K values updated on a set-up core
    espreso_feti_stats st{};
    int ncoarse = 0;
    ~FETI4IStructInstance() {
        if (h) espreso_feti_destroy(h);
    }
};

namespace {

template <class F>
void guarded(F &&fn) {
    try {
        fn();
        g_status = 0;
        g_error.clear();
    } catch (const F4Error &e) {
        g_status = e.code;
        g_error = e.what();
    } catch (const std::exception &e) {
        g_status = ESPRESO_ERR_ARG;
        g_error = e.what();
    }
}

// ---------------------------------------------------------- partitioning ----
// Element graph: two elements are neighbours when they share a face -- at
// least 3 nodes (at least 2 for elements of up to 4 nodes).  Elements given
// without nodes use their DOFs instead.  Face adjacency keeps subdomains
// free of hinges and point contacts, which would add mechanisms to K_i.
struct ElementGraph {
    std::vector<int64_t> ptr;
    std::vector<int> adj;
    template <class V>
    void neighbours(int e, V &&visit) const {
        for (int64_t q = ptr[e]; q < ptr[e + 1]; q++) visit(adj[q]);
    }
};

void element_graph(const FETI4IStructMatrix &K, int ndofs, ElementGraph &G) {
    const int ne = K.nelem();
    const bool with_nodes = K.nptr.back() > 0;
    auto items = [&](int e, const int *&p, int &k) {
        if (with_nodes) { p = &K.nodes[K.nptr[e]]; k = (int)(K.nptr[e + 1] - K.nptr[e]); }
        else { p = &K.dofs[K.dptr[e]]; k = (int)(K.dptr[e + 1] - K.dptr[e]); }
    };
    int nitem = with_nodes ? 0 : ndofs;
    if (with_nodes) for (int v : K.nodes) nitem = std::max(nitem, v + 1);
    std::vector<int64_t> ip(nitem + 1, 0);
    for (int e = 0; e < ne; e++) {
        const int *p; int k;
        items(e, p, k);
        for (int a = 0; a < k; a++) ip[p[a] + 1]++;
    }
    for (int v = 0; v < nitem; v++) ip[v + 1] += ip[v];
    std::vector<int> ie(ip[nitem]);
    {
        std::vector<int64_t> fill(ip.begin(), ip.end() - 1);
        for (int e = 0; e < ne; e++) {
            const int *p; int k;
            items(e, p, k);
            for (int a = 0; a < k; a++) ie[fill[p[a]]++] = e;
        }
    }
    G.ptr.assign(ne + 1, 0);
    G.adj.clear();
    std::vector<int> count(ne, 0), touched;
    for (int e = 0; e < ne; e++) {
        const int *p; int k;
        items(e, p, k);
        const int thr = with_nodes ? (k <= 4 ? 2 : 3) : 3;
        touched.clear();
        for (int a = 0; a < k; a++)
            for (int64_t q = ip[p[a]]; q < ip[p[a] + 1]; q++) {
                const int e2 = ie[q];
                if (e2 == e) continue;
                if (count[e2]++ == 0) touched.push_back(e2);
            }
        for (int e2 : touched) {
            if (count[e2] >= std::min(thr, k)) G.adj.push_back(e2);
            count[e2] = 0;
        }
        G.ptr[e + 1] = (int64_t)G.adj.size();
    }
}

struct Bisector {
    const ElementGraph &G;
    std::vector<int> &part;
    std::vector<int> inset, seen;
    int stamp = 0, seen_stamp = 0;
    Bisector(const ElementGraph &g, std::vector<int> &p, int ne) : G(g), part(p), inset(ne, -1), seen(ne, -1) {}

    // BFS order of `elems` (restricted to the set), restarting on disconnected pieces
    void bfs(const std::vector<int> &elems, int start, std::vector<int> &order) {
        seen_stamp++;
        order.clear();
        std::vector<int> queue;
        auto run = [&](int s) {
            size_t head = queue.size();
            queue.push_back(s);
            seen[s] = seen_stamp;
            while (head < queue.size()) {
                const int e = queue[head++];
                order.push_back(e);
                G.neighbours(e, [&](int nb) {
                    if (inset[nb] == stamp && seen[nb] != seen_stamp) {
                        seen[nb] = seen_stamp;
                        queue.push_back(nb);
                    }
                });
            }
        };
        run(start);
        for (int e : elems)
            if (seen[e] != seen_stamp) run(e);
    }

    void split(std::vector<int> elems, int parts, int first) {
        if (parts <= 1 || elems.size() <= 1) {
            for (int e : elems) part[e] = first;
            return;
        }
        stamp++;
        for (int e : elems) inset[e] = stamp;
        std::vector<int> order;
        bfs(elems, elems[0], order);
        bfs(elems, order.back(), order);  // from a pseudo-peripheral element
        const int pl = parts / 2;
        const size_t nl = (size_t)std::llround((double)elems.size() * pl / parts);
        std::vector<int> left(order.begin(), order.begin() + nl), right(order.begin() + nl, order.end());
        split(std::move(left), pl, first);
        split(std::move(right), parts - pl, first + pl);
    }
};

// Move every piece of a part that is not its largest face-connected component
// to the neighbouring part it shares most faces with.
void merge_stray_pieces(const ElementGraph &G, std::vector<int> &part, int nparts) {
    const int ne = (int)part.size();
    for (int pass = 0; pass < 4; pass++) {
        std::vector<int> comp(ne, -1), csize, cpart;
        for (int e = 0; e < ne; e++) {
            if (comp[e] >= 0) continue;
            const int id = (int)csize.size();
            csize.push_back(0);
            cpart.push_back(part[e]);
            std::vector<int> q{e};
            comp[e] = id;
            for (size_t h = 0; h < q.size(); h++) {
                csize[id]++;
                G.neighbours(q[h], [&](int nb) {
                    if (comp[nb] < 0 && part[nb] == part[q[h]]) { comp[nb] = id; q.push_back(nb); }
                });
            }
        }
        std::vector<int> best(nparts, -1);
        for (size_t c = 0; c < csize.size(); c++)
            if (best[cpart[c]] < 0 || csize[c] > csize[best[cpart[c]]]) best[cpart[c]] = (int)c;
        bool moved = false;
        for (size_t c = 0; c < csize.size(); c++) {
            if ((int)c == best[cpart[c]]) continue;
            std::vector<int> votes(nparts, 0);
            for (int e = 0; e < ne; e++)
                if (comp[e] == (int)c) G.neighbours(e, [&](int nb) { if (part[nb] != cpart[c]) votes[part[nb]]++; });
            const int to = (int)(std::max_element(votes.begin(), votes.end()) - votes.begin());
            if (votes[to] == 0) continue;  // isolated piece: leave it (reported later)
            for (int e = 0; e < ne; e++)
                if (comp[e] == (int)c) part[e] = to;
            moved = true;
        }
        if (!moved) break;
    }
}

// -------------------------------------------------------- algebraic kernel ----
// reverse Cuthill-McKee order (new -> old) of a symmetric CSR graph
std::vector<int> rcm(int n, const std::vector<int> &rp, const std::vector<int> &ci) {
    std::vector<int> deg(n), order, seen(n, 0);
    for (int i = 0; i < n; i++) deg[i] = rp[i + 1] - rp[i];
    order.reserve(n);
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return deg[a] < deg[b]; });
    std::vector<int> nb;
    for (int s0 : idx) {
        if (seen[s0]) continue;
        // pseudo-peripheral start: last vertex of a BFS from s0
        int s = s0;
        for (int pass = 0; pass < 2; pass++) {
            std::vector<int> q{s}, mark{s};
            std::vector<char> in(n, 0);
            in[s] = 1;
            for (size_t h = 0; h < q.size(); h++)
                for (int k = rp[q[h]]; k < rp[q[h] + 1]; k++)
                    if (!in[ci[k]] && !seen[ci[k]]) { in[ci[k]] = 1; q.push_back(ci[k]); }
            s = q.back();
        }
        size_t head = order.size();
        order.push_back(s);
        seen[s] = 1;
        while (head < order.size()) {
            const int v = order[head++];
            nb.clear();
            for (int k = rp[v]; k < rp[v + 1]; k++)
                if (!seen[ci[k]]) { seen[ci[k]] = 1; nb.push_back(ci[k]); }
            std::sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] < deg[b]; });
            order.insert(order.end(), nb.begin(), nb.end());
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

// kernel basis of a symmetric positive semidefinite CSR matrix (orthonormal columns)
void algebraic_kernel(Subdomain &s) {
    const int n = (int)s.gdofs.size();
    std::vector<int> perm = rcm(n, s.rowptr, s.colidx), ip(n);
    for (int i = 0; i < n; i++) ip[perm[i]] = i;
    std::vector<int> first(n);
    for (int i = 0; i < n; i++) {
        const int o = perm[i];
        int f = i;
        for (int k = s.rowptr[o]; k < s.rowptr[o + 1]; k++) f = std::min(f, ip[s.colidx[k]]);
        first[i] = f;
    }
    std::vector<int64_t> ptr(n + 1, 0);
    for (int i = 0; i < n; i++) ptr[i + 1] = ptr[i] + (i - first[i] + 1);
    std::vector<double> L(ptr[n], 0.0), diag(n, 0.0);
    for (int i = 0; i < n; i++) {
        const int o = perm[i];
        for (int k = s.rowptr[o]; k < s.rowptr[o + 1]; k++) {
            const int j = ip[s.colidx[k]];
            if (j <= i) L[ptr[i] + j - first[i]] += s.vals[k];
            if (j == i) diag[i] = s.vals[k];
        }
    }
    std::vector<char> fixed(n, 0);
    for (int i = 0; i < n; i++) {
        double *Li = &L[ptr[i]];
        const int fi = first[i];
        for (int j = fi; j < i; j++) {
            if (fixed[j]) { Li[j - fi] = 0.0; continue; }
            const double *Lj = &L[ptr[j]];
            const int fj = first[j], k0 = std::max(fi, fj);
            double acc = 0;
            for (int k = k0; k < j; k++) acc += Li[k - fi] * Lj[k - fj];
            Li[j - fi] = (Li[j - fi] - acc) / Lj[j - fj];
        }
        double acc = 0;
        for (int k = fi; k < i; k++) acc += Li[k - fi] * Li[k - fi];
        const double d = Li[i - fi] - acc;
        if (!(diag[i] > 0)) throw F4Error(ESPRESO_ERR_NUMERIC, "stiffness matrix has a non-positive diagonal entry");
        if (d <= 1e-8 * diag[i]) {
            fixed[i] = 1;
            for (int k = fi; k < i; k++) Li[k - fi] = 0.0;
            Li[i - fi] = 1.0;
        } else {
            Li[i - fi] = std::sqrt(d);
        }
    }
    std::vector<int> fix;
    for (int i = 0; i < n; i++)
        if (fixed[i]) fix.push_back(i);
    const int k = (int)fix.size();
    F4_CHECK(k >= 1, ESPRESO_ERR_NUMERIC, "subdomain stiffness matrix is not singular (no floating kernel)");
    F4_CHECK(k <= 12, ESPRESO_ERR_NUMERIC, "subdomain kernel dimension above 12 (disconnected subdomain?)");
    s.kdim = k;
    s.R.assign((size_t)n * k, 0.0);
    std::vector<double> x(n);
    for (int a = 0; a < k; a++) {
        const int p = fix[a], op = perm[p];
        std::fill(x.begin(), x.end(), 0.0);
        // rhs = -K(:, p) on the free rows (K symmetric: row p of the CSR)
        for (int q = s.rowptr[op]; q < s.rowptr[op + 1]; q++) {
            const int j = ip[s.colidx[q]];
            if (!fixed[j]) x[j] = -s.vals[q];
        }
        for (int i = 0; i < n; i++) {  // forward
            const double *Li = &L[ptr[i]];
            double acc = x[i];
            for (int kk = first[i]; kk < i; kk++) acc -= Li[kk - first[i]] * x[kk];
            x[i] = acc / Li[i - first[i]];
        }
        for (int i = n - 1; i >= 0; i--) {  // backward
            const double *Li = &L[ptr[i]];
            x[i] /= Li[i - first[i]];
            for (int kk = first[i]; kk < i; kk++) x[kk] -= Li[kk - first[i]] * x[i];
        }
        for (int j : fix) x[j] = 0.0;
        x[p] = 1.0;
        for (int i = 0; i < n; i++) s.R[(size_t)a * n + perm[i]] = x[i];
    }
    // orthonormalise (two passes of modified Gram-Schmidt)
    for (int pass = 0; pass < 2; pass++)
        for (int a = 0; a < k; a++) {
            double *v = &s.R[(size_t)a * n];
            for (int b = 0; b < a; b++) {
                const double *u = &s.R[(size_t)b * n];
                double d = 0;
                for (int i = 0; i < n; i++) d += u[i] * v[i];
                for (int i = 0; i < n; i++) v[i] -= d * u[i];
            }
            double nr = 0;
            for (int i = 0; i < n; i++) nr += v[i] * v[i];
            F4_CHECK(nr > 0, ESPRESO_ERR_NUMERIC, "kernel detection produced dependent vectors");
            nr = 1.0 / std::sqrt(nr);
            for (int i = 0; i < n; i++) v[i] *= nr;
        }
    // check K R = 0
    double kr = 0, kn = 0;
    for (int i = 0; i < n; i++)
        for (int q = s.rowptr[i]; q < s.rowptr[i + 1]; q++) kn = std::max(kn, std::fabs(s.vals[q]));
    for (int a = 0; a < k; a++)
        for (int i = 0; i < n; i++) {
            double acc = 0;
            for (int q = s.rowptr[i]; q < s.rowptr[i + 1]; q++) acc += s.vals[q] * s.R[(size_t)a * n + s.colidx[q]];
            kr = std::max(kr, std::fabs(acc));
        }
    F4_CHECK(kr <= 1e-7 * kn, ESPRESO_ERR_NUMERIC, "algebraic kernel detection failed (K R != 0)");
}

// ---------------------------------------------------------- B0 corners ----
// HYBRID_FETI: the instance is one cluster (one cluster per process, as the
// reference's library interface has one per MPI rank).  Corners follow the
// reference's CORNERS B0 (fetidata.cpp:146-187, equalityconstraints.cpp:
// 504-552): interface nodes are grouped by the full set of subdomains owning
// them, each group gives its middle node (by node number), and a corner glues
// consecutive pairs of its owners, one B0 row per DOF of the node (+1 in the
// first subdomain, -1 in the second).  Elements given without nodes group
// DOFs instead of nodes.
void corners(FETI4IStructInstance &I, const std::vector<int> &part) {
    for (auto &S : I.sub) { S.b0row.clear(); S.b0dof.clear(); S.b0sign.clear(); }
    I.n0 = 0;
    if (I.method != 1) return;
    const FETI4IStructMatrix &K = I.K;
    const int ne = K.nelem(), ns = (int)I.sub.size();
    // entities: nodes with their DOFs (node-major in the element's DOF list),
    // or single DOFs when the elements carry no nodes or the DOF count is not
    // a multiple of the node count
    bool with_nodes = K.nptr.back() > 0;
    for (int e = 0; e < ne && with_nodes; e++) {
        const int64_t kn = K.nptr[e + 1] - K.nptr[e], kd = K.dptr[e + 1] - K.dptr[e];
        with_nodes = kn > 0 && kd % kn == 0 && kd / kn == I.ndof_node;
    }
    const int nd = with_nodes ? I.ndof_node : 1;
    int nent = 0;
    if (with_nodes) for (int v : K.nodes) nent = std::max(nent, v + 1);
    else nent = I.size;
    std::vector<int> edof((size_t)nent * nd, -1);
    std::vector<std::vector<int>> own(nent);
    for (int e = 0; e < ne; e++) {
        const int64_t kn = with_nodes ? K.nptr[e + 1] - K.nptr[e] : K.dptr[e + 1] - K.dptr[e];
        for (int64_t a = 0; a < kn; a++) {
            const int v = with_nodes ? K.nodes[K.nptr[e] + a] : K.dofs[K.dptr[e] + a];
            for (int c = 0; c < nd; c++) edof[(size_t)v * nd + c] = K.dofs[K.dptr[e] + a * nd + c];
            own[v].push_back(part[e]);
        }
    }
    std::vector<std::pair<std::vector<int>, int>> keyed;  // (owner set, entity)
    for (int v = 0; v < nent; v++) {
        auto &o = own[v];
        std::sort(o.begin(), o.end());
        o.erase(std::unique(o.begin(), o.end()), o.end());
        if (o.size() >= 2) keyed.push_back({o, v});
    }
    std::stable_sort(keyed.begin(), keyed.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    auto ldof = [&](int s, int g) {
        const auto &v = I.sub[s].gdofs;
        const auto it = std::lower_bound(v.begin(), v.end(), g);
        F4_CHECK(it != v.end() && *it == g, ESPRESO_ERR_ARG, "corner DOF missing from its subdomain");
        return (int)(it - v.begin());
    };
    int row = 0;
    for (size_t g0 = 0; g0 < keyed.size();) {
        size_t g1 = g0;
        while (g1 < keyed.size() && keyed[g1].first == keyed[g0].first) g1++;
        const std::vector<int> &ow = keyed[g0].first;
        const int v = keyed[g0 + (g1 - g0) / 2].second;  // middle entity of the group (ascending number)
        for (size_t a = 0; a + 1 < ow.size(); a++)
            for (int c = 0; c < nd; c++) {
                const int g = edof[(size_t)v * nd + c];
                F4_CHECK(g >= 0, ESPRESO_ERR_ARG, "corner node without DOFs");
                Subdomain &s1 = I.sub[ow[a]], &s2 = I.sub[ow[a + 1]];
                s1.b0row.push_back(row);
                s1.b0dof.push_back(ldof(ow[a], g));
                s1.b0sign.push_back(1.0);
                s2.b0row.push_back(row);
                s2.b0dof.push_back(ldof(ow[a + 1], g));
                s2.b0sign.push_back(-1.0);
                row++;
            }
        g0 = g1;
    }
    (void)ns;
    I.n0 = row;
}

// K_i values of subdomain s from its elements on the existing CSR pattern
// (the element structure is unchanged); false when the kernel R_i no longer
// annihilates the new K_i.
bool reassemble_values(FETI4IStructInstance &I, int s) {
    const FETI4IStructMatrix &K = I.K;
    Subdomain &S = I.sub[s];
    std::fill(S.vals.begin(), S.vals.end(), 0.0);
    auto loc = [&](int g) { return (int)(std::lower_bound(S.gdofs.begin(), S.gdofs.end(), g) - S.gdofs.begin()); };
    std::vector<int> l;
    for (int e : I.elems[s]) {
        const int k = (int)(K.dptr[e + 1] - K.dptr[e]);
        const int *ed = &K.dofs[K.dptr[e]];
        const double *ev = &K.vals[K.vptr[e]];
        l.resize(k);
        for (int a = 0; a < k; a++) l[a] = loc(ed[a]);
        for (int a = 0; a < k; a++) {
            const int *c0 = &S.colidx[S.rowptr[l[a]]], *c1 = &S.colidx[S.rowptr[l[a] + 1]];
            for (int b = 0; b < k; b++) {
                const int q = (int)(std::lower_bound(c0, c1, l[b]) - &S.colidx[0]);
                S.vals[q] += ev[(size_t)a * k + b];
            }
        }
    }
    const int n = (int)S.gdofs.size();
    double kr = 0, kn = 0;
    for (int i = 0; i < n; i++)
        for (int q = S.rowptr[i]; q < S.rowptr[i + 1]; q++) kn = std::max(kn, std::fabs(S.vals[q]));
    for (int a = 0; a < S.kdim; a++)
        for (int i = 0; i < n; i++) {
            double acc = 0;
            for (int q = S.rowptr[i]; q < S.rowptr[i + 1]; q++) acc += S.vals[q] * S.R[(size_t)a * n + S.colidx[q]];
            kr = std::max(kr, std::fabs(acc));
        }
    return kr <= 1e-7 * kn;
}

// --------------------------------------------------------- decomposition ----
void decompose(FETI4IStructInstance &I) {
    const FETI4IStructMatrix &K = I.K;
    const int ne = K.nelem(), n = I.size;
    F4_CHECK(ne > 0, ESPRESO_ERR_ARG, "stiffness matrix has no elements");
    F4_CHECK(I.nparts >= 1 && I.nparts <= ne, ESPRESO_ERR_ARG, "FETI4I_SUBDOMAINS must be in [1, number of elements]");
    for (int d : K.dofs) F4_CHECK(d >= 0 && d < n, ESPRESO_ERR_ARG, "element DOF outside [0, size)");
    // DOFs per node (for the report only)
    if (K.nptr.back() > 0) I.ndof_node = std::max<int>(1, (int)(K.dptr[1] / std::max<int64_t>(1, K.nptr[1])));
    ElementGraph G;
    element_graph(K, n, G);
    std::vector<int> part(ne, 0);
    {
        Bisector B(G, part, ne);
        std::vector<int> all(ne);
        std::iota(all.begin(), all.end(), 0);
        B.split(std::move(all), I.nparts, 0);
    }
    merge_stray_pieces(G, part, I.nparts);
    const int ns = I.nparts;
    I.sub.assign(ns, Subdomain());
    // subdomain DOF sets
    std::vector<std::vector<int>> dset(ns);
    for (int e = 0; e < ne; e++)
        for (int64_t q = K.dptr[e]; q < K.dptr[e + 1]; q++) dset[part[e]].push_back(K.dofs[q]);
    std::vector<int> mult(n, 0);
    for (int s = 0; s < ns; s++) {
        auto &v = dset[s];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        F4_CHECK(!v.empty(), ESPRESO_ERR_ARG, "empty subdomain after partitioning");
        for (int d : v) mult[d]++;
        I.sub[s].gdofs = v;
    }
    for (int d = 0; d < n; d++) F4_CHECK(mult[d] > 0, ESPRESO_ERR_ARG, "DOF not covered by any element");
    // element lists per subdomain
    std::vector<std::vector<int>> &elems = I.elems;
    elems.assign(ns, {});
    for (int e = 0; e < ne; e++) elems[part[e]].push_back(e);
    std::string err;
#pragma omp parallel for schedule(dynamic)
    for (int s = 0; s < ns; s++) {
        try {
            Subdomain &S = I.sub[s];
            const int nl = (int)S.gdofs.size();
            auto loc = [&](int g) { return (int)(std::lower_bound(S.gdofs.begin(), S.gdofs.end(), g) - S.gdofs.begin()); };
            std::vector<std::vector<std::pair<int, double>>> rows(nl);
            for (int e : elems[s]) {
                const int k = (int)(K.dptr[e + 1] - K.dptr[e]);
                const int *ed = &K.dofs[K.dptr[e]];
                const double *ev = &K.vals[K.vptr[e]];
                std::vector<int> l(k);
                for (int a = 0; a < k; a++) l[a] = loc(ed[a]);
                for (int a = 0; a < k; a++)
                    for (int b = 0; b < k; b++) rows[l[a]].push_back({l[b], ev[(size_t)a * k + b]});
            }
            S.rowptr.assign(nl + 1, 0);
            S.colidx.clear();
            S.vals.clear();
            for (int r = 0; r < nl; r++) {
                auto &v = rows[r];
                std::sort(v.begin(), v.end(), [](auto &x, auto &y) { return x.first < y.first; });
                for (size_t q = 0; q < v.size(); q++) {
                    if (!S.colidx.empty() && (int)S.colidx.size() > S.rowptr[r] && S.colidx.back() == v[q].first)
                        S.vals.back() += v[q].second;
                    else {
                        S.colidx.push_back(v[q].first);
                        S.vals.push_back(v[q].second);
                    }
                }
                S.rowptr[r + 1] = (int)S.colidx.size();
                std::vector<std::pair<int, double>>().swap(v);
            }
            S.f.resize(nl);
            for (int r = 0; r < nl; r++) S.f[r] = I.rhs[S.gdofs[r]] / mult[S.gdofs[r]];
            algebraic_kernel(S);
        } catch (const std::exception &e) {
#pragma omp critical
            err = e.what();
        }
    }
    if (!err.empty()) throw F4Error(ESPRESO_ERR_NUMERIC, err);
    I.max_kdim = 0;
    for (auto &S : I.sub) I.max_kdim = std::max(I.max_kdim, S.kdim);
    // B1: Dirichlet rows first (by DOF, then subdomain), then redundant gluing
    std::vector<std::vector<int>> owners(n);
    for (int s = 0; s < ns; s++)
        for (int d : I.sub[s].gdofs) owners[d].push_back(s);
    std::vector<double> dval(n, 0.0);
    std::vector<char> isdir(n, 0);
    for (size_t q = 0; q < I.dir_idx.size(); q++) {
        const int d = I.dir_idx[q];
        F4_CHECK(d >= 0 && d < n, ESPRESO_ERR_ARG, "Dirichlet index outside [0, size)");
        if (!isdir[d]) { isdir[d] = 1; dval[d] = I.dir_val[q]; }
    }
    auto ldof = [&](int s, int g) {
        const auto &v = I.sub[s].gdofs;
        return (int)(std::lower_bound(v.begin(), v.end(), g) - v.begin());
    };
    auto kdiag = [&](int s, int l) {
        const Subdomain &S = I.sub[s];
        for (int q = S.rowptr[l]; q < S.rowptr[l + 1]; q++)
            if (S.colidx[q] == l) return S.vals[q];
        return 0.0;
    };
    for (auto &S : I.sub) { S.lam.clear(); S.dof.clear(); S.sign.clear(); S.scale.clear(); }
    I.c.clear();
    int lam = 0;
    auto add = [&](int s, int l, double sg, double sc) {
        Subdomain &S = I.sub[s];
        S.lam.push_back(lam);
        S.dof.push_back(l);
        S.sign.push_back(sg);
        S.scale.push_back(sc);
    };
    for (int d = 0; d < n; d++)
        if (isdir[d])
            for (int s : owners[d]) {
                add(s, ldof(s, d), 1.0, 1.0);
                I.c.push_back(dval[d]);
                lam++;
            }
    for (int d = 0; d < n; d++) {
        if (isdir[d] || owners[d].size() < 2) continue;
        const auto &ow = owners[d];
        const int k = (int)ow.size();
        std::vector<double> D(k);
        std::vector<int> l(k);
        double sum = 0;
        for (int q = 0; q < k; q++) {
            l[q] = ldof(ow[q], d);
            D[q] = kdiag(ow[q], l[q]);
            sum += D[q];
        }
        for (int a = 0; a < k; a++)
            for (int b = a + 1; b < k; b++) {
                add(ow[a], l[a], 1.0, D[b] / sum);
                add(ow[b], l[b], -1.0, D[a] / sum);
                I.c.push_back(0.0);
                lam++;
            }
    }
    I.nlam = lam;
    for (int s = 0; s < ns; s++)
        F4_CHECK(!I.sub[s].lam.empty(), ESPRESO_ERR_ARG, "subdomain without multipliers (isolated part of the mesh)");
    corners(I, part);
    I.built = true;
    I.gpu_ready = false;
}

void ensure_gpu(FETI4IStructInstance &I) {
    if (!I.built) decompose(I);
    if (I.gpu_ready) return;
    I.needs_setup = false;
    if (I.h) {
        espreso_feti_destroy(I.h);
        I.h = nullptr;
    }
    espreso_feti_options o{};
    o.method = I.method;
    o.preconditioner = I.prec;
    o.max_iterations = I.max_it;
    o.scaling = 1;
    o.sc_precision = 0;
    o.device = 0;
    o.precision = I.precision;
    int rc = espreso_feti_create(&I.h, (int)I.sub.size(), &o);
    F4_CHECK(rc == 0, rc, espreso_last_error());
    for (size_t s = 0; s < I.sub.size(); s++) {
        const Subdomain &S = I.sub[s];
        rc = espreso_feti_set_subdomain(I.h, (int)s, (int)S.gdofs.size(), S.rowptr.data(), S.colidx.data(),
                                        S.vals.data(), S.f.data(), S.kdim, S.R.data(), (int)S.lam.size(),
                                        S.lam.data(), S.dof.data(), S.sign.data(), S.scale.data());
        F4_CHECK(rc == 0, rc, espreso_last_error());
        if (I.method == 1) {
            rc = espreso_feti_set_cluster(I.h, (int)s, 0, (int)S.b0row.size(), S.b0row.data(), S.b0dof.data(),
                                          S.b0sign.data());
            F4_CHECK(rc == 0, rc, espreso_last_error());
        }
    }
    rc = espreso_feti_set_dual(I.h, I.nlam, I.c.data());
    F4_CHECK(rc == 0, rc, espreso_last_error());
    rc = espreso_feti_setup(I.h);
    F4_CHECK(rc == 0, rc, espreso_last_error());
    int32_t sz[3];
    espreso_feti_sizes(I.h, sz);
    I.ncoarse = sz[2];
    I.gpu_ready = true;
}

void set_options(FETI4IStructInstance &I, const FETI4IInt *io, const FETI4IReal *ro) {
    if (io) {
        F4_CHECK(io[FETI4I_SUBDOMAINS] >= 1, ESPRESO_ERR_ARG, "FETI4I_SUBDOMAINS must be positive");
        F4_CHECK(io[FETI4I_MAX_ITERATIONS] >= 1, ESPRESO_ERR_ARG, "FETI4I_MAX_ITERATIONS must be positive");
        F4_CHECK(io[FETI4I_FETI_METHOD] == 0 || io[FETI4I_FETI_METHOD] == 1, ESPRESO_ERR_ARG,
                 "FETI4I_FETI_METHOD: TOTAL_FETI (0) or HYBRID_FETI (1)");
        F4_CHECK(io[FETI4I_PRECONDITIONER] >= 0 && io[FETI4I_PRECONDITIONER] <= 4, ESPRESO_ERR_ARG,
                 "FETI4I_PRECONDITIONER: NONE, LUMPED, WEIGHT_FUNCTION, DIRICHLET or SUPER_DIRICHLET (0..4)");
        F4_CHECK(io[FETI4I_CGSOLVER] == 0, ESPRESO_ERR_ARG, "FETI4I_CGSOLVER: only PCG (0)");
        I.nparts = (int)io[FETI4I_SUBDOMAINS];
        I.max_it = (int)io[FETI4I_MAX_ITERATIONS];
        I.method = (int)io[FETI4I_FETI_METHOD];
        I.prec = (int)io[FETI4I_PRECONDITIONER];
        I.cg = (int)io[FETI4I_CGSOLVER];
        I.verbose = (int)io[FETI4I_VERBOSE_LEVEL];
    }
    if (ro) {
        F4_CHECK(ro[FETI4I_PRECISION] > 0, ESPRESO_ERR_ARG, "FETI4I_PRECISION must be positive");
        I.precision = ro[FETI4I_PRECISION];
    }
}

bool is_instance(void *p) {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_instances.count(p) != 0;
}
bool is_matrix(void *p) {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_matrices.count(p) != 0;
}

}  // namespace

extern "C" {

void FETI4ISetDefaultIntegerOptions(FETI4IInt *options) {
    guarded([&] {
        F4_CHECK(options, ESPRESO_ERR_ARG, "null options");
        options[FETI4I_SUBDOMAINS] = 4;
        options[FETI4I_MAX_ITERATIONS] = 200;
        options[FETI4I_FETI_METHOD] = 0;      // TOTAL_FETI
        options[FETI4I_PRECONDITIONER] = 3;   // DIRICHLET
        options[FETI4I_CGSOLVER] = 0;         // PCG
        options[FETI4I_N_MICS] = 2;
        options[FETI4I_VERBOSE_LEVEL] = 1;
        options[FETI4I_MEASURE_LEVEL] = 0;
        options[FETI4I_PRINT_MATRICES] = 0;
        options[FETI4I_SC_SIZE] = 200;
    });
}

void FETI4ISetDefaultRealOptions(FETI4IReal *options) {
    guarded([&] {
        F4_CHECK(options, ESPRESO_ERR_ARG, "null options");
        options[FETI4I_PRECISION] = 1e-5;
    });
}

void FETI4ICreateStiffnessMatrix(FETI4IMatrix *matrix, FETI4IInt type, FETI4IInt indexBase) {
    guarded([&] {
        F4_CHECK(matrix, ESPRESO_ERR_ARG, "null matrix pointer");
        F4_CHECK(type >= 0 && type <= 2, ESPRESO_ERR_ARG, "unknown matrix type");
        F4_CHECK(type == FETI4I_REAL_SYMMETRIC_POSITIVE_DEFINITE, ESPRESO_ERR_ARG,
                 "only REAL_SYMMETRIC_POSITIVE_DEFINITE stiffness matrices are served");
        auto *m = new FETI4IStructMatrix();
        m->type = (int)type;
        m->base = (int)indexBase;
        std::lock_guard<std::mutex> lk(g_mu);
        g_matrices.insert(m);
        *matrix = m;
    });
}

void FETI4IAddElement(FETI4IMatrix matrix, FETI4IInt type, FETI4IInt nodesSize, FETI4IInt *nodes, FETI4IInt dofsSize,
                      FETI4IInt *dofs, FETI4IReal *values) {
    guarded([&] {
        F4_CHECK(is_matrix(matrix), ESPRESO_ERR_ARG, "unknown matrix handle");
        F4_CHECK(dofsSize > 0 && dofs && values && nodesSize >= 0, ESPRESO_ERR_ARG, "invalid element");
        (void)type;
        const size_t k = (size_t)dofsSize;
        bool zero = true;
        for (size_t q = 0; q < k * k && zero; q++) zero = values[q] == 0.0;
        if (zero) return;  // elements without stiffness are skipped
        for (size_t q = 0; q < k * k; q++) F4_CHECK(std::isfinite(values[q]), ESPRESO_ERR_NUMERIC, "non-finite element value");
        for (FETI4IInt q = 0; q < nodesSize; q++) matrix->nodes.push_back((int)(nodes[q] - matrix->base));
        for (size_t q = 0; q < k; q++) matrix->dofs.push_back((int)(dofs[q] - matrix->base));
        matrix->vals.insert(matrix->vals.end(), values, values + k * k);
        matrix->nptr.push_back((int64_t)matrix->nodes.size());
        matrix->dptr.push_back((int64_t)matrix->dofs.size());
        matrix->vptr.push_back((int64_t)matrix->vals.size());
    });
}

void FETI4ICreateInstance(FETI4IInstance *instance, FETI4IMatrix matrix, FETI4IInt size, FETI4IReal *rhs,
                          FETI4IInt *l2g, FETI4IMPIInt neighbours_size, FETI4IMPIInt *neighbours,
                          FETI4IInt dirichlet_size, FETI4IInt *dirichlet_indices, FETI4IReal *dirichlet_values,
                          FETI4IInt *integer_options, FETI4IReal *real_options) {
    guarded([&] {
        F4_CHECK(instance && is_matrix(matrix), ESPRESO_ERR_ARG, "invalid matrix handle");
        F4_CHECK(size > 0 && rhs, ESPRESO_ERR_ARG, "invalid right-hand side");
        F4_CHECK(neighbours_size == 0, ESPRESO_ERR_ARG,
                 "one cluster per process: neighbouring MPI ranks are not served by this library");
        (void)neighbours;
        F4_CHECK(dirichlet_size >= 0 && (dirichlet_size == 0 || (dirichlet_indices && dirichlet_values)),
                 ESPRESO_ERR_ARG, "invalid Dirichlet arrays");
        auto inst = std::make_unique<FETI4IStructInstance>();
        inst->K = *matrix;
        inst->size = (int)size;
        inst->rhs.assign(rhs, rhs + size);
        for (double v : inst->rhs) F4_CHECK(std::isfinite(v), ESPRESO_ERR_NUMERIC, "non-finite right-hand side");
        if (l2g) inst->l2g.assign(l2g, l2g + size);
        for (FETI4IInt q = 0; q < dirichlet_size; q++) {
            inst->dir_idx.push_back((int)(dirichlet_indices[q] - matrix->base));
            inst->dir_val.push_back(dirichlet_values[q]);
        }
        set_options(*inst, integer_options, real_options);
        decompose(*inst);
        std::lock_guard<std::mutex> lk(g_mu);
        g_instances.insert(inst.get());
        *instance = inst.release();
    });
}

void FETI4ISolve(FETI4IInstance instance, FETI4IInt solution_size, FETI4IReal *solution) {
    guarded([&] {
        F4_CHECK(is_instance(instance), ESPRESO_ERR_ARG, "unknown instance handle");
        F4_CHECK(solution && solution_size == instance->size, ESPRESO_ERR_ARG, "solution_size must equal the instance size");
        FETI4IStructInstance &I = *instance;
        ensure_gpu(I);
        int rc = 0;
        if (I.needs_setup) {
            rc = espreso_feti_setup(I.h);
            F4_CHECK(rc == 0, rc, espreso_last_error());
            I.needs_setup = false;
        }
        rc = espreso_feti_solve(I.h, &I.st);
        F4_CHECK(rc == 0 || rc == ESPRESO_NOT_CONVERGED, rc, espreso_last_error());
        const bool converged = rc == 0;
        std::vector<double> acc(I.size, 0.0), u;
        std::vector<int> cnt(I.size, 0);
        for (size_t s = 0; s < I.sub.size(); s++) {
            const Subdomain &S = I.sub[s];
            u.resize(S.gdofs.size());
            rc = espreso_feti_primal(I.h, (int)s, u.data());
            F4_CHECK(rc == 0, rc, espreso_last_error());
            for (size_t l = 0; l < S.gdofs.size(); l++) {
                acc[S.gdofs[l]] += u[l];
                cnt[S.gdofs[l]]++;
            }
        }
        for (int d = 0; d < I.size; d++) solution[d] = acc[d] / std::max(cnt[d], 1);
        // the solution is returned either way; the status says whether the stop test was met
        F4_CHECK(converged, ESPRESO_NOT_CONVERGED, "FETI: stop test not met within FETI4I_MAX_ITERATIONS");
    });
}

// New element values on the same element structure keep the decomposition,
// B1 and its scaling, the kernels and the GPU instance: only K_i values are
// re-assembled and handed to espreso_feti_update_values, and the next solve
// re-assembles the dual operators.  A changed structure, or values whose
// kernel differs, take the full decomposition again.
void FETI4IUpdateStiffnessMatrix(FETI4IInstance instance, FETI4IMatrix stiffnessMatrix) {
    guarded([&] {
        F4_CHECK(is_instance(instance) && is_matrix(stiffnessMatrix), ESPRESO_ERR_ARG, "invalid handle");
        FETI4IStructInstance &I = *instance;
        const FETI4IStructMatrix &M = *stiffnessMatrix;
        const bool same = I.built && M.nelem() == I.K.nelem() && M.dptr == I.K.dptr && M.vptr == I.K.vptr &&
                          M.dofs == I.K.dofs && M.base == I.K.base;
        I.K = M;
        if (same) {
            bool ok = true;
            for (size_t s = 0; s < I.sub.size() && ok; s++) ok = reassemble_values(I, (int)s);
            if (ok) {
                if (I.gpu_ready && I.h) {
                    for (size_t s = 0; s < I.sub.size(); s++) {
                        const int rc = espreso_feti_update_values(I.h, (int)s, I.sub[s].vals.data());
                        F4_CHECK(rc == 0, rc, espreso_last_error());
                    }
                    I.needs_setup = true;
                }
                return;
            }
        }
        decompose(I);
    });
}

void FETI4IUpdateRhs(FETI4IInstance instance, FETI4IInt size, FETI4IReal *values) {
    guarded([&] {
        F4_CHECK(is_instance(instance), ESPRESO_ERR_ARG, "unknown instance handle");
        F4_CHECK(size == instance->size && values, ESPRESO_ERR_ARG, "right-hand side size differs from the instance");
        instance->rhs.assign(values, values + size);
        decompose(*instance);
    });
}

void FETI4IUpdateDirichlet(FETI4IInstance instance, FETI4IInt size, FETI4IInt *indices, FETI4IReal *values) {
    guarded([&] {
        F4_CHECK(is_instance(instance), ESPRESO_ERR_ARG, "unknown instance handle");
        F4_CHECK(size >= 0 && (size == 0 || (indices && values)), ESPRESO_ERR_ARG, "invalid Dirichlet arrays");
        instance->dir_idx.clear();
        instance->dir_val.clear();
        for (FETI4IInt q = 0; q < size; q++) {
            instance->dir_idx.push_back((int)(indices[q] - instance->K.base));
            instance->dir_val.push_back(values[q]);
        }
        decompose(*instance);
    });
}

void FETI4IDestroy(void *ptr) {
    guarded([&] {
        bool inst = false, mat = false;
        {
            std::lock_guard<std::mutex> lk(g_mu);
            inst = g_instances.erase(ptr) != 0;
            mat = !inst && g_matrices.erase(ptr) != 0;
        }
        if (inst) delete (FETI4IStructInstance *)ptr;
        else if (mat) delete (FETI4IStructMatrix *)ptr;
        else F4_CHECK(false, ESPRESO_ERR_ARG, "FETI4IDestroy: unknown handle");
    });
}

int FETI4IGetLastStatus(void) { return g_status; }
const char *FETI4IGetLastError(void) { return g_error.c_str(); }

int FETI4IInstanceInfo(FETI4IInstance instance, FETI4IInt *info) {
    guarded([&] {
        F4_CHECK(is_instance(instance) && info, ESPRESO_ERR_ARG, "invalid arguments");
        info[0] = (FETI4IInt)instance->sub.size();
        info[1] = instance->nlam;
        info[2] = instance->ncoarse;
        info[3] = instance->st.iterations;
        info[4] = instance->ndof_node;
        info[5] = instance->max_kdim;
        info[6] = instance->n0;
        info[7] = instance->method;
    });
    return g_status;
}

int FETI4IExportSubdomain(FETI4IInstance instance, FETI4IInt i, FETI4IInt *sizes, FETI4IInt *gdofs, FETI4IInt *rowptr,
                          FETI4IInt *colidx, FETI4IReal *values, FETI4IReal *f, FETI4IReal *R, FETI4IInt *lambda,
                          FETI4IInt *dof, FETI4IReal *sign, FETI4IReal *scale) {
    guarded([&] {
        F4_CHECK(is_instance(instance) && sizes, ESPRESO_ERR_ARG, "invalid arguments");
        F4_CHECK(i >= 0 && i < (FETI4IInt)instance->sub.size(), ESPRESO_ERR_ARG, "subdomain out of range");
        const Subdomain &S = instance->sub[i];
        const size_t n = S.gdofs.size();
        sizes[0] = (FETI4IInt)n;
        sizes[1] = S.rowptr[n];
        sizes[2] = S.kdim;
        sizes[3] = (FETI4IInt)S.lam.size();
        if (gdofs) std::copy(S.gdofs.begin(), S.gdofs.end(), gdofs);
        if (rowptr) std::copy(S.rowptr.begin(), S.rowptr.end(), rowptr);
        if (colidx) std::copy(S.colidx.begin(), S.colidx.end(), colidx);
        if (values) std::copy(S.vals.begin(), S.vals.end(), values);
        if (f) std::copy(S.f.begin(), S.f.end(), f);
        if (R) std::copy(S.R.begin(), S.R.end(), R);
        if (lambda) std::copy(S.lam.begin(), S.lam.end(), lambda);
        if (dof) std::copy(S.dof.begin(), S.dof.end(), dof);
        if (sign) std::copy(S.sign.begin(), S.sign.end(), sign);
        if (scale) std::copy(S.scale.begin(), S.scale.end(), scale);
    });
    return g_status;
}

int FETI4IExportCorners(FETI4IInstance instance, FETI4IInt i, FETI4IInt *size, FETI4IInt *row, FETI4IInt *dof,
                        FETI4IReal *sign) {
    guarded([&] {
        F4_CHECK(is_instance(instance) && size, ESPRESO_ERR_ARG, "invalid arguments");
        F4_CHECK(i >= 0 && i < (FETI4IInt)instance->sub.size(), ESPRESO_ERR_ARG, "subdomain out of range");
        const Subdomain &S = instance->sub[i];
        *size = (FETI4IInt)S.b0row.size();
        if (row) std::copy(S.b0row.begin(), S.b0row.end(), row);
        if (dof) std::copy(S.b0dof.begin(), S.b0dof.end(), dof);
        if (sign) std::copy(S.b0sign.begin(), S.b0sign.end(), sign);
    });
    return g_status;
}

int FETI4IGetStats(FETI4IInstance instance, FETI4IReal *stats) {
    guarded([&] {
        F4_CHECK(is_instance(instance) && stats, ESPRESO_ERR_ARG, "invalid arguments");
        stats[0] = instance->st.iterations;
        stats[1] = instance->st.final_norm;
        stats[2] = instance->st.tol;
        stats[3] = instance->st.time_assembly;
        stats[4] = instance->st.time_pcg;
        stats[5] = instance->st.time_total;
    });
    return g_status;
}

}  // extern "C"
