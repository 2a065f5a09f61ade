// FETI oracle: CPU restatement of the reference's TFETI path.
//
// TEST INFRASTRUCTURE ONLY -- the checker for the HIP product.  Loaded by
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg only.
// Reference citations (file:line) are relative to the reference snapshot.
//
// Parity status of each part (how it is pinned) is listed in oracle.h and in
// the test files: global solve -> closed-form patch tests + reference .emr
// values; FETI solve -> the global solve; operators -> algebraic invariants.

#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;

// ---------------------------------------------------------------- RNG ----
// splitmix64 stream; uniform in [0,1) from the top 53 bits.
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ------------------------------------------------------------- grid ----
// Global grid of NE[d] elements and NN[d] = NE[d]+1 nodes per direction.
// Subdomains are numbered cluster-major (cluster id, then domain inside the
// cluster, x fastest), as the reference generator walks clusters
// (gridgenerator.cpp:44-69).
// Subdomain slabs may differ in element count (slab_elements; the reference's
// generator takes one count per direction, gridgenerator.cpp:44-69): slab S of
// direction d holds global elements soff[d][S] .. soff[d][S+1]-1.
struct Grid {
    oracle_grid_problem p;
    int ndof;
    int NS[3];      // subdomains per direction
    int NE[3];      // elements per direction (global)
    int NN[3];      // nodes per direction (global)
    double h[3];
    int nsub;
    std::vector<int> soff[3];

    explicit Grid(const oracle_grid_problem &pp) : p(pp) {
        ndof = p.physics == 0 ? 3 : 1;
        nsub = 1;
        int q = 0;
        for (int d = 0; d < 3; d++) {
            NS[d] = p.clusters[d] * p.domains[d];
            soff[d].assign(NS[d] + 1, 0);
            for (int S = 0; S < NS[d]; S++) soff[d][S + 1] = soff[d][S] + (p.slab_elements ? p.slab_elements[q++] : p.elements[d]);
            NE[d] = soff[d][NS[d]];
            NN[d] = NE[d] + 1;
            h[d] = p.length[d] / NE[d];
            nsub *= NS[d];
        }
        p.slab_elements = nullptr;  // not owned: the offsets hold the layout
    }
    // slab of global element e along d
    int slab_of(int d, int e) const {
        return (int)(std::upper_bound(soff[d].begin(), soff[d].end(), e) - soff[d].begin()) - 1;
    }
    // elements per direction and first global element of the subdomain at slab position S
    void sub_box(const int *S, int *ex, int *org) const {
        for (int d = 0; d < 3; d++) {
            org[d] = soff[d][S[d]];
            ex[d] = soff[d][S[d] + 1] - org[d];
        }
    }
    int nnodes() const { return NN[0] * NN[1] * NN[2]; }
    int node(int i, int j, int k) const { return i + NN[0] * (j + NN[1] * k); }
    // subdomain id -> subdomain grid position
    void subpos(int s, int *S) const {
        int dpc = p.domains[0] * p.domains[1] * p.domains[2];
        int c = s / dpc, d = s % dpc;
        int ci[3] = {c % p.clusters[0], (c / p.clusters[0]) % p.clusters[1], c / (p.clusters[0] * p.clusters[1])};
        int di[3] = {d % p.domains[0], (d / p.domains[0]) % p.domains[1], d / (p.domains[0] * p.domains[1])};
        for (int q = 0; q < 3; q++) S[q] = ci[q] * p.domains[q] + di[q];
    }
    int subid(const int *S) const {
        int ci[3], di[3];
        for (int q = 0; q < 3; q++) { ci[q] = S[q] / p.domains[q]; di[q] = S[q] % p.domains[q]; }
        int c = ci[0] + p.clusters[0] * (ci[1] + p.clusters[1] * ci[2]);
        int d = di[0] + p.domains[0] * (di[1] + p.domains[1] * di[2]);
        return c * (p.domains[0] * p.domains[1] * p.domains[2]) + d;
    }
    // subdomains containing global node (i,j,k), ascending id
    void node_subdomains(int i, int j, int k, std::vector<int> &out) const {
        out.clear();
        int g[3] = {i, j, k};
        int lo[3], hi[3];
        for (int q = 0; q < 3; q++) {
            hi[q] = std::min(NS[q] - 1, slab_of(q, g[q]));
            lo[q] = (hi[q] > 0 && soff[q][hi[q]] == g[q]) ? hi[q] - 1 : hi[q];
        }
        int S[3];
        for (S[2] = lo[2]; S[2] <= hi[2]; S[2]++)
            for (S[1] = lo[1]; S[1] <= hi[1]; S[1]++)
                for (S[0] = lo[0]; S[0] <= hi[0]; S[0]++) out.push_back(subid(S));
        std::sort(out.begin(), out.end());
    }
};

// ------------------------------------------------------ HEXA8 element ----
// Node order (hexahedron8.cpp:37-44): (-,-,-),(+,-,-),(+,+,-),(-,+,-),(-,-,+),(+,-,+),(+,+,+),(-,+,+)
const int HEX[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

// Box element hx*hy*hz, 2x2x2 Gauss.  Elasticity: DOFs node-major (3a+c),
// strain order (xx,yy,zz,xy,yz,xz), isotropic D with E = 1 (scaled later).
void element_matrices(int physics, const double *h, double nu, std::vector<double> &Ke, std::vector<double> &Nint) {
    const double gp = 0.577350269189626;
    int nd = physics == 0 ? 3 : 1;
    int ne = 8 * nd;
    Ke.assign(ne * ne, 0.0);
    Nint.assign(8, 0.0);
    double detJ = h[0] * h[1] * h[2] / 8.0;
    double D[6][6] = {{0}};
    if (physics == 0) {
        double EE = 1.0 / ((1 + nu) * (1 - 2 * nu));
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) D[a][b] = (a == b) ? EE * (1 - nu) : EE * nu;
        for (int a = 3; a < 6; a++) D[a][a] = EE * (0.5 - nu);
    }
    for (int g = 0; g < 8; g++) {
        double r = (g & 4) ? gp : -gp, s = (g & 2) ? gp : -gp, t = (g & 1) ? gp : -gp;
        double N[8], dN[8][3];
        for (int a = 0; a < 8; a++) {
            double ra = 2 * HEX[a][0] - 1, sa = 2 * HEX[a][1] - 1, ta = 2 * HEX[a][2] - 1;
            N[a] = 0.125 * (1 + r * ra) * (1 + s * sa) * (1 + t * ta);
            dN[a][0] = 0.125 * ra * (1 + s * sa) * (1 + t * ta) * 2.0 / h[0];
            dN[a][1] = 0.125 * (1 + r * ra) * sa * (1 + t * ta) * 2.0 / h[1];
            dN[a][2] = 0.125 * (1 + r * ra) * (1 + s * sa) * ta * 2.0 / h[2];
            Nint[a] += N[a] * detJ;
        }
        if (physics == 0) {
            double B[6][24] = {{0}};
            for (int a = 0; a < 8; a++) {
                B[0][3 * a + 0] = dN[a][0];
                B[1][3 * a + 1] = dN[a][1];
                B[2][3 * a + 2] = dN[a][2];
                B[3][3 * a + 0] = dN[a][1]; B[3][3 * a + 1] = dN[a][0];
                B[4][3 * a + 1] = dN[a][2]; B[4][3 * a + 2] = dN[a][1];
                B[5][3 * a + 0] = dN[a][2]; B[5][3 * a + 2] = dN[a][0];
            }
            double DB[6][24];
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 24; j++) {
                    double acc = 0;
                    for (int k = 0; k < 6; k++) acc += D[i][k] * B[k][j];
                    DB[i][j] = acc;
                }
            for (int i = 0; i < 24; i++)
                for (int j = 0; j < 24; j++) {
                    double acc = 0;
                    for (int k = 0; k < 6; k++) acc += B[k][i] * DB[k][j];
                    Ke[i * 24 + j] += acc * detJ;
                }
        } else {
            for (int i = 0; i < 8; i++)
                for (int j = 0; j < 8; j++)
                    Ke[i * 8 + j] += (dN[i][0] * dN[j][0] + dN[i][1] * dN[j][1] + dN[i][2] * dN[j][2]) * detJ;
        }
    }
}

// ---------------------------------------------- loads of the problem ----
// Per-element material scale (E_e or k_s), body force, face loads.
struct Loads {
    std::vector<double> elem_scale;     // per global element
    std::vector<double> traction;       // per x1-face node (j + NN1*k) * ndof
};

Loads make_loads(const Grid &G) {
    Loads L;
    int nel = G.NE[0] * G.NE[1] * G.NE[2];
    double base = G.p.physics == 0 ? G.p.young : G.p.conductivity;
    L.elem_scale.assign(nel, base);
    if (G.p.physics == 0 && G.p.young_seed) {
        Rng r(G.p.young_seed);
        for (int e = 0; e < nel; e++) L.elem_scale[e] = G.p.young * (0.1 + 0.9 * r.uniform());
    }
    if (G.p.physics == 1 && G.p.conductivity_seed) {
        Rng r(G.p.conductivity_seed);
        std::vector<double> ks(G.nsub);
        for (int s = 0; s < G.nsub; s++) ks[s] = G.p.conductivity * std::pow(10.0, 2.0 * r.uniform() - 1.0);
        for (int k = 0; k < G.NE[2]; k++)
            for (int j = 0; j < G.NE[1]; j++)
                for (int i = 0; i < G.NE[0]; i++) {
                    int S[3] = {G.slab_of(0, i), G.slab_of(1, j), G.slab_of(2, k)};
                    L.elem_scale[i + G.NE[0] * (j + G.NE[1] * k)] = ks[G.subid(S)];
                }
    }
    if (G.p.material_elem)
        for (int e = 0; e < nel; e++) L.elem_scale[e] = G.p.material_elem[e];
    if (G.p.traction_seed) {
        Rng r(G.p.traction_seed);
        int nf = G.NN[1] * G.NN[2];
        L.traction.assign((size_t)nf * G.ndof, 0.0);
        for (int n = 0; n < nf; n++)
            for (int c = 0; c < G.ndof; c++) L.traction[(size_t)n * G.ndof + c] = G.p.traction_scale * (2.0 * r.uniform() - 1.0);
    }
    return L;
}

// Element (i,j,k) right-hand side (8*ndof), node-major.  Includes body force,
// heat source, face pressure and x1 traction on boundary faces of the element.
void element_rhs(const Grid &G, const Loads &L, const std::vector<double> &Nint, int i, int j, int k, double *fe) {
    int nd = G.ndof;
    std::fill(fe, fe + 8 * nd, 0.0);
    if (G.p.physics == 0) {
        for (int a = 0; a < 8; a++)
            for (int c = 0; c < 3; c++) fe[3 * a + c] += G.p.density * G.p.accel[c] * Nint[a];
    } else {
        for (int a = 0; a < 8; a++) fe[a] += G.p.heat_source * Nint[a];
    }
    int e[3] = {i, j, k};
    for (int f = 0; f < 6; f++) {
        int dir = f / 2, side = f % 2;
        bool on = side == 0 ? e[dir] == 0 : e[dir] == G.NE[dir] - 1;
        if (!on) continue;
        int t1 = (dir + 1) % 3, t2 = (dir + 2) % 3;
        double area = G.h[t1] * G.h[t2];
        // face nodes of the element in cyclic order in the (t1,t2) plane
        int loc[4];
        int q = 0;
        const int cyc[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        for (int c = 0; c < 4; c++) {
            for (int a = 0; a < 8; a++)
                if (HEX[a][dir] == side && HEX[a][t1] == cyc[c][0] && HEX[a][t2] == cyc[c][1]) loc[q++] = a;
        }
        if (G.p.pressure[f] != 0.0 && G.p.physics == 0) {
            double nrm = side == 0 ? -1.0 : 1.0;   // outward normal component along dir
            for (int c = 0; c < 4; c++) fe[3 * loc[c] + dir] += -G.p.pressure[f] * nrm * area / 4.0;
        }
        if (f == 1 && !L.traction.empty()) {
            // traction field interpolated from nodal values: M = A/36 [4 2 1 2; ...]
            double tv[4][3];
            for (int c = 0; c < 4; c++) {
                int a = loc[c];
                int nj = j + HEX[a][1], nk = k + HEX[a][2];
                for (int d = 0; d < nd; d++) tv[c][d] = L.traction[(size_t)(nj + G.NN[1] * nk) * nd + d];
            }
            for (int c = 0; c < 4; c++)
                for (int c2 = 0; c2 < 4; c2++) {
                    int dd = (c2 - c + 4) % 4;
                    double m = (dd == 0 ? 4.0 : (dd == 2 ? 1.0 : 2.0)) * area / 36.0;
                    for (int d = 0; d < nd; d++) fe[nd * loc[c] + d] += m * tv[c2][d];
                }
        }
    }
}

// Dirichlet value of global DOF (node, comp), first face in order x0..z1.
bool dirichlet_of(const Grid &G, int i, int j, int k, int c, double *val) {
    int e[3] = {i, j, k};
    for (int f = 0; f < 6; f++) {
        int dir = f / 2, side = f % 2;
        bool on = side == 0 ? e[dir] == 0 : e[dir] == G.NN[dir] - 1;
        if (on && (G.p.fixed[f] >> c & 1)) { *val = G.p.fixed_value[f]; return true; }
    }
    return false;
}

// ------------------------------------------------- profile Cholesky ----
// Row-envelope (skyline) storage: row i holds columns first[i]..i.
struct Profile {
    int n = 0;
    std::vector<int> first;
    std::vector<size_t> ptr;
    std::vector<double> v;
    void init(const std::vector<int> &fst) {
        n = (int)fst.size();
        first = fst;
        ptr.assign(n + 1, 0);
        for (int i = 0; i < n; i++) ptr[i + 1] = ptr[i] + (i - first[i] + 1);
        v.assign(ptr[n], 0.0);
    }
    double &at(int i, int j) { return v[ptr[i] + (j - first[i])]; }   // j in [first[i], i]
    bool factor() {
        for (int i = 0; i < n; i++) {
            double *Li = &v[ptr[i]];
            int fi = first[i];
            for (int j = fi; j < i; j++) {
                double *Lj = &v[ptr[j]];
                int fj = first[j];
                int k0 = std::max(fi, fj);
                double s = 0;
                const double *a = Li + (k0 - fi), *b = Lj + (k0 - fj);
                int len = j - k0;
#pragma omp simd reduction(+ : s)
                for (int k = 0; k < len; k++) s += a[k] * b[k];
                Li[j - fi] = (Li[j - fi] - s) / Lj[j - fj];
            }
            double s = 0;
            int len = i - fi;
#pragma omp simd reduction(+ : s)
            for (int k = 0; k < len; k++) s += Li[k] * Li[k];
            double d = Li[i - fi] - s;
            if (!(d > 0)) return false;
            Li[i - fi] = std::sqrt(d);
        }
        return true;
    }
    void solve(double *x) const {
        for (int i = 0; i < n; i++) {
            const double *Li = &v[ptr[i]];
            int fi = first[i];
            double s = 0;
            for (int k = fi; k < i; k++) s += Li[k - fi] * x[k];
            x[i] = (x[i] - s) / Li[i - fi];
        }
        for (int i = n - 1; i >= 0; i--) {
            const double *Li = &v[ptr[i]];
            int fi = first[i];
            x[i] /= Li[i - fi];
            double xi = x[i];
            for (int k = fi; k < i; k++) x[k] -= Li[k - fi] * xi;
        }
    }
};

// ------------------------------------------------------- dense helpers ----
// Column-major n x n.
// Blocked right-looking Cholesky (lower triangle, column-major), OpenMP over
// the panel rows and the trailing columns: the coarse G G^T of a thousand
// subdomains is 6000 x 6000.
bool dense_cholesky(int n, double *A) {
    const int NB = 96;
    std::vector<double> P;
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = std::min(NB, n - k0), r0 = k0 + kb, nr = n - r0;
        // diagonal block
        for (int j = k0; j < r0; j++) {
            double d = A[j + (size_t)j * n];
            for (int k = k0; k < j; k++) d -= A[j + (size_t)k * n] * A[j + (size_t)k * n];
            if (!(d > 0)) return false;
            d = std::sqrt(d);
            A[j + (size_t)j * n] = d;
            for (int i = j + 1; i < r0; i++) {
                double s = A[i + (size_t)j * n];
                for (int k = k0; k < j; k++) s -= A[i + (size_t)k * n] * A[j + (size_t)k * n];
                A[i + (size_t)j * n] = s / d;
            }
        }
        if (nr <= 0) break;
        // panel rows below the block, row-major copy P (nr x kb): P = A_rk L_kk^-T
        P.assign((size_t)nr * kb, 0.0);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nr; i++) {
            double *pi = &P[(size_t)i * kb];
            for (int t = 0; t < kb; t++) pi[t] = A[(r0 + i) + (size_t)(k0 + t) * n];
            for (int j = 0; j < kb; j++) {
                double s = pi[j];
                for (int t = 0; t < j; t++) s -= pi[t] * A[(k0 + j) + (size_t)(k0 + t) * n];
                pi[j] = s / A[(k0 + j) + (size_t)(k0 + j) * n];
            }
            for (int t = 0; t < kb; t++) A[(r0 + i) + (size_t)(k0 + t) * n] = pi[t];
        }
        // trailing update of the lower triangle: A_ij -= P_i . P_j
#pragma omp parallel for schedule(dynamic, 4)
        for (int j = 0; j < nr; j++) {
            const double *pj = &P[(size_t)j * kb];
            double *col = A + (size_t)(r0 + j) * n + r0;
            for (int i = j; i < nr; i++) {
                const double *pi = &P[(size_t)i * kb];
                double s = 0;
#pragma omp simd reduction(+ : s)
                for (int t = 0; t < kb; t++) s += pi[t] * pj[t];
                col[i] -= s;
            }
        }
    }
    return true;
}
// L L^T x = b in place (L from dense_cholesky), column-oriented sweeps with
// the row ranges split over threads
void dense_chol_solve(int n, const double *L, double *x) {
    const int NB = 64;
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = std::min(NB, n - k0), r0 = k0 + kb;
        for (int k = k0; k < r0; k++) {
            x[k] /= L[k + (size_t)k * n];
            for (int i = k + 1; i < r0; i++) x[i] -= L[i + (size_t)k * n] * x[k];
        }
        if (r0 >= n) break;
#pragma omp parallel for schedule(static) if (n - r0 > 4096)
        for (int c = r0; c < n; c += 512) {
            const int ce = std::min(n, c + 512);
            for (int k = k0; k < r0; k++) {
                const double xk = x[k];
                const double *lk = L + (size_t)k * n;
                for (int i = c; i < ce; i++) x[i] -= lk[i] * xk;
            }
        }
    }
    for (int k1 = n; k1 > 0; k1 -= NB) {
        const int k0 = std::max(0, k1 - NB);
        if (k1 < n) {
#pragma omp parallel for schedule(static) if (n - k1 > 4096)
            for (int i = k0; i < k1; i++) {
                const double *li = L + (size_t)i * n;
                double s = 0;
                for (int k = k1; k < n; k++) s += li[k] * x[k];
                x[i] -= s;
            }
        }
        for (int i = k1 - 1; i >= k0; i--) {
            double s = x[i];
            for (int k = i + 1; k < k1; k++) s -= L[k + (size_t)i * n] * x[k];
            x[i] = s / L[i + (size_t)i * n];
        }
    }
}
// SPD inverse via Cholesky: A <- A^-1 (full symmetric)
bool dense_spd_inverse(int n, std::vector<double> &A) {
    std::vector<double> L = A;
    if (!dense_cholesky(n, L.data())) return false;
    std::vector<double> col(n);
    for (int j = 0; j < n; j++) {
        std::fill(col.begin(), col.end(), 0.0);
        col[j] = 1.0;
        dense_chol_solve(n, L.data(), col.data());
        for (int i = 0; i < n; i++) A[i + (size_t)j * n] = col[i];
    }
    for (int j = 0; j < n; j++)
        for (int i = j + 1; i < n; i++) {
            double s = 0.5 * (A[i + (size_t)j * n] + A[j + (size_t)i * n]);
            A[i + (size_t)j * n] = A[j + (size_t)i * n] = s;
        }
    return true;
}

// Modified Gram-Schmidt (two passes) on the k columns of the n x k matrix V.
void mgs(int n, int k, double *V) {
    for (int pass = 0; pass < 2; pass++)
        for (int j = 0; j < k; j++) {
            double *vj = V + (size_t)j * n;
            for (int i = 0; i < j; i++) {
                const double *vi = V + (size_t)i * n;
                double d = 0;
                for (int r = 0; r < n; r++) d += vi[r] * vj[r];
                for (int r = 0; r < n; r++) vj[r] -= d * vi[r];
            }
            double nr = 0;
            for (int r = 0; r < n; r++) nr += vj[r] * vj[r];
            nr = std::sqrt(nr);
            for (int r = 0; r < n; r++) vj[r] /= nr;
        }
}

// ------------------------------------------------------------ subdomain ----
struct LamEntry { int lam; int dof; double sign; double scale; };

struct Subdomain {
    int S[3];
    int nn[3];                 // local nodes per direction
    int n;                     // local DOFs
    std::vector<int> l2g;      // local node -> global node
    // CSR of full symmetric K (unregularised)
    std::vector<int> rp, ci;
    std::vector<double> kv;
    std::vector<double> f;
    int kdim = 0;
    std::vector<double> R;     // n x kdim, orthonormal columns
    std::vector<LamEntry> lam; // sorted by lambda id
    std::vector<int> bdofs;    // boundary DOFs (touched by B1), ascending
    std::vector<int> bpos;     // local dof -> boundary index or -1
    // fix-point regularisation
    std::vector<int> fixdofs;
    std::vector<double> regmat;   // |fixdofs|^2 dense block
    Profile Kreg;                 // factor of K + RegMat
    // boundary-kernel mode
    std::vector<int> idofs;       // interior DOFs ascending
    std::vector<int> ipos;
    Profile Kii;
    std::vector<double> SD;       // m x m Schur of K on the boundary
    std::vector<double> SSD;      // m x m K_bb - K_bi diag(K_ii)^-1 K_ib (SUPER_DIRICHLET)
    std::vector<double> Fb;       // m x m (K^+)_bb for the active mode
    std::vector<double> u;        // primal of last solve

    double kdiag(int d) const {
        for (int q = rp[d]; q < rp[d + 1]; q++)
            if (ci[q] == d) return kv[q];
        return 0;
    }
    void matvec(const double *x, double *y) const {
        for (int i = 0; i < n; i++) {
            double s = 0;
            for (int q = rp[i]; q < rp[i + 1]; q++) s += kv[q] * x[ci[q]];
            y[i] = s;
        }
    }
};

struct Oracle {
    Grid G;
    oracle_options o;
    Loads L;
    std::vector<double> Ke1, Nint;
    std::vector<Subdomain> sub;
    int nlam = 0, ncoarse = 0;
    std::vector<double> cvec;
    std::vector<int> coff;         // coarse offset of subdomain
    std::vector<double> GGt;       // Cholesky factor, ncoarse^2
    // last solve
    std::vector<double> lambda;
    double t_setup = 0;

    Oracle(const oracle_grid_problem &p, const oracle_options &oo) : G(p), o(oo) {}

    // ---- subdomain assembly (physics.cpp:147-228 element loop) ----
    void assemble(Subdomain &sd, int s) {
        G.subpos(s, sd.S);
        int ex[3], org[3];
        G.sub_box(sd.S, ex, org);
        int nd = G.ndof;
        for (int q = 0; q < 3; q++) sd.nn[q] = ex[q] + 1;
        int nnode = sd.nn[0] * sd.nn[1] * sd.nn[2];
        sd.n = nnode * nd;
        sd.l2g.resize(nnode);
        for (int c = 0; c < sd.nn[2]; c++)
            for (int b = 0; b < sd.nn[1]; b++)
                for (int a = 0; a < sd.nn[0]; a++)
                    sd.l2g[a + sd.nn[0] * (b + sd.nn[1] * c)] = G.node(org[0] + a, org[1] + b, org[2] + c);
        // CSR pattern: 27-point node stencil
        sd.rp.assign(sd.n + 1, 0);
        std::vector<std::vector<int>> cols(sd.n);
        for (int c = 0; c < sd.nn[2]; c++)
            for (int b = 0; b < sd.nn[1]; b++)
                for (int a = 0; a < sd.nn[0]; a++) {
                    int ln = a + sd.nn[0] * (b + sd.nn[1] * c);
                    for (int dc = -1; dc <= 1; dc++)
                        for (int db = -1; db <= 1; db++)
                            for (int da = -1; da <= 1; da++) {
                                int a2 = a + da, b2 = b + db, c2 = c + dc;
                                if (a2 < 0 || b2 < 0 || c2 < 0 || a2 >= sd.nn[0] || b2 >= sd.nn[1] || c2 >= sd.nn[2]) continue;
                                int ln2 = a2 + sd.nn[0] * (b2 + sd.nn[1] * c2);
                                for (int d1 = 0; d1 < nd; d1++)
                                    for (int d2 = 0; d2 < nd; d2++) cols[ln * nd + d1].push_back(ln2 * nd + d2);
                            }
                }
        for (int i = 0; i < sd.n; i++) {
            std::sort(cols[i].begin(), cols[i].end());
            sd.rp[i + 1] = sd.rp[i] + (int)cols[i].size();
        }
        sd.ci.resize(sd.rp[sd.n]);
        for (int i = 0; i < sd.n; i++) std::copy(cols[i].begin(), cols[i].end(), sd.ci.begin() + sd.rp[i]);
        sd.kv.assign(sd.ci.size(), 0.0);
        sd.f.assign(sd.n, 0.0);
        int ne = 8 * nd;
        std::vector<double> fe(ne);
        for (int c = 0; c < ex[2]; c++)
            for (int b = 0; b < ex[1]; b++)
                for (int a = 0; a < ex[0]; a++) {
                    int gi = org[0] + a, gj = org[1] + b, gk = org[2] + c;
                    double sc = L.elem_scale[gi + G.NE[0] * (gj + G.NE[1] * gk)];
                    int dofs[24];
                    for (int v = 0; v < 8; v++) {
                        int ln = (a + HEX[v][0]) + sd.nn[0] * ((b + HEX[v][1]) + sd.nn[1] * (c + HEX[v][2]));
                        for (int d = 0; d < nd; d++) dofs[v * nd + d] = ln * nd + d;
                    }
                    for (int r = 0; r < ne; r++) {
                        int row = dofs[r];
                        for (int cc = 0; cc < ne; cc++) {
                            int col = dofs[cc];
                            int q = (int)(std::lower_bound(sd.ci.begin() + sd.rp[row], sd.ci.begin() + sd.rp[row + 1], col) - sd.ci.begin());
                            sd.kv[q] += sc * Ke1[r * ne + cc];
                        }
                    }
                    element_rhs(G, L, Nint, gi, gj, gk, fe.data());
                    for (int r = 0; r < ne; r++) sd.f[dofs[r]] += fe[r];
                }
    }

    // ---- kernels (structuralmechanics3d.cpp:34-99): orthonormal RBMs ----
    void kernels(Subdomain &sd) {
        int nd = G.ndof;
        int nnode = sd.n / nd;
        sd.kdim = nd == 3 ? 6 : 1;
        sd.R.assign((size_t)sd.n * sd.kdim, 0.0);
        if (nd == 1) {
            for (int i = 0; i < sd.n; i++) sd.R[i] = 1.0;
        } else {
            double cen[3] = {0, 0, 0};
            std::vector<double> X((size_t)nnode * 3);
            for (int ln = 0; ln < nnode; ln++) {
                int g = sd.l2g[ln];
                int gi = g % G.NN[0], gj = (g / G.NN[0]) % G.NN[1], gk = g / (G.NN[0] * G.NN[1]);
                X[3 * ln] = gi * G.h[0]; X[3 * ln + 1] = gj * G.h[1]; X[3 * ln + 2] = gk * G.h[2];
                for (int q = 0; q < 3; q++) cen[q] += X[3 * ln + q];
            }
            for (int q = 0; q < 3; q++) cen[q] /= nnode;
            int n = sd.n;
            for (int ln = 0; ln < nnode; ln++) {
                double x = X[3 * ln] - cen[0], y = X[3 * ln + 1] - cen[1], z = X[3 * ln + 2] - cen[2];
                for (int c = 0; c < 3; c++) sd.R[(size_t)c * n + 3 * ln + c] = 1.0;
                sd.R[(size_t)3 * n + 3 * ln + 0] = -y; sd.R[(size_t)3 * n + 3 * ln + 1] = x;
                sd.R[(size_t)4 * n + 3 * ln + 0] = -z; sd.R[(size_t)4 * n + 3 * ln + 2] = x;
                sd.R[(size_t)5 * n + 3 * ln + 1] = -z; sd.R[(size_t)5 * n + 3 * ln + 2] = y;
            }
        }
        mgs(sd.n, sd.kdim, sd.R.data());
    }

    // ---- fix-point regularisation (structuralmechanics3d.cpp:101-183) ----
    void regularise(Subdomain &sd) {
        int nd = G.ndof;
        int ex[3], org[3];
        G.sub_box(sd.S, ex, org);
        int pa[3][2];
        for (int q = 0; q < 3; q++) { pa[q][0] = ex[q] / 4; pa[q][1] = ex[q] - ex[q] / 4; }
        std::vector<int> fixnodes;
        for (int c = 0; c < 2; c++)
            for (int b = 0; b < 2; b++)
                for (int a = 0; a < 2; a++) fixnodes.push_back(pa[0][a] + sd.nn[0] * (pa[1][b] + sd.nn[1] * pa[2][c]));
        std::sort(fixnodes.begin(), fixnodes.end());
        fixnodes.erase(std::unique(fixnodes.begin(), fixnodes.end()), fixnodes.end());
        int nf = (int)fixnodes.size();
        int k = nd == 3 ? 6 : 1;
        int m = nf * nd;
        sd.fixdofs.resize(m);
        for (int i = 0; i < nf; i++)
            for (int d = 0; d < nd; d++) sd.fixdofs[i * nd + d] = fixnodes[i] * nd + d;
        // N: m x k (raw coordinates, as the reference's Nt rows)
        std::vector<double> N((size_t)m * k, 0.0);
        for (int i = 0; i < nf; i++) {
            int g = sd.l2g[fixnodes[i]];
            double x = (g % G.NN[0]) * G.h[0], y = ((g / G.NN[0]) % G.NN[1]) * G.h[1], z = (g / (G.NN[0] * G.NN[1])) * G.h[2];
            if (nd == 1) { N[i] = 1.0; continue; }
            for (int c = 0; c < 3; c++) N[(size_t)c * m + 3 * i + c] = 1.0;
            N[(size_t)3 * m + 3 * i + 0] = -y; N[(size_t)3 * m + 3 * i + 1] = x;
            N[(size_t)4 * m + 3 * i + 0] = -z; N[(size_t)4 * m + 3 * i + 2] = x;
            N[(size_t)5 * m + 3 * i + 1] = -z; N[(size_t)5 * m + 3 * i + 2] = y;
        }
        std::vector<double> NtN((size_t)k * k, 0.0);
        for (int a = 0; a < k; a++)
            for (int b = 0; b < k; b++) {
                double s = 0;
                for (int r = 0; r < m; r++) s += N[(size_t)a * m + r] * N[(size_t)b * m + r];
                NtN[a + (size_t)b * k] = s;
            }
        dense_spd_inverse(k, NtN);
        double maxd = 0;
        for (int i = 0; i < sd.n; i++) maxd = std::max(maxd, sd.kdiag(i));
        sd.regmat.assign((size_t)m * m, 0.0);
        for (int r = 0; r < m; r++)
            for (int c = 0; c < m; c++) {
                double s = 0;
                for (int a = 0; a < k; a++)
                    for (int b = 0; b < k; b++) s += N[(size_t)a * m + r] * NtN[a + (size_t)b * k] * N[(size_t)b * m + c];
                sd.regmat[r + (size_t)c * m] = s * maxd;
            }
    }

    void factor_kreg(Subdomain &sd) {
        int n = sd.n;
        std::vector<int> fst(n);
        for (int i = 0; i < n; i++) fst[i] = sd.ci[sd.rp[i]];
        int m = (int)sd.fixdofs.size();
        for (int r = 0; r < m; r++)
            for (int c = 0; c < m; c++) {
                int i = sd.fixdofs[r], j = sd.fixdofs[c];
                if (j < i) fst[i] = std::min(fst[i], j);
            }
        sd.Kreg.init(fst);
        for (int i = 0; i < n; i++)
            for (int q = sd.rp[i]; q < sd.rp[i + 1]; q++)
                if (sd.ci[q] <= i) sd.Kreg.at(i, sd.ci[q]) += sd.kv[q];
        for (int r = 0; r < m; r++)
            for (int c = 0; c < m; c++) {
                int i = sd.fixdofs[r], j = sd.fixdofs[c];
                if (j <= i) sd.Kreg.at(i, j) += sd.regmat[r + (size_t)c * m];
            }
        if (!sd.Kreg.factor()) throw std::string("K_reg factorisation failed");
    }

    // interior factor K_ii and dense Schur S_D = K_bb - K_bi K_ii^-1 K_ib
    // (clustercpu.cpp:167-333 with K_modif = K - RegMat = K)
    void schur(Subdomain &sd) {
        int n = sd.n;
        sd.bpos.assign(n, -1);
        sd.ipos.assign(n, -1);
        sd.idofs.clear();
        for (size_t q = 0; q < sd.bdofs.size(); q++) sd.bpos[sd.bdofs[q]] = (int)q;
        for (int i = 0; i < n; i++)
            if (sd.bpos[i] < 0) { sd.ipos[i] = (int)sd.idofs.size(); sd.idofs.push_back(i); }
        int ni = (int)sd.idofs.size(), m = (int)sd.bdofs.size();
        std::vector<int> fst(ni);
        for (int r = 0; r < ni; r++) {
            int i = sd.idofs[r];
            int f = r;
            for (int q = sd.rp[i]; q < sd.rp[i + 1]; q++)
                if (sd.ipos[sd.ci[q]] >= 0) f = std::min(f, sd.ipos[sd.ci[q]]);
            fst[r] = f;
        }
        sd.Kii.init(fst);
        for (int r = 0; r < ni; r++) {
            int i = sd.idofs[r];
            for (int q = sd.rp[i]; q < sd.rp[i + 1]; q++) {
                int c = sd.ipos[sd.ci[q]];
                if (c >= 0 && c <= r) sd.Kii.at(r, c) += sd.kv[q];
            }
        }
        if (ni > 0 && !sd.Kii.factor()) throw std::string("K_ii factorisation failed");
        sd.SD.assign((size_t)m * m, 0.0);
        std::vector<double> x(ni);
        for (int bj = 0; bj < m; bj++) {
            int j = sd.bdofs[bj];
            std::fill(x.begin(), x.end(), 0.0);
            for (int q = sd.rp[j]; q < sd.rp[j + 1]; q++) {
                int c = sd.ci[q];
                if (sd.ipos[c] >= 0) x[sd.ipos[c]] = sd.kv[q];          // K_ib(:, j) = K(i, j) by symmetry
                else sd.SD[sd.bpos[c] + (size_t)bj * m] += sd.kv[q];      // K_bb
            }
            if (ni > 0) sd.Kii.solve(x.data());
            // S(:, j) -= K_bi x
            for (int bi = 0; bi < m; bi++) {
                int i = sd.bdofs[bi];
                double s = 0;
                for (int q = sd.rp[i]; q < sd.rp[i + 1]; q++) {
                    int c = sd.ipos[sd.ci[q]];
                    if (c >= 0) s += sd.kv[q] * x[c];
                }
                sd.SD[bi + (size_t)bj * m] -= s;
            }
        }
        for (int j = 0; j < m; j++)
            for (int i = j + 1; i < m; i++) {
                double s = 0.5 * (sd.SD[i + (size_t)j * m] + sd.SD[j + (size_t)i * m]);
                sd.SD[i + (size_t)j * m] = sd.SD[j + (size_t)i * m] = s;
            }
    }

    // SUPER_DIRICHLET block (clustercpu.cpp:243-314 with diagonalized_K_rr):
    // K_bb - K_bi diag(K_ii)^-1 K_ib, K the unregularised stiffness
    void super_dirichlet(Subdomain &sd) {
        int n = sd.n, m = (int)sd.bdofs.size();
        std::vector<int> bpos(n, -1);
        for (int q = 0; q < m; q++) bpos[sd.bdofs[q]] = q;
        sd.SSD.assign((size_t)m * m, 0.0);
        for (int bj = 0; bj < m; bj++) {
            int j = sd.bdofs[bj];
            for (int q = sd.rp[j]; q < sd.rp[j + 1]; q++) {
                int c = sd.ci[q];
                if (bpos[c] >= 0) { sd.SSD[bpos[c] + (size_t)bj * m] += sd.kv[q]; continue; }
                // interior neighbour c of boundary DOF j: K_bc K_cc^-1 K_cj for every boundary b next to c
                double w = sd.kv[q] / sd.kdiag(c);
                for (int q2 = sd.rp[c]; q2 < sd.rp[c + 1]; q2++) {
                    int b = bpos[sd.ci[q2]];
                    if (b >= 0) sd.SSD[b + (size_t)bj * m] -= sd.kv[q2] * w;
                }
            }
        }
    }

    // boundary-kernel generalised inverse: F_b = (S_D + rho Q Q^T)^-1,
    // Q = orth(R_b), rho = max diag S_D.
    void boundary_inverse(Subdomain &sd) {
        int m = (int)sd.bdofs.size(), k = sd.kdim;
        std::vector<double> Q((size_t)m * k);
        for (int a = 0; a < k; a++)
            for (int bi = 0; bi < m; bi++) Q[(size_t)a * m + bi] = sd.R[(size_t)a * sd.n + sd.bdofs[bi]];
        mgs(m, k, Q.data());
        double rho = 0;
        for (int i = 0; i < m; i++) rho = std::max(rho, sd.SD[i + (size_t)i * m]);
        sd.Fb = sd.SD;
        for (int j = 0; j < m; j++)
            for (int i = 0; i < m; i++) {
                double s = 0;
                for (int a = 0; a < k; a++) s += Q[(size_t)a * m + i] * Q[(size_t)a * m + j];
                sd.Fb[i + (size_t)j * m] += rho * s;
            }
        if (!dense_spd_inverse(m, sd.Fb)) throw std::string("boundary Schur factorisation failed");
    }

    // K^+ x (in place, length n)
    void kplus(const Subdomain &sd, double *x) const {
        if (o.kplus == 0) { sd.Kreg.solve(x); return; }
        int ni = (int)sd.idofs.size(), m = (int)sd.bdofs.size();
        std::vector<double> t(ni), h(m), ub(m);
        for (int r = 0; r < ni; r++) t[r] = x[sd.idofs[r]];
        if (ni) sd.Kii.solve(t.data());
        for (int bi = 0; bi < m; bi++) {
            int i = sd.bdofs[bi];
            double s = x[i];
            for (int q = sd.rp[i]; q < sd.rp[i + 1]; q++) {
                int c = sd.ipos[sd.ci[q]];
                if (c >= 0) s -= sd.kv[q] * t[c];
            }
            h[bi] = s;
        }
        for (int bi = 0; bi < m; bi++) {
            double s = 0;
            for (int bj = 0; bj < m; bj++) s += sd.Fb[bi + (size_t)bj * m] * h[bj];
            ub[bi] = s;
        }
        std::vector<double> w(ni, 0.0);
        for (int r = 0; r < ni; r++) {
            int i = sd.idofs[r];
            double s = 0;
            for (int q = sd.rp[i]; q < sd.rp[i + 1]; q++) {
                int c = sd.bpos[sd.ci[q]];
                if (c >= 0) s += sd.kv[q] * ub[c];
            }
            w[r] = s;
        }
        if (ni) sd.Kii.solve(w.data());
        for (int r = 0; r < ni; r++) x[sd.idofs[r]] = t[r] - w[r];
        for (int bi = 0; bi < m; bi++) x[sd.bdofs[bi]] = ub[bi];
    }

    // ---- B1 (equalityconstraints.cpp:25-423), canonical lambda order ----
    void build_b1() {
        int nd = G.ndof;
        std::vector<int> subs;
        // subdomain-local dof of a global node inside subdomain s
        auto localdof = [&](int s, int gi, int gj, int gk, int c) {
            int S[3], ex[3], org[3];
            G.subpos(s, S);
            G.sub_box(S, ex, org);
            int a = gi - org[0], b = gj - org[1], cc = gk - org[2];
            return (a + (ex[0] + 1) * (b + (ex[1] + 1) * cc)) * nd + c;
        };
        cvec.clear();
        int lam = 0;
        // Dirichlet rows: every subdomain copy gets its own multiplier
        for (int gk = 0; gk < G.NN[2]; gk++)
            for (int gj = 0; gj < G.NN[1]; gj++)
                for (int gi = 0; gi < G.NN[0]; gi++)
                    for (int c = 0; c < nd; c++) {
                        double val;
                        if (!dirichlet_of(G, gi, gj, gk, c, &val)) continue;
                        G.node_subdomains(gi, gj, gk, subs);
                        for (int s : subs) {
                            sub[s].lam.push_back({lam, localdof(s, gi, gj, gk, c), 1.0, 1.0});
                            cvec.push_back(val);
                            lam++;
                        }
                    }
        // redundant gluing: n(n-1)/2 rows per shared free DOF, +1 on the lower subdomain
        for (int gk = 0; gk < G.NN[2]; gk++)
            for (int gj = 0; gj < G.NN[1]; gj++)
                for (int gi = 0; gi < G.NN[0]; gi++) {
                    G.node_subdomains(gi, gj, gk, subs);
                    int ns = (int)subs.size();
                    if (ns < 2) continue;
                    for (int c = 0; c < nd; c++) {
                        double val;
                        if (dirichlet_of(G, gi, gj, gk, c, &val)) continue;
                        std::vector<double> D(ns);
                        std::vector<int> ld(ns);
                        double sum = 0;
                        for (int q = 0; q < ns; q++) {
                            ld[q] = localdof(subs[q], gi, gj, gk, c);
                            D[q] = sub[subs[q]].kdiag(ld[q]);
                            sum += D[q];
                        }
                        for (int a = 0; a < ns; a++)
                            for (int b = a + 1; b < ns; b++) {
                                double sa = o.scaling ? D[b] / sum : 1.0 / ns;
                                double sb = o.scaling ? D[a] / sum : 1.0 / ns;
                                sub[subs[a]].lam.push_back({lam, ld[a], 1.0, sa});
                                sub[subs[b]].lam.push_back({lam, ld[b], -1.0, sb});
                                cvec.push_back(0.0);
                                lam++;
                            }
                    }
                }
        nlam = lam;
        for (auto &sd : sub) {
            std::vector<int> b;
            for (auto &e : sd.lam) b.push_back(e.dof);
            std::sort(b.begin(), b.end());
            b.erase(std::unique(b.begin(), b.end()), b.end());
            sd.bdofs = b;
        }
    }

    // ---- coarse problem: G = -(B R)^T, GG^T factor (itersolver.cpp:4931-5121) ----
    void coarse() {
        coff.assign(sub.size() + 1, 0);
        for (size_t s = 0; s < sub.size(); s++) coff[s + 1] = coff[s] + sub[s].kdim;
        ncoarse = coff.back();
        // lambda -> entries
        std::vector<std::vector<std::pair<int, int>>> ent(nlam);
        for (size_t s = 0; s < sub.size(); s++)
            for (size_t q = 0; q < sub[s].lam.size(); q++) ent[sub[s].lam[q].lam].push_back({(int)s, (int)q});
        GGt.assign((size_t)ncoarse * ncoarse, 0.0);
        for (int l = 0; l < nlam; l++)
            for (auto &e1 : ent[l])
                for (auto &e2 : ent[l]) {
                    const Subdomain &s1 = sub[e1.first], &s2 = sub[e2.first];
                    const LamEntry &a = s1.lam[e1.second], &b = s2.lam[e2.second];
                    for (int i = 0; i < s1.kdim; i++)
                        for (int j = 0; j < s2.kdim; j++)
                            GGt[(coff[e1.first] + i) + (size_t)(coff[e2.first] + j) * ncoarse] +=
                                a.sign * s1.R[(size_t)i * s1.n + a.dof] * b.sign * s2.R[(size_t)j * s2.n + b.dof];
                }
        if (!dense_cholesky(ncoarse, GGt.data())) throw std::string("GG^T factorisation failed");
    }

    void build() {
        double t0 = now();
        Loads LL = make_loads(G);
        L = LL;
        element_matrices(G.p.physics, G.h, G.p.poisson, Ke1, Nint);
        sub.assign(G.nsub, Subdomain());
#pragma omp parallel for schedule(dynamic)
        for (int s = 0; s < G.nsub; s++) {
            assemble(sub[s], s);
            kernels(sub[s]);
            regularise(sub[s]);
        }
        build_b1();
        std::string err;
#pragma omp parallel for schedule(dynamic)
        for (int s = 0; s < G.nsub; s++) {
            try {
                factor_kreg(sub[s]);
                if (o.preconditioner == 3 || o.kplus == 1) schur(sub[s]);

                if (o.kplus == 1) boundary_inverse(sub[s]);
            } catch (std::string &e) {
#pragma omp critical
                err = e;
            }
        }
        if (!err.empty()) throw err;
        coarse();
        t_setup = now() - t0;
    }

    // ---- dual operators ----
    void apply_F(const double *x, double *y) const {
        std::vector<std::vector<double>> part(sub.size());
#pragma omp parallel for schedule(dynamic)
        for (size_t s = 0; s < sub.size(); s++) {
            const Subdomain &sd = sub[s];
            std::vector<double> v(sd.n, 0.0);
            for (auto &e : sd.lam) v[e.dof] += e.sign * x[e.lam];
            kplus(sd, v.data());
            part[s].resize(sd.lam.size());
            for (size_t q = 0; q < sd.lam.size(); q++) part[s][q] = sd.lam[q].sign * v[sd.lam[q].dof];
        }
        std::fill(y, y + nlam, 0.0);
        for (size_t s = 0; s < sub.size(); s++)
            for (size_t q = 0; q < sub[s].lam.size(); q++) y[sub[s].lam[q].lam] += part[s][q];
    }

    void apply_prec(const double *x, double *y) const {
        if (o.preconditioner == 0) { std::copy(x, x + nlam, y); return; }
        std::vector<std::vector<double>> part(sub.size());
#pragma omp parallel for schedule(dynamic)
        for (size_t s = 0; s < sub.size(); s++) {
            const Subdomain &sd = sub[s];
            std::vector<double> v(sd.n, 0.0), w(sd.n, 0.0);
            for (auto &e : sd.lam) v[e.dof] += e.sign * e.scale * x[e.lam];
            if (o.preconditioner == 1) {
                sd.matvec(v.data(), w.data());
            } else if (o.preconditioner == 2) {
                w = v;
            } else if (o.preconditioner == 4) {
                // K_bb v_b - K_bi diag(K_ii)^-1 K_ib v_b straight from the sparse K
                // (the same block super_dirichlet() forms densely)
                std::vector<char> isb(sd.n, 0);
                for (int b : sd.bdofs) isb[b] = 1;
                std::vector<double> t(sd.n, 0.0);
                for (int i = 0; i < sd.n; i++) {
                    if (isb[i]) continue;
                    double acc = 0;
                    for (int q = sd.rp[i]; q < sd.rp[i + 1]; q++)
                        if (isb[sd.ci[q]]) acc += sd.kv[q] * v[sd.ci[q]];
                    t[i] = acc / sd.kdiag(i);
                }
                for (int b : sd.bdofs) {
                    double acc = 0;
                    for (int q = sd.rp[b]; q < sd.rp[b + 1]; q++) {
                        const int c = sd.ci[q];
                        acc += isb[c] ? sd.kv[q] * v[c] : -sd.kv[q] * t[c];
                    }
                    w[b] = acc;
                }
            } else {
                int m = (int)sd.bdofs.size();
                for (int bi = 0; bi < m; bi++) {
                    double acc = 0;
                    for (int bj = 0; bj < m; bj++) acc += sd.SD[bi + (size_t)bj * m] * v[sd.bdofs[bj]];
                    w[sd.bdofs[bi]] = acc;
                }
            }
            part[s].resize(sd.lam.size());
            for (size_t q = 0; q < sd.lam.size(); q++) part[s][q] = sd.lam[q].sign * sd.lam[q].scale * w[sd.lam[q].dof];
        }
        std::fill(y, y + nlam, 0.0);
        for (size_t s = 0; s < sub.size(); s++)
            for (size_t q = 0; q < sub[s].lam.size(); q++) y[sub[s].lam[q].lam] += part[s][q];
    }

    void apply_G(const double *x, double *d) const {
        for (size_t s = 0; s < sub.size(); s++) {
            const Subdomain &sd = sub[s];
            for (int j = 0; j < sd.kdim; j++) {
                double acc = 0;
                for (auto &e : sd.lam) acc += -e.sign * sd.R[(size_t)j * sd.n + e.dof] * x[e.lam];
                d[coff[s] + j] = acc;
            }
        }
    }
    void apply_Gt(const double *d, double *y) const {
        std::fill(y, y + nlam, 0.0);
        for (size_t s = 0; s < sub.size(); s++) {
            const Subdomain &sd = sub[s];
            for (auto &e : sd.lam) {
                double acc = 0;
                for (int j = 0; j < sd.kdim; j++) acc += -e.sign * sd.R[(size_t)j * sd.n + e.dof] * d[coff[s] + j];
                y[e.lam] += acc;
            }
        }
    }
    // Projector_Inv (itersolver.cpp:5541-5623): mode 0 y = x - G^T (GG^T)^-1 G x;
    // mode 1 (x in coarse space) y = G^T (GG^T)^-1 x; mode 2 y = (GG^T)^-1 G x
    void project(const double *x, double *y, int mode) const {
        std::vector<double> d(ncoarse);
        if (mode == 1) std::copy(x, x + ncoarse, d.begin());
        else apply_G(x, d.data());
        dense_chol_solve(ncoarse, GGt.data(), d.data());
        if (mode == 2) { std::copy(d.begin(), d.end(), y); return; }
        std::vector<double> t(nlam);
        apply_Gt(d.data(), t.data());
        if (mode == 0) for (int i = 0; i < nlam; i++) y[i] = x[i] - t[i];
        else std::copy(t.begin(), t.end(), y);
    }

    void vec_b(double *b) const {
        std::vector<std::vector<double>> part(sub.size());
#pragma omp parallel for schedule(dynamic)
        for (size_t s = 0; s < sub.size(); s++) {
            const Subdomain &sd = sub[s];
            std::vector<double> v = sd.f;
            kplus(sd, v.data());
            part[s].resize(sd.lam.size());
            for (size_t q = 0; q < sd.lam.size(); q++) part[s][q] = sd.lam[q].sign * v[sd.lam[q].dof];
        }
        for (int i = 0; i < nlam; i++) b[i] = -cvec[i];
        for (size_t s = 0; s < sub.size(); s++)
            for (size_t q = 0; q < sub[s].lam.size(); q++) b[sub[s].lam[q].lam] += part[s][q];
    }
    void vec_e(double *e) const {
        for (size_t s = 0; s < sub.size(); s++)
            for (int j = 0; j < sub[s].kdim; j++) {
                double acc = 0;
                for (int i = 0; i < sub[s].n; i++) acc += sub[s].R[(size_t)j * sub[s].n + i] * sub[s].f[i];
                e[coff[s] + j] = -acc;
            }
    }

    static double dot(const std::vector<double> &a, const std::vector<double> &b) {
        double s = 0;
        for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
        return s;
    }

    // ---- Solve_RegCG (itersolver.cpp:1406-1724) + primal (:372-431) ----
    void solve(oracle_stats *st) {
        double t0 = now();
        int n = nlam;
        std::vector<double> x(n), Ax(n), Ap(n), r(n), w(n, 0.0), wp(n), y(n, 0.0), yp(n), z(n), p(n), u(n), b(n);
        vec_b(b.data());
        std::vector<double> e(ncoarse);
        vec_e(e.data());
        project(e.data(), x.data(), 1);
        apply_F(x.data(), Ax.data());
        for (int i = 0; i < n; i++) r[i] = b[i] - Ax[i];
        project(r.data(), u.data(), 0);
        double tol1 = o.precision * std::sqrt(dot(u, u));
        double tol2 = o.precision * std::sqrt(dot(b, b));
        double tol = std::min(tol1, tol2);
        double norm = 0;
        int it = 0;
        for (int iter = 0; iter < o.max_iterations; iter++) {
            wp = w;
            yp = y;
            project(r.data(), w.data(), 0);
            if (o.preconditioner == 0) {
                y = w;
            } else {
                apply_prec(w.data(), z.data());
                project(z.data(), y.data(), 0);
            }
            if (iter == 0) {
                p = y;
            } else {
                double beta = dot(y, w) / dot(yp, wp);
                for (int i = 0; i < n; i++) p[i] = y[i] + beta * p[i];
            }
            apply_F(p.data(), Ap.data());
            double alpha = dot(y, w) / dot(p, Ap);
            for (int i = 0; i < n; i++) {
                x[i] += alpha * p[i];
                r[i] -= alpha * Ap[i];
            }
            norm = std::sqrt(dot(w, w));
            it = iter + 1;
            if (norm < tol) break;
        }
        lambda = x;
        std::vector<double> amp(ncoarse);
        project(r.data(), amp.data(), 2);
        // primal: u_i = K^+ (f_i - B_i^T lambda) + R_i amp_i
        double kp_num = 0, kp_den = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : kp_num, kp_den)
        for (size_t s = 0; s < sub.size(); s++) {
            Subdomain &sd = sub[s];
            std::vector<double> g = sd.f;
            for (auto &ent : sd.lam) g[ent.dof] -= ent.sign * x[ent.lam];
            std::vector<double> v = g;
            kplus(sd, v.data());
            for (int j = 0; j < sd.kdim; j++)
                for (int i = 0; i < sd.n; i++) v[i] += sd.R[(size_t)j * sd.n + i] * amp[coff[s] + j];
            sd.u = v;
            std::vector<double> Ku(sd.n);
            sd.matvec(v.data(), Ku.data());
            for (int i = 0; i < sd.n; i++) {
                kp_num += (Ku[i] - g[i]) * (Ku[i] - g[i]);
                kp_den += g[i] * g[i];
            }
        }
        double bu = 0, cc = 0;
        {
            std::vector<double> Bu(n, 0.0);
            for (auto &sd : sub)
                for (auto &ent : sd.lam) Bu[ent.lam] += ent.sign * sd.u[ent.dof];
            for (int i = 0; i < n; i++) { bu += (Bu[i] - cvec[i]) * (Bu[i] - cvec[i]); cc += cvec[i] * cvec[i]; }
        }
        if (st) {
            st->iterations = it;
            st->n_lambda = nlam;
            st->n_coarse = ncoarse;
            st->n_subdomains = (int)sub.size();
            st->final_norm = norm;
            st->tol = tol;
            st->norm_b = std::sqrt(dot(b, b));
            st->kkt_primal = kp_den > 0 ? std::sqrt(kp_num / kp_den) : std::sqrt(kp_num);
            double nc = std::sqrt(cc);
            if (std::fabs(nc) < 1) nc += 1.0;   // itersolver.cpp:330-332
            st->kkt_dual = std::sqrt(bu) / nc;
            st->time_setup = t_setup;
            st->time_solve = now() - t0;
        }
    }

    // new loads: each node's load split evenly over the subdomains sharing it
    void set_nodal_load(const double *f_nodal) {
        int nd = G.ndof;
        std::vector<int> mult(G.nnodes(), 0);
        for (auto &sd : sub)
            for (int g : sd.l2g) mult[g]++;
#pragma omp parallel for schedule(dynamic)
        for (size_t s = 0; s < sub.size(); s++) {
            Subdomain &sd = sub[s];
            for (size_t ln = 0; ln < sd.l2g.size(); ln++) {
                int g = sd.l2g[ln];
                for (int d = 0; d < nd; d++) sd.f[ln * nd + d] = f_nodal[(size_t)g * nd + d] / mult[g];
            }
        }
    }

    void nodal(double *u_nodal) const {
        int nd = G.ndof;
        int nn = G.nnodes();
        std::vector<int> cnt(nn, 0);
        std::fill(u_nodal, u_nodal + (size_t)nn * nd, 0.0);
        for (auto &sd : sub)
            for (size_t ln = 0; ln < sd.l2g.size(); ln++) {
                int g = sd.l2g[ln];
                cnt[g]++;
                for (int d = 0; d < nd; d++) u_nodal[(size_t)g * nd + d] += sd.u[ln * nd + d];
            }
        for (int g = 0; g < nn; g++)
            for (int d = 0; d < nd; d++) u_nodal[(size_t)g * nd + d] /= cnt[g];
    }
};

// ------------------------------------------------- global direct solve ----
int global_solve(const oracle_grid_problem &pp, double *u_nodal, double *rel_res_check, const double *u_in,
                 const double *f_nodal = nullptr, double *res_out = nullptr) {
    Grid G(pp);
    Loads L = make_loads(G);
    std::vector<double> Ke1, Nint;
    element_matrices(pp.physics, G.h, pp.poisson, Ke1, Nint);
    int nd = G.ndof;
    int n = G.nnodes() * nd;
    std::vector<char> fixed(n, 0);
    std::vector<double> uval(n, 0.0);
    for (int k = 0; k < G.NN[2]; k++)
        for (int j = 0; j < G.NN[1]; j++)
            for (int i = 0; i < G.NN[0]; i++)
                for (int c = 0; c < nd; c++) {
                    double v;
                    if (dirichlet_of(G, i, j, k, c, &v)) {
                        int g = G.node(i, j, k) * nd + c;
                        fixed[g] = 1;
                        uval[g] = v;
                    }
                }
    std::vector<int> fst(n);
    for (int k = 0; k < G.NN[2]; k++)
        for (int j = 0; j < G.NN[1]; j++)
            for (int i = 0; i < G.NN[0]; i++) {
                int g0 = G.node(std::max(i - 1, 0), std::max(j - 1, 0), std::max(k - 1, 0)) * nd;
                for (int c = 0; c < nd; c++) fst[G.node(i, j, k) * nd + c] = g0;
            }
    for (int g = 0; g < n; g++)
        if (fixed[g]) fst[g] = g;
    Profile P;
    if (!u_in) P.init(fst);  // the residual mode assembles K u only
    std::vector<double> f(n, 0.0);
    if (f_nodal) std::copy(f_nodal, f_nodal + n, f.begin());
    int ne = 8 * nd;
    std::vector<double> fe(ne);
    // residual mode: y = K u - f on free rows
    std::vector<double> Ku;
    if (u_in) Ku.assign(n, 0.0);
    for (int k = 0; k < G.NE[2]; k++)
        for (int j = 0; j < G.NE[1]; j++)
            for (int i = 0; i < G.NE[0]; i++) {
                double sc = L.elem_scale[i + G.NE[0] * (j + G.NE[1] * k)];
                int dofs[24];
                for (int v = 0; v < 8; v++)
                    for (int d = 0; d < nd; d++) dofs[v * nd + d] = G.node(i + HEX[v][0], j + HEX[v][1], k + HEX[v][2]) * nd + d;
                if (!f_nodal) element_rhs(G, L, Nint, i, j, k, fe.data());
                for (int r = 0; r < ne; r++) {
                    int row = dofs[r];
                    if (!f_nodal) f[row] += fe[r];
                    for (int c = 0; c < ne; c++) {
                        int col = dofs[c];
                        double kv = sc * Ke1[r * ne + c];
                        if (u_in) { Ku[row] += kv * u_in[col]; continue; }
                        if (fixed[row]) continue;
                        if (fixed[col]) { f[row] -= kv * uval[col]; continue; }
                        if (col <= row) P.at(row, col) += kv;
                    }
                }
            }
    if (u_in) {
        double num = 0, den = 0;
        for (int g = 0; g < n; g++) {
            if (res_out) res_out[g] = fixed[g] ? 0.0 : Ku[g] - f[g];
            if (fixed[g]) continue;
            num += (Ku[g] - f[g]) * (Ku[g] - f[g]);
            den += f[g] * f[g];
        }
        *rel_res_check = den > 0 ? std::sqrt(num / den) : std::sqrt(num);
        return 0;
    }
    for (int g = 0; g < n; g++)
        if (fixed[g]) { P.at(g, g) = 1.0; f[g] = uval[g]; }
    if (!P.factor()) { g_err = "global factorisation failed"; return 1; }
    P.solve(f.data());
    std::copy(f.begin(), f.end(), u_nodal);
    return 0;
}

}  // namespace

// ================================================================ C ABI ====
extern "C" {

const char *oracle_last_error(void) { return g_err.c_str(); }

int oracle_global_sizes(const oracle_grid_problem *p, int32_t *n_nodes, int32_t *ndof) {
    Grid G(*p);
    *n_nodes = G.nnodes();
    *ndof = G.ndof;
    return 0;
}

// unit-modulus element matrix of the grid (row-major, node-major DOFs)
int oracle_unit_element(const oracle_grid_problem *p, double *Ke) {
    Grid G(*p);
    std::vector<double> K, Nint;
    element_matrices(p->physics, G.h, p->poisson, K, Nint);
    std::copy(K.begin(), K.end(), Ke);
    return 0;
}

int oracle_global_solve(const oracle_grid_problem *p, double *u_nodal) {
    return global_solve(*p, u_nodal, nullptr, nullptr);
}

int oracle_global_solve_load(const oracle_grid_problem *p, const double *f_nodal, double *u_nodal) {
    return global_solve(*p, u_nodal, nullptr, nullptr, f_nodal);
}

int oracle_global_residual(const oracle_grid_problem *p, const double *u_nodal, double *rel_res) {
    return global_solve(*p, nullptr, rel_res, u_nodal);
}

// the per-subdomain conductivities that conductivity_seed draws (subdomain ids
// cluster-major, x fastest), for meshes whose subdomains are not uniform boxes
void oracle_seeded_conductivity(uint64_t seed, int32_t nsub, double base, double *out) {
    Rng r(seed);
    for (int s = 0; s < nsub; s++) out[s] = base * std::pow(10.0, 2.0 * r.uniform() - 1.0);
}

// K u - f per DOF (0 on fixed DOFs), nodal layout
int oracle_global_residual_vector(const oracle_grid_problem *p, const double *u_nodal, double *res) {
    double r = 0;
    return global_solve(*p, nullptr, &r, u_nodal, nullptr, res);
}

void *oracle_create(const oracle_grid_problem *p, const oracle_options *o) {
#ifdef _OPENMP
    if (o->threads > 0) omp_set_num_threads(o->threads);
#endif
    Oracle *h = new Oracle(*p, *o);
    try {
        h->build();
    } catch (std::string &e) {
        g_err = e;
        delete h;
        return nullptr;
    }
    return h;
}

void oracle_destroy(void *h) { delete (Oracle *)h; }

void oracle_sizes(void *hh, int32_t *s) {
    Oracle *h = (Oracle *)hh;
    s[0] = (int)h->sub.size();
    s[1] = h->nlam;
    s[2] = h->ncoarse;
    s[3] = h->G.nnodes();
    s[4] = h->G.ndof;
    s[5] = h->sub.empty() ? 0 : h->sub[0].kdim;
}

void oracle_subdomain_info(void *hh, int32_t i, int32_t *info) {
    Oracle *h = (Oracle *)hh;
    info[0] = h->sub[i].n;
    info[1] = (int)h->sub[i].bdofs.size();
    info[2] = (int)h->sub[i].lam.size();
}

void oracle_subdomain_boundary(void *hh, int32_t i, int32_t *b) {
    Oracle *h = (Oracle *)hh;
    std::copy(h->sub[i].bdofs.begin(), h->sub[i].bdofs.end(), b);
}

void oracle_subdomain_lambdas(void *hh, int32_t i, int32_t *lam, int32_t *dof, double *sign, double *scale) {
    Oracle *h = (Oracle *)hh;
    auto &v = h->sub[i].lam;
    for (size_t q = 0; q < v.size(); q++) {
        lam[q] = v[q].lam; dof[q] = v[q].dof; sign[q] = v[q].sign; scale[q] = v[q].scale;
    }
}

int oracle_subdomain_dense(void *hh, int32_t i, int32_t which, double *out) {
    Oracle *h = (Oracle *)hh;
    Subdomain &sd = h->sub[i];
    int n = sd.n, m = (int)sd.bdofs.size();
    if (which == 0 || which == 1) {
        std::fill(out, out + (size_t)n * n, 0.0);
        for (int r = 0; r < n; r++)
            for (int q = sd.rp[r]; q < sd.rp[r + 1]; q++) out[r + (size_t)sd.ci[q] * n] = sd.kv[q];
        if (which == 1) {
            int f = (int)sd.fixdofs.size();
            for (int a = 0; a < f; a++)
                for (int b = 0; b < f; b++) out[sd.fixdofs[a] + (size_t)sd.fixdofs[b] * n] += sd.regmat[a + (size_t)b * f];
        }
        return 0;
    }
    if (which == 2) { std::copy(sd.R.begin(), sd.R.end(), out); return 0; }
    if (which == 3) {
        if (sd.SD.empty()) h->schur(sd);
        std::copy(sd.SD.begin(), sd.SD.end(), out);
        return 0;
    }
    if (which == 4) {
        if (h->o.kplus == 1) { std::copy(sd.Fb.begin(), sd.Fb.end(), out); return 0; }
        std::vector<double> v(n);
        for (int bj = 0; bj < m; bj++) {
            std::fill(v.begin(), v.end(), 0.0);
            v[sd.bdofs[bj]] = 1.0;
            sd.Kreg.solve(v.data());
            for (int bi = 0; bi < m; bi++) out[bi + (size_t)bj * m] = v[sd.bdofs[bi]];
        }
        return 0;
    }
    if (which == 5) {
        if (sd.SSD.empty()) h->super_dirichlet(sd);
        std::copy(sd.SSD.begin(), sd.SSD.end(), out);
        return 0;
    }
    return 1;
}

void oracle_subdomain_rhs(void *hh, int32_t i, double *f) {
    Oracle *h = (Oracle *)hh;
    std::copy(h->sub[i].f.begin(), h->sub[i].f.end(), f);
}

void oracle_subdomain_l2g(void *hh, int32_t i, int32_t *l2g) {
    Oracle *h = (Oracle *)hh;
    std::copy(h->sub[i].l2g.begin(), h->sub[i].l2g.end(), l2g);
}

void oracle_vec_c(void *hh, double *c) {
    Oracle *h = (Oracle *)hh;
    std::copy(h->cvec.begin(), h->cvec.end(), c);
}
void oracle_vec_b(void *hh, double *b) { ((Oracle *)hh)->vec_b(b); }
void oracle_vec_e(void *hh, double *e) { ((Oracle *)hh)->vec_e(e); }
void oracle_apply_F(void *hh, const double *x, double *y) { ((Oracle *)hh)->apply_F(x, y); }
void oracle_apply_prec(void *hh, const double *x, double *y) { ((Oracle *)hh)->apply_prec(x, y); }
void oracle_project(void *hh, const double *x, double *y) { ((Oracle *)hh)->project(x, y, 0); }
void oracle_apply_G(void *hh, const double *x, double *y) { ((Oracle *)hh)->apply_G(x, y); }

int oracle_solve(void *hh, double *u_nodal, double *lambda, oracle_stats *st) {
    Oracle *h = (Oracle *)hh;
    try {
        h->solve(st);
    } catch (std::string &e) {
        g_err = e;
        return 1;
    }
    if (u_nodal) h->nodal(u_nodal);
    if (lambda) std::copy(h->lambda.begin(), h->lambda.end(), lambda);
    return 0;
}

void oracle_set_nodal_load(void *hh, const double *f_nodal) { ((Oracle *)hh)->set_nodal_load(f_nodal); }

void oracle_subdomain_primal(void *hh, int32_t i, double *u) {
    Oracle *h = (Oracle *)hh;
    std::copy(h->sub[i].u.begin(), h->sub[i].u.end(), u);
}

int oracle_cpu_sample(const oracle_grid_problem *p, const oracle_options *o, int32_t sample, double *times) {
#ifdef _OPENMP
    if (o->threads > 0) omp_set_num_threads(o->threads);
#endif
    try {
        Oracle h(*p, *o);
        h.L = make_loads(h.G);
        element_matrices(p->physics, h.G.h, p->poisson, h.Ke1, h.Nint);
        h.sub.assign(h.G.nsub, Subdomain());
#pragma omp parallel for schedule(dynamic)
        for (int s = 0; s < h.G.nsub; s++) {
            h.assemble(h.sub[s], s);
            h.kernels(h.sub[s]);
            h.regularise(h.sub[s]);
        }
        h.build_b1();
        int ns = std::min<int>(sample, h.G.nsub);
        double t0 = now();
#pragma omp parallel for schedule(dynamic)
        for (int s = 0; s < ns; s++) {
            h.factor_kreg(h.sub[s]);
            h.schur(h.sub[s]);
        }
        double t1 = now();
        int reps = 3;
#pragma omp parallel for schedule(dynamic)
        for (int s = 0; s < ns; s++) {
            std::vector<double> v(h.sub[s].n, 1.0);
            for (int r = 0; r < reps; r++) h.sub[s].Kreg.solve(v.data());
        }
        double t2 = now();
#pragma omp parallel for schedule(dynamic)
        for (int s = 0; s < ns; s++) {
            const Subdomain &sd = h.sub[s];
            int m = (int)sd.bdofs.size();
            std::vector<double> v(m, 1.0), w(m);
            for (int r = 0; r < reps; r++)
                for (int bi = 0; bi < m; bi++) {
                    double acc = 0;
                    for (int bj = 0; bj < m; bj++) acc += sd.SD[bi + (size_t)bj * m] * v[bj];
                    w[bi] = acc;
                }
        }
        double t3 = now();
        times[0] = (t1 - t0) / ns;
        times[1] = (t2 - t1) / (ns * reps);
        times[2] = (t3 - t2) / (ns * reps);
    } catch (std::string &e) {
        g_err = e;
        return 1;
    }
    return 0;
}

// CPU baseline with explicit dense S_i (the reference's Create_SC_w_Mat
// semantics: the dense Schur complement of each subdomain and its generalised
// inverse F_i, plus the Dirichlet block S_D from the same Schur step):
//   times[0]  seconds per subdomain of the explicit set-up (sample, threaded)
//   times[1]  seconds per subdomain of one dense F_i apply (sample)
//   times[2]  seconds per subdomain of one dense S_D apply (sample)
//   times[3]  coarse set-up at full size: G G^T formed and factorised (s)
//   times[4]  per PCG iteration at full size, everything but the subdomain
//             matrices: two projections, the B gathers/scatters of the F and
//             preconditioner applies, three dots and three vector updates (s)
//   times[5]  subdomains in the sample
int oracle_cpu_baseline(const oracle_grid_problem *p, const oracle_options *o, int32_t sample, double *times) {
#ifdef _OPENMP
    if (o->threads > 0) omp_set_num_threads(o->threads);
#endif
    try {
        Oracle h(*p, *o);
        h.L = make_loads(h.G);
        element_matrices(p->physics, h.G.h, p->poisson, h.Ke1, h.Nint);
        h.sub.assign(h.G.nsub, Subdomain());
#pragma omp parallel for schedule(dynamic)
        for (int s = 0; s < h.G.nsub; s++) {
            h.assemble(h.sub[s], s);
            h.kernels(h.sub[s]);
        }
        h.build_b1();
        const int ns = std::max(1, std::min<int>(sample, h.G.nsub));
        double t0 = now();
#pragma omp parallel for schedule(dynamic)
        for (int s = 0; s < ns; s++) {
            h.schur(h.sub[s]);
            h.boundary_inverse(h.sub[s]);
        }
        double t1 = now();
        const int reps = 3;
        auto gemv_time = [&](bool fi) {
            double ta = now();
#pragma omp parallel for schedule(dynamic)
            for (int s = 0; s < ns; s++) {
                const Subdomain &sd = h.sub[s];
                const int m = (int)sd.bdofs.size();
                const std::vector<double> &A = fi ? sd.Fb : sd.SD;
                std::vector<double> v(m, 1.0), w(m, 0.0);
                for (int r = 0; r < reps; r++) {
                    std::fill(w.begin(), w.end(), 0.0);
                    for (int j = 0; j < m; j++) {
                        const double *col = &A[(size_t)j * m];
                        const double vj = v[j];
                        for (int i = 0; i < m; i++) w[i] += col[i] * vj;
                    }
                    v[0] = w[0] * 1e-30 + 1.0;
                }
            }
            return (now() - ta) / (ns * reps);
        };
        times[0] = (t1 - t0) / ns;
        times[1] = gemv_time(true);
        times[2] = gemv_time(false);
        for (int s = 0; s < ns; s++) {  // release the dense blocks before the full-size steps
            std::vector<double>().swap(h.sub[s].SD);
            std::vector<double>().swap(h.sub[s].Fb);
            h.sub[s].Kii = Profile();
        }
        double t2 = now();
        h.coarse();
        double t3 = now();
        // one iteration's work outside the subdomain matrices
        const int n = h.nlam;
        std::vector<double> x(n), y(n), z(n), r(n);
        Rng rng(7);
        for (int i = 0; i < n; i++) x[i] = 2.0 * rng.uniform() - 1.0;
        auto gather_scatter = [&](const double *in, double *out, bool scaled) {
            std::vector<std::vector<double>> part(h.sub.size());
#pragma omp parallel for schedule(dynamic)
            for (size_t s = 0; s < h.sub.size(); s++) {
                const Subdomain &sd = h.sub[s];
                std::vector<double> v(sd.n, 0.0);
                for (auto &e : sd.lam) v[e.dof] += e.sign * (scaled ? e.scale : 1.0) * in[e.lam];
                part[s].resize(sd.lam.size());
                for (size_t q = 0; q < sd.lam.size(); q++)
                    part[s][q] = sd.lam[q].sign * (scaled ? sd.lam[q].scale : 1.0) * v[sd.lam[q].dof];
            }
            std::fill(out, out + n, 0.0);
            for (size_t s = 0; s < h.sub.size(); s++)
                for (size_t q = 0; q < h.sub[s].lam.size(); q++) out[h.sub[s].lam[q].lam] += part[s][q];
        };
        double t4 = now();
        h.project(x.data(), y.data(), 0);
        h.project(y.data(), z.data(), 0);
        gather_scatter(z.data(), r.data(), false);
        gather_scatter(r.data(), y.data(), true);
        double acc = 0;
        for (int k = 0; k < 3; k++) {
            double d = 0;
            for (int i = 0; i < n; i++) d += x[i] * y[i];
            acc += d;
            for (int i = 0; i < n; i++) x[i] += 1e-30 * d * y[i];
        }
        double t5 = now();
        times[3] = t3 - t2;
        times[4] = (t5 - t4) + 0.0 * acc;
        times[5] = ns;
    } catch (std::string &e) {
        g_err = e;
        return 1;
    }
    return 0;
}

}  // extern "C"
