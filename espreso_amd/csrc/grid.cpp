// Structured-grid problem builder (host side of the benchmark inputs).
//
// Element matrices use the closed-form tensor-product integrals of the
// trilinear box element (1D mass/stiffness/first-moment matrices combined with
// Lame constants), which equal the reference's 2x2x2 Gauss quadrature
// (hexahedron8.cpp:25-52, structuralmechanics3d.cpp:353-509) for box elements.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <tuple>

#include "grid.h"

namespace esp {

namespace {

struct SplitMix {
    uint64_t s;
    explicit SplitMix(uint64_t seed) : s(seed) {}
    double next01() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        return (double)(z >> 11) * 0x1.0p-53;
    }
};

// corner (x,y,z) bits of the 8 element nodes, reference ordering
const int CORNER[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

struct Box {
    const espreso_grid_problem &p;
    int nd;
    int NS[3], NE[3], NN[3];
    double h[3];
    int dpc;
    std::vector<int> slab[3], off[3];  // elements per subdomain slab, prefix sums
    explicit Box(const espreso_grid_problem &pp) : p(pp) {
        nd = p.physics == 0 ? 3 : 1;
        int q = 0;
        for (int d = 0; d < 3; d++) {
            if (p.clusters[d] < 1 || p.domains[d] < 1 || (!p.slab_elements && p.elements[d] < 1))
                throw std::invalid_argument("grid: clusters, domains and elements must be >= 1");
            NS[d] = p.clusters[d] * p.domains[d];
            slab[d].assign(NS[d], p.elements[d]);
            off[d].assign(NS[d] + 1, 0);
        }
        if (p.slab_elements)
            for (int d = 0; d < 3; d++)
                for (int i = 0; i < NS[d]; i++) {
                    slab[d][i] = p.slab_elements[q++];
                    if (slab[d][i] < 1) throw std::invalid_argument("grid: slab element counts must be >= 1");
                }
        for (int d = 0; d < 3; d++) {
            for (int i = 0; i < NS[d]; i++) off[d][i + 1] = off[d][i] + slab[d][i];
            NE[d] = off[d][NS[d]];
            NN[d] = NE[d] + 1;
            h[d] = p.length[d] / NE[d];
        }
        dpc = p.domains[0] * p.domains[1] * p.domains[2];
    }
    int nsub() const { return NS[0] * NS[1] * NS[2]; }
    // subdomain grid coordinate -> id (cluster-major, x fastest)
    int id_of(int sx, int sy, int sz) const {
        int cx = sx / p.domains[0], cy = sy / p.domains[1], cz = sz / p.domains[2];
        int dx = sx % p.domains[0], dy = sy % p.domains[1], dz = sz % p.domains[2];
        int cl = cx + p.clusters[0] * (cy + p.clusters[1] * cz);
        return cl * dpc + dx + p.domains[0] * (dy + p.domains[1] * dz);
    }
    void coord_of(int id, int *S) const {
        int cl = id / dpc, dd = id % dpc;
        int c[3] = {cl % p.clusters[0], (cl / p.clusters[0]) % p.clusters[1], cl / (p.clusters[0] * p.clusters[1])};
        int d[3] = {dd % p.domains[0], (dd / p.domains[0]) % p.domains[1], dd / (p.domains[0] * p.domains[1])};
        for (int q = 0; q < 3; q++) S[q] = c[q] * p.domains[q] + d[q];
    }
    // slab holding element index e along direction d
    int slab_of(int d, int e) const {
        return (int)(std::upper_bound(off[d].begin(), off[d].end(), e) - off[d].begin()) - 1;
    }
    // subdomain range owning node index g along direction d
    void owners(int d, int g, int &lo, int &hi) const {
        hi = std::min(slab_of(d, g), NS[d] - 1);
        lo = (g == off[d][hi] && hi > 0) ? hi - 1 : hi;
    }
    bool dirichlet(const int *g, int comp, double &val) const {
        for (int f = 0; f < 6; f++) {
            int d = f >> 1;
            bool on = (f & 1) ? g[d] == NN[d] - 1 : g[d] == 0;
            if (on && ((p.fixed[f] >> comp) & 1)) {
                val = p.fixed_value[f];
                return true;
            }
        }
        return false;
    }
};

// G_kl(a,b) = int dN_a/dx_k dN_b/dx_l over the box element, via 1D integrals.
void gradient_products(const double *h, double G[3][3][8][8]) {
    double M1[3][2][2], K1[3][2][2], C1[3][2][2];
    for (int d = 0; d < 3; d++)
        for (int p = 0; p < 2; p++)
            for (int q = 0; q < 2; q++) {
                M1[d][p][q] = h[d] * (p == q ? 2.0 : 1.0) / 6.0;
                K1[d][p][q] = (p == q ? 1.0 : -1.0) / h[d];
                C1[d][p][q] = p == 1 ? 0.5 : -0.5;  // int phi_p' phi_q
            }
    for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++)
            for (int a = 0; a < 8; a++)
                for (int b = 0; b < 8; b++) {
                    double v = 1.0;
                    for (int d = 0; d < 3; d++) {
                        int pa = CORNER[a][d], pb = CORNER[b][d];
                        if (k == l) v *= (d == k) ? K1[d][pa][pb] : M1[d][pa][pb];
                        else if (d == k) v *= C1[d][pa][pb];
                        else if (d == l) v *= C1[d][pb][pa];
                        else v *= M1[d][pa][pb];
                    }
                    G[k][l][a][b] = v;
                }
}

// unit-material element matrix (E = 1 or conductivity = 1), DOFs node-major
void unit_element(const espreso_grid_problem &p, const double *h, std::vector<double> &Ke) {
    static thread_local double G[3][3][8][8];
    gradient_products(h, G);
    if (p.physics == 0) {
        const double nu = p.poisson;
        const double lam = nu / ((1 + nu) * (1 - 2 * nu)), mu = 1.0 / (2 * (1 + nu));
        Ke.assign(24 * 24, 0.0);
        for (int a = 0; a < 8; a++)
            for (int b = 0; b < 8; b++) {
                double tr = G[0][0][a][b] + G[1][1][a][b] + G[2][2][a][b];
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++)
                        Ke[(3 * a + i) * 24 + 3 * b + j] = lam * G[i][j][a][b] + mu * G[j][i][a][b] + (i == j ? mu * tr : 0.0);
            }
    } else {
        Ke.assign(64, 0.0);
        for (int a = 0; a < 8; a++)
            for (int b = 0; b < 8; b++) Ke[a * 8 + b] = G[0][0][a][b] + G[1][1][a][b] + G[2][2][a][b];
    }
}

// consistent loads of element (ei,ej,ek) in global element coordinates
void element_load(const Box &B, const std::vector<double> &trac, int ei, int ej, int ek, double *fe) {
    const espreso_grid_problem &p = B.p;
    const int nd = B.nd;
    std::fill(fe, fe + 8 * nd, 0.0);
    const double vol8 = B.h[0] * B.h[1] * B.h[2] / 8.0;
    for (int a = 0; a < 8; a++) {
        if (nd == 3)
            for (int c = 0; c < 3; c++) fe[3 * a + c] = p.density * p.accel[c] * vol8;
        else
            fe[a] = p.heat_source * vol8;
    }
    const int e[3] = {ei, ej, ek};
    for (int face = 0; face < 6; face++) {
        const int d = face >> 1, side = face & 1;
        if (e[d] != (side ? B.NE[d] - 1 : 0)) continue;
        const int t1 = (d + 1) % 3, t2 = (d + 2) % 3;
        const double area = B.h[t1] * B.h[t2];
        // element nodes on this face, cyclic order in the (t1,t2) plane
        int ring[4];
        const int cyc[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        for (int q = 0; q < 4; q++)
            for (int a = 0; a < 8; a++)
                if (CORNER[a][d] == side && CORNER[a][t1] == cyc[q][0] && CORNER[a][t2] == cyc[q][1]) ring[q] = a;
        if (nd == 3 && p.pressure[face] != 0.0) {
            const double outward = side ? 1.0 : -1.0;
            for (int q = 0; q < 4; q++) fe[3 * ring[q] + d] -= p.pressure[face] * outward * area * 0.25;
        }
        if (face == 1 && !trac.empty()) {
            for (int q = 0; q < 4; q++) {
                for (int q2 = 0; q2 < 4; q2++) {
                    const int a2 = ring[q2];
                    const int node = (ej + CORNER[a2][1]) + B.NN[1] * (ek + CORNER[a2][2]);
                    const int dist = (q2 - q + 4) & 3;
                    const double w = area / 36.0 * (dist == 0 ? 4.0 : dist == 2 ? 1.0 : 2.0);
                    for (int c = 0; c < nd; c++) fe[nd * ring[q] + c] += w * trac[(size_t)node * nd + c];
                }
            }
        }
    }
}

void mgs_columns(int n, int k, double *V) {
    for (int pass = 0; pass < 2; pass++)
        for (int j = 0; j < k; j++) {
            double *v = V + (size_t)j * n;
            for (int i = 0; i < j; i++) {
                const double *u = V + (size_t)i * n;
                double dot = 0;
                for (int r = 0; r < n; r++) dot += u[r] * v[r];
                for (int r = 0; r < n; r++) v[r] -= dot * u[r];
            }
            double nrm = 0;
            for (int r = 0; r < n; r++) nrm += v[r] * v[r];
            nrm = 1.0 / std::sqrt(nrm);
            for (int r = 0; r < n; r++) v[r] *= nrm;
        }
}

}  // namespace

int grid_subdomain_count(const espreso_grid_problem &p) { return Box(p).nsub(); }
void grid_unit_element(const espreso_grid_problem &p, std::vector<double> &Ke) {
    Box B(p);
    unit_element(p, B.h, Ke);
}
int grid_node_count(const espreso_grid_problem &p) {
    Box b(p);
    return b.NN[0] * b.NN[1] * b.NN[2];
}

void grid_build(const espreso_grid_problem &p, int scaling, GridProblem &out, bool values) {
    Box B(p);
    const int nd = B.nd, ns = B.nsub();
    out.ndof = nd;
    out.nnodes = B.NN[0] * B.NN[1] * B.NN[2];
    for (int d = 0; d < 3; d++) out.NN[d] = B.NN[d];
    out.sub.assign(ns, GridSubdomain());

    // seeded material and traction fields
    const int nel = B.NE[0] * B.NE[1] * B.NE[2];
    std::vector<double> escale(nel, nd == 3 ? p.young : p.conductivity);
    if (nd == 3 && p.young_seed) {
        SplitMix r(p.young_seed);
        for (int e = 0; e < nel; e++) escale[e] = p.young * (0.1 + 0.9 * r.next01());
    }
    if (nd == 1 && p.conductivity_seed) {
        SplitMix r(p.conductivity_seed);
        std::vector<double> ks(ns);
        for (int s = 0; s < ns; s++) ks[s] = p.conductivity * std::pow(10.0, 2.0 * r.next01() - 1.0);
        for (int e = 0; e < nel; e++) {
            int ei = e % B.NE[0], ej = (e / B.NE[0]) % B.NE[1], ek = e / (B.NE[0] * B.NE[1]);
            escale[e] = ks[B.id_of(B.slab_of(0, ei), B.slab_of(1, ej), B.slab_of(2, ek))];
        }
    }
    if (p.material_elem)
        for (int e = 0; e < nel; e++) escale[e] = p.material_elem[e];
    out.escale = escale;
    std::vector<double> trac;
    if (p.traction_seed) {
        SplitMix r(p.traction_seed);
        trac.resize((size_t)B.NN[1] * B.NN[2] * nd);
        for (auto &t : trac) t = p.traction_scale * (2.0 * r.next01() - 1.0);
    }
    std::vector<double> Ke;
    unit_element(p, B.h, Ke);
    const int ne = 8 * nd;
    out.Ke = Ke;

#pragma omp parallel for schedule(dynamic)
    for (int s = 0; s < ns; s++) {
        GridSubdomain &sd = out.sub[s];
        int S[3];
        B.coord_of(s, S);
        const int ex = B.slab[0][S[0]], ey = B.slab[1][S[1]], ez = B.slab[2][S[2]];
        const int lx = ex + 1, ly = ey + 1, lz = ez + 1;
        const int nloc = lx * ly * lz;
        const int gx0 = B.off[0][S[0]], gy0 = B.off[1][S[1]], gz0 = B.off[2][S[2]];
        sd.box[0] = gx0;
        sd.box[1] = gy0;
        sd.box[2] = gz0;
        sd.box[3] = ex;
        sd.box[4] = ey;
        sd.box[5] = ez;
        sd.n = nloc * nd;
        sd.l2g.resize(nloc);
        // neighbour position table: for node ln, slot of the 27 offsets among its existing neighbours
        std::vector<int> nbr(27 * (size_t)nloc, -1), nnbr(nloc, 0);
        for (int z = 0; z < lz; z++)
            for (int y = 0; y < ly; y++)
                for (int x = 0; x < lx; x++) {
                    const int ln = x + lx * (y + ly * z);
                    sd.l2g[ln] = (gx0 + x) + B.NN[0] * ((gy0 + y) + B.NN[1] * (gz0 + z));
                    int cnt = 0;
                    for (int o = 0; o < 27; o++) {
                        int ox = o % 3 - 1, oy = (o / 3) % 3 - 1, oz = o / 9 - 1;
                        int x2 = x + ox, y2 = y + oy, z2 = z + oz;
                        if (x2 < 0 || y2 < 0 || z2 < 0 || x2 >= lx || y2 >= ly || z2 >= lz) continue;
                        nbr[27 * (size_t)ln + o] = cnt++;
                    }
                    nnbr[ln] = cnt;
                }
        sd.rowptr.assign(sd.n + 1, 0);
        for (int ln = 0; ln < nloc; ln++)
            for (int c = 0; c < nd; c++) sd.rowptr[ln * nd + c + 1] = nnbr[ln] * nd;
        for (int r = 0; r < sd.n; r++) sd.rowptr[r + 1] += sd.rowptr[r];
        sd.colidx.resize(sd.rowptr[sd.n]);
        if (values) sd.vals.assign(sd.rowptr[sd.n], 0.0);
        for (int z = 0; z < lz; z++)
            for (int y = 0; y < ly; y++)
                for (int x = 0; x < lx; x++) {
                    const int ln = x + lx * (y + ly * z);
                    for (int o = 0; o < 27; o++) {
                        int slot = nbr[27 * (size_t)ln + o];
                        if (slot < 0) continue;
                        int ln2 = (x + o % 3 - 1) + lx * ((y + (o / 3) % 3 - 1) + ly * (z + o / 9 - 1));
                        for (int c = 0; c < nd; c++)
                            for (int c2 = 0; c2 < nd; c2++)
                                sd.colidx[sd.rowptr[ln * nd + c] + slot * nd + c2] = ln2 * nd + c2;
                    }
                }
        sd.f.assign(sd.n, 0.0);
        std::vector<double> fe(ne);
        for (int z = 0; z < ez; z++)
            for (int y = 0; y < ey; y++)
                for (int x = 0; x < ex; x++) {
                    const int ei = gx0 + x, ej = gy0 + y, ek = gz0 + z;
                    const double sc = escale[ei + B.NE[0] * (ej + B.NE[1] * ek)];
                    int ln[8];
                    for (int a = 0; a < 8; a++)
                        ln[a] = (x + CORNER[a][0]) + lx * ((y + CORNER[a][1]) + ly * (z + CORNER[a][2]));
                    for (int a = 0; a < 8 && values; a++)
                        for (int b = 0; b < 8; b++) {
                            const int da = CORNER[b][0] - CORNER[a][0] + 1, db = CORNER[b][1] - CORNER[a][1] + 1,
                                      dc = CORNER[b][2] - CORNER[a][2] + 1;
                            const int slot = nbr[27 * (size_t)ln[a] + da + 3 * db + 9 * dc];
                            for (int i = 0; i < nd; i++)
                                for (int j = 0; j < nd; j++)
                                    sd.vals[sd.rowptr[ln[a] * nd + i] + slot * nd + j] += sc * Ke[(nd * a + i) * ne + nd * b + j];
                        }
                    element_load(B, trac, ei, ej, ek, fe.data());
                    for (int a = 0; a < 8; a++)
                        for (int i = 0; i < nd; i++) sd.f[ln[a] * nd + i] += fe[nd * a + i];
                }
        // kernel: rigid body modes about the subdomain centroid, orthonormalised
        sd.kdim = nd == 3 ? 6 : 1;
        sd.R.assign((size_t)sd.n * sd.kdim, 0.0);
        if (nd == 1) {
            std::fill(sd.R.begin(), sd.R.end(), 1.0);
        } else {
            const double cx = (gx0 + 0.5 * (lx - 1)) * B.h[0], cy = (gy0 + 0.5 * (ly - 1)) * B.h[1],
                         cz = (gz0 + 0.5 * (lz - 1)) * B.h[2];
            const size_t n = sd.n;
            for (int z = 0; z < lz; z++)
                for (int y = 0; y < ly; y++)
                    for (int x = 0; x < lx; x++) {
                        const size_t q = 3 * (size_t)(x + lx * (y + ly * z));
                        const double X = (gx0 + x) * B.h[0] - cx, Y = (gy0 + y) * B.h[1] - cy, Z = (gz0 + z) * B.h[2] - cz;
                        sd.R[q] = 1; sd.R[n + q + 1] = 1; sd.R[2 * n + q + 2] = 1;
                        sd.R[3 * n + q] = -Y; sd.R[3 * n + q + 1] = X;
                        sd.R[4 * n + q] = -Z; sd.R[4 * n + q + 2] = X;
                        sd.R[5 * n + q + 1] = -Z; sd.R[5 * n + q + 2] = Y;
                    }
        }
        mgs_columns(sd.n, sd.kdim, sd.R.data());
    }

    // ---- B1: Dirichlet rows first, then redundant gluing (both by global DOF) ----
    auto local_dof = [&](int s, const int *g, int c) {
        int S[3];
        B.coord_of(s, S);
        const int lx = B.slab[0][S[0]] + 1, ly = B.slab[1][S[1]] + 1;
        int x = g[0] - B.off[0][S[0]], y = g[1] - B.off[1][S[1]], z = g[2] - B.off[2][S[2]];
        return (x + lx * (y + ly * z)) * nd + c;
    };
    auto diag_of = [&](int s, int dof) {
        const GridSubdomain &sd = out.sub[s];
        if (values) {
            for (int q = sd.rowptr[dof]; q < sd.rowptr[dof + 1]; q++)
                if (sd.colidx[q] == dof) return sd.vals[q];
            return 0.0;
        }
        // the same sum as the assembled diagonal: elements around the node in
        // (z, y, x) order, each adding scale * Ke
        const int ln = dof / nd, i = dof % nd;
        const int lx = sd.box[3] + 1, ly = sd.box[4] + 1;
        const int x = ln % lx, y = (ln / lx) % ly, z = ln / (lx * ly);
        double v = 0.0;
        for (int ez0 = std::max(z - 1, 0); ez0 <= std::min(z, sd.box[5] - 1); ez0++)
            for (int ey0 = std::max(y - 1, 0); ey0 <= std::min(y, sd.box[4] - 1); ey0++)
                for (int ex0 = std::max(x - 1, 0); ex0 <= std::min(x, sd.box[3] - 1); ex0++) {
                    int a = 0;
                    while (CORNER[a][0] != x - ex0 || CORNER[a][1] != y - ey0 || CORNER[a][2] != z - ez0) a++;
                    const double sc = out.escale[(sd.box[0] + ex0) + B.NE[0] * ((sd.box[1] + ey0) + B.NE[1] * (sd.box[2] + ez0))];
                    v += sc * Ke[(nd * a + i) * ne + nd * a + i];
                }
        return v;
    };
    out.c.clear();
    int lam = 0;
    std::vector<int> owners;
    auto collect = [&](const int *g) {
        owners.clear();
        int lo[3], hi[3];
        for (int d = 0; d < 3; d++) B.owners(d, g[d], lo[d], hi[d]);
        for (int z = lo[2]; z <= hi[2]; z++)
            for (int y = lo[1]; y <= hi[1]; y++)
                for (int x = lo[0]; x <= hi[0]; x++) owners.push_back(B.id_of(x, y, z));
        std::sort(owners.begin(), owners.end());
    };
    auto add = [&](int s, int dof, double sg, double sc) {
        GridSubdomain &sd = out.sub[s];
        sd.lam.push_back(lam);
        sd.dof.push_back(dof);
        sd.sign.push_back(sg);
        sd.scale.push_back(sc);
    };
    int g[3];
    for (g[2] = 0; g[2] < B.NN[2]; g[2]++)
        for (g[1] = 0; g[1] < B.NN[1]; g[1]++)
            for (g[0] = 0; g[0] < B.NN[0]; g[0]++)
                for (int c = 0; c < nd; c++) {
                    double val;
                    if (!B.dirichlet(g, c, val)) continue;
                    collect(g);
                    for (int s : owners) {
                        add(s, local_dof(s, g, c), 1.0, 1.0);
                        out.c.push_back(val);
                        lam++;
                    }
                }
    std::vector<double> D;
    std::vector<int> ld;
    for (g[2] = 0; g[2] < B.NN[2]; g[2]++)
        for (g[1] = 0; g[1] < B.NN[1]; g[1]++)
            for (g[0] = 0; g[0] < B.NN[0]; g[0]++) {
                collect(g);
                const int k = (int)owners.size();
                if (k < 2) continue;
                for (int c = 0; c < nd; c++) {
                    double val;
                    if (B.dirichlet(g, c, val)) continue;
                    D.assign(k, 0.0);
                    ld.assign(k, 0);
                    double sum = 0;
                    for (int q = 0; q < k; q++) {
                        ld[q] = local_dof(owners[q], g, c);
                        D[q] = diag_of(owners[q], ld[q]);
                        sum += D[q];
                    }
                    for (int a = 0; a < k; a++)
                        for (int b = a + 1; b < k; b++) {
                            add(owners[a], ld[a], 1.0, scaling ? D[b] / sum : 1.0 / k);
                            add(owners[b], ld[b], -1.0, scaling ? D[a] / sum : 1.0 / k);
                            out.c.push_back(0.0);
                            lam++;
                        }
                }
            }
    out.nlam = lam;

    // ---- B0 corners per cluster (fetidata.cpp:146-187, equalityconstraints.cpp:504-552) ----
    // Interface nodes of a cluster are grouped by the full set of subdomains
    // owning them; a group touching other clusters ("external") gives a corner
    // every size/3 nodes, any other group its middle node.  Each corner glues
    // consecutive pairs of the cluster's owners, one row per DOF.
    const int ncl = p.clusters[0] * p.clusters[1] * p.clusters[2];
    for (int s = 0; s < ns; s++) out.sub[s].cluster = s / B.dpc;
    if (p.b0_type == 1) {
        // ---- B0 KERNELS (equalityconstraints.cpp:425-502; face pairs as fetidata.cpp:40-140) ----
        // Every pair (d, nd), d < nd, of subdomains of one cluster that share a
        // face gives one row per kernel column c of the master (the pair's
        // subdomain with more kernel columns, d on a tie): sum over the face
        // nodes n (ascending global id) and components k of
        // R_master(n, k, c) (u_d(n, k) - u_nd(n, k)).  Rows per cluster in the
        // order of the pairs.
#pragma omp parallel for schedule(dynamic)
        for (int cl = 0; cl < ncl; cl++) {
            std::vector<std::pair<int, int>> pairs;  // (d, axis of the face towards nd)
            for (int d = cl * B.dpc; d < (cl + 1) * B.dpc && d < ns; d++) {
                int S[3];
                B.coord_of(d, S);
                for (int a = 0; a < 3; a++) {
                    int T[3] = {S[0], S[1], S[2]};
                    T[a]++;
                    if (T[a] >= B.NS[a] || T[a] / p.domains[a] != S[a] / p.domains[a]) continue;  // other cluster
                    pairs.push_back({d, a});
                }
            }
            std::vector<std::tuple<int, int, int>> pl;  // (d, nd, axis), sorted by (d, nd)
            for (auto &pr : pairs) {
                int S[3];
                B.coord_of(pr.first, S);
                S[pr.second]++;
                pl.emplace_back(pr.first, B.id_of(S[0], S[1], S[2]), pr.second);
            }
            std::sort(pl.begin(), pl.end());
            int row = 0;
            for (auto &t : pl) {
                const int d = std::get<0>(t), nd_ = std::get<1>(t), a = std::get<2>(t);
                const int master = out.sub[nd_].kdim > out.sub[d].kdim ? nd_ : d;
                const GridSubdomain &M = out.sub[master];
                int S[3];
                B.coord_of(d, S);
                int lo[3], hi[3];
                for (int q = 0; q < 3; q++) {
                    lo[q] = B.off[q][S[q]];
                    hi[q] = B.off[q][S[q] + 1];
                }
                lo[a] = hi[a];  // the face between d and nd
                for (int c = 0; c < M.kdim; c++) {
                    for (int side = 0; side < 2; side++) {
                        const int s2 = side ? nd_ : d;
                        const double sg = side ? -1.0 : 1.0;
                        GridSubdomain &sd = out.sub[s2];
                        int g3[3];
                        for (g3[2] = lo[2]; g3[2] <= hi[2]; g3[2]++)
                            for (g3[1] = lo[1]; g3[1] <= hi[1]; g3[1]++)
                                for (g3[0] = lo[0]; g3[0] <= hi[0]; g3[0]++)
                                    for (int k = 0; k < nd; k++) {
                                        sd.b0row.push_back(row + c);
                                        sd.b0dof.push_back(local_dof(s2, g3, k));
                                        sd.b0sign.push_back(sg * M.R[local_dof(master, g3, k) + (size_t)c * M.n]);
                                    }
                    }
                }
                row += M.kdim;
            }
        }
        return;
    }
#pragma omp parallel for schedule(dynamic)
    for (int cl = 0; cl < ncl; cl++) {
        const int cc[3] = {cl % p.clusters[0], (cl / p.clusters[0]) % p.clusters[1], cl / (p.clusters[0] * p.clusters[1])};
        int lo3[3], hi3[3];
        for (int d = 0; d < 3; d++) {
            lo3[d] = B.off[d][cc[d] * p.domains[d]];
            hi3[d] = B.off[d][(cc[d] + 1) * p.domains[d]];
        }
        std::vector<std::vector<int>> keys;
        std::vector<std::vector<int>> members;
        std::vector<int> own, loc;
        int gg[3];
        for (gg[2] = lo3[2]; gg[2] <= hi3[2]; gg[2]++)
            for (gg[1] = lo3[1]; gg[1] <= hi3[1]; gg[1]++)
                for (gg[0] = lo3[0]; gg[0] <= hi3[0]; gg[0]++) {
                    own.clear();
                    int lo[3], hi[3];
                    for (int d = 0; d < 3; d++) B.owners(d, gg[d], lo[d], hi[d]);
                    for (int z = lo[2]; z <= hi[2]; z++)
                        for (int y = lo[1]; y <= hi[1]; y++)
                            for (int x = lo[0]; x <= hi[0]; x++) own.push_back(B.id_of(x, y, z));
                    std::sort(own.begin(), own.end());
                    int nloc_own = 0;
                    for (int o : own) nloc_own += o / B.dpc == cl;
                    if (nloc_own < 2) continue;
                    size_t gi = 0;
                    while (gi < keys.size() && keys[gi] != own) gi++;
                    if (gi == keys.size()) {
                        keys.push_back(own);
                        members.emplace_back();
                    }
                    members[gi].push_back(gg[0] + B.NN[0] * (gg[1] + B.NN[1] * gg[2]));
                }
        int row = 0;
        for (size_t gi = 0; gi < keys.size(); gi++) {
            loc.clear();
            for (int o : keys[gi])
                if (o / B.dpc == cl) loc.push_back(o);
            const bool external = loc.size() != keys[gi].size();
            const std::vector<int> &nodes = members[gi];
            const int sz = (int)nodes.size();
            std::vector<int> picks;
            if (external) {
                const int inc = std::max(1, sz / 3);
                for (int q = 0; q < sz; q += inc) picks.push_back(nodes[q]);
            } else {
                picks.push_back(nodes[sz / 2]);
            }
            for (int node : picks) {
                const int g3[3] = {node % B.NN[0], (node / B.NN[0]) % B.NN[1], node / (B.NN[0] * B.NN[1])};
                for (size_t a = 0; a + 1 < loc.size(); a++)
                    for (int c = 0; c < nd; c++) {
                        GridSubdomain &s1 = out.sub[loc[a]], &s2 = out.sub[loc[a + 1]];
                        s1.b0row.push_back(row);
                        s1.b0dof.push_back(local_dof(loc[a], g3, c));
                        s1.b0sign.push_back(1.0);
                        s2.b0row.push_back(row);
                        s2.b0dof.push_back(local_dof(loc[a + 1], g3, c));
                        s2.b0sign.push_back(-1.0);
                        row++;
                    }
            }
        }
    }
}

}  // namespace esp
