// Structured-grid problem builder (host).  Stands in for the reference's GRID
// generator and assembler: src/input/meshgenerator/composition/gridgenerator.cpp:44-69,
// src/assembler/physics/physics.cpp:147-228 (element loop),
// src/assembler/constraints/equalityconstraints.cpp:25-423 (B1),
// src/assembler/physics/structuralmechanics3d.cpp:34-99 (kernels).
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/espreso_amd.h"

namespace esp {

struct GridSubdomain {
    int n = 0;                       // local DOFs
    std::vector<int> l2g;            // local node -> global node
    std::vector<int> rowptr, colidx; // full symmetric CSR
    std::vector<double> vals;
    std::vector<double> f;
    int kdim = 0;
    std::vector<double> R;           // n x kdim, orthonormal columns
    std::vector<int> lam, dof;       // lambda rows touching the subdomain, ascending
    std::vector<double> sign, scale;
    int box[6] = {0, 0, 0, 0, 0, 0};  // first global element (x, y, z) and elements per direction
    // hybrid FETI: cluster id and the B0 corner entries (cluster-local row, DOF, +1/-1)
    int cluster = 0;
    std::vector<int> b0row, b0dof;
    std::vector<double> b0sign;
};

struct GridProblem {
    int ndof = 0, nnodes = 0, nlam = 0;
    int NN[3] = {0, 0, 0};
    std::vector<double> escale;      // material scale of every element (ex + NEx (ey + NEy ez))
    std::vector<double> Ke;          // unit element matrix (8 nd x 8 nd, row-major)
    std::vector<GridSubdomain> sub;
    std::vector<double> c;
};

int grid_subdomain_count(const espreso_grid_problem &p);
int grid_node_count(const espreso_grid_problem &p);
// unit-modulus element matrix (row-major, node-major DOFs, reference corner order)
void grid_unit_element(const espreso_grid_problem &p, std::vector<double> &Ke);
// scaling: 1 = K-diagonal weights, 0 = 1/multiplicity; values = false leaves
// GridSubdomain::vals empty (structure only: the K_i values are assembled on
// the device from escale and Ke, in the same element order)
void grid_build(const espreso_grid_problem &p, int scaling, GridProblem &out, bool values = true);

}  // namespace esp
