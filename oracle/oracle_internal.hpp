// oracle_internal.hpp — TEST INFRASTRUCTURE (checker only; see oracle.h).
// Shared helpers of the CPU FP64 restatement.
#pragma once
#include <array>
#include <cstdint>
#include <vector>

#include "oracle.h"

namespace orc {

struct Map2 {
    double J[2][2];
    double invT[2][2];
    double det;
};
struct Map3 {
    double J[3][3];
    double invT[3][3];
    double det;
};
struct Rule {
    std::vector<std::array<double, 3>> pts;
    std::vector<double> w;
};

bool affine2(const double* v0, const double* v1, const double* v2, Map2& m);
bool affine3(const double* v0, const double* v1, const double* v2, const double* v3, Map3& m);
void grad2(const Map2& m, const double* gr, double* g);
void grad3(const Map3& m, const double* gr, double* g);
void map2(const Map2& m, const double* v0, const double* xi, double* x);
void map3(const Map3& m, const double* v0, const double* xi, double* x);
Rule rule(int dim, int degree);
void p1_tri_basis(const double* xi, double* phi);
void p1_tet_basis(const double* xi, double* phi);
void p2_tet_basis(const double* xi, double* phi);
void p2_tet_grads(const double* xi, double g[10][3]);
extern const double P1_TRI_GRAD[3][2];
extern const double P1_TET_GRAD[4][3];

struct Trip {
    uint32_t row, col;
    double val;
};

}  // namespace orc

// faer SparseColMat<usize, f64> raw parts.
struct or_csc {
    uint64_t n = 0;
    std::vector<uint64_t> col_ptr;
    std::vector<uint32_t> row_idx;
    std::vector<double> vals;
};

struct or_p2 {
    uint64_t nv = 0, nc = 0, ne = 0;
    std::vector<double> verts;       // corner vertices [nv][3]
    std::vector<uint64_t> cells;     // corner connectivity [nc][4]
    std::vector<uint64_t> cell_dofs; // [nc][10]
    std::vector<std::array<uint64_t, 2>> edges;  // sorted (min, max)
    std::vector<uint64_t> erank;  // edge dof number - nv, by position in edges
};

namespace orc {
or_csc* csc_from_trips(uint64_t n, std::vector<Trip>& t);
int nthreads_or_default(int nthreads);
}  // namespace orc
