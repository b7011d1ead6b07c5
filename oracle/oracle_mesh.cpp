// oracle_mesh.cpp — TEST INFRASTRUCTURE (checker only; see oracle.h).
//
// Mesh generators, boundary facets, affine maps, reference elements and
// quadrature rules, restated from fenics-rs:
//   crates/fenics-mesh/src/generators.rs:24-117, mesh.rs:77-253,
//   crates/fenics-element/src/jacobian.rs:16-100, lagrange.rs:51-114,273-347,
//   crates/fenics-element/src/quadrature.rs:29-219.
#include "oracle_internal.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

namespace orc {

// jacobian.rs:18-41.  J = [v1-v0, v2-v0] (columns), det by cofactor
// expansion, J^{-1} = adj(J)/det (entry-wise division), invT = (J^{-1})^T.
// nalgebra's own rounding of det/inverse is parity unpinned (third-party).
bool affine2(const double* v0, const double* v1, const double* v2, Map2& m) {
    m.J[0][0] = v1[0] - v0[0];
    m.J[1][0] = v1[1] - v0[1];
    m.J[0][1] = v2[0] - v0[0];
    m.J[1][1] = v2[1] - v0[1];
    const double det = m.J[0][0] * m.J[1][1] - m.J[1][0] * m.J[0][1];
    m.det = det;
    if (std::fabs(det) < 1e-14) return false;
    const double i00 = m.J[1][1] / det;
    const double i01 = -m.J[0][1] / det;
    const double i10 = -m.J[1][0] / det;
    const double i11 = m.J[0][0] / det;
    m.invT[0][0] = i00;
    m.invT[0][1] = i10;
    m.invT[1][0] = i01;
    m.invT[1][1] = i11;
    return true;
}

// jacobian.rs:60-85, 3x3 adjugate over the determinant.
bool affine3(const double* v0, const double* v1, const double* v2, const double* v3, Map3& m) {
    for (int i = 0; i < 3; ++i) {
        m.J[i][0] = v1[i] - v0[i];
        m.J[i][1] = v2[i] - v0[i];
        m.J[i][2] = v3[i] - v0[i];
    }
    const double m11 = m.J[0][0], m12 = m.J[0][1], m13 = m.J[0][2];
    const double m21 = m.J[1][0], m22 = m.J[1][1], m23 = m.J[1][2];
    const double m31 = m.J[2][0], m32 = m.J[2][1], m33 = m.J[2][2];
    const double c00 = m22 * m33 - m32 * m23;
    const double c01 = m21 * m33 - m31 * m23;
    const double c02 = m21 * m32 - m31 * m22;
    const double det = m11 * c00 - m12 * c01 + m13 * c02;
    m.det = det;
    if (std::fabs(det) < 1e-14) return false;
    double inv[3][3];
    inv[0][0] = c00 / det;
    inv[0][1] = (m13 * m32 - m33 * m12) / det;
    inv[0][2] = (m12 * m23 - m22 * m13) / det;
    inv[1][0] = -c01 / det;
    inv[1][1] = (m11 * m33 - m31 * m13) / det;
    inv[1][2] = (m13 * m21 - m23 * m11) / det;
    inv[2][0] = c02 / det;
    inv[2][1] = (m12 * m31 - m32 * m11) / det;
    inv[2][2] = (m11 * m22 - m21 * m12) / det;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m.invT[i][j] = inv[j][i];
    return true;
}

// jacobian.rs:51-55 / 95-99: g = invT * g_ref, summed left to right.
void grad2(const Map2& m, const double* gr, double* g) {
    g[0] = m.invT[0][0] * gr[0] + m.invT[0][1] * gr[1];
    g[1] = m.invT[1][0] * gr[0] + m.invT[1][1] * gr[1];
}
void grad3(const Map3& m, const double* gr, double* g) {
    for (int i = 0; i < 3; ++i)
        g[i] = m.invT[i][0] * gr[0] + m.invT[i][1] * gr[1] + m.invT[i][2] * gr[2];
}

// jacobian.rs:44-48 / 88-92: x = v0 + J xi.
void map2(const Map2& m, const double* v0, const double* xi, double* x) {
    for (int i = 0; i < 2; ++i) x[i] = v0[i] + (m.J[i][0] * xi[0] + m.J[i][1] * xi[1]);
}
void map3(const Map3& m, const double* v0, const double* xi, double* x) {
    for (int i = 0; i < 3; ++i)
        x[i] = v0[i] + (m.J[i][0] * xi[0] + m.J[i][1] * xi[1] + m.J[i][2] * xi[2]);
}

// lagrange.rs:72-74 (P1Triangle) and 106-113 (P1Tetrahedron) reference gradients.
const double P1_TRI_GRAD[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
const double P1_TET_GRAD[4][3] = {
    {-1.0, -1.0, -1.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};

// quadrature.rs:83-219.
Rule rule(int dim, int degree) {
    Rule r;
    if (dim == 1) {
        if (degree <= 1) {
            r.pts = {{0.5, 0, 0}};
            r.w = {1.0};
        } else if (degree <= 3) {
            const double c = 1.0 / (2.0 * std::sqrt(3.0));
            r.pts = {{0.5 - c, 0, 0}, {0.5 + c, 0, 0}};
            r.w = {0.5, 0.5};
        } else {
            const double c = 0.5 * std::sqrt(3.0 / 5.0);
            r.pts = {{0.5 - c, 0, 0}, {0.5, 0, 0}, {0.5 + c, 0, 0}};
            r.w = {5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0};
        }
    } else if (dim == 2) {
        if (degree <= 1) {
            r.pts = {{1.0 / 3.0, 1.0 / 3.0, 0}};
            r.w = {0.5};
        } else if (degree == 2) {
            r.pts = {{0.5, 0.0, 0}, {0.5, 0.5, 0}, {0.0, 0.5, 0}};
            r.w = {1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0};
        } else {
            const double wc = -27.0 / 96.0, wo = 25.0 / 96.0;
            r.pts = {{1.0 / 3.0, 1.0 / 3.0, 0}, {0.6, 0.2, 0}, {0.2, 0.6, 0}, {0.2, 0.2, 0}};
            r.w = {wc, wo, wo, wo};
        }
    } else {
        if (degree <= 1) {
            r.pts = {{0.25, 0.25, 0.25}};
            r.w = {1.0 / 6.0};
        } else if (degree == 2) {
            const double s5 = std::sqrt(5.0);
            const double a = (5.0 - s5) / 20.0;
            const double b = (5.0 + 3.0 * s5) / 20.0;
            const double w = 1.0 / 24.0;
            r.pts = {{a, a, a}, {b, a, a}, {a, b, a}, {a, a, b}};
            r.w = {w, w, w, w};
        } else {
            const double wc = -2.0 / 15.0, wo = 3.0 / 40.0;
            r.pts = {{0.25, 0.25, 0.25},
                     {0.5, 1.0 / 6.0, 1.0 / 6.0},
                     {1.0 / 6.0, 0.5, 1.0 / 6.0},
                     {1.0 / 6.0, 1.0 / 6.0, 0.5},
                     {1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0}};
            r.w = {wc, wo, wo, wo, wo};
        }
    }
    return r;
}

// lagrange.rs:63-68 / 96-101 P1 basis values.
void p1_tri_basis(const double* xi, double* phi) {
    phi[0] = 1.0 - xi[0] - xi[1];
    phi[1] = xi[0];
    phi[2] = xi[1];
}
void p1_tet_basis(const double* xi, double* phi) {
    phi[0] = 1.0 - xi[0] - xi[1] - xi[2];
    phi[1] = xi[0];
    phi[2] = xi[1];
    phi[3] = xi[2];
}

// lagrange.rs:301-346, P2Tetrahedron (nodes: 4 corners, edges 01,12,02,03,13,23).
void p2_tet_basis(const double* xi, double* phi) {
    const double x = xi[0], y = xi[1], z = xi[2];
    const double l0 = 1.0 - x - y - z, l1 = x, l2 = y, l3 = z;
    phi[0] = l0 * (2.0 * l0 - 1.0);
    phi[1] = l1 * (2.0 * l1 - 1.0);
    phi[2] = l2 * (2.0 * l2 - 1.0);
    phi[3] = l3 * (2.0 * l3 - 1.0);
    phi[4] = 4.0 * l0 * l1;
    phi[5] = 4.0 * l1 * l2;
    phi[6] = 4.0 * l0 * l2;
    phi[7] = 4.0 * l0 * l3;
    phi[8] = 4.0 * l1 * l3;
    phi[9] = 4.0 * l2 * l3;
}
void p2_tet_grads(const double* xi, double g[10][3]) {
    const double x = xi[0], y = xi[1], z = xi[2];
    const double l0 = 1.0 - x - y - z, l1 = x, l2 = y, l3 = z;
    const double a0 = -(4.0 * l0 - 1.0);
    const double G[10][3] = {{a0, a0, a0},
                             {4.0 * l1 - 1.0, 0.0, 0.0},
                             {0.0, 4.0 * l2 - 1.0, 0.0},
                             {0.0, 0.0, 4.0 * l3 - 1.0},
                             {4.0 * (l0 - l1), -4.0 * l1, -4.0 * l1},
                             {4.0 * l2, 4.0 * l1, 0.0},
                             {-4.0 * l2, 4.0 * (l0 - l2), -4.0 * l2},
                             {-4.0 * l3, -4.0 * l3, 4.0 * (l0 - l3)},
                             {4.0 * l3, 0.0, 4.0 * l1},
                             {0.0, 4.0 * l3, 4.0 * l2}};
    for (int i = 0; i < 10; ++i)
        for (int k = 0; k < 3; ++k) g[i][k] = G[i][k];
}

}  // namespace orc

using namespace orc;

extern "C" {

// generators.rs:24-55: x-fastest vertices at (i/nx, j/ny); per square two
// triangles (v00,v10,v11), (v00,v11,v01), squares in j-then-i order.
int or_unit_square(uint64_t nx, uint64_t ny, double* verts, uint64_t* cells) {
    if (nx == 0 || ny == 0) return 3;
    uint64_t k = 0;
    for (uint64_t j = 0; j <= ny; ++j) {
        const double y = (double)j / (double)ny;
        for (uint64_t i = 0; i <= nx; ++i) {
            verts[2 * k] = (double)i / (double)nx;
            verts[2 * k + 1] = y;
            ++k;
        }
    }
    auto vid = [&](uint64_t i, uint64_t j) { return j * (nx + 1) + i; };
    uint64_t c = 0;
    for (uint64_t j = 0; j < ny; ++j)
        for (uint64_t i = 0; i < nx; ++i) {
            const uint64_t v00 = vid(i, j), v10 = vid(i + 1, j), v01 = vid(i, j + 1),
                           v11 = vid(i + 1, j + 1);
            uint64_t* t = cells + 6 * c;
            t[0] = v00; t[1] = v10; t[2] = v11;
            t[3] = v00; t[4] = v11; t[5] = v01;
            ++c;
        }
    return 0;
}

// generators.rs:58-111: vertices k,j,i order with coordinates (i/nx)*L ...;
// six Kuhn tetrahedra per hex sharing the p0-p7 diagonal.
int or_box_beam(double L, double W, double H, uint64_t nx, uint64_t ny, uint64_t nz,
                double* verts, uint64_t* cells) {
    if (nx == 0 || ny == 0 || nz == 0) return 3;
    uint64_t v = 0;
    for (uint64_t k = 0; k <= nz; ++k) {
        const double z = ((double)k / (double)nz) * H;
        for (uint64_t j = 0; j <= ny; ++j) {
            const double y = ((double)j / (double)ny) * W;
            for (uint64_t i = 0; i <= nx; ++i) {
                verts[3 * v] = ((double)i / (double)nx) * L;
                verts[3 * v + 1] = y;
                verts[3 * v + 2] = z;
                ++v;
            }
        }
    }
    auto vid = [&](uint64_t i, uint64_t j, uint64_t k) {
        return k * (nx + 1) * (ny + 1) + j * (nx + 1) + i;
    };
    uint64_t c = 0;
    for (uint64_t k = 0; k < nz; ++k)
        for (uint64_t j = 0; j < ny; ++j)
            for (uint64_t i = 0; i < nx; ++i) {
                const uint64_t p0 = vid(i, j, k), p1 = vid(i + 1, j, k), p2 = vid(i, j + 1, k),
                               p3 = vid(i + 1, j + 1, k), p4 = vid(i, j, k + 1),
                               p5 = vid(i + 1, j, k + 1), p6 = vid(i, j + 1, k + 1),
                               p7 = vid(i + 1, j + 1, k + 1);
                const uint64_t t[6][4] = {{p0, p1, p3, p7}, {p0, p3, p2, p7}, {p0, p2, p6, p7},
                                          {p0, p6, p4, p7}, {p0, p4, p5, p7}, {p0, p5, p1, p7}};
                for (int q = 0; q < 6; ++q)
                    for (int a = 0; a < 4; ++a) cells[4 * (6 * c + q) + a] = t[q][a];
                ++c;
            }
    return 0;
}

// mesh.rs:79-113: edges (1,2),(2,0),(0,1) keyed by the sorted pair; a facet
// is on the boundary when exactly one cell holds it.  The reference returns
// them in HashMap order; here they are sorted (order never matters to callers).
int or_boundary_facets_2d(uint64_t nc, const uint64_t* cells, uint64_t* facets_out,
                          uint64_t* n_out) {
    std::vector<std::array<uint64_t, 2>> e;
    e.reserve(3 * nc);
    for (uint64_t c = 0; c < nc; ++c) {
        const uint64_t* t = cells + 3 * c;
        const uint64_t loc[3][2] = {{t[1], t[2]}, {t[2], t[0]}, {t[0], t[1]}};
        for (auto& p : loc) e.push_back({std::min(p[0], p[1]), std::max(p[0], p[1])});
    }
    std::sort(e.begin(), e.end());
    uint64_t n = 0;
    for (size_t i = 0; i < e.size();) {
        size_t j = i;
        while (j < e.size() && e[j] == e[i]) ++j;
        if (j - i == 1) {
            if (facets_out) {
                facets_out[2 * n] = e[i][0];
                facets_out[2 * n + 1] = e[i][1];
            }
            ++n;
        }
        i = j;
    }
    *n_out = n;
    return 0;
}

// mesh.rs:145-210 (orientation is not needed by any caller here; faces are
// returned as sorted triples).
int or_boundary_facets_3d(uint64_t nc, const uint64_t* cells, uint64_t* facets_out,
                          uint64_t* n_out) {
    std::vector<std::array<uint64_t, 3>> f;
    f.reserve(4 * nc);
    for (uint64_t c = 0; c < nc; ++c) {
        const uint64_t* t = cells + 4 * c;
        const uint64_t loc[4][3] = {
            {t[1], t[2], t[3]}, {t[0], t[2], t[3]}, {t[0], t[1], t[3]}, {t[0], t[1], t[2]}};
        for (auto& q : loc) {
            std::array<uint64_t, 3> s = {q[0], q[1], q[2]};
            std::sort(s.begin(), s.end());
            f.push_back(s);
        }
    }
    std::sort(f.begin(), f.end());
    uint64_t n = 0;
    for (size_t i = 0; i < f.size();) {
        size_t j = i;
        while (j < f.size() && f[j] == f[i]) ++j;
        if (j - i == 1) {
            if (facets_out)
                for (int a = 0; a < 3; ++a) facets_out[3 * n + a] = f[i][a];
            ++n;
        }
        i = j;
    }
    *n_out = n;
    return 0;
}

// mesh.rs:79-113 / 145-210 with each facet's cell, local index and vertex
// order: 2D edges as the sorted pair (the HashMap key), 3D faces in the
// cell's local order, swapped to [f0, f2, f1] when the normal
// (f1 - f0) x (f2 - f0) points towards the opposite vertex.  Sorted by the
// sorted vertex tuple (the reference's HashMap order is random).
int or_boundary_facets_full(int dim, const double* V, uint64_t nc, const uint64_t* cells,
                            uint64_t* verts_out, uint64_t* cell_out, uint64_t* local_out,
                            uint64_t* n_out) {
    struct F {
        std::array<uint64_t, 3> key;
        uint64_t cell, local;
    };
    const int K = dim + 1;
    std::vector<F> all;
    all.reserve(K * nc);
    for (uint64_t c = 0; c < nc; ++c) {
        const uint64_t* t = cells + K * c;
        for (int a = 0; a < K; ++a) {
            std::array<uint64_t, 3> k = {0, 0, 0};
            int m = 0;
            if (dim == 2) {  // local facet a: (1,2), (2,0), (0,1)
                k[0] = t[(a + 1) % 3];
                k[1] = t[(a + 2) % 3];
                m = 2;
            } else {
                for (int b = 0; b < 4; ++b)
                    if (b != a) k[m++] = t[b];
            }
            std::sort(k.begin(), k.begin() + m);
            all.push_back({k, c, (uint64_t)a});
        }
    }
    std::sort(all.begin(), all.end(), [](const F& x, const F& y) {
        return x.key != y.key ? x.key < y.key : x.cell < y.cell;
    });
    uint64_t n = 0;
    for (size_t i = 0; i < all.size();) {
        size_t j = i;
        while (j < all.size() && all[j].key == all[i].key) ++j;
        if (j - i == 1) {
            const F& f = all[i];
            if (verts_out) {
                if (dim == 2) {
                    verts_out[2 * n] = f.key[0];
                    verts_out[2 * n + 1] = f.key[1];
                } else {
                    const uint64_t* t = cells + 4 * f.cell;
                    uint64_t fv[3];
                    int m = 0;
                    for (int b = 0; b < 4; ++b)
                        if (b != (int)f.local) fv[m++] = t[b];
                    const double* va = V + 3 * fv[0];
                    const double* vb = V + 3 * fv[1];
                    const double* vc = V + 3 * fv[2];
                    const double* vo = V + 3 * t[f.local];
                    const double ab[3] = {vb[0] - va[0], vb[1] - va[1], vb[2] - va[2]};
                    const double ac[3] = {vc[0] - va[0], vc[1] - va[1], vc[2] - va[2]};
                    const double nrm[3] = {ab[1] * ac[2] - ab[2] * ac[1],
                                           ab[2] * ac[0] - ab[0] * ac[2],
                                           ab[0] * ac[1] - ab[1] * ac[0]};
                    const double tf[3] = {va[0] - vo[0], va[1] - vo[1], va[2] - vo[2]};
                    const double dot = nrm[0] * tf[0] + nrm[1] * tf[1] + nrm[2] * tf[2];
                    verts_out[3 * n] = fv[0];
                    verts_out[3 * n + 1] = dot >= 0.0 ? fv[1] : fv[2];
                    verts_out[3 * n + 2] = dot >= 0.0 ? fv[2] : fv[1];
                }
            }
            if (cell_out) cell_out[n] = f.cell;
            if (local_out) local_out[n] = f.local;
            ++n;
        }
        i = j;
    }
    *n_out = n;
    return 0;
}

// poisson.rs:266-287: facet vertices, sorted and deduplicated.
static int boundary_dofs(int dim, uint64_t nc, const uint64_t* cells, uint64_t* out,
                         uint64_t* n_out) {
    uint64_t nf = 0;
    const int k = dim;  // vertices per facet
    if (dim == 2) or_boundary_facets_2d(nc, cells, nullptr, &nf);
    else or_boundary_facets_3d(nc, cells, nullptr, &nf);
    std::vector<uint64_t> f(nf * k);
    if (dim == 2) or_boundary_facets_2d(nc, cells, f.data(), &nf);
    else or_boundary_facets_3d(nc, cells, f.data(), &nf);
    std::sort(f.begin(), f.end());
    f.erase(std::unique(f.begin(), f.end()), f.end());
    if (out) std::copy(f.begin(), f.end(), out);
    *n_out = f.size();
    return 0;
}
int or_boundary_dofs_2d(uint64_t nc, const uint64_t* cells, uint64_t* out, uint64_t* n_out) {
    return boundary_dofs(2, nc, cells, out, n_out);
}
int or_boundary_dofs_3d(uint64_t nc, const uint64_t* cells, uint64_t* out, uint64_t* n_out) {
    return boundary_dofs(3, nc, cells, out, n_out);
}

// mesh.rs:116-126.
double or_total_volume_2d(const double* V, uint64_t nc, const uint64_t* cells) {
    double s = -0.0;
    for (uint64_t c = 0; c < nc; ++c) {
        const double* v0 = V + 2 * cells[3 * c];
        const double* v1 = V + 2 * cells[3 * c + 1];
        const double* v2 = V + 2 * cells[3 * c + 2];
        s += 0.5 * std::fabs((v1[0] - v0[0]) * (v2[1] - v0[1]) - (v1[1] - v0[1]) * (v2[0] - v0[0]));
    }
    return s;
}
// mesh.rs:213-231.
double or_total_volume_3d(const double* V, uint64_t nc, const uint64_t* cells) {
    double s = -0.0;
    for (uint64_t c = 0; c < nc; ++c) {
        const double* v0 = V + 3 * cells[4 * c];
        const double* v1 = V + 3 * cells[4 * c + 1];
        const double* v2 = V + 3 * cells[4 * c + 2];
        const double* v3 = V + 3 * cells[4 * c + 3];
        const double a[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
        const double b[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
        const double cc[3] = {v3[0] - v0[0], v3[1] - v0[1], v3[2] - v0[2]};
        const double det = a[0] * (b[1] * cc[2] - b[2] * cc[1]) - a[1] * (b[0] * cc[2] - b[2] * cc[0]) +
                           a[2] * (b[0] * cc[1] - b[1] * cc[0]);
        s += std::fabs(det) / 6.0;
    }
    return s;
}
// mesh.rs:129-140.
double or_boundary_measure_2d(const double* V, uint64_t nc, const uint64_t* cells) {
    uint64_t nf = 0;
    or_boundary_facets_2d(nc, cells, nullptr, &nf);
    std::vector<uint64_t> f(2 * nf);
    or_boundary_facets_2d(nc, cells, f.data(), &nf);
    double s = -0.0;
    for (uint64_t i = 0; i < nf; ++i) {
        const double* p0 = V + 2 * f[2 * i];
        const double* p1 = V + 2 * f[2 * i + 1];
        const double dx = p1[0] - p0[0], dy = p1[1] - p0[1];
        s += std::sqrt(dx * dx + dy * dy);
    }
    return s;
}
// mesh.rs:234-253.
double or_boundary_measure_3d(const double* V, uint64_t nc, const uint64_t* cells) {
    uint64_t nf = 0;
    or_boundary_facets_3d(nc, cells, nullptr, &nf);
    std::vector<uint64_t> f(3 * nf);
    or_boundary_facets_3d(nc, cells, f.data(), &nf);
    double s = -0.0;
    for (uint64_t i = 0; i < nf; ++i) {
        const double* p0 = V + 3 * f[3 * i];
        const double* p1 = V + 3 * f[3 * i + 1];
        const double* p2 = V + 3 * f[3 * i + 2];
        const double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        const double b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        const double cr[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                              a[0] * b[1] - a[1] * b[0]};
        s += 0.5 * std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    }
    return s;
}

int or_affine_map_2d(const double* v, double* J, double* invT, double* det) {
    Map2 m;
    const bool ok = affine2(v, v + 2, v + 4, m);
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            J[2 * i + j] = m.J[i][j];
            invT[2 * i + j] = ok ? m.invT[i][j] : 0.0;
        }
    *det = m.det;
    return ok ? 0 : 3;
}
int or_affine_map_3d(const double* v, double* J, double* invT, double* det) {
    Map3 m;
    const bool ok = affine3(v, v + 3, v + 6, v + 9, m);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            J[3 * i + j] = m.J[i][j];
            invT[3 * i + j] = ok ? m.invT[i][j] : 0.0;
        }
    *det = m.det;
    return ok ? 0 : 3;
}
int or_quadrature(int dim, int degree, double* points, double* weights, int* npts) {
    const Rule r = rule(dim, degree);
    *npts = (int)r.w.size();
    for (size_t q = 0; q < r.w.size(); ++q) {
        for (int d = 0; d < dim; ++d) points[q * dim + d] = r.pts[q][d];
        weights[q] = r.w[q];
    }
    return 0;
}
void or_p2_tet_basis(const double* xi, double* out10) { p2_tet_basis(xi, out10); }
void or_p2_tet_grads(const double* xi, double* out30) {
    double g[10][3];
    p2_tet_grads(xi, g);
    for (int i = 0; i < 10; ++i)
        for (int k = 0; k < 3; ++k) out30[3 * i + k] = g[i][k];
}

}  // extern "C"
