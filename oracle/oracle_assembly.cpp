// oracle_assembly.cpp — TEST INFRASTRUCTURE (checker only; see oracle.h).
//
// Cell-wise assembly into canonical CSC, restated from fenics-rs:
//   crates/fenics-assembly/src/poisson.rs:23-263   (stiffness, rhs, Dirichlet)
//   crates/fenics-assembly/src/elasticity.rs:14-227 (elasticity, body force, BC)
//   crates/fenics-assembly/src/mass.rs:10-44        (P1 mass, 2D)
//   crates/fenics-solver/src/poisson_solver.rs:62-96,135-172 (L2 error)
// The triplet -> CSC step restates faer 0.20.2 SparseColMat::try_new_from_triplets
// (third party, not in the reference tree): entries sorted by (col, row),
// duplicates summed, explicit zeros kept.  Duplicates are summed in triplet
// order (the order rayon's collect preserves, i.e. ascending cell index);
// faer's own summation order is parity unpinned at the last ulp.
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "oracle_internal.hpp"

namespace orc {

int nthreads_or_default(int nthreads) {
    if (nthreads > 0) return nthreads;
    const unsigned h = std::thread::hardware_concurrency();
    return h ? (int)h : 1;
}

template <class F>
static void parallel_ranges(uint64_t n, int nth, F&& f) {
    nth = std::max(1, std::min<int>(nth, (int)std::max<uint64_t>(1, n)));
    std::vector<std::thread> th;
    for (int t = 0; t < nth; ++t) {
        const uint64_t a = n * t / nth, b = n * (t + 1) / nth;
        th.emplace_back([&, t, a, b] { f(t, a, b); });
    }
    for (auto& x : th) x.join();
}

or_csc* csc_from_trips(uint64_t n, std::vector<Trip>& t) {
    for (const Trip& e : t)
        if (e.row >= n || e.col >= n) return nullptr;
    std::vector<uint64_t> start(n + 1, 0);
    for (const Trip& e : t) start[e.col + 1]++;
    for (uint64_t c = 0; c < n; ++c) start[c + 1] += start[c];
    std::vector<uint32_t> rows(t.size());
    std::vector<double> vals(t.size());
    {
        std::vector<uint64_t> cur(start.begin(), start.end() - 1);
        for (const Trip& e : t) {
            const uint64_t p = cur[e.col]++;
            rows[p] = e.row;
            vals[p] = e.val;
        }
    }
    std::vector<Trip>().swap(t);
    // per column: stable sort by row, then sum runs of equal rows in order
    std::vector<uint64_t> ucount(n, 0);
    const int nth = nthreads_or_default(0);
    parallel_ranges(n, nth, [&](int, uint64_t c0, uint64_t c1) {
        std::vector<uint32_t> idx;
        std::vector<uint32_t> r2;
        std::vector<double> v2;
        for (uint64_t c = c0; c < c1; ++c) {
            const uint64_t a = start[c], b = start[c + 1];
            const uint64_t m = b - a;
            idx.resize(m);
            for (uint64_t i = 0; i < m; ++i) idx[i] = (uint32_t)i;
            std::stable_sort(idx.begin(), idx.end(),
                             [&](uint32_t x, uint32_t y) { return rows[a + x] < rows[a + y]; });
            r2.clear();
            v2.clear();
            for (uint64_t i = 0; i < m;) {
                const uint32_t r = rows[a + idx[i]];
                double s = -0.0;
                uint64_t j = i;
                while (j < m && rows[a + idx[j]] == r) s += vals[a + idx[j++]];
                r2.push_back(r);
                v2.push_back(s);
                i = j;
            }
            for (size_t i = 0; i < r2.size(); ++i) {
                rows[a + i] = r2[i];
                vals[a + i] = v2[i];
            }
            ucount[c] = r2.size();
        }
    });
    or_csc* A = new or_csc;
    A->n = n;
    A->col_ptr.assign(n + 1, 0);
    for (uint64_t c = 0; c < n; ++c) A->col_ptr[c + 1] = A->col_ptr[c] + ucount[c];
    A->row_idx.resize(A->col_ptr[n]);
    A->vals.resize(A->col_ptr[n]);
    for (uint64_t c = 0; c < n; ++c) {
        std::copy(rows.begin() + start[c], rows.begin() + start[c] + ucount[c],
                  A->row_idx.begin() + A->col_ptr[c]);
        std::copy(vals.begin() + start[c], vals.begin() + start[c] + ucount[c],
                  A->vals.begin() + A->col_ptr[c]);
    }
    return A;
}

// Parallel triplet generation over contiguous cell ranges, concatenated in
// cell order (rayon par_iter().flat_map().collect() preserves this order).
template <class CellFn>
static std::vector<Trip> gen_trips(uint64_t nc, int nthreads, size_t per_cell, CellFn&& fn) {
    const int nth = nthreads_or_default(nthreads);
    std::vector<std::vector<Trip>> parts(nth);
    parallel_ranges(nc, nth, [&](int t, uint64_t a, uint64_t b) {
        parts[t].reserve((b - a) * per_cell);
        for (uint64_t c = a; c < b; ++c) fn(c, parts[t]);
    });
    size_t tot = 0;
    for (auto& p : parts) tot += p.size();
    std::vector<Trip> all;
    all.reserve(tot);
    for (auto& p : parts) {
        all.insert(all.end(), p.begin(), p.end());
        std::vector<Trip>().swap(p);
    }
    return all;
}

// elasticity.rs:35-47.
static inline double lame_lambda(double e, double nu) {
    return (e * nu) / ((1.0 + nu) * (1.0 - 2.0 * nu));
}
static inline double lame_mu(double e, double nu) { return e / (2.0 * (1.0 + nu)); }

// elasticity.rs:113-142: the 12x12 element matrix entry (i alpha, j beta).
static inline double elas_entry(const double* gi, const double* gj, int alpha, int beta,
                                double lambda, double mu, double vol) {
    const double gigj = gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2];
    const double delta = (alpha == beta) ? 1.0 : 0.0;
    return vol * (lambda * gj[beta] * gi[alpha] + mu * gigj * delta + mu * gj[alpha] * gi[beta]);
}

static or_csc* elasticity_impl(uint64_t nv, const double* V, uint64_t nc, const uint64_t* C,
                               const double* Ecell, double E, double nu, int nthreads) {
    if (!(nu > -1.0 && nu < 0.5)) return nullptr;  // elasticity.rs:23-26 assert
    auto tr = gen_trips(nc, nthreads, 144, [&](uint64_t c, std::vector<Trip>& out) {
        const uint64_t* t = C + 4 * c;
        Map3 m;
        if (!affine3(V + 3 * t[0], V + 3 * t[1], V + 3 * t[2], V + 3 * t[3], m)) return;
        const double e = Ecell ? Ecell[c] : E;
        const double lambda = lame_lambda(e, nu), mu = lame_mu(e, nu);
        const double vol = (1.0 / 6.0) * std::fabs(m.det);
        double g[4][3];
        for (int i = 0; i < 4; ++i) grad3(m, P1_TET_GRAD[i], g[i]);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        out.push_back({(uint32_t)(3 * t[i] + a), (uint32_t)(3 * t[j] + b),
                                       elas_entry(g[i], g[j], a, b, lambda, mu, vol)});
    });
    return csc_from_trips(3 * nv, tr);
}

}  // namespace orc

using namespace orc;

extern "C" {

or_csc* or_csc_from_triplets(uint64_t n, uint64_t nt, const uint64_t* rows, const uint64_t* cols,
                             const double* vals) {
    std::vector<Trip> t(nt);
    for (uint64_t i = 0; i < nt; ++i) {
        if (rows[i] >= n || cols[i] >= n) return nullptr;
        t[i] = {(uint32_t)rows[i], (uint32_t)cols[i], vals[i]};
    }
    return csc_from_trips(n, t);
}
or_csc* or_csc_from_parts(uint64_t n, const uint64_t* col_ptr, const uint64_t* row_idx,
                          const double* vals) {
    or_csc* A = new or_csc;
    A->n = n;
    A->col_ptr.assign(col_ptr, col_ptr + n + 1);
    const uint64_t nnz = col_ptr[n];
    A->row_idx.resize(nnz);
    for (uint64_t i = 0; i < nnz; ++i) {
        if (row_idx[i] >= n) {
            delete A;
            return nullptr;
        }
        A->row_idx[i] = (uint32_t)row_idx[i];
    }
    A->vals.assign(vals, vals + nnz);
    return A;
}
uint64_t or_csc_n(const or_csc* a) { return a->n; }
uint64_t or_csc_nnz(const or_csc* a) { return a->col_ptr[a->n]; }
void or_csc_export(const or_csc* a, uint64_t* col_ptr, uint64_t* row_idx, double* vals) {
    if (col_ptr) std::copy(a->col_ptr.begin(), a->col_ptr.end(), col_ptr);
    if (row_idx)
        for (size_t i = 0; i < a->row_idx.size(); ++i) row_idx[i] = a->row_idx[i];
    if (vals) std::copy(a->vals.begin(), a->vals.end(), vals);
}
void or_csc_free(or_csc* a) { delete a; }

// poisson.rs:23-69: K_e(i,j) = (g_i . g_j) * area.
or_csc* or_assemble_stiffness_2d(uint64_t nv, const double* V, uint64_t nc, const uint64_t* C,
                                 int nthreads) {
    auto tr = gen_trips(nc, nthreads, 9, [&](uint64_t c, std::vector<Trip>& out) {
        const uint64_t* t = C + 3 * c;
        Map2 m;
        if (!affine2(V + 2 * t[0], V + 2 * t[1], V + 2 * t[2], m)) return;
        const double area = 0.5 * std::fabs(m.det);
        double g[3][2];
        for (int i = 0; i < 3; ++i) grad2(m, P1_TRI_GRAD[i], g[i]);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                out.push_back({(uint32_t)t[i], (uint32_t)t[j],
                               (g[i][0] * g[j][0] + g[i][1] * g[j][1]) * area});
    });
    return csc_from_trips(nv, tr);
}

// poisson.rs:124-165: K_e(i,j) = (g_i . g_j) * vol.
or_csc* or_assemble_stiffness_3d(uint64_t nv, const double* V, uint64_t nc, const uint64_t* C,
                                 int nthreads) {
    auto tr = gen_trips(nc, nthreads, 16, [&](uint64_t c, std::vector<Trip>& out) {
        const uint64_t* t = C + 4 * c;
        Map3 m;
        if (!affine3(V + 3 * t[0], V + 3 * t[1], V + 3 * t[2], V + 3 * t[3], m)) return;
        const double vol = (1.0 / 6.0) * std::fabs(m.det);
        double g[4][3];
        for (int i = 0; i < 4; ++i) grad3(m, P1_TET_GRAD[i], g[i]);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                out.push_back({(uint32_t)t[i], (uint32_t)t[j],
                               (g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2]) * vol});
    });
    return csc_from_trips(nv, tr);
}

// mass.rs:10-44: area/6 on the diagonal, area/12 off it.
or_csc* or_assemble_mass_2d(uint64_t nv, const double* V, uint64_t nc, const uint64_t* C,
                            int nthreads) {
    auto tr = gen_trips(nc, nthreads, 9, [&](uint64_t c, std::vector<Trip>& out) {
        const uint64_t* t = C + 3 * c;
        Map2 m;
        if (!affine2(V + 2 * t[0], V + 2 * t[1], V + 2 * t[2], m)) return;
        const double area = 0.5 * std::fabs(m.det);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double factor = (i == j) ? 1.0 / 6.0 : 1.0 / 12.0;
                out.push_back({(uint32_t)t[i], (uint32_t)t[j], area * factor});
            }
    });
    return csc_from_trips(nv, tr);
}

// mass.rs:86-126: M_{(i,a),(j,a)} = density * vol * (1/10 if i == j else 1/20),
// the same scalar mass on each displacement component.
or_csc* or_assemble_elasticity_mass_3d(uint64_t nv, const double* V, uint64_t nc,
                                       const uint64_t* C, double density, int nthreads) {
    auto tr = gen_trips(nc, nthreads, 48, [&](uint64_t c, std::vector<Trip>& out) {
        const uint64_t* t = C + 4 * c;
        Map3 m;
        if (!affine3(V + 3 * t[0], V + 3 * t[1], V + 3 * t[2], V + 3 * t[3], m)) return;
        const double vol = (1.0 / 6.0) * std::fabs(m.det);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                const double factor = (i == j) ? 1.0 / 10.0 : 1.0 / 20.0;
                const double mv = density * vol * factor;
                for (int a = 0; a < 3; ++a)
                    out.push_back({(uint32_t)(3 * t[i] + a), (uint32_t)(3 * t[j] + a), mv});
            }
    });
    return csc_from_trips(3 * nv, tr);
}

// mass.rs:190-250: exact reference-tetrahedron products of the P2 basis
// (4 corners lambda_i (2 lambda_i - 1), then edges 01, 12, 02, 03, 13, 23 as
// 4 lambda_i lambda_j), from prod m_i! / (sum m_i + 3)!, term by term.
void or_p2_mass_products(double* out) {
    static const double F[12] = {1.0,   1.0,    2.0,     6.0,      24.0,      120.0,
                                 720.0, 5040.0, 40320.0, 362880.0, 3628800.0, 39916800.0};
    struct Term {
        std::array<int, 4> e;
        double c;
    };
    std::vector<std::vector<Term>> T;
    for (int i = 0; i < 4; ++i) {
        std::array<int, 4> two = {0, 0, 0, 0}, one = {0, 0, 0, 0};
        two[i] = 2;
        one[i] = 1;
        T.push_back({{two, 2.0}, {one, -1.0}});
    }
    const int E[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
    for (auto& ed : E) {
        std::array<int, 4> m = {0, 0, 0, 0};
        m[ed[0]] = 1;
        m[ed[1]] = 1;
        T.push_back({{m, 4.0}});
    }
    for (int a = 0; a < 10; ++a)
        for (int b = 0; b < 10; ++b) {
            double acc = 0.0;
            for (const Term& ta : T[a])
                for (const Term& tb : T[b]) {
                    std::array<int, 4> m;
                    int deg = 0;
                    for (int q = 0; q < 4; ++q) {
                        m[q] = ta.e[q] + tb.e[q];
                        deg += m[q];
                    }
                    acc += ta.c * tb.c * F[m[0]] * F[m[1]] * F[m[2]] * F[m[3]] / F[deg + 3];
                }
            out[a * 10 + b] = acc;
        }
}

// mass.rs:136-180: triplets per cell (cell order), density |det J| I_ab on
// each displacement component.
or_csc* or_assemble_elasticity_mass_p2_3d(uint64_t n_nodes, const double* V, uint64_t nc,
                                          const uint64_t* C, double density, int nthreads) {
    double I[100];
    or_p2_mass_products(I);
    auto tr = gen_trips(nc, nthreads, 300, [&](uint64_t c, std::vector<Trip>& out) {
        const uint64_t* t = C + 10 * c;
        Map3 m;
        if (!affine3(V + 3 * t[0], V + 3 * t[1], V + 3 * t[2], V + 3 * t[3], m)) return;
        const double det = std::fabs(m.det);
        for (int i = 0; i < 10; ++i)
            for (int j = 0; j < 10; ++j) {
                const double mv = density * det * I[i * 10 + j];
                for (int a = 0; a < 3; ++a)
                    out.push_back({(uint32_t)(3 * t[i] + a), (uint32_t)(3 * t[j] + a), mv});
            }
    });
    return csc_from_trips(3 * n_nodes, tr);
}

// elasticity.rs:73-149.
or_csc* or_assemble_elasticity_3d(uint64_t nv, const double* V, uint64_t nc, const uint64_t* C,
                                  double E, double nu, int nthreads) {
    return elasticity_impl(nv, V, nc, C, nullptr, E, nu, nthreads);
}
or_csc* or_assemble_elasticity_3d_cellE(uint64_t nv, const double* V, uint64_t nc,
                                        const uint64_t* C, const double* E_cell, double nu,
                                        int nthreads) {
    return elasticity_impl(nv, V, nc, C, E_cell, 0.0, nu, nthreads);
}

// poisson.rs:98-100 / 194-196: x_q = map_to_physical(v0, xi_q).
int or_quad_points_2d(const double* V, uint64_t nc, const uint64_t* C, int degree, double* xq) {
    const Rule r = rule(2, degree);
    const size_t nq = r.w.size();
    for (uint64_t c = 0; c < nc; ++c) {
        const uint64_t* t = C + 3 * c;
        Map2 m;
        affine2(V + 2 * t[0], V + 2 * t[1], V + 2 * t[2], m);
        for (size_t q = 0; q < nq; ++q) map2(m, V + 2 * t[0], r.pts[q].data(), xq + 2 * (c * nq + q));
    }
    return 0;
}
int or_quad_points_3d(const double* V, uint64_t nc, const uint64_t* C, int degree, double* xq) {
    const Rule r = rule(3, degree);
    const size_t nq = r.w.size();
    for (uint64_t c = 0; c < nc; ++c) {
        const uint64_t* t = C + 4 * c;
        Map3 m;
        affine3(V + 3 * t[0], V + 3 * t[1], V + 3 * t[2], V + 3 * t[3], m);
        for (size_t q = 0; q < nq; ++q) map3(m, V + 3 * t[0], r.pts[q].data(), xq + 3 * (c * nq + q));
    }
    return 0;
}

// poisson.rs:72-121: local_b[i] += w * |det| * f(x_q) * phi_i(xi_q);
// scatter b[dof] += local_b sequentially in cell order.
int or_assemble_rhs_2d(uint64_t nv, const double* V, uint64_t nc, const uint64_t* C,
                       const double* fq, double* b) {
    const Rule r = rule(2, 2);
    for (uint64_t i = 0; i < nv; ++i) b[i] = 0.0;
    for (uint64_t c = 0; c < nc; ++c) {
        const uint64_t* t = C + 3 * c;
        Map2 m;
        if (!affine2(V + 2 * t[0], V + 2 * t[1], V + 2 * t[2], m)) continue;
        const double detj = std::fabs(m.det);
        double lb[3] = {0.0, 0.0, 0.0};
        for (size_t q = 0; q < r.w.size(); ++q) {
            double phi[3];
            p1_tri_basis(r.pts[q].data(), phi);
            const double f = fq[c * 3 + q];
            for (int i = 0; i < 3; ++i) lb[i] += r.w[q] * detj * f * phi[i];
        }
        for (int i = 0; i < 3; ++i) b[t[i]] += lb[i];
    }
    return 0;
}
// poisson.rs:168-218.
int or_assemble_rhs_3d(uint64_t nv, const double* V, uint64_t nc, const uint64_t* C,
                       const double* fq, double* b) {
    const Rule r = rule(3, 2);
    for (uint64_t i = 0; i < nv; ++i) b[i] = 0.0;
    for (uint64_t c = 0; c < nc; ++c) {
        const uint64_t* t = C + 4 * c;
        Map3 m;
        if (!affine3(V + 3 * t[0], V + 3 * t[1], V + 3 * t[2], V + 3 * t[3], m)) continue;
        const double detj = std::fabs(m.det);
        double lb[4] = {0.0, 0.0, 0.0, 0.0};
        for (size_t q = 0; q < r.w.size(); ++q) {
            double phi[4];
            p1_tet_basis(r.pts[q].data(), phi);
            const double f = fq[c * 4 + q];
            for (int i = 0; i < 4; ++i) lb[i] += r.w[q] * detj * f * phi[i];
        }
        for (int i = 0; i < 4; ++i) b[t[i]] += lb[i];
    }
    return 0;
}

// elasticity.rs:153-185: f * (vol/4) to each vertex, sequential in cell order.
int or_assemble_body_force_3d(uint64_t nv, const double* V, uint64_t nc, const uint64_t* C,
                              const double* f3, double* b) {
    for (uint64_t i = 0; i < 3 * nv; ++i) b[i] = 0.0;
    for (uint64_t c = 0; c < nc; ++c) {
        const uint64_t* t = C + 4 * c;
        Map3 m;
        if (!affine3(V + 3 * t[0], V + 3 * t[1], V + 3 * t[2], V + 3 * t[3], m)) continue;
        const double vol = (1.0 / 6.0) * std::fabs(m.det);
        const double load[3] = {f3[0] * (vol * 0.25), f3[1] * (vol * 0.25), f3[2] * (vol * 0.25)};
        for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 3; ++d) b[3 * t[a] + d] += load[d];
    }
    return 0;
}

// poisson.rs:222-263 == elasticity.rs:188-227: keep (r,c) iff both free;
// b_r -= A_rc g_c for fixed c, free r (column-major scan); then (d,d,1.0) and
// b_d = g_d in the order the prescribed dofs are listed.
or_csc* or_apply_dirichlet_bc(const or_csc* A, const double* rhs, uint64_t nfix,
                              const uint64_t* dofs, const double* vals, double* rhs_out) {
    const uint64_t n = A->n;
    std::vector<char> isd(n, 0);
    std::vector<double> g(n, 0.0);
    for (uint64_t k = 0; k < nfix; ++k) {
        if (dofs[k] >= n) return nullptr;
        isd[dofs[k]] = 1;
        g[dofs[k]] = vals[k];
    }
    for (uint64_t i = 0; i < n; ++i) rhs_out[i] = rhs[i];
    std::vector<Trip> tr;
    tr.reserve(A->vals.size() + nfix);
    for (uint64_t c = 0; c < n; ++c)
        for (uint64_t p = A->col_ptr[c]; p < A->col_ptr[c + 1]; ++p) {
            const uint32_t r = A->row_idx[p];
            const double v = A->vals[p];
            if (isd[c] && !isd[r]) rhs_out[r] -= v * g[c];
            else if (!isd[c] && !isd[r]) tr.push_back({r, (uint32_t)c, v});
        }
    for (uint64_t k = 0; k < nfix; ++k) {
        tr.push_back({(uint32_t)dofs[k], (uint32_t)dofs[k], 1.0});
        rhs_out[dofs[k]] = vals[k];
    }
    return csc_from_trips(n, tr);
}

// poisson_solver.rs:62-96 with the degree-3 rule; uq = u_exact at its points.
double or_l2_error_2d(const double* V, uint64_t nc, const uint64_t* C, const double* u,
                      const double* uq) {
    const Rule r = rule(2, 3);
    const size_t nq = r.w.size();
    double tot = 0.0;
    for (uint64_t c = 0; c < nc; ++c) {
        const uint64_t* t = C + 3 * c;
        Map2 m;
        affine2(V + 2 * t[0], V + 2 * t[1], V + 2 * t[2], m);
        const double detj = std::fabs(m.det);
        for (size_t q = 0; q < nq; ++q) {
            double phi[3];
            p1_tri_basis(r.pts[q].data(), phi);
            const double ap = u[t[0]] * phi[0] + u[t[1]] * phi[1] + u[t[2]] * phi[2];
            const double err = ap - uq[c * nq + q];
            tot += r.w[q] * detj * err * err;
        }
    }
    return std::sqrt(tot);
}
// poisson_solver.rs:135-172.
double or_l2_error_3d(const double* V, uint64_t nc, const uint64_t* C, const double* u,
                      const double* uq) {
    const Rule r = rule(3, 3);
    const size_t nq = r.w.size();
    double tot = 0.0;
    for (uint64_t c = 0; c < nc; ++c) {
        const uint64_t* t = C + 4 * c;
        Map3 m;
        affine3(V + 3 * t[0], V + 3 * t[1], V + 3 * t[2], V + 3 * t[3], m);
        const double detj = std::fabs(m.det);
        for (size_t q = 0; q < nq; ++q) {
            double phi[4];
            p1_tet_basis(r.pts[q].data(), phi);
            const double ap =
                u[t[0]] * phi[0] + u[t[1]] * phi[1] + u[t[2]] * phi[2] + u[t[3]] * phi[3];
            const double err = ap - uq[c * nq + q];
            tot += r.w[q] * detj * err * err;
        }
    }
    return std::sqrt(tot);
}

// Adjoint of the compliance (new; pinned by finite differences, the
// homogeneity identity sum_c E_c dJ/dE_c = -J and exact rational arithmetic):
// out[c] = -w_c^T K0_c u_c with K0 the elasticity.rs:133-136 element matrix at
// E = 1, evaluated in strain form, vol (lambda tr(e_w) tr(e_u) + 2 mu e_w:e_u)
// with e = sym(G), G[a][b] = sum_{i=1..3} (v_i[a] - v_0[a]) g_i[b] (the same
// contraction; the differences remove the rigid translation before rounding).
static void or_cell_grad(const double v[4][3], const double g[4][3], double G[3][3]) {
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            G[a][b] = ((v[1][a] - v[0][a]) * g[1][b] + (v[2][a] - v[0][a]) * g[2][b]) +
                      (v[3][a] - v[0][a]) * g[3][b];
}

int or_cell_bilinear_k0(const double* V, uint64_t nc, const uint64_t* C, double nu,
                        const double* w, const double* u, double* out) {
    const double lambda = lame_lambda(1.0, nu), mu = lame_mu(1.0, nu);
    for (uint64_t c = 0; c < nc; ++c) {
        const uint64_t* t = C + 4 * c;
        Map3 m;
        if (!affine3(V + 3 * t[0], V + 3 * t[1], V + 3 * t[2], V + 3 * t[3], m)) {
            out[c] = 0.0;
            continue;
        }
        const double vol = (1.0 / 6.0) * std::fabs(m.det);
        double g[4][3];
        for (int i = 0; i < 4; ++i) grad3(m, P1_TET_GRAD[i], g[i]);
        double wl[4][3], ul[4][3];
        for (int i = 0; i < 4; ++i)
            for (int a = 0; a < 3; ++a) {
                wl[i][a] = w[3 * t[i] + a];
                ul[i][a] = u[3 * t[i] + a];
            }
        double Gw[3][3], Gu[3][3];
        or_cell_grad(wl, g, Gw);
        or_cell_grad(ul, g, Gu);
        double ee = 0.0;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                ee += (0.5 * (Gw[a][b] + Gw[b][a])) * (0.5 * (Gu[a][b] + Gu[b][a]));
        const double trw = (Gw[0][0] + Gw[1][1]) + Gw[2][2];
        const double tru = (Gu[0][0] + Gu[1][1]) + Gu[2][2];
        out[c] = -(vol * (lambda * trw * tru + 2.0 * mu * ee));
    }
    return 0;
}

// elasticity.rs:230-320.
int or_element_stresses(const double* V, uint64_t nc, const uint64_t* C, const double* u,
                        double E, double nu, double* out) {
    const double lambda = lame_lambda(E, nu), mu = lame_mu(E, nu);
    for (uint64_t c = 0; c < nc; ++c) {
        const uint64_t* t = C + 4 * c;
        double* o = out + 20 * c;
        for (int k = 0; k < 20; ++k) o[k] = 0.0;
        Map3 m;
        if (!affine3(V + 3 * t[0], V + 3 * t[1], V + 3 * t[2], V + 3 * t[3], m)) continue;
        double g[4][3];
        for (int i = 0; i < 4; ++i) grad3(m, P1_TET_GRAD[i], g[i]);
        double gu[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int nd = 0; nd < 4; ++nd) gu[i][j] += u[3 * t[nd] + i] * g[nd][j];
        double eps[3][3], sig[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) eps[i][j] = 0.5 * (gu[i][j] + gu[j][i]);
        const double tr = eps[0][0] + eps[1][1] + eps[2][2];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double delta = (i == j) ? 1.0 : 0.0;
                sig[i][j] = lambda * tr * delta + 2.0 * mu * eps[i][j];
            }
        const double s00 = sig[0][0], s11 = sig[1][1], s22 = sig[2][2], s01 = sig[0][1],
                     s12 = sig[1][2], s20 = sig[2][0];
        const double vm2 = 0.5 * ((s00 - s11) * (s00 - s11) + (s11 - s22) * (s11 - s22) +
                                  (s22 - s00) * (s22 - s00) +
                                  6.0 * (s01 * s01 + s12 * s12 + s20 * s20));
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                o[3 * i + j] = sig[i][j];
                o[9 + 3 * i + j] = eps[i][j];
            }
        o[18] = std::sqrt(std::max(vm2, 0.0));
        o[19] = s00 + s11 + s22;
    }
    return 0;
}

// elasticity.rs:323-360.
int or_vertex_von_mises(uint64_t nv, const double* V, uint64_t nc, const uint64_t* C,
                        const double* vm_cell, double* out) {
    std::vector<double> vs(nv, 0.0);
    for (uint64_t i = 0; i < nv; ++i) out[i] = 0.0;
    for (uint64_t c = 0; c < nc; ++c) {
        const uint64_t* t = C + 4 * c;
        Map3 m;
        if (!affine3(V + 3 * t[0], V + 3 * t[1], V + 3 * t[2], V + 3 * t[3], m)) continue;
        const double vol = (1.0 / 6.0) * std::fabs(m.det);
        for (int a = 0; a < 4; ++a) {
            out[t[a]] += vm_cell[c] * vol;
            vs[t[a]] += vol;
        }
    }
    for (uint64_t i = 0; i < nv; ++i)
        if (vs[i] > 1e-15) out[i] /= vs[i];
    return 0;
}

// boundary.rs:8-46 (the 2-point Gauss rule of quadrature.rs:41-56 on [0,1]):
// the flux is sampled by the caller at or_neumann_points_2d's points g[f][q].
static double gauss2_point(int q) {
    const double c = 1.0 / (2.0 * std::sqrt(3.0));
    return q == 0 ? 0.5 - c : 0.5 + c;
}
int or_neumann_points_2d(const double* V, uint64_t nf, const uint64_t* facets, double* xq) {
    for (uint64_t f = 0; f < nf; ++f) {
        const double* v0 = V + 2 * facets[2 * f];
        const double* v1 = V + 2 * facets[2 * f + 1];
        for (int q = 0; q < 2; ++q) {
            const double s = gauss2_point(q);
            xq[4 * f + 2 * q] = (1.0 - s) * v0[0] + s * v1[0];
            xq[4 * f + 2 * q + 1] = (1.0 - s) * v0[1] + s * v1[1];
        }
    }
    return 0;
}
int or_assemble_neumann_2d(uint64_t nv, const double* V, uint64_t nf, const uint64_t* facets,
                           const double* g, double* rhs) {
    for (uint64_t i = 0; i < nv; ++i) rhs[i] = 0.0;
    for (uint64_t f = 0; f < nf; ++f) {
        const uint64_t a = facets[2 * f], b = facets[2 * f + 1];
        const double* v0 = V + 2 * a;
        const double* v1 = V + 2 * b;
        const double dx = v1[0] - v0[0], dy = v1[1] - v0[1];
        const double length = std::sqrt(dx * dx + dy * dy);
        for (int q = 0; q < 2; ++q) {
            const double s = gauss2_point(q);
            const double w = 0.5 * length;
            const double gv = g[2 * f + q];
            rhs[a] += w * gv * (1.0 - s);
            rhs[b] += w * gv * s;
        }
    }
    return 0;
}

// boundary.rs:49-108: edge-midpoint rule on each facet triangle; the caller
// samples the traction at or_traction_points_3d's midpoints t[f][m][3].
int or_traction_points_3d(const double* V, uint64_t nf, const uint64_t* facets, double* mids) {
    for (uint64_t f = 0; f < nf; ++f) {
        const double* v[3] = {V + 3 * facets[3 * f], V + 3 * facets[3 * f + 1],
                              V + 3 * facets[3 * f + 2]};
        for (int m = 0; m < 3; ++m)
            for (int a = 0; a < 3; ++a)
                mids[9 * f + 3 * m + a] = 0.5 * (v[m][a] + v[(m + 1) % 3][a]);
    }
    return 0;
}
int or_assemble_surface_traction_3d(uint64_t nv, const double* V, uint64_t nf,
                                    const uint64_t* facets, const double* t, double* rhs) {
    for (uint64_t i = 0; i < 3 * nv; ++i) rhs[i] = 0.0;
    for (uint64_t f = 0; f < nf; ++f) {
        const uint64_t* fv = facets + 3 * f;
        const double* v0 = V + 3 * fv[0];
        const double* v1 = V + 3 * fv[1];
        const double* v2 = V + 3 * fv[2];
        const double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
        const double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
        const double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                             e1[0] * e2[1] - e1[1] * e2[0]};
        const double area = 0.5 * std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        const double w = area / 3.0;
        for (int m = 0; m < 3; ++m) {
            const double* tm = t + 9 * f + 3 * m;
            for (int k = 0; k < 3; ++k)
                for (int a = 0; a < 3; ++a) rhs[3 * fv[k] + a] += (w / 3.0) * tm[a];
        }
    }
    return 0;
}

// boundary.rs:111-139: gamma times the 1D P1 mass [L/3 L/6; L/6 L/3] of each
// facet edge, as triplets summed per position.
or_csc* or_assemble_robin_matrix_2d(uint64_t nv, const double* V, uint64_t nf,
                                    const uint64_t* facets, double gamma) {
    std::vector<Trip> tr;
    tr.reserve(4 * nf);
    for (uint64_t f = 0; f < nf; ++f) {
        const uint64_t n0 = facets[2 * f], n1 = facets[2 * f + 1];
        const double* v0 = V + 2 * n0;
        const double* v1 = V + 2 * n1;
        const double dx = v1[0] - v0[0], dy = v1[1] - v0[1];
        const double length = std::sqrt(dx * dx + dy * dy);
        tr.push_back({(uint32_t)n0, (uint32_t)n0, gamma * length / 3.0});
        tr.push_back({(uint32_t)n0, (uint32_t)n1, gamma * length / 6.0});
        tr.push_back({(uint32_t)n1, (uint32_t)n0, gamma * length / 6.0});
        tr.push_back({(uint32_t)n1, (uint32_t)n1, gamma * length / 3.0});
    }
    return csc_from_trips(nv, tr);
}

}  // extern "C"
