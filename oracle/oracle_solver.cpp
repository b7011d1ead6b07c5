// oracle_solver.cpp — TEST INFRASTRUCTURE (checker only; see oracle.h).
//
// Jacobi-PCG and the CSC->CSR conversion, restated from fenics-rs:
//   crates/fenics-solver/src/cg.rs:7-149   (spmv, dot, norm2, solve_cg)
//   crates/fenics-solver/src/gpu.rs:1552-1583 (CSC -> CSR + inverse diagonal)
// and the P2 tetrahedral dofmap / stiffness that the reference lacks:
// vertices first, then edges (the vertices-then-edges convention of
// mass.rs:132-135), the edges ordered by the number of cells around them,
// most first (counts above 255 count as 255), then by the vertex difference
// max - min (edges of one direction of a structured mesh), then by the
// smaller vertex; per-cell dofs in
// P2Tetrahedron order (lagrange.rs:284-297: corners, then edges 01, 12, 02,
// 03, 13, 23); K_e = sum_q (w_q |det J|) (g_i^q . g_j^q) with the degree-2
// tetrahedron rule (quadrature.rs:157-184), exact for P2 x P2 gradients.
#include <algorithm>
#include <chrono>
#include <thread>
#include <cmath>
#include <vector>

#include "oracle_internal.hpp"

using namespace orc;

namespace {

// cg.rs:27-29: sequential left-to-right sum of products.
double dot(const double* a, const double* b, uint64_t n) {
    double s = -0.0;
    for (uint64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

// cg.rs:7-23: CSC column scatter that skips x[col] == 0.
void spmv(const or_csc* A, const double* x, double* y) {
    const uint64_t n = A->n;
    for (uint64_t i = 0; i < n; ++i) y[i] = 0.0;
    for (uint64_t c = 0; c < n; ++c) {
        const double xv = x[c];
        if (xv != 0.0)
            for (uint64_t p = A->col_ptr[c]; p < A->col_ptr[c + 1]; ++p)
                y[A->row_idx[p]] += A->vals[p] * xv;
    }
}

// cg.rs:51-64: 1/d when |d| > 1e-15, else 1 (also when the diagonal is absent).
std::vector<double> inv_diagonal(const or_csc* A) {
    std::vector<double> d(A->n, 1.0);
    for (uint64_t c = 0; c < A->n; ++c)
        for (uint64_t p = A->col_ptr[c]; p < A->col_ptr[c + 1]; ++p)
            if (A->row_idx[p] == c) {
                if (std::fabs(A->vals[p]) > 1e-15) d[c] = 1.0 / A->vals[p];
                break;
            }
    return d;
}

// eigen.rs:22-45: mean |diagonal| over the columns that store a diagonal entry.
double mean_abs_diag(const or_csc* A) {
    if (A->n == 0) return 0.0;
    double acc = 0.0;
    uint64_t found = 0;
    for (uint64_t c = 0; c < A->n; ++c)
        for (uint64_t p = A->col_ptr[c]; p < A->col_ptr[c + 1]; ++p)
            if (A->row_idx[p] == c) {
                acc += std::fabs(A->vals[p]);
                ++found;
                break;
            }
    return found ? acc / (double)found : 0.0;
}

}  // namespace

extern "C" {

void or_spmv(const or_csc* a, const double* x, double* y) { spmv(a, x, y); }

// eigen.rs:51-78, column by column: drop constrained rows/columns and exact
// zeros; append the unit diagonal of a constrained column after its entries.
or_csc* or_fold_dirichlet(const or_csc* a, uint64_t nfix, const uint64_t* fixed, int unit_diag) {
    const uint64_t n = a->n;
    std::vector<char> isf(n, 0);
    for (uint64_t i = 0; i < nfix; ++i)
        if (fixed[i] < n) isf[fixed[i]] = 1;
    std::vector<Trip> tr;
    for (uint64_t c = 0; c < n; ++c) {
        for (uint64_t p = a->col_ptr[c]; p < a->col_ptr[c + 1]; ++p) {
            const uint64_t r = a->row_idx[p];
            if (isf[r] || isf[c] || a->vals[p] == 0.0) continue;
            tr.push_back({(uint32_t)r, (uint32_t)c, a->vals[p]});
        }
        if (unit_diag && isf[c]) tr.push_back({(uint32_t)c, (uint32_t)c, 1.0});
    }
    return csc_from_trips(n, tr);
}

// eigen.rs:94-266: inverse power iteration with M-deflation against the
// converged modes; mass rescaled by s = mean|diag K| / mean|diag M| (exact:
// eigenvalues are recovered with the unscaled M).
int or_vibration_modes(const or_csc* k, const or_csc* m, uint64_t nfix, const uint64_t* fixed,
                       uint64_t num_modes, uint64_t max_power_iters, double tol,
                       double inner_tol, uint64_t inner_max_iter, double* freq_hz,
                       double* omega_out, double* eig_out, double* shapes) {
    or_csc* kd = or_fold_dirichlet(k, nfix, fixed, 1);
    or_csc* md = or_fold_dirichlet(m, nfix, fixed, 0);
    const uint64_t n = kd->n;
    std::vector<char> isf(n, 0);
    for (uint64_t i = 0; i < nfix; ++i)
        if (fixed[i] < n) isf[fixed[i]] = 1;
    uint64_t nfree = 0;
    for (uint64_t i = 0; i < n; ++i) nfree += isf[i] ? 0 : 1;
    int code = 0;
    if (nfree == 0) code = 4;
    const double kdd = mean_abs_diag(kd), mdd = mean_abs_diag(md);
    const double s = (kdd > 0.0 && mdd > 0.0) ? kdd / mdd : 1.0;
    std::vector<std::vector<double>> modes;
    std::vector<double> v(n), mv(n), w(n), kw(n), mw(n);
    const double pi = 3.14159265358979323846;
    for (uint64_t mi = 0; mi < num_modes && code == 0; ++mi) {
        for (uint64_t i = 0; i < n; ++i) {
            v[i] = 0.0;
            if (isf[i]) continue;
            const uint64_t q = ((i + 1) * 31 + (mi + 1) * 97) % 1000;
            v[i] = (double)q / 1000.0 - 0.5;
        }
        std::vector<std::vector<double>> mprev;
        for (const auto& p : modes) {
            std::vector<double> y(n);
            spmv(md, p.data(), y.data());
            for (auto& t : y) t *= s;
            mprev.push_back(std::move(y));
        }
        for (size_t q = 0; q < modes.size(); ++q) {
            const double coeff = dot(v.data(), mprev[q].data(), n);
            for (uint64_t i = 0; i < n; ++i) v[i] -= coeff * modes[q][i];
        }
        spmv(md, v.data(), mv.data());
        for (auto& t : mv) t *= s;
        const double norm_m = std::sqrt(dot(v.data(), mv.data(), n));
        if (norm_m > 1e-14)
            for (auto& t : v) t /= norm_m;
        double prev = 0.0;
        for (uint64_t it = 0; it < max_power_iters; ++it) {
            spmv(md, v.data(), mv.data());
            for (auto& t : mv) t *= s;
            uint64_t iters = 0;
            double res = 0.0;
            const int st = or_solve_cg(kd, mv.data(), v.data(), inner_tol, inner_max_iter,
                                       w.data(), &iters, &res);
            if (st != 0) {
                code = st;
                break;
            }
            for (uint64_t i = 0; i < n; ++i)
                if (isf[i]) w[i] = 0.0;
            for (size_t q = 0; q < modes.size(); ++q) {
                const double coeff = dot(w.data(), mprev[q].data(), n);
                for (uint64_t i = 0; i < n; ++i) w[i] -= coeff * modes[q][i];
            }
            spmv(md, w.data(), mw.data());
            const double mwn = std::sqrt(dot(w.data(), mw.data(), n)) * std::sqrt(s);
            if (mwn < 1e-15) break;
            for (auto& t : w) t /= mwn;
            spmv(kd, w.data(), kw.data());
            const double ray = dot(w.data(), kw.data(), n);
            v = w;
            const double denom = std::max(std::fabs(ray), std::fabs(prev));
            if (denom == 0.0 || std::fabs(ray - prev) / denom < tol) break;
            prev = ray;
        }
        if (code) break;
        spmv(kd, v.data(), kw.data());
        spmv(md, v.data(), mw.data());
        const double lam = std::max(dot(v.data(), kw.data(), n) / dot(v.data(), mw.data(), n), 0.0);
        const double om = std::sqrt(lam);
        freq_hz[mi] = om / (2.0 * pi);
        omega_out[mi] = om;
        eig_out[mi] = lam;
        if (shapes) std::copy(v.begin(), v.end(), shapes + mi * n);
        modes.push_back(v);
    }
    or_csc_free(kd);
    or_csc_free(md);
    return code;
}

// cg.rs:40-149, control flow and stopping rule kept exactly:
// early exit when ||r0|| < tol or ||r0||/||b|| < tol; |pAp| < 1e-30 is a
// breakdown (status 1); stop when the recursive ||r||/||b|| < tol; status 2
// after max_iter.  *res is the absolute recursive residual norm.
int or_solve_cg(const or_csc* A, const double* b, const double* x0, double tol,
                uint64_t max_iter, double* x, uint64_t* iters, double* res) {
    const uint64_t n = A->n;
    const std::vector<double> inv = inv_diagonal(A);
    for (uint64_t i = 0; i < n; ++i) x[i] = x0 ? x0[i] : 0.0;
    std::vector<double> r(n), z(n), p(n), ap(n);
    spmv(A, x, ap.data());
    for (uint64_t i = 0; i < n; ++i) r[i] = b[i] - ap[i];
    const double bnorm = std::sqrt(dot(b, b, n));
    const double r0 = std::sqrt(dot(r.data(), r.data(), n));
    *iters = 0;
    *res = r0;
    if (r0 < tol || (bnorm > 0.0 && r0 / bnorm < tol)) return 0;
    for (uint64_t i = 0; i < n; ++i) z[i] = r[i] * inv[i];
    p = z;
    double rz_old = dot(r.data(), z.data(), n);
    for (uint64_t it = 1; it <= max_iter; ++it) {
        spmv(A, p.data(), ap.data());
        const double pap = dot(p.data(), ap.data(), n);
        *iters = it;
        if (std::fabs(pap) < 1e-30) {
            *res = std::sqrt(dot(r.data(), r.data(), n));
            return 1;
        }
        const double alpha = rz_old / pap;
        for (uint64_t i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * ap[i];
        }
        const double rn = std::sqrt(dot(r.data(), r.data(), n));
        const double rel = bnorm > 0.0 ? rn / bnorm : rn;
        *res = rn;
        if (rel < tol) return 0;
        for (uint64_t i = 0; i < n; ++i) z[i] = r[i] * inv[i];
        const double rz_new = dot(r.data(), z.data(), n);
        const double beta = rz_new / rz_old;
        rz_old = rz_new;
        for (uint64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    return 2;
}

// The same loop body run for exactly `iters` iterations (no stopping test),
// single-threaded like cg.rs; returns the wall time of the iterations only.
double or_time_cg_iterations(const or_csc* A, const double* b, uint64_t iters) {
    const uint64_t n = A->n;
    const std::vector<double> inv = inv_diagonal(A);
    std::vector<double> x(n, 0.0), r(b, b + n), z(n), p(n), ap(n);
    for (uint64_t i = 0; i < n; ++i) z[i] = r[i] * inv[i];
    p = z;
    double rz_old = dot(r.data(), z.data(), n);
    volatile double sink = 0.0;
    const auto t0 = std::chrono::steady_clock::now();
    for (uint64_t it = 1; it <= iters; ++it) {
        spmv(A, p.data(), ap.data());
        const double pap = dot(p.data(), ap.data(), n);
        const double alpha = rz_old / pap;
        for (uint64_t i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * ap[i];
        }
        sink = sink + std::sqrt(dot(r.data(), r.data(), n));
        for (uint64_t i = 0; i < n; ++i) z[i] = r[i] * inv[i];
        const double rz_new = dot(r.data(), z.data(), n);
        const double beta = rz_new / rz_old;
        rz_old = rz_new;
        for (uint64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// gpu.rs:1552-1583: CSR rows filled in column order, inv_diag from |d| > 1e-15.
void or_csc_to_csr(const or_csc* A, uint32_t* row_ptr, uint32_t* col_idx, double* vals,
                   double* inv_diag) {
    const uint64_t n = A->n;
    std::vector<uint32_t> cnt(n, 0);
    for (uint64_t i = 0; i < n; ++i) inv_diag[i] = 1.0;
    for (uint64_t c = 0; c < n; ++c)
        for (uint64_t p = A->col_ptr[c]; p < A->col_ptr[c + 1]; ++p) {
            const uint32_t r = A->row_idx[p];
            cnt[r]++;
            if (r == c && std::fabs(A->vals[p]) > 1e-15) inv_diag[r] = 1.0 / A->vals[p];
        }
    row_ptr[0] = 0;
    for (uint64_t i = 0; i < n; ++i) row_ptr[i + 1] = row_ptr[i] + cnt[i];
    std::vector<uint32_t> cur(row_ptr, row_ptr + n);
    for (uint64_t c = 0; c < n; ++c)
        for (uint64_t p = A->col_ptr[c]; p < A->col_ptr[c + 1]; ++p) {
            const uint32_t r = A->row_idx[p];
            const uint32_t pos = cur[r]++;
            col_idx[pos] = (uint32_t)c;
            vals[pos] = A->vals[p];
        }
}

// ---------------- P2 ----------------

static const int P2_EDGE[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};

// position of edge {a, b} in the lexicographic list p->edges
static uint64_t p2_edge_index(const or_p2* p, uint64_t a, uint64_t b) {
    const std::array<uint64_t, 2> key = {std::min(a, b), std::max(a, b)};
    return (uint64_t)(std::lower_bound(p->edges.begin(), p->edges.end(), key) - p->edges.begin());
}

or_p2* or_p2_build(uint64_t nv, const double* verts, uint64_t nc, const uint64_t* cells) {
    or_p2* p = new or_p2;
    p->nv = nv;
    p->nc = nc;
    p->verts.assign(verts, verts + 3 * nv);
    p->cells.assign(cells, cells + 4 * nc);
    p->edges.reserve(6 * nc);
    for (uint64_t c = 0; c < nc; ++c)
        for (auto& e : P2_EDGE) {
            const uint64_t a = cells[4 * c + e[0]], b = cells[4 * c + e[1]];
            p->edges.push_back({std::min(a, b), std::max(a, b)});
        }
    std::sort(p->edges.begin(), p->edges.end());
    // cells around each edge: the run length of its key in the sorted list
    std::vector<uint64_t> around;
    {
        std::vector<std::array<uint64_t, 2>> uniq;
        for (size_t i = 0; i < p->edges.size();) {
            size_t j = i;
            while (j < p->edges.size() && p->edges[j] == p->edges[i]) ++j;
            uniq.push_back(p->edges[i]);
            around.push_back(std::min<uint64_t>(j - i, 255));
            i = j;
        }
        p->edges.swap(uniq);
    }
    p->ne = p->edges.size();
    // edge dof numbers: most cells around first, then by max - min, then by
    // min (the edges are in lexicographic order, so the stable sort keeps
    // ascending min among equal keys)
    std::vector<uint64_t> order(p->ne);
    for (uint64_t e = 0; e < p->ne; ++e) order[e] = e;
    auto span = [&](uint64_t e) { return p->edges[e][1] - p->edges[e][0]; };
    std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) {
        if (around[a] != around[b]) return around[a] > around[b];
        return span(a) < span(b);
    });
    p->erank.assign(p->ne, 0);
    for (uint64_t r = 0; r < p->ne; ++r) p->erank[order[r]] = r;
    p->cell_dofs.resize(10 * nc);
    for (uint64_t c = 0; c < nc; ++c) {
        for (int a = 0; a < 4; ++a) p->cell_dofs[10 * c + a] = cells[4 * c + a];
        for (int k = 0; k < 6; ++k) {
            const uint64_t a = cells[4 * c + P2_EDGE[k][0]], b = cells[4 * c + P2_EDGE[k][1]];
            p->cell_dofs[10 * c + 4 + k] = nv + p->erank[p2_edge_index(p, a, b)];
        }
    }
    return p;
}
uint64_t or_p2_ndofs(const or_p2* p) { return p->nv + p->ne; }
uint64_t or_p2_nedges(const or_p2* p) { return p->ne; }
void or_p2_cell_dofs(const or_p2* p, uint64_t* out) {
    std::copy(p->cell_dofs.begin(), p->cell_dofs.end(), out);
}
void or_p2_dof_coords(const or_p2* p, double* out) {
    std::copy(p->verts.begin(), p->verts.end(), out);
    for (uint64_t e = 0; e < p->ne; ++e)
        for (int d = 0; d < 3; ++d)
            out[3 * (p->nv + p->erank[e]) + d] =
                (p->verts[3 * p->edges[e][0] + d] + p->verts[3 * p->edges[e][1] + d]) * 0.5;
}
// Boundary P2 dofs: vertices and edges of the boundary facets (mesh.rs:145-210).
int or_p2_boundary_dofs(const or_p2* p, uint64_t* out, uint64_t* n_out) {
    uint64_t nf = 0;
    or_boundary_facets_3d(p->nc, p->cells.data(), nullptr, &nf);
    std::vector<uint64_t> f(3 * nf);
    or_boundary_facets_3d(p->nc, p->cells.data(), f.data(), &nf);
    std::vector<uint64_t> d;
    d.reserve(6 * nf);
    for (uint64_t i = 0; i < nf; ++i) {
        const uint64_t* t = &f[3 * i];
        for (int a = 0; a < 3; ++a) d.push_back(t[a]);
        const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (auto& pr : pairs) {
            d.push_back(p->nv + p->erank[p2_edge_index(p, t[pr[0]], t[pr[1]])]);
        }
    }
    std::sort(d.begin(), d.end());
    d.erase(std::unique(d.begin(), d.end()), d.end());
    if (out) std::copy(d.begin(), d.end(), out);
    *n_out = d.size();
    return 0;
}

or_csc* or_assemble_p2_stiffness_3d(const or_p2* p, int nthreads) {
    const Rule r = rule(3, 2);
    double G[4][10][3];
    for (int q = 0; q < 4; ++q) p2_tet_grads(r.pts[q].data(), G[q]);
    const int nth = nthreads_or_default(nthreads);
    std::vector<std::vector<Trip>> parts(nth);
    std::vector<std::thread> th;
    for (int t = 0; t < nth; ++t)
        th.emplace_back([&, t] {
            const uint64_t a = p->nc * t / nth, b = p->nc * (t + 1) / nth;
            auto& out = parts[t];
            out.reserve((b - a) * 100);
            for (uint64_t c = a; c < b; ++c) {
                const uint64_t* cv = &p->cells[4 * c];
                const double* V = p->verts.data();
                Map3 m;
                if (!affine3(V + 3 * cv[0], V + 3 * cv[1], V + 3 * cv[2], V + 3 * cv[3], m))
                    continue;
                const double detj = std::fabs(m.det);
                double K[10][10];
                for (int i = 0; i < 10; ++i)
                    for (int j = 0; j < 10; ++j) K[i][j] = 0.0;
                for (int q = 0; q < 4; ++q) {
                    double g[10][3];
                    for (int i = 0; i < 10; ++i) grad3(m, G[q][i], g[i]);
                    const double s = r.w[q] * detj;
                    for (int i = 0; i < 10; ++i)
                        for (int j = 0; j < 10; ++j)
                            K[i][j] += s * (g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2]);
                }
                const uint64_t* d = &p->cell_dofs[10 * c];
                for (int i = 0; i < 10; ++i)
                    for (int j = 0; j < 10; ++j)
                        out.push_back({(uint32_t)d[i], (uint32_t)d[j], K[i][j]});
            }
        });
    for (auto& x : th) x.join();
    std::vector<Trip> all;
    for (auto& pp : parts) all.insert(all.end(), pp.begin(), pp.end());
    return csc_from_trips(p->nv + p->ne, all);
}

// poisson.rs:168-218 restated for P2 with a per-cell constant source:
// local_b[i] += w * |det| * f_c * phi_i(xi_q), scattered in cell order.
int or_assemble_p2_rhs_cellconst(const or_p2* p, const double* f_cell, double* b) {
    const Rule r = rule(3, 2);
    const uint64_t nd = p->nv + p->ne;
    for (uint64_t i = 0; i < nd; ++i) b[i] = 0.0;
    for (uint64_t c = 0; c < p->nc; ++c) {
        const uint64_t* cv = &p->cells[4 * c];
        const double* V = p->verts.data();
        Map3 m;
        if (!affine3(V + 3 * cv[0], V + 3 * cv[1], V + 3 * cv[2], V + 3 * cv[3], m)) continue;
        const double detj = std::fabs(m.det);
        double lb[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int q = 0; q < 4; ++q) {
            double phi[10];
            p2_tet_basis(r.pts[q].data(), phi);
            for (int i = 0; i < 10; ++i) lb[i] += r.w[q] * detj * f_cell[c] * phi[i];
        }
        for (int i = 0; i < 10; ++i) b[p->cell_dofs[10 * c + i]] += lb[i];
    }
    return 0;
}
void or_p2_free(or_p2* p) { delete p; }

}  // extern "C"
