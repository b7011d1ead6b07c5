// fh_transient.hip — transient heat conduction on the GPU: the time-marching
// consumer of the CG (transient.rs:32-97 TransientHeatSolver::step and
// add_scaled_matrices, heat_solver.rs:23-108 solve_transient_heat_2d).
//
// Each implicit step solves (M + theta dt K) u_{n+1} = rhs with theta = 1
// (Backward Euler) or 1/2 (Crank-Nicolson).  The reference factors that
// matrix densely with partial pivoting (direct.rs:26-53); for a mass matrix
// and a non-negative dt K it is symmetric positive definite, so here it goes
// to the device Jacobi-CG, warm-started from u_n, to ||r||/||b|| < tol.  The
// step matrices share the mesh pattern, so M + alpha K is one elementwise pass
// over the SELL values; the right-hand side uses the column-ordered FP64 SpMV
// (bitwise the reference's spmv), and the whole march stays on the device.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <vector>

#include "fh_internal.hpp"

namespace fh {

// out = a + alpha * b entry by entry (transient.rs:81-95: the triplets of a and
// alpha * b summed per position); padding entries are 0 + alpha * 0 = 0.
__global__ void k_add_scaled(const double* __restrict__ a, const double* __restrict__ b,
                             double alpha, double* __restrict__ out, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = a[i] + alpha * b[i];
}

// heat_solver.rs:49-58: every stored stiffness value times the diffusivity
// (in place allowed)
__global__ void k_scale_vals(const double* a, double c, double* out, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = c * a[i];
}

// transient.rs:47-51 (BE, cn = 0): rhs = M u_n + dt f_{n+1};
// transient.rs:60-64 (CN, cn = 1): rhs = (M - dt/2 K) u_n + dt/2 (f_n + f_{n+1})
__global__ void k_step_rhs(const double* __restrict__ y, const double* __restrict__ fn,
                           const double* __restrict__ fnext, double h, int cn,
                           double* __restrict__ rhs, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        rhs[i] = cn ? y[i] + h * (fn[i] + fnext[i]) : y[i] + h * fnext[i];
}

// u[dofs[k]] = vals[k]; the list holds each dof once (heat_solver.rs:82-86
// assigns in list order, so the host keeps the last value of a repeated dof)
__global__ void k_scatter_bc(const uint32_t* __restrict__ dofs, const double* __restrict__ vals,
                             uint64_t m, double* __restrict__ u) {
    const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (k < m) u[dofs[k]] = vals[k];
}

namespace {

using MatU = std::unique_ptr<fh_matrix>;

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

unsigned egrid(uint64_t n) { return std::max(1u, std::min(cdiv(n, 256), 8192u)); }

template <class T>
bool same_dev_array(const DBuf<T>& a, const DBuf<T>& b, size_t n, hipStream_t s) {
    if (n == 0) return true;
    std::vector<T> ha(n), hb(n);
    d2h(ha.data(), a.p, n, s);
    d2h(hb.data(), b.p, n, s);
    FH_HIP(hipStreamSynchronize(s));
    return std::equal(ha.begin(), ha.end(), hb.begin());
}

// Same block pattern and the same SELL placement of every entry.
bool same_layout(const fh_matrix* a, const fh_matrix* b, hipStream_t s) {
    if (a->pat == b->pat) return true;
    const Pattern &p = *a->pat, &q = *b->pat;
    if (a->bs != b->bs || p.nrows != q.nrows || p.nnzb != q.nnzb || p.nslices != q.nslices ||
        p.sell_entries != q.sell_entries)
        return false;
    return same_dev_array(p.row_ptr, q.row_ptr, p.nrows + 1, s) &&
           same_dev_array(p.cols, q.cols, p.nnzb, s) &&
           same_dev_array(p.slice_off, q.slice_off, p.nslices + 1, s) &&
           same_dev_array(p.sell_cols, q.sell_cols, p.sell_entries, s);
}

}  // namespace

// transient.rs:72-97.  Operands on one pattern (the mass and stiffness of one
// mesh, the only pairing the reference's callers make) are combined on the
// device in place of the triplet round trip; any other pair goes through the
// host union of the exported CSC parts, with the reference's summation.
fh_matrix* add_scaled(const fh_matrix* a, const fh_matrix* b, double alpha) {
    FH_CHECK(a->n == b->n, "add_scaled_matrices needs operands of one size");
    FH_CHECK(a->dev == b->dev, "add_scaled_matrices needs operands on one device");
    DeviceCtx& d = device(a->dev);
    hipStream_t s = d.stream;
    // A condensed or folded operand exports fewer entries than it stores, so
    // only plain assembled operators share the stored pattern with the result.
    if (!a->condensed && !b->condensed && !a->folded && !b->folded && same_layout(a, b, s)) {
        std::unique_ptr<fh_matrix> o(new fh_matrix);
        o->dev = a->dev;
        o->bs = a->bs;
        o->pat = a->pat;
        o->n = a->n;
        o->comp_diag = a->comp_diag && b->comp_diag;
        o->nnz_scalar = o->comp_diag ? a->pat->nnzb * (uint64_t)a->bs
                                     : a->pat->nnzb * (uint64_t)a->bs * a->bs;
        const uint64_t nv = sell_value_count(a);
        o->vals.alloc(a->dev, std::max<uint64_t>(nv, 1));
        if (nv) {
            hipLaunchKernelGGL(k_add_scaled, dim3(egrid(nv)), dim3(256), 0, s, a->vals.p,
                               b->vals.p, alpha, o->vals.p, nv);
            FH_LAUNCH_CHECK("add_scaled");
        }
        FH_HIP(hipStreamSynchronize(s));
        return o.release();
    }
    // General pair: triplets of a, then of alpha * b, summed per position in
    // that order (a + alpha b where both are stored).
    const HostCsr ha = export_scalar_csr(a), hb = export_scalar_csr(b);
    const uint64_t n = a->n;
    std::vector<uint32_t> rp(n + 1, 0), ci;
    std::vector<double> v;
    std::vector<std::pair<uint32_t, double>> row;
    for (uint64_t r = 0; r < n; ++r) {
        row.clear();
        for (uint64_t p = ha.rp[r]; p < ha.rp[r + 1]; ++p) row.push_back({ha.ci[p], ha.v[p]});
        for (uint64_t p = hb.rp[r]; p < hb.rp[r + 1]; ++p)
            row.push_back({hb.ci[p], alpha * hb.v[p]});
        std::stable_sort(row.begin(), row.end(),
                         [](const auto& x, const auto& y) { return x.first < y.first; });
        for (size_t k = 0; k < row.size(); ++k) {
            if (k > 0 && row[k].first == row[k - 1].first) {
                v.back() = v.back() + row[k].second;
            } else {
                ci.push_back(row[k].first);
                v.push_back(row[k].second);
            }
        }
        FH_CHECK(ci.size() < 0xffffffffull, "too many nonzeros");
        rp[r + 1] = (uint32_t)ci.size();
    }
    return matrix_from_csr(a->dev, (uint32_t)n, 1, rp.data(), ci.data(), v.data());
}

namespace {

struct StepBufs {
    DBuf<double> u, un, fc, fn, y, rhs;
    StepBufs(int dev, uint64_t n)
        : u(dev, n), un(dev, n), fc(dev, n), fn(dev, n), y(dev, n), rhs(dev, n) {}
};

// One implicit step on device vectors: rhs from R (M for BE, M - dt/2 K for
// CN), then lhs u_next = rhs by CG from x0 = u.
int device_step(const fh_matrix* lhs, const fh_matrix* rmat, int cn, double h, StepBufs& b,
                double tol, uint64_t max_iter, int precision, fh_solve_stats* st,
                hipStream_t s) {
    const uint64_t n = lhs->n;
    spmv_f64(rmat, b.u.p, b.y.p, FH_ORDER_ORDERED, s);
    hipLaunchKernelGGL(k_step_rhs, dim3(egrid(n)), dim3(256), 0, s, b.y.p, b.fc.p, b.fn.p, h, cn,
                       b.rhs.p, n);
    FH_LAUNCH_CHECK("step_rhs");
    SolveArgs sa{b.rhs.p, b.u.p, b.un.p, tol, max_iter, precision, FH_ORDER_FAST};
    return solve_device(lhs, sa, st);
}

void check_scheme(int scheme) {
    FH_CHECK(scheme == FH_SCHEME_BACKWARD_EULER || scheme == FH_SCHEME_CRANK_NICOLSON,
             "unknown time stepping scheme");
}

[[noreturn]] void solve_failed(int code) {
    throw Error(code, code == FH_ERR_BREAKDOWN
                          ? "Breakdown in CG: curvature p^T A p is near zero"
                          : "CG failed to converge within the iteration limit");
}

}  // namespace

}  // namespace fh

using namespace fh;

extern "C" {

int fh_add_scaled_matrices(const fh_matrix* a, const fh_matrix* b, double alpha,
                           fh_matrix** out) {
    return guard([&] {
        FH_CHECK(a && b && out, "null argument");
        *out = add_scaled(a, b, alpha);
    });
}

int fh_transient_heat_step(const fh_matrix* M, const fh_matrix* K, int scheme, double dt,
                           const double* u_n, const double* f_n, const double* f_next,
                           double tol, uint64_t max_iter, int precision, double* u_next,
                           fh_solve_stats* stats) {
    return guard([&] {
        FH_CHECK(M && K && u_n && f_n && f_next && u_next, "null argument");
        check_scheme(scheme);
        FH_CHECK(M->n == K->n, "mass and stiffness sizes differ");
        DeviceCtx& d = device(M->dev);
        hipStream_t s = d.stream;
        const uint64_t n = M->n;
        const bool cn = scheme == FH_SCHEME_CRANK_NICOLSON;
        const double h = cn ? 0.5 * dt : dt;
        MatU lhs(add_scaled(M, K, h)), rmat(cn ? add_scaled(M, K, -h) : nullptr);
        StepBufs b(M->dev, n);
        h2d(b.u.p, u_n, n, s);
        h2d(b.fc.p, f_n, n, s);
        h2d(b.fn.p, f_next, n, s);
        const int code = device_step(lhs.get(), cn ? rmat.get() : M, cn ? 1 : 0, h, b, tol,
                                     max_iter, precision, stats, s);
        if (code != FH_OK) solve_failed(code);
        d2h(u_next, b.un.p, n, s);
        FH_HIP(hipStreamSynchronize(s));
    });
}

int fh_solve_transient_heat_2d(fh_mesh* m, double diffusivity, const double* u_initial,
                               uint64_t nbc, const uint64_t* bc_dofs, const double* bc_vals,
                               double dt, uint64_t num_steps, int scheme, fh_heat_source_fn source,
                               void* source_ctx, double tol, uint64_t max_iter, int precision,
                               double* history, double* final_out, fh_transient_stats* stats) {
    return guard([&] {
        FH_CHECK(m && u_initial && final_out, "null argument");
        FH_CHECK(nbc == 0 || (bc_dofs && bc_vals), "null boundary arrays");
        FH_CHECK(m->dim == 2, "solve_transient_heat_2d needs a triangle mesh");
        check_scheme(scheme);
        DeviceCtx& d = device(m->dev);
        hipStream_t s = d.stream;
        const uint64_t n = m->nv;
        const auto t0 = now_ms();
        // heat_solver.rs:82-86 / 92-96: later entries of the list win
        std::map<uint64_t, double> bc;
        for (uint64_t k = 0; k < nbc; ++k) {
            FH_CHECK(bc_dofs[k] < n, "Dirichlet dof out of range");
            bc[bc_dofs[k]] = bc_vals[k];
        }
        std::vector<uint32_t> hd;
        std::vector<double> hv;
        for (const auto& kv : bc) {
            hd.push_back((uint32_t)kv.first);
            hv.push_back(kv.second);
        }
        const uint64_t nb = hd.size();
        DBuf<uint32_t> ddofs(m->dev, std::max<uint64_t>(nb, 1));
        DBuf<double> dvals(m->dev, std::max<uint64_t>(nb, 1));
        h2d(ddofs.p, hd.data(), nb, s);
        h2d(dvals.p, hv.data(), nb, s);
        // heat_solver.rs:45-58: M, K scaled by the diffusivity; the step
        // matrices are built once (transient.rs:46 / 57-58 rebuild them per step)
        MatU mass(assemble_p1(m, 1)), stiff(assemble_p1(m, 0));
        const uint64_t nvals = sell_value_count(stiff.get());
        if (nvals)
            hipLaunchKernelGGL(k_scale_vals, dim3(egrid(nvals)), dim3(256), 0, s,
                               stiff->vals.p, diffusivity, stiff->vals.p, nvals);
        FH_LAUNCH_CHECK("scale_vals");
        FH_HIP(hipStreamSynchronize(s));
        const bool cn = scheme == FH_SCHEME_CRANK_NICOLSON;
        const double h = cn ? 0.5 * dt : dt;
        MatU lhs(add_scaled(mass.get(), stiff.get(), h));
        MatU rmat(cn ? add_scaled(mass.get(), stiff.get(), -h) : nullptr);
        const fh_matrix* R = cn ? rmat.get() : mass.get();

        StepBufs b(m->dev, n);
        h2d(b.u.p, u_initial, n, s);
        std::vector<double> verts, fc, fnx;
        if (source) {
            verts.resize(n * 2);
            d2h(verts.data(), m->verts.p, n * 2, s);
            fc.resize(n);
            fnx.resize(n);
        } else {
            fill_f64(b.fc.p, n, 0.0, s);
            fill_f64(b.fn.p, n, 0.0, s);
        }
        FH_HIP(hipStreamSynchronize(s));
        if (history) std::copy(u_initial, u_initial + n, history);  // heat_solver.rs:66
        uint64_t it_total = 0, it_max = 0, it_low = 0;
        double solve_ms = 0.0;
        for (uint64_t step = 1; step <= num_steps; ++step) {
            const double t_curr = (double)(step - 1) * dt, t_next = (double)step * dt;
            if (source) {  // heat_solver.rs:73-79: both times, vertex by vertex
                for (uint64_t i = 0; i < n; ++i) {
                    fc[i] = source(t_curr, &verts[2 * i], source_ctx);
                    fnx[i] = source(t_next, &verts[2 * i], source_ctx);
                }
                h2d(b.fc.p, fc.data(), n, s);
                h2d(b.fn.p, fnx.data(), n, s);
            }
            if (nb)
                hipLaunchKernelGGL(k_scatter_bc, dim3(cdiv(nb, 256)), dim3(256), 0, s, ddofs.p,
                                   dvals.p, nb, b.u.p);
            fh_solve_stats st{};
            const int code =
                device_step(lhs.get(), R, cn ? 1 : 0, h, b, tol, max_iter, precision, &st, s);
            if (code != FH_OK) solve_failed(code);
            it_total += st.iterations;
            it_low += st.iterations_low;
            it_max = std::max<uint64_t>(it_max, st.iterations);
            solve_ms += st.wall_ms;
            if (nb)
                hipLaunchKernelGGL(k_scatter_bc, dim3(cdiv(nb, 256)), dim3(256), 0, s, ddofs.p,
                                   dvals.p, nb, b.un.p);
            std::swap(b.u.p, b.un.p);
            if (history) d2h(history + step * n, b.u.p, n, s);
        }
        d2h(final_out, b.u.p, n, s);
        FH_HIP(hipStreamSynchronize(s));
        if (stats) {
            stats->steps = num_steps;
            stats->cg_iterations = it_total;
            stats->cg_iterations_low = it_low;
            stats->cg_iterations_max = it_max;
            stats->solve_ms = solve_ms;
            stats->wall_ms = now_ms() - t0;
            stats->device = m->dev;
            stats->precision = precision;
        }
    });
}

}  // extern "C"
