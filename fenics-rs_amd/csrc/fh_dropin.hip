// fh_dropin.hip — the drop-in entry points for fenics-rs's GPU seam and its
// solver drivers:
//   solve_cg / solve_cg_gpu_or_cpu on faer CSC raw parts (cg.rs:40-46,
//   gpu.rs:1542-1602), solve_csr_cg_gpu (gpu.rs:1145-1442),
//   solve_bsr3_pcg_gpu (gpu.rs:775-1141), GpuBsr3Context (gpu.rs:432-658),
//   spmv_bsr3_multi_gpu (gpu.rs:662-771), solve_poisson_2d/3d
//   (poisson_solver.rs:26-59,99-132) and solve_elasticity_3d
//   (elasticity_solver.rs:31-44).
// Host arrays are laid out into SELL-64 on the host once and uploaded; all
// arithmetic runs on the device.
#include <algorithm>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

#include "fh_internal.hpp"

namespace fh {

// CSR (u32 row_ptr/cols, bs*bs row-major values per block) -> device SELL-64.
fh_matrix* matrix_from_csr(int dev, uint32_t nrows, int bs, const uint32_t* row_ptr,
                           const uint32_t* cols, const double* vals) {
    FH_CHECK(row_ptr && (row_ptr[nrows] == 0 || (cols && vals)), "null CSR arrays");
    FH_CHECK(row_ptr[0] == 0, "row_ptr[0] must be 0");
    const int bs2 = bs * bs;
    for (uint32_t r = 0; r < nrows; ++r) FH_CHECK(row_ptr[r + 1] >= row_ptr[r], "row_ptr not monotone");
    const uint64_t nnzb = row_ptr[nrows];
    for (uint64_t i = 0; i < nnzb; ++i) FH_CHECK(cols[i] < nrows, "column index out of range");
    DeviceCtx& d = device(dev);
    hipStream_t s = d.stream;
    auto p = std::make_shared<Pattern>();
    p->dev = dev;
    p->nrows = nrows;
    p->nnzb = nnzb;
    p->nslices = cdiv(nrows, SLICE);
    std::vector<uint32_t> soff(p->nslices + 1, 0), rlen(std::max<uint32_t>(nrows, 1), 0);
    uint32_t mw = 0;
    for (uint32_t sl = 0; sl < p->nslices; ++sl) {
        uint32_t w = 0;
        for (uint32_t l = 0; l < SLICE; ++l) {
            const uint32_t r = sl * SLICE + l;
            if (r < nrows) {
                rlen[r] = row_ptr[r + 1] - row_ptr[r];
                w = std::max(w, rlen[r]);
            }
        }
        mw = std::max(mw, w);
        const uint64_t next = (uint64_t)soff[sl] + (uint64_t)w * SLICE;
        FH_CHECK(next < 0xffffffffull, "matrix too large for 32-bit SELL offsets");
        soff[sl + 1] = (uint32_t)next;
    }
    p->max_width = mw;
    p->sell_entries = soff[p->nslices];
    std::vector<uint32_t> scols(std::max<uint64_t>(p->sell_entries, 1), 0);
    std::vector<double> svals(std::max<uint64_t>(p->sell_entries * bs2, 1), 0.0);
    for (uint32_t r = 0; r < nrows; ++r) {
        const uint32_t sl = r / SLICE, l = r % SLICE;
        const uint32_t w = (soff[sl + 1] - soff[sl]) / SLICE;
        for (uint32_t k = 0; k < w; ++k) {
            const uint64_t e = (uint64_t)soff[sl] + (uint64_t)k * SLICE;
            if (k < rlen[r]) {
                const uint64_t src = row_ptr[r] + k;
                scols[e + l] = cols[src];
                for (int j = 0; j < bs2; ++j) svals[vpos(e, j, l, bs2)] = vals[src * bs2 + j];
            } else {
                scols[e + l] = r;
            }
        }
    }
    p->row_ptr.alloc(dev, nrows + 1);
    p->cols.alloc(dev, std::max<uint64_t>(nnzb, 1));
    p->slice_off.alloc(dev, p->nslices + 1);
    p->row_len.alloc(dev, rlen.size());
    p->sell_cols.alloc(dev, scols.size());
    h2d(p->row_ptr.p, row_ptr, nrows + 1, s);
    h2d(p->cols.p, cols, nnzb, s);
    h2d(p->slice_off.p, soff.data(), soff.size(), s);
    h2d(p->row_len.p, rlen.data(), rlen.size(), s);
    h2d(p->sell_cols.p, scols.data(), scols.size(), s);
    auto* a = new fh_matrix;
    a->dev = dev;
    a->bs = bs;
    a->pat = p;
    a->n = nrows * bs;
    a->nnz_scalar = nnzb * bs2;
    a->vals.alloc(dev, svals.size());
    h2d(a->vals.p, svals.data(), svals.size(), s);
    FH_HIP(hipStreamSynchronize(s));
    return a;
}

// faer CSC raw parts -> device SELL-64 (gpu.rs:1552-1583: rows filled in
// ascending column order, so every row's columns stay sorted).
fh_matrix* matrix_from_csc(int dev, uint64_t n, const uint64_t* col_ptr, const uint64_t* row_idx,
                           const double* vals) {
    FH_CHECK(col_ptr, "null CSC arrays");
    FH_CHECK(n < 0xffffffffull, "matrix too large");
    FH_CHECK(col_ptr[0] == 0, "col_ptr[0] must be 0");
    const uint64_t nnz = col_ptr[n];
    FH_CHECK(nnz < 0xffffffffull, "too many nonzeros");
    FH_CHECK(nnz == 0 || (row_idx && vals), "null CSC arrays");
    for (uint64_t c = 0; c < n; ++c) FH_CHECK(col_ptr[c + 1] >= col_ptr[c], "col_ptr not monotone");
    for (uint64_t i = 0; i < nnz; ++i) FH_CHECK(row_idx[i] < n, "row index out of range");
    std::vector<uint32_t> rp(n + 1, 0), ci(nnz);
    std::vector<double> v(nnz);
    for (uint64_t i = 0; i < nnz; ++i) rp[row_idx[i] + 1]++;
    for (uint64_t r = 0; r < n; ++r) rp[r + 1] += rp[r];
    std::vector<uint32_t> cur(rp.begin(), rp.end() - 1);
    for (uint64_t c = 0; c < n; ++c)
        for (uint64_t p = col_ptr[c]; p < col_ptr[c + 1]; ++p) {
            const uint32_t q = cur[row_idx[p]]++;
            ci[q] = (uint32_t)c;
            v[q] = vals[p];
        }
    return matrix_from_csr(dev, (uint32_t)n, 1, rp.data(), ci.data(), v.data());
}

struct MatPtr {
    fh_matrix* p = nullptr;
    ~MatPtr() { delete p; }
};

static int solve_host(fh_matrix* a, const double* b, const double* x0, double tol, uint64_t max_iter,
                      int precision, double* x_out, fh_solve_stats* stats, const double* inv_diag,
                      const double* block_dinv) {
    DeviceCtx& d = device(a->dev);
    hipStream_t s = d.stream;
    DBuf<double> db(a->dev, a->n), dx(a->dev, a->n), dx0, dinv;
    h2d(db.p, b, a->n, s);
    if (x0) {
        dx0.alloc(a->dev, a->n);
        h2d(dx0.p, x0, a->n, s);
    }
    SolveArgs sa{db.p, dx0.p, dx.p, tol, max_iter, precision, FH_ORDER_FAST};
    if (const char* o = std::getenv("FH_ORDER")) sa.order = std::atoi(o);
    if (inv_diag) {
        dinv.alloc(a->dev, a->n);
        h2d(dinv.p, inv_diag, a->n, s);
        sa.inv_diag_override = dinv.p;
    } else if (block_dinv) {
        dinv.alloc(a->dev, (size_t)(a->n / 3) * 9);
        h2d(dinv.p, block_dinv, dinv.n, s);
        sa.block_dinv = dinv.p;
    }
    const int code = solve_device(a, sa, stats);
    d2h(x_out, dx.p, a->n, s);
    FH_HIP(hipStreamSynchronize(s));
    return code;
}

static void report(int code) {
    if (code == FH_ERR_NOT_CONVERGED) set_last_error("CG failed to converge within the iteration limit");
    else if (code == FH_ERR_BREAKDOWN) set_last_error("Breakdown in CG: curvature p^T A p is near zero");
}

__global__ void k_apply_dinv3(uint32_t nn, const double* __restrict__ D, const double* __restrict__ r,
                              double* __restrict__ z) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    const double* M = D + (uint64_t)v * 9;
    const double r0 = r[3 * v], r1 = r[3 * v + 1], r2 = r[3 * v + 2];
#pragma unroll
    for (int a = 0; a < 3; ++a) z[3 * v + a] = M[3 * a] * r0 + M[3 * a + 1] * r1 + M[3 * a + 2] * r2;
}
// x += omega D^{-1} (r - y)   (gpu.rs:627-651 bsr3_jacobi_step_f64)
__global__ void k_jacobi3(uint32_t nn, const double* __restrict__ D, const double* __restrict__ r,
                          const double* __restrict__ y, double* __restrict__ x, double omega) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    const double* M = D + (uint64_t)v * 9;
    const double e0 = r[3 * v] - y[3 * v], e1 = r[3 * v + 1] - y[3 * v + 1],
                 e2 = r[3 * v + 2] - y[3 * v + 2];
#pragma unroll
    for (int a = 0; a < 3; ++a)
        x[3 * v + a] = x[3 * v + a] + omega * (M[3 * a] * e0 + M[3 * a + 1] * e1 + M[3 * a + 2] * e2);
}

}  // namespace fh

struct fh_bsr3_ctx {
    fh_matrix* a = nullptr;
    fh::DBuf<double> dinv, x, y, r;
    std::string name;
    ~fh_bsr3_ctx() { delete a; }
};

using namespace fh;

extern "C" {

int fh_solve_csc_cg(uint64_t n, const uint64_t* col_ptr, const uint64_t* row_idx,
                    const double* vals, const double* b, const double* x0, double tol,
                    uint64_t max_iter, int device_idx, double* x_out, fh_solve_stats* stats) {
    int code = FH_OK;
    const int g = guard([&] {
        FH_CHECK(col_ptr && b && x_out, "null argument");
        MatPtr a{matrix_from_csc(resolve_device(device_idx), n, col_ptr, row_idx, vals)};
        code = solve_host(a.p, b, x0, tol, max_iter, FH_PREC_FP64, x_out, stats, nullptr, nullptr);
        report(code);
    });
    return g != FH_OK ? g : code;
}

int fh_matrix_from_csc(uint64_t n, const uint64_t* col_ptr, const uint64_t* row_idx,
                       const double* vals, int device_idx, fh_matrix** out) {
    return guard([&] {
        FH_CHECK(out && col_ptr, "null argument");
        *out = matrix_from_csc(resolve_device(device_idx), n, col_ptr, row_idx, vals);
    });
}

int fh_solve_csr_cg(uint64_t n, const uint32_t* row_ptr, const uint32_t* col_idx,
                    const double* vals, const double* inv_diag, const double* b,
                    const double* x0, double rtol, uint64_t max_iter, int device_idx,
                    double* x_out, fh_solve_stats* stats) {
    int code = FH_OK;
    const int g = guard([&] {
        FH_CHECK(row_ptr && inv_diag && b && x_out, "null argument");
        FH_CHECK(n < 0xffffffffull, "matrix too large");
        const int dev = resolve_device(device_idx);
        MatPtr a{matrix_from_csr(dev, (uint32_t)n, 1, row_ptr, col_idx, vals)};
        code = solve_host(a.p, b, x0, rtol, max_iter, FH_PREC_FP64, x_out, stats, inv_diag, nullptr);
        report(code);
    });
    return g != FH_OK ? g : code;
}

int fh_solve_bsr3_pcg(uint64_t n_nodes, const uint32_t* row_ptr, const uint32_t* cols,
                      const double* vals_flat, const double* dinv_flat, const double* b,
                      const double* x0, double rtol, uint64_t max_iter, int device_idx,
                      double* x_out, fh_solve_stats* stats) {
    int code = FH_OK;
    const int g = guard([&] {
        FH_CHECK(row_ptr && dinv_flat && b && x_out, "null argument");
        FH_CHECK(n_nodes < 0x3fffffffull, "matrix too large");
        const int dev = resolve_device(device_idx);
        MatPtr a{matrix_from_csr(dev, (uint32_t)n_nodes, 3, row_ptr, cols, vals_flat)};
        code = solve_host(a.p, b, x0, rtol, max_iter, FH_PREC_FP64, x_out, stats, nullptr, dinv_flat);
        report(code);
    });
    return g != FH_OK ? g : code;
}

int fh_bsr3_ctx_create(uint64_t n_nodes, const uint32_t* row_ptr, const uint32_t* cols,
                       const double* vals_flat, const double* dinv_flat, int device_idx,
                       fh_bsr3_ctx** out) {
    return guard([&] {
        FH_CHECK(out && row_ptr && dinv_flat, "null argument");
        const int dev = resolve_device(device_idx);
        DeviceCtx& d = device(dev);
        std::unique_ptr<fh_bsr3_ctx> c(new fh_bsr3_ctx);
        c->a = matrix_from_csr(dev, (uint32_t)n_nodes, 3, row_ptr, cols, vals_flat);
        c->dinv.alloc(dev, n_nodes * 9);
        h2d(c->dinv.p, dinv_flat, n_nodes * 9, d.stream);
        c->x.alloc(dev, n_nodes * 3);
        c->y.alloc(dev, n_nodes * 3);
        c->r.alloc(dev, n_nodes * 3);
        c->name = d.prop.name;
        FH_HIP(hipStreamSynchronize(d.stream));
        *out = c.release();
    });
}

int fh_bsr3_ctx_spmv(fh_bsr3_ctx* c, const double* x, double* y) {
    return guard([&] {
        FH_CHECK(c && x && y, "null argument");
        DeviceCtx& d = device(c->a->dev);
        h2d(c->x.p, x, c->a->n, d.stream);
        spmv_f64(c->a, c->x.p, c->y.p, FH_ORDER_FAST, d.stream);
        d2h(y, c->y.p, c->a->n, d.stream);
        FH_HIP(hipStreamSynchronize(d.stream));
    });
}

// gpu.rs:544-657: z <- z + omega D^{-1} (r - A z), `sweeps` times; with
// zero_initial_z the first sweep is z = D^{-1} r (one SpMV saved).
int fh_bsr3_ctx_jacobi_smooth(fh_bsr3_ctx* c, const double* r, double* z, uint64_t sweeps,
                              double omega, int zero_initial_z) {
    return guard([&] {
        FH_CHECK(c && r && z, "null argument");
        if (sweeps == 0) return;
        DeviceCtx& d = device(c->a->dev);
        const uint32_t nn = c->a->n / 3;
        h2d(c->r.p, r, c->a->n, d.stream);
        uint64_t start = 0;
        if (zero_initial_z) {
            hipLaunchKernelGGL(k_apply_dinv3, dim3(cdiv(nn, 256)), dim3(256), 0, d.stream, nn,
                               c->dinv.p, c->r.p, c->x.p);
            start = 1;
        } else {
            h2d(c->x.p, z, c->a->n, d.stream);
        }
        for (uint64_t k = start; k < sweeps; ++k) {
            spmv_f64(c->a, c->x.p, c->y.p, FH_ORDER_FAST, d.stream);
            hipLaunchKernelGGL(k_jacobi3, dim3(cdiv(nn, 256)), dim3(256), 0, d.stream, nn,
                               c->dinv.p, c->r.p, c->y.p, c->x.p, omega);
        }
        FH_LAUNCH_CHECK("jacobi_smooth");
        d2h(z, c->x.p, c->a->n, d.stream);
        FH_HIP(hipStreamSynchronize(d.stream));
    });
}

int fh_bsr3_ctx_device_name(const fh_bsr3_ctx* c, char* buf, size_t len) {
    return guard([&] {
        FH_CHECK(c && buf && len, "null argument");
        std::snprintf(buf, len, "%s", c->name.c_str());
    });
}

int fh_bsr3_ctx_nnz_blocks(const fh_bsr3_ctx* c, uint64_t* nnzb) {
    return guard([&] {
        FH_CHECK(c && nnzb, "null argument");
        *nnzb = c->a->pat->nnzb;
    });
}

int fh_bsr3_ctx_destroy(fh_bsr3_ctx* c) {
    return guard([&] { delete c; });
}

// gpu.rs:662-771: contiguous block-row ranges, one per usable device (>= 256 MiB
// free), x replicated, local y slices concatenated.
int fh_spmv_bsr3_multi(uint64_t n_nodes, const uint32_t* row_ptr, const uint32_t* cols,
                       const double* vals_flat, const double* x, uint64_t max_gpus, double* y,
                       int* n_used) {
    return guard([&] {
        FH_CHECK(row_ptr && x && y, "null argument");
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
            throw Error(FH_ERR_NO_DEVICE, "no HIP device available");
        std::vector<int> devs;
        for (int i = 0; i < count && devs.size() < std::max<uint64_t>(max_gpus, 1); ++i) {
            device(i);
            size_t fr = 0, tot = 0;
            FH_HIP(hipMemGetInfo(&fr, &tot));
            if ((fr >> 20) >= 256) devs.push_back(i);
        }
        if (devs.empty()) devs.push_back(0);
        const uint64_t ng = devs.size(), chunk = (n_nodes + ng - 1) / ng;
        int used = 0;
        for (uint64_t sl = 0; sl < ng; ++sl) {
            const uint64_t r0 = sl * chunk, r1 = std::min(n_nodes, (sl + 1) * chunk);
            if (r0 >= r1) continue;
            const uint32_t b0 = row_ptr[r0];
            const int dev = devs[sl];
            DeviceCtx& d = device(dev);
            // this device's block rows in a square frame whose other rows are
            // empty (columns stay global: x is replicated)
            std::vector<uint32_t> frp(n_nodes + 1, 0);
            for (uint64_t r = 0; r < n_nodes; ++r) {
                const uint32_t len = (r >= r0 && r < r1) ? row_ptr[r + 1] - row_ptr[r] : 0;
                frp[r + 1] = frp[r] + len;
            }
            MatPtr a{matrix_from_csr(dev, (uint32_t)n_nodes, 3, frp.data(), cols + b0,
                                     vals_flat + (uint64_t)b0 * 9)};
            DBuf<double> dx(dev, n_nodes * 3), dy(dev, n_nodes * 3);
            h2d(dx.p, x, n_nodes * 3, d.stream);
            spmv_f64(a.p, dx.p, dy.p, FH_ORDER_FAST, d.stream);
            d2h(y + r0 * 3, dy.p + r0 * 3, (r1 - r0) * 3, d.stream);
            FH_HIP(hipStreamSynchronize(d.stream));
            ++used;
        }
        if (n_used) *n_used = used;
    });
}

int fh_solve_poisson(fh_mesh* m, const double* fq, const double* bvals, double tol,
                     uint64_t max_iter, int precision, double* u_out, fh_solve_stats* stats) {
    int code = FH_OK;
    const int g = guard([&] {
        FH_CHECK(m && fq && u_out, "null argument");
        DeviceCtx& d = device(m->dev);
        hipStream_t s = d.stream;
        const int nq = m->dim == 2 ? 3 : 4;
        // poisson_solver.rs:37-52
        MatPtr A{assemble_p1(m, 0)};
        DBuf<double> dfq(m->dev, (size_t)m->nc * nq), b(m->dev, m->nv), bc(m->dev, m->nv),
            u(m->dev, m->nv);
        h2d(dfq.p, fq, dfq.n, s);
        rhs_p1_dev(m, dfq.p, b.p, s);
        DBuf<uint8_t> mask(m->dev, m->nv);
        boundary_vertex_mask(m, mask.p, s, nullptr, nullptr);
        const std::vector<uint64_t> bnd = mask_to_sorted(m->dev, mask.p, m->nv, s);
        std::vector<double> gv(bnd.size(), 0.0);
        if (bvals) std::copy(bvals, bvals + bnd.size(), gv.begin());
        MatPtr Ac{apply_dirichlet(A.p, b.p, bnd.size(), bnd.data(), gv.data(), bc.p, s)};
        delete A.p;
        A.p = nullptr;
        SolveArgs sa{bc.p, nullptr, u.p, tol, max_iter, precision, FH_ORDER_FAST};
        if (const char* o = std::getenv("FH_ORDER")) sa.order = std::atoi(o);
        code = solve_device(Ac.p, sa, stats);
        d2h(u_out, u.p, m->nv, s);
        FH_HIP(hipStreamSynchronize(s));
        report(code);
    });
    return g != FH_OK ? g : code;
}

int fh_solve_elasticity_3d(fh_mesh* m, double E, double nu, const double* E_cell,
                           const double f[3], uint64_t nfix, const uint64_t* dofs,
                           const double* vals, double tol, uint64_t max_iter, int precision,
                           double* u_out, double* u_dev, fh_solve_stats* stats) {
    int code = FH_OK;
    const int g = guard([&] {
        FH_CHECK(m && f && (u_out || u_dev), "null argument");
        FH_CHECK(nfix == 0 || (dofs && vals), "null Dirichlet list");
        DeviceCtx& d = device(m->dev);
        hipStream_t s = d.stream;
        const size_t n = 3 * (size_t)m->nv;
        DBuf<double> ec;
        if (E_cell) {
            ec.alloc(m->dev, m->nc);
            h2d(ec.p, E_cell, m->nc, s);
        }
        // elasticity_solver.rs:39-44
        MatPtr K{assemble_elasticity(m, E, nu, ec.p)};
        DBuf<double> b(m->dev, n), bc(m->dev, n), u;
        body_force_dev(m, f, b.p, s);
        MatPtr Kc{apply_dirichlet(K.p, b.p, nfix, dofs, vals, bc.p, s)};
        delete K.p;
        K.p = nullptr;
        double* ux = u_dev;
        if (!ux) {
            u.alloc(m->dev, n);
            ux = u.p;
        }
        SolveArgs sa{bc.p, nullptr, ux, tol, max_iter, precision, FH_ORDER_FAST};
        if (const char* o = std::getenv("FH_ORDER")) sa.order = std::atoi(o);
        code = solve_device(Kc.p, sa, stats);
        if (u_out) d2h(u_out, ux, n, s);
        FH_HIP(hipStreamSynchronize(s));
        report(code);
    });
    return g != FH_OK ? g : code;
}

}  // extern "C"
