// fh_batched.hip — many independent Poisson P1 problems, several right-hand
// sides each, solved together (new; no counterpart in the reference, which
// solves one system per call through solve_poisson_2d, poisson_solver.rs:26).
//
// Layout: the instances form one block-diagonal system over the disjoint union
// of their unit_square meshes (generators.rs:24-55), assembled and condensed by
// the ordinary single-system kernels.  Each instance's rows are padded to a
// multiple of 256 (four 64-row slices = one workgroup) with isolated, fixed
// rows, so every workgroup of the batched kernels belongs to exactly one
// instance.  Vectors hold the NR right-hand sides interleaved per row
// ([row][NR]): one SpMM pass reads each matrix entry once for all of them.
//
// Multi-RHS CG: every (instance, rhs) column runs its own Jacobi-PCG
// recurrences (alpha, beta per column — block CG with block size 1 per column,
// sharing every SpMM); a converged column is frozen, a finished instance is
// skipped.  One iteration = 4 launches over all instances (SpMM, per-instance
// reduction, update, per-instance reduction), all arguments fixed, so a chunk
// of iterations is captured once into a HIP graph and replayed.  Reductions
// are fixed-order (deterministic).  Mixed precision: FP32 matrix and vectors in
// the loop, FP64 residual replacement per instance (b - A64 (x64 + x32)),
// convergence declared on the true FP64 residual.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <vector>

#include "fh_internal.hpp"

namespace fh {

constexpr int NR = 4;

struct BatchState {
    // per column (instance * NR + k)
    double* rz;      // r.z of the current iterate
    double* alpha;
    double* beta;
    double* bnorm;
    double* r_rep;   // true residual norm at the last replacement
    double* res;     // last residual norm
    uint32_t* iters; // iterations when the column finished
    uint8_t* done;   // 0 running, 1 converged, 2 not converged, 3 breakdown
    // per instance
    uint32_t* it;        // completed iterations
    uint8_t* inst_done;
    uint8_t* need_rep;
    const uint32_t* wg_begin;  // workgroup range of the instance
    const uint32_t* wg_end;
    const uint32_t* wg_inst;   // instance of each workgroup
    double tol, delta;
    uint32_t max_iter;
    uint32_t mixed, restart;
};

template <class T>
struct SellB {
    const uint32_t* soff;
    const uint32_t* rlen;
    const uint32_t* cols;
    const T* vals;
    uint32_t nrows;
    uint32_t nslices;
};

// Ap = A (z + beta p_old) for the columns of one instance; partial p.Ap.
template <class T>
__global__ __launch_bounds__(256) void k_b_spmm(SellB<T> A, const T* __restrict__ z, T* p0, T* p1,
                                                T* __restrict__ Ap, BatchState S,
                                                double* __restrict__ part) {
    __shared__ double sh[4];
    const uint32_t w = blockIdx.x, inst = S.wg_inst[w];
    if (S.inst_done[inst] || S.need_rep[inst]) return;
    const uint32_t j = S.it[inst] + 1;
    const T* pold = (j & 1) ? p0 : p1;
    T* pnew = (j & 1) ? p1 : p0;
    T bt[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) bt[k] = (T)S.beta[inst * NR + k];
    const uint32_t s = w * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    double dot[NR] = {0, 0, 0, 0};
    if (s < A.nslices) {
        const uint32_t row = s * SLICE + lane;
        const uint32_t base = A.soff[s], wd = (A.soff[s + 1] - base) / SLICE;
        const uint32_t len = row < A.nrows ? A.rlen[row] : 0;
        T acc[NR] = {0, 0, 0, 0};
        for (uint32_t q = 0; q < wd; ++q) {
            if (q < len) {
                const uint64_t e = (uint64_t)base + (uint64_t)q * SLICE + lane;
                const uint32_t c = A.cols[e];
                const T v = A.vals[e];
#pragma unroll
                for (int k = 0; k < NR; ++k) {
                    const T pc = fma(bt[k], pold[(uint64_t)c * NR + k], z[(uint64_t)c * NR + k]);
                    acc[k] = fma(v, pc, acc[k]);
                }
            }
        }
        if (row < A.nrows) {
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                const uint64_t i = (uint64_t)row * NR + k;
                const T pn = fma(bt[k], pold[i], z[i]);
                pnew[i] = pn;
                Ap[i] = acc[k];
                dot[k] = (double)pn * (double)acc[k];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const double d = fhd::block_sum256(dot[k], sh);
        if (threadIdx.x == 0) part[(uint64_t)w * NR + k] = d;
    }
}

// alpha per column from the instance's p.Ap partials (one block per instance).
__global__ __launch_bounds__(256) void k_b_red1(BatchState S, const double* __restrict__ part) {
    __shared__ double sh[4];
    const uint32_t inst = blockIdx.x;
    if (S.inst_done[inst] || S.need_rep[inst]) return;
    const uint32_t w0 = S.wg_begin[inst], w1 = S.wg_end[inst];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        double sacc = 0.0;
        for (uint32_t w = w0 + threadIdx.x; w < w1; w += 256) sacc += part[(uint64_t)w * NR + k];
        const double pap = fhd::block_sum256(sacc, sh);
        if (threadIdx.x == 0) {
            const uint32_t col = inst * NR + k;
            double a = 0.0;
            if (!S.done[col]) {
                if (fabs(pap) < 1e-30) {  // cg.rs:102-107, per column
                    S.done[col] = 3;
                    S.iters[col] = S.it[inst] + 1;
                } else {
                    a = S.rz[col] / pap;
                }
            }
            S.alpha[col] = a;
        }
    }
}

// x += alpha p, r -= alpha Ap, z = D^{-1} r; partials (r.r, r.z) per column.
template <class T>
__global__ __launch_bounds__(256) void k_b_update(const T* p0, const T* p1, const T* __restrict__ Ap,
                                                  const T* __restrict__ dinv, T* __restrict__ x,
                                                  T* __restrict__ r, T* __restrict__ z,
                                                  BatchState S, uint32_t nrows,
                                                  double* __restrict__ part) {
    __shared__ double sh[4];
    const uint32_t w = blockIdx.x, inst = S.wg_inst[w];
    if (S.inst_done[inst] || S.need_rep[inst]) return;
    const uint32_t j = S.it[inst] + 1;
    const T* p = (j & 1) ? p1 : p0;
    T al[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) al[k] = (T)S.alpha[inst * NR + k];
    const uint32_t row = w * 256 + threadIdx.x;
    double rr[NR] = {0, 0, 0, 0}, rz[NR] = {0, 0, 0, 0};
    if (row < nrows) {
        const T d = dinv[row];
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const uint64_t i = (uint64_t)row * NR + k;
            x[i] = fma(al[k], p[i], x[i]);
            const T ri = fma(-al[k], Ap[i], r[i]);
            const T zi = ri * d;
            r[i] = ri;
            z[i] = zi;
            rr[k] = (double)ri * (double)ri;
            rz[k] = (double)ri * (double)zi;
        }
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const double a = fhd::block_sum256(rr[k], sh);
        const double b = fhd::block_sum256(rz[k], sh);
        if (threadIdx.x == 0) {
            part[((uint64_t)w * NR + k) * 2] = a;
            part[((uint64_t)w * NR + k) * 2 + 1] = b;
        }
    }
}

// Per instance: residual test per column, beta, iteration count.  `init`
// handles the first residual (cg.rs:80-84 initial test); `rep` a replaced
// (true FP64) residual.
__global__ __launch_bounds__(256) void k_b_red2(BatchState S, const double* __restrict__ part,
                                                const double* __restrict__ bpart, int mode) {
    __shared__ double sh[4];
    const uint32_t inst = blockIdx.x;
    // mode 0: after an update; 1: initial residual; 2: after a replacement
    if (S.inst_done[inst]) return;
    if (mode == 0 && S.need_rep[inst]) return;
    if (mode == 2 && !S.need_rep[inst]) return;
    const uint32_t w0 = S.wg_begin[inst], w1 = S.wg_end[inst];
    const uint32_t it = S.it[inst] + (mode == 0 ? 1u : 0u);
    double rr[NR], rz[NR], bb[NR];
    for (int k = 0; k < NR; ++k) {
        double a = 0.0, b = 0.0, c = 0.0;
        for (uint32_t w = w0 + threadIdx.x; w < w1; w += 256) {
            a += part[((uint64_t)w * NR + k) * 2];
            b += part[((uint64_t)w * NR + k) * 2 + 1];
            if (mode == 1) c += bpart[(uint64_t)w * NR + k];
        }
        rr[k] = fhd::block_sum256(a, sh);
        rz[k] = fhd::block_sum256(b, sh);
        bb[k] = mode == 1 ? fhd::block_sum256(c, sh) : 0.0;
    }
    if (threadIdx.x != 0) return;
    // pass 1: stopping tests; in mixed mode a recursive residual that reached
    // tol or dropped by delta asks for an FP64 replacement of the instance
    int want_rep = 0;
    bool conv[NR];
    for (int k = 0; k < NR; ++k) {
        const uint32_t col = inst * NR + k;
        if (mode == 1) S.bnorm[col] = sqrt(bb[k]);
        conv[k] = false;
        if (S.done[col]) continue;
        const double bn = S.bnorm[col];
        const double res = sqrt(rr[k]);
        const double rel = bn > 0.0 ? res / bn : res;
        S.res[col] = res;
        conv[k] = mode == 1 ? (res < S.tol || (bn > 0.0 && rel < S.tol)) : rel < S.tol;
        if (mode == 0 && S.mixed && (conv[k] || res < S.delta * S.r_rep[col])) want_rep = 1;
    }
    if (want_rep) {  // the recurrences resume from the replaced residual
        S.it[inst] = it;  // iteration `it` was completed (x, r updated)
        S.need_rep[inst] = 1;
        return;
    }
    // pass 2: finish converged columns, advance the others
    int all_done = 1;
    for (int k = 0; k < NR; ++k) {
        const uint32_t col = inst * NR + k;
        if (S.done[col]) continue;
        if (conv[k] && (mode != 0 || !S.mixed)) {
            S.done[col] = 1;
            S.iters[col] = it;
            S.alpha[col] = 0.0;
            S.beta[col] = 0.0;
            continue;
        }
        if (mode == 0 && it >= S.max_iter) {
            S.done[col] = 2;
            S.iters[col] = it;
            S.beta[col] = 0.0;
            continue;
        }
        all_done = 0;
        if (mode != 0) S.r_rep[col] = sqrt(rr[k]);
        const bool restart = (mode == 1) || (mode == 2 && S.restart);
        S.beta[col] = restart ? 0.0 : rz[k] / S.rz[col];
        S.rz[col] = rz[k];
    }
    if (mode == 0) S.it[inst] = it;
    if (mode == 2) S.need_rep[inst] = 0;
    if (all_done) S.inst_done[inst] = 1;
}

// FP64 residual of the instances that asked for it (or all, at init):
// t = b - A64 (x64 + x32); r = t, z = D^{-1} r; partials (t.t, r.z) and b.b.
template <class T>
__global__ __launch_bounds__(256) void k_b_residual(SellB<double> A, const double* __restrict__ b,
                                                    const double* __restrict__ x64,
                                                    const float* __restrict__ x32,
                                                    const T* __restrict__ dinv, T* __restrict__ r,
                                                    T* __restrict__ z, BatchState S, int only_rep,
                                                    double* __restrict__ part,
                                                    double* __restrict__ bpart) {
    __shared__ double sh[4];
    const uint32_t w = blockIdx.x, inst = S.wg_inst[w];
    if (S.inst_done[inst]) return;
    if (only_rep && !S.need_rep[inst]) return;
    const uint32_t s = w * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    double tt[NR] = {0, 0, 0, 0}, rz[NR] = {0, 0, 0, 0}, bb[NR] = {0, 0, 0, 0};
    if (s < A.nslices) {
        const uint32_t row = s * SLICE + lane;
        const uint32_t base = A.soff[s], wd = (A.soff[s + 1] - base) / SLICE;
        const uint32_t len = row < A.nrows ? A.rlen[row] : 0;
        double acc[NR] = {0, 0, 0, 0};
        for (uint32_t q = 0; q < wd; ++q)
            if (q < len) {
                const uint64_t e = (uint64_t)base + (uint64_t)q * SLICE + lane;
                const uint32_t c = A.cols[e];
                const double v = A.vals[e];
#pragma unroll
                for (int k = 0; k < NR; ++k) {
                    double xv = x64[(uint64_t)c * NR + k];
                    if (x32) xv = xv + (double)x32[(uint64_t)c * NR + k];
                    acc[k] = fma(v, xv, acc[k]);
                }
            }
        if (row < A.nrows) {
            const T d = dinv[row];
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                const uint64_t i = (uint64_t)row * NR + k;
                const double t = b[i] - acc[k];
                const T rv = (T)t, zv = rv * d;
                r[i] = rv;
                z[i] = zv;
                tt[k] = t * t;
                rz[k] = (double)rv * (double)zv;
                bb[k] = b[i] * b[i];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const double a = fhd::block_sum256(tt[k], sh);
        const double c = fhd::block_sum256(rz[k], sh);
        const double d = fhd::block_sum256(bb[k], sh);
        if (threadIdx.x == 0) {
            part[((uint64_t)w * NR + k) * 2] = a;
            part[((uint64_t)w * NR + k) * 2 + 1] = c;
            if (bpart) bpart[(uint64_t)w * NR + k] = d;
        }
    }
}

__global__ void k_b_commit(double* __restrict__ x64, float* __restrict__ x32,
                           const uint32_t* __restrict__ wg_inst, const uint8_t* __restrict__ need,
                           uint32_t nrows, int all) {
    const uint32_t row = blockIdx.x * 256 + threadIdx.x;
    if (row >= nrows) return;
    if (!all && !need[wg_inst[blockIdx.x]]) return;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const uint64_t i = (uint64_t)row * NR + k;
        x64[i] = x64[i] + (double)x32[i];
        x32[i] = 0.0f;
    }
}

// b = M f for NR columns (FP64).
__global__ __launch_bounds__(256) void k_spmm_f64(SellB<double> A, const double* __restrict__ f,
                                                  double* __restrict__ y) {
    const uint32_t s = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (s >= A.nslices) return;
    const uint32_t row = s * SLICE + lane;
    const uint32_t base = A.soff[s], wd = (A.soff[s + 1] - base) / SLICE;
    const uint32_t len = row < A.nrows ? A.rlen[row] : 0;
    double acc[NR] = {0, 0, 0, 0};
    for (uint32_t q = 0; q < wd; ++q)
        if (q < len) {
            const uint64_t e = (uint64_t)base + (uint64_t)q * SLICE + lane;
            const uint32_t c = A.cols[e];
            const double v = A.vals[e];
#pragma unroll
            for (int k = 0; k < NR; ++k) acc[k] = fma(v, f[(uint64_t)c * NR + k], acc[k]);
        }
    if (row < A.nrows)
#pragma unroll
        for (int k = 0; k < NR; ++k) y[(uint64_t)row * NR + k] = acc[k];
}

__global__ void k_zero_fixed_rows(const uint8_t* __restrict__ fixed, double* __restrict__ b,
                                  uint32_t nrows) {
    const uint32_t row = blockIdx.x * 256 + threadIdx.x;
    if (row < nrows && fixed[row])
#pragma unroll
        for (int k = 0; k < NR; ++k) b[(uint64_t)row * NR + k] = 0.0;
}

// Union of unit_square(N_i, N_i) meshes; vertex offsets voff, cell offsets coff.
__global__ void k_union_square(uint32_t N, uint32_t voff, uint32_t coff, double* __restrict__ V,
                               uint32_t* __restrict__ C) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t nv = (uint64_t)(N + 1) * (N + 1);
    if (t < nv) {
        const uint32_t i = t % (N + 1), j = t / (N + 1);
        V[2 * (voff + t)] = (double)i / (double)N;
        V[2 * (voff + t) + 1] = (double)j / (double)N;
    }
    if (t < (uint64_t)N * N) {
        const uint32_t i = t % N, j = t / N;
        const uint32_t v00 = voff + j * (N + 1) + i, v10 = v00 + 1, v01 = v00 + N + 1, v11 = v01 + 1;
        uint32_t* c = C + 3 * ((uint64_t)coff + 2 * t);
        c[0] = v00; c[1] = v10; c[2] = v11;
        c[3] = v00; c[4] = v11; c[5] = v01;
    }
}

template <class T>
static SellB<T> sellb(const fh_matrix* a, const T* vals) {
    const Pattern& p = *a->pat;
    return {p.slice_off.p, p.row_len.p, p.sell_cols.p, vals, p.nrows, p.nslices};
}

template <class T>
static int run_batched(uint64_t n_inst, const uint64_t* sizes, const double* f_nodal, double tol,
                       uint64_t max_iter, int precision, int dev, double* u_out,
                       uint64_t* iters_out, fh_solve_stats* stats) {
    constexpr bool MIX = std::is_same<T, float>::value;
    DeviceCtx& d = device(dev);
    hipStream_t s = d.stream;
    const auto tw0 = std::chrono::steady_clock::now();
    // ---- union mesh with per-instance padding to 256 rows
    std::vector<uint32_t> voff(n_inst + 1), coff(n_inst + 1), nvi(n_inst);
    uint64_t V = 0, Cn = 0;
    for (uint64_t i = 0; i < n_inst; ++i) {
        FH_CHECK(sizes[i] >= 1 && sizes[i] < 20000, "instance size out of range");
        voff[i] = (uint32_t)V;
        coff[i] = (uint32_t)Cn;
        nvi[i] = (uint32_t)((sizes[i] + 1) * (sizes[i] + 1));
        V += (nvi[i] + 255) / 256 * 256;
        Cn += 2 * sizes[i] * sizes[i];
        FH_CHECK(V < 0x7fffffffull, "batch too large");
    }
    voff[n_inst] = (uint32_t)V;
    std::unique_ptr<fh_mesh> m(new fh_mesh);
    m->dev = dev;
    m->dim = 2;
    m->k = 3;
    m->nv = (uint32_t)V;
    m->nc = (uint32_t)Cn;
    m->verts.alloc(dev, V * 2);
    m->cells.alloc(dev, Cn * 3);
    fill_f64(m->verts.p, V * 2, 0.0, s);
    for (uint64_t i = 0; i < n_inst; ++i) {
        const uint64_t nt = std::max<uint64_t>(nvi[i], sizes[i] * sizes[i]);
        hipLaunchKernelGGL(k_union_square, dim3(cdiv(nt, 256)), dim3(256), 0, s, (uint32_t)sizes[i],
                           voff[i], coff[i], m->verts.p, m->cells.p);
    }
    FH_LAUNCH_CHECK("union_square");
    // ---- assembly (shared single-system kernels) and condensation
    std::unique_ptr<fh_matrix> A(assemble_p1(m.get(), 0)), M(assemble_p1(m.get(), 1));
    DBuf<uint8_t> mask(dev, V);
    boundary_vertex_mask(m.get(), mask.p, s, nullptr, nullptr);
    std::vector<uint8_t> hmask(V);
    d2h(hmask.data(), mask.p, V, s);
    FH_HIP(hipStreamSynchronize(s));
    std::vector<uint64_t> fixed;
    for (uint64_t i = 0; i < n_inst; ++i)
        for (uint32_t r = voff[i]; r < voff[i + 1]; ++r)
            if (r >= voff[i] + nvi[i] || hmask[r]) fixed.push_back(r);
    const std::vector<double> zeros(fixed.size(), 0.0);
    DBuf<double> dummy(dev, V), dummy2(dev, V);
    fill_f64(dummy.p, V, 0.0, s);
    std::unique_ptr<fh_matrix> Ac(
        apply_dirichlet(A.get(), dummy.p, fixed.size(), fixed.data(), zeros.data(), dummy2.p, s));
    A.reset();
    // ---- right-hand sides b = M f (f given per real vertex)
    const uint64_t nvec = V * NR;
    DBuf<double> f(dev, nvec), b(dev, nvec), x64(dev, nvec);
    {
        // each instance's rows are contiguous on both sides: one copy per
        // instance into its padded range, the padding rows zero
        fill_f64(f.p, nvec, 0.0, s);
        uint64_t src = 0;
        for (uint64_t i = 0; i < n_inst; ++i) {
            h2d(f.p + (uint64_t)voff[i] * NR, f_nodal + src * NR, (size_t)nvi[i] * NR, s);
            src += nvi[i];
        }
        FH_HIP(hipStreamSynchronize(s));
    }
    const unsigned gslices = cdiv(Ac->pat->nslices, 4);
    hipLaunchKernelGGL(k_spmm_f64, dim3(gslices), dim3(256), 0, s, sellb(M.get(), M->vals.p), f.p, b.p);
    hipLaunchKernelGGL(k_zero_fixed_rows, dim3(cdiv(V, 256)), dim3(256), 0, s, Ac->fixed.p, b.p,
                       (uint32_t)V);
    M.reset();
    // ---- workgroup -> instance tables (256 rows per workgroup)
    const uint32_t nwg = (uint32_t)(V / 256);
    std::vector<uint32_t> wg_inst(nwg), wg_begin(n_inst), wg_end(n_inst);
    for (uint64_t i = 0; i < n_inst; ++i) {
        wg_begin[i] = voff[i] / 256;
        wg_end[i] = voff[i + 1] / 256;
        for (uint32_t w = wg_begin[i]; w < wg_end[i]; ++w) wg_inst[w] = (uint32_t)i;
    }
    DBuf<uint32_t> dwi(dev, nwg), dwb(dev, n_inst), dwe(dev, n_inst), dit(dev, n_inst),
        diters(dev, n_inst * NR);
    h2d(dwi.p, wg_inst.data(), nwg, s);
    h2d(dwb.p, wg_begin.data(), n_inst, s);
    h2d(dwe.p, wg_end.data(), n_inst, s);
    const uint64_t ncol = n_inst * NR;
    DBuf<double> rz(dev, ncol), alpha(dev, ncol), beta(dev, ncol), bnorm(dev, ncol), rrep(dev, ncol),
        res(dev, ncol);
    DBuf<uint8_t> done(dev, ncol), idone(dev, n_inst), need(dev, n_inst);
    fill_f64(beta.p, ncol, 0.0, s);
    fill_f64(alpha.p, ncol, 0.0, s);
    fill_f64(rz.p, ncol, 1.0, s);
    fill_u8(done.p, ncol, 0, s);
    fill_u8(idone.p, n_inst, 0, s);
    fill_u8(need.p, n_inst, 0, s);
    fill_u32(dit.p, n_inst, 0, s);
    fill_u32(diters.p, ncol, 0, s);
    BatchState S{rz.p, alpha.p, beta.p, bnorm.p, rrep.p, res.p, diters.p, done.p, dit.p, idone.p,
                 need.p, dwb.p, dwe.p, dwi.p, tol,
                 std::atof(std::getenv("FH_RR_DELTA") ? std::getenv("FH_RR_DELTA") : "0.1"),
                 (uint32_t)std::min<uint64_t>(max_iter, 0xffffffffu), MIX ? 1u : 0u,
                 precision == FH_PREC_MIXED_IR ? 1u : 0u};
    // ---- Krylov vectors
    if (MIX) ensure_vals32(Ac.get());
    DBuf<T> r(dev, nvec), z(dev, nvec), p0(dev, nvec), p1(dev, nvec), Ap(dev, nvec), dinv(dev, V);
    DBuf<float> x32;
    T* xk;
    fill_f64(x64.p, nvec, 0.0, s);
    if constexpr (MIX) {
        x32.alloc(dev, nvec);
        fill_f32(x32.p, nvec, 0.0f, s);
        fill_f32((float*)p0.p, nvec, 0.0f, s);
        fill_f32((float*)p1.p, nvec, 0.0f, s);
        xk = (T*)x32.p;
    } else {
        fill_f64((double*)p0.p, nvec, 0.0, s);
        fill_f64((double*)p1.p, nvec, 0.0, s);
        xk = (T*)x64.p;
    }
    {
        // Jacobi: 1/d of the condensed diagonal (cg.rs:52-64)
        DBuf<double> dd(dev, V);
        inv_diag_f64(Ac.get(), dd.p, s);
        if constexpr (MIX) {
            std::vector<double> h(V);
            d2h(h.data(), dd.p, V, s);
            FH_HIP(hipStreamSynchronize(s));
            std::vector<float> hf(h.begin(), h.end());
            h2d((float*)dinv.p, hf.data(), V, s);
            FH_HIP(hipStreamSynchronize(s));
        } else {
            d2d((double*)dinv.p, dd.p, V, s);
            FH_HIP(hipStreamSynchronize(s));
        }
    }
    const SellB<double> A64 = sellb(Ac.get(), Ac->vals.p);
    const SellB<T> AK = MIX ? sellb<T>(Ac.get(), (const T*)(const void*)Ac->vals32.p)
                            : sellb<T>(Ac.get(), (const T*)(const void*)Ac->vals.p);
    DBuf<double> part1(dev, (size_t)nwg * NR), part2(dev, (size_t)nwg * NR * 2),
        bpart(dev, (size_t)nwg * NR);
    FH_HIP(hipStreamSynchronize(s));

    Timer tk(s);
    hipLaunchKernelGGL(k_b_residual<T>, dim3(nwg), dim3(256), 0, s, A64, b.p, x64.p, (const float*)nullptr,
                       dinv.p, r.p, z.p, S, 0, part2.p, bpart.p);
    hipLaunchKernelGGL(k_b_red2, dim3(n_inst), dim3(256), 0, s, S, part2.p, bpart.p, 1);
    FH_LAUNCH_CHECK("batched_init");
    const int Mch = std::max(1, std::atoi(std::getenv("FH_CG_CHUNK") ? std::getenv("FH_CG_CHUNK") : "16"));
    auto chunk = [&]() {
        for (int q = 0; q < Mch; ++q) {
            hipLaunchKernelGGL(k_b_spmm<T>, dim3(nwg), dim3(256), 0, s, AK, z.p, p0.p, p1.p, Ap.p, S,
                               part1.p);
            hipLaunchKernelGGL(k_b_red1, dim3(n_inst), dim3(256), 0, s, S, part1.p);
            hipLaunchKernelGGL(k_b_update<T>, dim3(nwg), dim3(256), 0, s, p0.p, p1.p, Ap.p, dinv.p, xk,
                               r.p, z.p, S, (uint32_t)V, part2.p);
            hipLaunchKernelGGL(k_b_red2, dim3(n_inst), dim3(256), 0, s, S, part2.p, bpart.p, 0);
        }
    };
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    const bool use_graph = !std::getenv("FH_CG_GRAPH") || std::atoi(std::getenv("FH_CG_GRAPH")) != 0;
    std::vector<uint8_t> hdone(n_inst), hneed(n_inst);
    uint64_t launched = 0, nrep = 0;
    for (;;) {
        if (use_graph) {
            if (!gexec) {
                FH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                chunk();
                FH_HIP(hipStreamEndCapture(s, &graph));
                FH_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
            }
            FH_HIP(hipGraphLaunch(gexec, s));
        } else {
            chunk();
            FH_LAUNCH_CHECK("batched_chunk");
        }
        launched += Mch;
        d2h(hdone.data(), idone.p, n_inst, s);
        d2h(hneed.data(), need.p, n_inst, s);
        FH_HIP(hipStreamSynchronize(s));
        bool all = true, any_need = false;
        for (uint64_t i = 0; i < n_inst; ++i) {
            all = all && hdone[i];
            any_need = any_need || (hneed[i] && !hdone[i]);
        }
        if (all) break;
        if (any_need) {
            hipLaunchKernelGGL(k_b_residual<T>, dim3(nwg), dim3(256), 0, s, A64, b.p, x64.p, x32.p,
                               dinv.p, r.p, z.p, S, 1, part2.p, (double*)nullptr);
            hipLaunchKernelGGL(k_b_commit, dim3(nwg), dim3(256), 0, s, x64.p, x32.p, dwi.p, need.p,
                               (uint32_t)V, 0);
            hipLaunchKernelGGL(k_b_red2, dim3(n_inst), dim3(256), 0, s, S, part2.p, bpart.p, 2);
            FH_LAUNCH_CHECK("batched_replacement");
            ++nrep;
        }
        if (launched > max_iter + (nrep + 2) * (uint64_t)Mch + 64) break;  // safety net
    }
    if constexpr (MIX)
        hipLaunchKernelGGL(k_b_commit, dim3(nwg), dim3(256), 0, s, x64.p, x32.p, dwi.p, need.p,
                           (uint32_t)V, 1);
    const double kms = tk.stop_ms();
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    // ---- outputs
    std::vector<uint32_t> hit(ncol);
    std::vector<uint8_t> hd(ncol);
    {
        uint64_t dst = 0;
        for (uint64_t i = 0; i < n_inst; ++i) {  // the caller's rows of each instance
            d2h(u_out + dst * NR, x64.p + (uint64_t)voff[i] * NR, (size_t)nvi[i] * NR, s);
            dst += nvi[i];
        }
    }
    d2h(hit.data(), diters.p, ncol, s);
    d2h(hd.data(), done.p, ncol, s);
    FH_HIP(hipStreamSynchronize(s));
    double dof_iters = 0.0;
    uint64_t max_it = 0;
    int code = FH_OK;
    for (uint64_t i = 0; i < n_inst; ++i) {
        for (int k = 0; k < NR; ++k) {
            const uint64_t col = i * NR + k;
            if (iters_out) iters_out[col] = hit[col];
            dof_iters += (double)nvi[i] * hit[col];
            max_it = std::max<uint64_t>(max_it, hit[col]);
            if (hd[col] == 3) code = FH_ERR_BREAKDOWN;
            else if (hd[col] != 1 && code == FH_OK) code = FH_ERR_NOT_CONVERGED;
        }
    }
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->iterations = max_it;
        stats->replacements = nrep;
        stats->kernel_ms = kms;
        stats->wall_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tw0).count();
        // DOF * iterations per second over all columns
        stats->gflops = dof_iters / (kms * 1e-3) / 1e9;
        stats->spmv_bytes = (double)Ac->pat->nnzb * (sizeof(T) + 4.0) + 4.0 * (V + 1) +
                            2.0 * (double)nvec * sizeof(T);
        stats->device = dev;
        stats->precision = precision;
        stats->status = code;
    }
    return code;
}

}  // namespace fh

using namespace fh;

extern "C" int fh_batched_poisson_block_cg(uint64_t n_inst, const uint64_t* sizes, uint64_t nrhs,
                                           const double* f_nodal, double tol, uint64_t max_iter,
                                           int precision, int device_idx, double* u_out,
                                           uint64_t* iters_out, fh_solve_stats* stats) {
    int code = FH_OK;
    const int g = guard([&] {
        FH_CHECK(sizes && f_nodal && u_out && n_inst > 0, "null argument");
        FH_CHECK(nrhs == (uint64_t)NR, "this build solves 4 right-hand sides per instance");
        const int dev = resolve_device(device_idx);
        if (precision == FH_PREC_FP64)
            code = run_batched<double>(n_inst, sizes, f_nodal, tol, max_iter, precision, dev, u_out,
                                       iters_out, stats);
        else
            code = run_batched<float>(n_inst, sizes, f_nodal, tol, max_iter, precision, dev, u_out,
                                      iters_out, stats);
        if (code == FH_ERR_NOT_CONVERGED) set_last_error("a batched column failed to converge");
        if (code == FH_ERR_BREAKDOWN) set_last_error("breakdown in a batched column");
    });
    return g != FH_OK ? g : code;
}
