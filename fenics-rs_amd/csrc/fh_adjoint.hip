// fh_adjoint.hip — per-cell adjoint contraction for compliance gradients and
// the stress recovery that follows the solve (elasticity.rs:230-360).
//
// For K = sum_c E_c K0_c (K0_c: the elasticity.rs:133-136 element matrix with
// E = 1) and J = g^T u with K u = f, the adjoint K lambda = g gives
// dJ/dE_c = -lambda_c^T K0_c u_c; for the compliance J = f^T u, lambda = u.
// k_cell_bilinear_k0 evaluates -w_c^T K0_c u_c for every cell (one thread per
// cell; reads 4 vertices and 2 x 12 nodal values, writes one double) in
// strain form, in the oracle's operation order, so the FP64 result is
// bit-identical to it.
#include "fh_geom.hpp"
#include "fh_internal.hpp"

namespace fh {

using namespace fhg;

// Displacement gradient G[a][b] = sum_{i=1..3} (v_i[a] - v_0[a]) g_i[b] of the
// nodal values v of one cell: sum_i g_i = 0, so subtracting vertex 0's value
// leaves the same gradient and removes the rigid translation before any
// rounding (the energy of a cell that mostly translates would otherwise be
// the small difference of large products).
__device__ __forceinline__ void cell_grad_u(const double v[4][3], const double g[4][3],
                                            double G[3][3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            G[a][b] = ((v[1][a] - v[0][a]) * g[1][b] + (v[2][a] - v[0][a]) * g[2][b]) +
                      (v[3][a] - v[0][a]) * g[3][b];
}

// -w_c^T K0_c u_c in strain form: w^T K0 u = vol (lambda tr(e_w) tr(e_u) +
// 2 mu e_w : e_u), e = (G + G^T) / 2, which is the elasticity.rs:133-136
// element matrix contracted with w and u (algebraically identical), without
// its cancellation.  Same operation order as the oracle.
__global__ __launch_bounds__(256) void k_cell_bilinear_k0(const double* __restrict__ V,
                                                          const uint32_t* __restrict__ C,
                                                          uint32_t nc, double nu,
                                                          const double* __restrict__ w,
                                                          const double* __restrict__ u,
                                                          double* __restrict__ out) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double x[4][3];
    uint32_t t[4];
    load_cell<3, 4>(V, C, c, x, t);
    Map3 m;
    if (!affine3(x[0], x[1], x[2], x[3], m)) {
        out[c] = 0.0;
        return;
    }
    const double lambda = lame_lambda(1.0, nu), mu = lame_mu(1.0, nu);
    const double vol = (1.0 / 6.0) * fabs(m.det);
    double g[4][3];
    p1_tet_grads(m, g);
    double wl[4][3], ul[4][3];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            wl[i][a] = w[3 * (uint64_t)t[i] + a];
            ul[i][a] = u[3 * (uint64_t)t[i] + a];
        }
    double Gw[3][3], Gu[3][3];
    cell_grad_u(wl, g, Gw);
    cell_grad_u(ul, g, Gu);
    double ee = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            ee += (0.5 * (Gw[a][b] + Gw[b][a])) * (0.5 * (Gu[a][b] + Gu[b][a]));
    const double trw = (Gw[0][0] + Gw[1][1]) + Gw[2][2];
    const double tru = (Gu[0][0] + Gu[1][1]) + Gu[2][2];
    out[c] = -(vol * (lambda * trw * tru + 2.0 * mu * ee));
}

// elasticity.rs:240-318: per cell stress[9], strain[9], von Mises, trace.
__global__ __launch_bounds__(256) void k_element_stresses(const double* __restrict__ V,
                                                          const uint32_t* __restrict__ C,
                                                          uint32_t nc, const double* __restrict__ u,
                                                          double lambda, double mu,
                                                          double* __restrict__ out) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double* o = out + 20 * (uint64_t)c;
    double x[4][3];
    uint32_t t[4];
    load_cell<3, 4>(V, C, c, x, t);
    Map3 m;
    if (!affine3(x[0], x[1], x[2], x[3], m)) {
        for (int k = 0; k < 20; ++k) o[k] = 0.0;
        return;
    }
    double g[4][3];
    p1_tet_grads(m, g);
    double gu[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
#pragma unroll
            for (int nd = 0; nd < 4; ++nd) s += u[3 * (uint64_t)t[nd] + i] * g[nd][j];
            gu[i][j] = s;
        }
    double eps[3][3], sig[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) eps[i][j] = 0.5 * (gu[i][j] + gu[j][i]);
    const double tr = eps[0][0] + eps[1][1] + eps[2][2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double delta = (i == j) ? 1.0 : 0.0;
            sig[i][j] = lambda * tr * delta + 2.0 * mu * eps[i][j];
        }
    const double s00 = sig[0][0], s11 = sig[1][1], s22 = sig[2][2], s01 = sig[0][1],
                 s12 = sig[1][2], s20 = sig[2][0];
    const double vm2 = 0.5 * ((s00 - s11) * (s00 - s11) + (s11 - s22) * (s11 - s22) +
                              (s22 - s00) * (s22 - s00) + 6.0 * (s01 * s01 + s12 * s12 + s20 * s20));
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            o[3 * i + j] = sig[i][j];
            o[9 + 3 * i + j] = eps[i][j];
        }
    o[18] = sqrt(fmax(vm2, 0.0));
    o[19] = s00 + s11 + s22;
}

// elasticity.rs:331-357: volume-weighted average over the incident cells, in
// ascending cell order (the reference's sequential scatter order).
__global__ __launch_bounds__(256) void k_vertex_vm(const double* __restrict__ V,
                                                   const uint32_t* __restrict__ C,
                                                   const uint32_t* __restrict__ aptr,
                                                   const uint32_t* __restrict__ aent, uint32_t nv,
                                                   const double* __restrict__ vm,
                                                   double* __restrict__ out) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double s = 0.0, vs = 0.0;
    for (uint32_t e = aptr[v]; e < aptr[v + 1]; ++e) {
        const uint32_t c = aent[e] / 4;
        double x[4][3];
        uint32_t t[4];
        load_cell<3, 4>(V, C, c, x, t);
        Map3 m;
        if (!affine3(x[0], x[1], x[2], x[3], m)) continue;
        const double vol = (1.0 / 6.0) * fabs(m.det);
        s += vm[c] * vol;
        vs += vol;
    }
    out[v] = vs > 1e-15 ? s / vs : s;
}

}  // namespace fh

using namespace fh;

extern "C" {

int fh_cell_bilinear_k0_dev(fh_mesh* m, double nu, const double* w_dev, const double* u_dev,
                            double* out_dev) {
    return guard([&] {
        FH_CHECK(m && w_dev && u_dev && out_dev, "null argument");
        FH_CHECK(m->dim == 3, "elasticity needs a tetrahedral mesh");
        DeviceCtx& d = device(m->dev);
        hipLaunchKernelGGL(k_cell_bilinear_k0, dim3(cdiv(m->nc, 256)), dim3(256), 0, d.stream,
                           m->verts.p, m->cells.p, m->nc, nu, w_dev, u_dev, out_dev);
        FH_LAUNCH_CHECK("cell_bilinear_k0");
        FH_HIP(hipStreamSynchronize(d.stream));
    });
}

int fh_cell_bilinear_k0(fh_mesh* m, double nu, const double* w, const double* u, double* out) {
    return guard([&] {
        FH_CHECK(m && w && u && out, "null argument");
        DeviceCtx& d = device(m->dev);
        const size_t n = 3 * (size_t)m->nv;
        DBuf<double> dw(m->dev, n), du(m->dev, n), dout(m->dev, m->nc);
        h2d(dw.p, w, n, d.stream);
        h2d(du.p, u, n, d.stream);
        if (fh_cell_bilinear_k0_dev(m, nu, dw.p, du.p, dout.p) != FH_OK)
            throw Error(FH_ERR_HIP, fh_last_error());
        d2h(out, dout.p, m->nc, d.stream);
        FH_HIP(hipStreamSynchronize(d.stream));
    });
}

int fh_element_stresses(fh_mesh* m, const double* u, double E, double nu, double* out) {
    return guard([&] {
        FH_CHECK(m && u && out, "null argument");
        FH_CHECK(m->dim == 3, "stresses need a tetrahedral mesh");
        DeviceCtx& d = device(m->dev);
        DBuf<double> du(m->dev, 3 * (size_t)m->nv), dout(m->dev, 20 * (size_t)m->nc);
        h2d(du.p, u, du.n, d.stream);
        hipLaunchKernelGGL(k_element_stresses, dim3(cdiv(m->nc, 256)), dim3(256), 0, d.stream,
                           m->verts.p, m->cells.p, m->nc, du.p, lame_lambda(E, nu), lame_mu(E, nu),
                           dout.p);
        FH_LAUNCH_CHECK("element_stresses");
        d2h(out, dout.p, dout.n, d.stream);
        FH_HIP(hipStreamSynchronize(d.stream));
    });
}

int fh_vertex_von_mises(fh_mesh* m, const double* vm_cell, double* out) {
    return guard([&] {
        FH_CHECK(m && vm_cell && out, "null argument");
        FH_CHECK(m->dim == 3, "stresses need a tetrahedral mesh");
        DeviceCtx& d = device(m->dev);
        Adjacency& adj = mesh_adjacency(m);
        DBuf<double> dvm(m->dev, m->nc), dout(m->dev, m->nv);
        h2d(dvm.p, vm_cell, m->nc, d.stream);
        hipLaunchKernelGGL(k_vertex_vm, dim3(cdiv(m->nv, 256)), dim3(256), 0, d.stream,
                           m->verts.p, m->cells.p, adj.ptr.p, adj.ent.p, m->nv, dvm.p, dout.p);
        FH_LAUNCH_CHECK("vertex_vm");
        d2h(out, dout.p, m->nv, d.stream);
        FH_HIP(hipStreamSynchronize(d.stream));
    });
}

// The batched multi-RHS solver lives in fh_batched.hip.

}  // extern "C"
