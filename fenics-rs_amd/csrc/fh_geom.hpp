// fh_geom.hpp — element geometry on the device, operation for operation the
// arithmetic of fenics-rs jacobian.rs:16-100 (affine map, J^{-T}, gradient
// transform) and the per-entry formulas of poisson.rs / elasticity.rs.  The
// library is compiled with -ffp-contract=off, so no multiply-add is fused and
// the FP64 results are bit-identical to any IEEE-754 evaluation of the same
// expressions in the same order.
#pragma once
#include <hip/hip_runtime.h>

namespace fhg {

struct Map2 {
    double J[2][2], invT[2][2], det;
};
struct Map3 {
    double J[3][3], invT[3][3], det;
};

// J = [v1-v0, v2-v0], det, J^{-1} = adj(J)/det, invT = J^{-T}.
__device__ __forceinline__ bool affine2(const double* v0, const double* v1, const double* v2,
                                        Map2& m) {
    m.J[0][0] = v1[0] - v0[0];
    m.J[1][0] = v1[1] - v0[1];
    m.J[0][1] = v2[0] - v0[0];
    m.J[1][1] = v2[1] - v0[1];
    const double det = m.J[0][0] * m.J[1][1] - m.J[1][0] * m.J[0][1];
    m.det = det;
    if (fabs(det) < 1e-14) return false;
    m.invT[0][0] = m.J[1][1] / det;
    m.invT[0][1] = -m.J[1][0] / det;
    m.invT[1][0] = -m.J[0][1] / det;
    m.invT[1][1] = m.J[0][0] / det;
    return true;
}

__device__ __forceinline__ bool affine3(const double* v0, const double* v1, const double* v2,
                                        const double* v3, Map3& m) {
#pragma unroll
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
    if (fabs(det) < 1e-14) return false;
    // invT[i][j] = inv[j][i]
    m.invT[0][0] = c00 / det;
    m.invT[1][0] = (m13 * m32 - m33 * m12) / det;
    m.invT[2][0] = (m12 * m23 - m22 * m13) / det;
    m.invT[0][1] = -c01 / det;
    m.invT[1][1] = (m11 * m33 - m31 * m13) / det;
    m.invT[2][1] = (m13 * m21 - m23 * m11) / det;
    m.invT[0][2] = c02 / det;
    m.invT[1][2] = (m12 * m31 - m32 * m11) / det;
    m.invT[2][2] = (m11 * m22 - m21 * m12) / det;
    return true;
}

// g = invT * g_ref, left to right.
__device__ __forceinline__ void grad2(const Map2& m, const double* gr, double* g) {
    g[0] = m.invT[0][0] * gr[0] + m.invT[0][1] * gr[1];
    g[1] = m.invT[1][0] * gr[0] + m.invT[1][1] * gr[1];
}
__device__ __forceinline__ void grad3(const Map3& m, const double* gr, double* g) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
        g[i] = m.invT[i][0] * gr[0] + m.invT[i][1] * gr[1] + m.invT[i][2] * gr[2];
}

// lagrange.rs:72-74 / 106-113 reference gradients of P1.
__device__ __forceinline__ void p1_tri_grads(const Map2& m, double g[3][2]) {
    const double G[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
#pragma unroll
    for (int i = 0; i < 3; ++i) grad2(m, G[i], g[i]);
}
__device__ __forceinline__ void p1_tet_grads(const Map3& m, double g[4][3]) {
    const double G[4][3] = {{-1.0, -1.0, -1.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
#pragma unroll
    for (int i = 0; i < 4; ++i) grad3(m, G[i], g[i]);
}

// elasticity.rs:35-47
__device__ __host__ __forceinline__ double lame_lambda(double e, double nu) {
    return (e * nu) / ((1.0 + nu) * (1.0 - 2.0 * nu));
}
__device__ __host__ __forceinline__ double lame_mu(double e, double nu) {
    return e / (2.0 * (1.0 + nu));
}

// elasticity.rs:120-136: K(i alpha, j beta) for gradients gi (row) and gj (col).
__device__ __forceinline__ double elas_entry(const double* gi, const double* gj, int alpha,
                                             int beta, double lambda, double mu, double vol) {
    const double gigj = gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2];
    const double delta = (alpha == beta) ? 1.0 : 0.0;
    return vol * (lambda * gj[beta] * gi[alpha] + mu * gigj * delta + mu * gj[alpha] * gi[beta]);
}

// Load the vertices of cell c.
template <int D, int K>
__device__ __forceinline__ void load_cell(const double* __restrict__ V,
                                          const uint32_t* __restrict__ C, uint32_t c,
                                          double x[K][D], uint32_t t[K]) {
#pragma unroll
    for (int a = 0; a < K; ++a) t[a] = C[(uint64_t)c * K + a];
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
        for (int d = 0; d < D; ++d) x[a][d] = V[(uint64_t)t[a] * D + d];
}

}  // namespace fhg
