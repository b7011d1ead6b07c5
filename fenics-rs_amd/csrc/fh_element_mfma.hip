// fh_element_mfma.hip — batched P2 tetrahedron element matrices on the MFMA
// units (FP32 in, FP32 accumulate), then a deterministic gather into the
// global SELL-64 matrix.
//
// The P2 Poisson element matrix K_e = sum_q w_q |det J| B_q^T B_q with
// B_q = J^{-T} grad phi(xi_q) (3 x 10) is a B^T D B contraction.  Because the
// map is affine, contracting the quadrature and spatial indices first leaves
//     K_e = sum_{k < 6} C_k(e) S_k,
//     C(e) = |det J| (G00, G11, G22, G01, G02, G12),  G = J^{-1} J^{-T},
//     S_k  = sum_q w_q (gh_q[:,a] gh_q[:,b]^T (+ transpose if a != b)),
// with S_k the six 10x10 reference matrices (gh = reference gradients).  For a
// batch of cells this is one GEMM  K[cells x 100] = C[cells x 6] . S[6 x 100]
// — M = cells (large), N = 100, K = 6 (padded to 8) — which maps onto
// v_mfma_f32_16x16x4f32: a wave owns 16 cells and sweeps 7 output tiles of 16
// entries with two K=4 steps each.  C is computed per cell in FP32 (16 lanes),
// staged through LDS; S lives in LDS for the block.  The element matrices go
// to an [nc][100] FP32 buffer; p2_gather then sums them per dof row in
// ascending cell order (no atomics).
#include <vector>

#include "fh_geom.hpp"
#include "fh_internal.hpp"

namespace fh {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int P2_NT = 7;    // 16-wide output tiles covering the 100 entries
constexpr int P2_NS = 112;  // padded entry count

__global__ __launch_bounds__(256) void k_p2_elem_mfma(const double* __restrict__ V,
                                                      const uint32_t* __restrict__ Cc, uint32_t nc,
                                                      const float* __restrict__ Sg,
                                                      float* __restrict__ Ke) {
    __shared__ float S[8 * P2_NS];
    __shared__ float coef[4][16][9];
    for (int i = threadIdx.x; i < 8 * P2_NS; i += 256) S[i] = Sg[i];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t ngroups = (nc + 15) / 16;
    // block-uniform trip count, so the __syncthreads below are reached by all waves
    for (uint32_t gbase = blockIdx.x * 4; gbase < ngroups; gbase += gridDim.x * 4) {
        const uint32_t grp = gbase + wave;
        const uint32_t c0 = grp * 16;  // cells >= nc are masked below
        if (lane < 16) {
            const uint32_t c = c0 + lane;
            float co[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (c < nc) {
                float x[4][3];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const uint32_t v = Cc[(uint64_t)c * 4 + a];
#pragma unroll
                    for (int d = 0; d < 3; ++d) x[a][d] = (float)V[(uint64_t)v * 3 + d];
                }
                float J[3][3];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) J[i][j] = x[j + 1][i] - x[0][i];
                const float c00 = J[1][1] * J[2][2] - J[2][1] * J[1][2];
                const float c01 = J[1][0] * J[2][2] - J[2][0] * J[1][2];
                const float c02 = J[1][0] * J[2][1] - J[2][0] * J[1][1];
                const float det = J[0][0] * c00 - J[0][1] * c01 + J[0][2] * c02;
                if (fabsf(det) >= 1e-14f) {
                    const float id = 1.0f / det;
                    float Ji[3][3];  // J^{-1}
                    Ji[0][0] = c00 * id;
                    Ji[0][1] = (J[0][2] * J[2][1] - J[2][2] * J[0][1]) * id;
                    Ji[0][2] = (J[0][1] * J[1][2] - J[1][1] * J[0][2]) * id;
                    Ji[1][0] = -c01 * id;
                    Ji[1][1] = (J[0][0] * J[2][2] - J[2][0] * J[0][2]) * id;
                    Ji[1][2] = (J[0][2] * J[1][0] - J[1][2] * J[0][0]) * id;
                    Ji[2][0] = c02 * id;
                    Ji[2][1] = (J[0][1] * J[2][0] - J[2][1] * J[0][0]) * id;
                    Ji[2][2] = (J[0][0] * J[1][1] - J[1][0] * J[0][1]) * id;
                    const float ad = fabsf(det);
                    auto G = [&](int a, int b) {
                        return ad * (Ji[a][0] * Ji[b][0] + Ji[a][1] * Ji[b][1] + Ji[a][2] * Ji[b][2]);
                    };
                    co[0] = G(0, 0);
                    co[1] = G(1, 1);
                    co[2] = G(2, 2);
                    co[3] = G(0, 1);
                    co[4] = G(0, 2);
                    co[5] = G(1, 2);
                }
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) coef[wave][lane][k] = co[k];
        }
        __syncthreads();
        // v_mfma_f32_16x16x4f32: lane holds A[m = lane%16][k = lane/16],
        // B[k = lane/16][n = lane%16]; result rows 4*(lane/16)+i, column lane%16
        const int m = lane & 15, kq = lane >> 4;
        const float a0 = coef[wave][m][kq], a1 = coef[wave][m][4 + kq];
#pragma unroll
        for (int t = 0; t < P2_NT; ++t) {
            const int n = t * 16 + m;
            v4f acc = {0.f, 0.f, 0.f, 0.f};
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, S[kq * P2_NS + n], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, S[(4 + kq) * P2_NS + n], acc, 0, 0, 0);
            if (n < 100) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint32_t c = c0 + 4 * kq + i;
                    if (c < nc) Ke[(uint64_t)c * 100 + n] = acc[i];
                }
            }
        }
        __syncthreads();
    }
}

// Reference tensors S_k [8][112] (rows 6,7 and columns >= 100 zero), FP64 on
// the host from the lagrange.rs:324-346 gradients at the quadrature.rs:157-184
// points, rounded to FP32.
static std::vector<float> p2_reference_tensors() {
    const double s5 = std::sqrt(5.0);
    const double a = (5.0 - s5) / 20.0, b = (5.0 + 3.0 * s5) / 20.0;
    const double P[4][3] = {{a, a, a}, {b, a, a}, {a, b, a}, {a, a, b}};
    const int pairs[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
    std::vector<double> S(8 * P2_NS, 0.0);
    for (int q = 0; q < 4; ++q) {
        const double x = P[q][0], y = P[q][1], z = P[q][2];
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
        const double w = 1.0 / 24.0;
        for (int k = 0; k < 6; ++k) {
            const int pa = pairs[k][0], pb = pairs[k][1];
            for (int i = 0; i < 10; ++i)
                for (int j = 0; j < 10; ++j) {
                    double v = G[i][pa] * G[j][pb];
                    if (pa != pb) v += G[i][pb] * G[j][pa];
                    S[k * P2_NS + i * 10 + j] += w * v;
                }
        }
    }
    return std::vector<float>(S.begin(), S.end());
}

// Fills a->vals32 (and, when fill64, a->vals as the widened FP32 values).
void assemble_p2_mfma(fh_mesh* m, fh_matrix* a, bool fill64, hipStream_t s) {
    MeshP2& p2 = mesh_p2(m);
    const std::vector<float> S = p2_reference_tensors();
    DBuf<float> dS(m->dev, S.size()), Ke(m->dev, (size_t)m->nc * 100);
    h2d(dS.p, S.data(), S.size(), s);
    const uint32_t ngroups = (m->nc + 15) / 16;
    const unsigned g = std::max(1u, std::min<unsigned>(cdiv(ngroups, 4), 4096));
    hipLaunchKernelGGL(k_p2_elem_mfma, dim3(g), dim3(256), 0, s, m->verts.p, m->cells.p, m->nc,
                       dS.p, Ke.p);
    FH_LAUNCH_CHECK("p2_elem_mfma");
    // the gather writes every SELL entry (padding as zeros), in FP32 and, for
    // fill64, in FP64
    if (!a->vals32.p) a->vals32.alloc(m->dev, std::max<uint64_t>(a->pat->sell_entries, 1));
    p2_gather(m, Ke.p, false, a->vals32.p, fill64 ? a->vals.p : nullptr, s);
    FH_HIP(hipStreamSynchronize(s));
}

// P1 tetrahedron elasticity element matrices on the MFMA units, FP32 in and
// FP32 accumulate: K_e = B^T (vol D) B, with B (6 x 12) the Voigt
// strain-displacement matrix of the cell's four gradients (rows e_xx, e_yy,
// e_zz, g_yz, g_xz, g_xy) and D (6 x 6) the isotropic stiffness (lambda + 2 mu
// and lambda on the normal block, mu on the shear diagonal).  Entry
// ((a, al), (b, be)) is vol (lambda g_a[al] g_b[be] + mu g_a[be] g_b[al] +
// mu (g_a . g_b) [al == be]), the elasticity.rs:133-136 entry.
//
// v_mfma_f32_16x16x1f32 computes four independent 16 x 16 outer products per
// instruction: block = cell, A = column k of B^T (16 rows (a, al), 12 used),
// B = row k of vol D B; K = 6 gives six instructions per four cells.  Lane l
// of a group owns cell l / 16 and index i = l % 16 of both operands, so it
// builds its B and vol D B entries from its own vertex gradient g_{i / 3}
// alone.  The inputs are the FP64 per-cell records of k_elas_cell_geom
// (gradients, vol, per-cell lambda and mu; vol < 0 marks a cell the
// reference skips, whose element matrix is zero), rounded to FP32.
// Result layout (16 registers): register r holds block r / 4, row
// 4 (l / 16) + r % 4, column l % 16.  The blocks go to Ke8 [nc][4][4][8]
// (components 0-7 of block (a, b)) and Ke9 [nc][4][4] (component 8),
// through LDS, as contiguous 16-byte stores.
typedef float v16f __attribute__((ext_vector_type(16)));

template <bool EC>
__global__ __launch_bounds__(256) void k_elas_elem_mfma(const double* __restrict__ G, uint32_t nc,
                                                        float lam0, float mu0,
                                                        float* __restrict__ Ke8,
                                                        float* __restrict__ Ke9) {
    // a wave's four element matrices, staged in the output order (Ke8 part,
    // then Ke9 part) so that they leave as contiguous 16-byte stores
    __shared__ __align__(16) float st[4][576];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t i = lane & 15, al = i % 3, ia = i / 3;
    const uint32_t ngroups = (nc + 3) / 4;
    float* sw = st[wave];
    // block-uniform trip count (the barriers below); groups past the end are
    // all-masked
    for (uint32_t gbase = blockIdx.x * 4; gbase < ngroups; gbase += gridDim.x * 4) {
        const uint32_t c4 = (gbase + wave) * 4, c = c4 + lane / 16;
        float g[3] = {0.f, 0.f, 0.f}, vol = 0.f, lam = lam0, mu = mu0;
        if (c < nc) {
            const double* rec = G + (uint64_t)c * 16;
            const double v = rec[3];
            if (v > 0.0 && i < 12) {
                vol = (float)v;
#pragma unroll
                for (int d = 0; d < 3; ++d) g[d] = (float)rec[4 * ia + d];
                if (EC) {
                    lam = (float)rec[7];
                    mu = (float)rec[11];
                }
            }
        }
        // column (a, al) of B: its entries in rows e_xx .. g_xy
        float Bk[6];
        Bk[0] = al == 0 ? g[0] : 0.f;
        Bk[1] = al == 1 ? g[1] : 0.f;
        Bk[2] = al == 2 ? g[2] : 0.f;
        Bk[3] = al == 1 ? g[2] : (al == 2 ? g[1] : 0.f);
        Bk[4] = al == 0 ? g[2] : (al == 2 ? g[0] : 0.f);
        Bk[5] = al == 0 ? g[1] : (al == 1 ? g[0] : 0.f);
        const float tr = Bk[0] + Bk[1] + Bk[2];
        const float vl = vol * lam, vm = vol * mu;
        float DBk[6];
#pragma unroll
        for (int k = 0; k < 3; ++k) DBk[k] = vl * tr + 2.f * vm * Bk[k];
#pragma unroll
        for (int k = 3; k < 6; ++k) DBk[k] = vm * Bk[k];
        v16f acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = 0.f;
#pragma unroll
        for (int k = 0; k < 6; ++k) acc = __builtin_amdgcn_mfma_f32_16x16x1f32(Bk[k], DBk[k], acc, 0, 0, 0);
        const uint32_t col = lane & 15, b = col / 3, be = col % 3;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const uint32_t cl = r / 4, row = 4 * (lane / 16) + r % 4;
            if (row < 12 && col < 12) {
                const uint32_t e = (row % 3) * 3 + be, blk = (row / 3) * 4 + b;
                if (e < 8)
                    sw[cl * 128 + blk * 8 + e] = acc[r];
                else
                    sw[512 + cl * 16 + blk] = acc[r];
            }
        }
        __syncthreads();
        // cells c4 .. c4 + 3: Ke8 [c4 * 128, + 512) as 128 float4, Ke9 [c4 * 16, + 64)
        typedef float f4 __attribute__((ext_vector_type(4)));
#pragma unroll
        for (uint32_t h = 0; h < 2; ++h) {
            const uint32_t f = lane + 64 * h;  // float4 index: cell c4 + f / 32
            if (c4 + f / 32 < nc)
                reinterpret_cast<f4*>(Ke8 + (uint64_t)c4 * 128)[f] =
                    reinterpret_cast<const f4*>(sw)[f];
        }
        if (c4 + lane / 16 < nc) Ke9[(uint64_t)c4 * 16 + lane] = sw[512 + lane];
        __syncthreads();
    }
}

void assemble_elasticity_mfma(fh_mesh* m, double E, double nu, const double* Ecell_dev,
                              fh_matrix* a, bool fill64, hipStream_t s) {
    const size_t nc = m->nc;
    DBuf<double> geom(m->dev, std::max<size_t>(nc * 16, 1));
    elas_cell_records(m, nu, Ecell_dev, geom.p, s);
    DBuf<float> Ke8(m->dev, std::max<size_t>(nc * 128, 1)), Ke9(m->dev, std::max<size_t>(nc * 16, 1));
    if (nc) {
        const uint32_t ngroups = (uint32_t)((nc + 3) / 4);
        const unsigned g = std::max(1u, std::min<unsigned>(cdiv(ngroups, 4), 8192));
        const float lam = (float)fhg::lame_lambda(E, nu), mu = (float)fhg::lame_mu(E, nu);
        if (Ecell_dev)
            hipLaunchKernelGGL((k_elas_elem_mfma<true>), dim3(g), dim3(256), 0, s, geom.p,
                               (uint32_t)nc, lam, mu, Ke8.p, Ke9.p);
        else
            hipLaunchKernelGGL((k_elas_elem_mfma<false>), dim3(g), dim3(256), 0, s, geom.p,
                               (uint32_t)nc, lam, mu, Ke8.p, Ke9.p);
        FH_LAUNCH_CHECK("elas_elem_mfma");
    }
    if (!a->vals32.p) a->vals32.alloc(m->dev, std::max<uint64_t>(sell_value_count(a), 1));
    elas_gather_f32(m, *a->pat, Ke8.p, Ke9.p, a->vals32.p, fill64 ? a->vals.p : nullptr, s);
    FH_HIP(hipStreamSynchronize(s));
}

}  // namespace fh
