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

}  // namespace fh
