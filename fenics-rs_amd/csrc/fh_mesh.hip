// fh_mesh.hip — device-resident meshes (replaces SimplicialMesh,
// crates/fenics-mesh/src/mesh.rs:10-45), the structured generators of
// generators.rs:24-117 built directly in HBM, boundary facets
// (mesh.rs:79-113,145-210) and quadrature points (jacobian.rs:44-48,88-92).
#include <algorithm>
#include <vector>

#include "fh_internal.hpp"

namespace fh {

// generators.rs:29-35: x-fastest vertices at (i/nx, j/ny).
__global__ void k_unit_square(uint32_t nx, uint32_t ny, double* __restrict__ V,
                              uint32_t* __restrict__ C) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t nv = (uint64_t)(nx + 1) * (ny + 1);
    if (t < nv) {
        const uint32_t i = t % (nx + 1), j = t / (nx + 1);
        V[2 * t] = (double)i / (double)nx;
        V[2 * t + 1] = (double)j / (double)ny;
    }
    // generators.rs:39-52: squares in j-then-i order, two triangles each
    const uint64_t nsq = (uint64_t)nx * ny;
    if (t < nsq) {
        const uint32_t i = t % nx, j = t / nx;
        const uint32_t v00 = j * (nx + 1) + i, v10 = v00 + 1, v01 = v00 + nx + 1, v11 = v01 + 1;
        uint32_t* c = C + 6 * t;
        c[0] = v00; c[1] = v10; c[2] = v11;
        c[3] = v00; c[4] = v11; c[5] = v01;
    }
}

// generators.rs:73-105: vertices k,j,i with ((i/nx)*L, (j/ny)*W, (k/nz)*H);
// six Kuhn tetrahedra per hex around the p0-p7 diagonal.
__global__ void k_box_beam(uint32_t nx, uint32_t ny, uint32_t nz, double L, double W, double H,
                           double* __restrict__ V, uint32_t* __restrict__ C) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t sx = nx + 1, sxy = (uint64_t)(nx + 1) * (ny + 1);
    const uint64_t nv = sxy * (nz + 1);
    if (t < nv) {
        const uint64_t i = t % sx, j = (t / sx) % (ny + 1), k = t / sxy;
        V[3 * t] = ((double)i / (double)nx) * L;
        V[3 * t + 1] = ((double)j / (double)ny) * W;
        V[3 * t + 2] = ((double)k / (double)nz) * H;
    }
    const uint64_t nh = (uint64_t)nx * ny * nz;
    if (t < nh) {
        const uint32_t i = t % nx, j = (t / nx) % ny, k = t / ((uint64_t)nx * ny);
        auto vid = [&](uint32_t a, uint32_t b, uint32_t c) {
            return (uint32_t)(c * sxy + b * sx + a);
        };
        const uint32_t p0 = vid(i, j, k), p1 = vid(i + 1, j, k), p2 = vid(i, j + 1, k),
                       p3 = vid(i + 1, j + 1, k), p4 = vid(i, j, k + 1), p5 = vid(i + 1, j, k + 1),
                       p6 = vid(i, j + 1, k + 1), p7 = vid(i + 1, j + 1, k + 1);
        uint32_t* c = C + 24 * t;
        const uint32_t q[24] = {p0, p1, p3, p7, p0, p3, p2, p7, p0, p2, p6, p7,
                                p0, p6, p4, p7, p0, p4, p5, p7, p0, p5, p1, p7};
#pragma unroll
        for (int a = 0; a < 24; ++a) c[a] = q[a];
    }
}

// ---------------------------------------------------------------- boundary
// A facet of cell c (opposite local vertex a) lies on the boundary when no
// other cell contains all of its vertices (the count == 1 rule of
// mesh.rs:99-112 / 194-209).  Marks the facet's vertices.
template <int K>
__global__ void k_boundary_mark(const uint32_t* __restrict__ C, uint32_t nc,
                                const uint32_t* __restrict__ aptr,
                                const uint32_t* __restrict__ aent, uint8_t* __restrict__ mask,
                                const uint32_t* __restrict__ cd10, uint8_t* __restrict__ mask2) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= (uint64_t)nc * K) return;
    const uint32_t c = t / K, a = t % K;
    uint32_t f[K - 1];
    int m = 0;
#pragma unroll
    for (int b = 0; b < K; ++b)
        if (b != (int)a) f[m++] = C[c * K + b];
    const uint32_t u = f[0];
    int shared = 0;
    for (uint32_t e = aptr[u]; e < aptr[u + 1]; ++e) {
        const uint32_t c2 = aent[e] / K;
        if (c2 == c) continue;
        int hits = 0;
#pragma unroll
        for (int b = 0; b < K; ++b) {
            const uint32_t v = C[c2 * K + b];
#pragma unroll
            for (int q = 1; q < K - 1; ++q) hits += (v == f[q]);
        }
        if (hits == K - 2) {
            shared = 1;
            break;
        }
    }
    if (!shared) {
#pragma unroll
        for (int q = 0; q < K - 1; ++q) mask[f[q]] = 1;
        if (K == 4 && cd10) {
            // P2: the face's three vertex dofs and the three edge dofs whose
            // local edge (01,12,02,03,13,23) avoids the opposite vertex a
            const int E[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
#pragma unroll
            for (int q = 0; q < K - 1; ++q) mask2[f[q]] = 1;
#pragma unroll
            for (int e = 0; e < 6; ++e)
                if (E[e][0] != (int)a && E[e][1] != (int)a) mask2[cd10[(uint64_t)c * 10 + 4 + e]] = 1;
        }
    }
}

// mask over vertices; when cd10/mask2 are given (P2), also a mask over P2 dofs
void boundary_vertex_mask(fh_mesh* m, uint8_t* mask, hipStream_t s, const uint32_t* cd10,
                          uint8_t* mask2) {
    Adjacency& adj = mesh_adjacency(m);
    fill_u8(mask, m->nv, 0, s);
    const uint64_t n = (uint64_t)m->nc * m->k;
    if (m->k == 3)
        hipLaunchKernelGGL(k_boundary_mark<3>, dim3(cdiv(n, 256)), dim3(256), 0, s, m->cells.p,
                           m->nc, adj.ptr.p, adj.ent.p, mask, nullptr, nullptr);
    else
        hipLaunchKernelGGL(k_boundary_mark<4>, dim3(cdiv(n, 256)), dim3(256), 0, s, m->cells.p,
                           m->nc, adj.ptr.p, adj.ent.p, mask, cd10, mask2);
    FH_LAUNCH_CHECK("boundary_mark");
}

__global__ void k_mask_to_u32(const uint8_t* __restrict__ m, uint32_t* __restrict__ o, size_t n) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) o[i] = m[i];
}
__global__ void k_compact(const uint8_t* __restrict__ m, const uint32_t* __restrict__ pos,
                          uint64_t* __restrict__ out, size_t n) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n && m[i]) out[pos[i]] = i;
}

// Sorted indices of the nonzero entries of a device mask (stream compaction).
std::vector<uint64_t> mask_to_sorted(int dev, const uint8_t* mask, size_t n, hipStream_t s) {
    DBuf<uint32_t> cnt(dev, std::max<size_t>(n, 1)), pos(dev, n + 1);
    hipLaunchKernelGGL(k_mask_to_u32, dim3(cdiv(n, 256)), dim3(256), 0, s, mask, cnt.p, n);
    const uint64_t tot = exclusive_scan_u32(cnt.p, pos.p, n, dev, s);
    DBuf<uint64_t> out(dev, std::max<uint64_t>(tot, 1));
    hipLaunchKernelGGL(k_compact, dim3(cdiv(n, 256)), dim3(256), 0, s, mask, pos.p, out.p, n);
    std::vector<uint64_t> h(tot);
    d2h(h.data(), out.p, tot, s);
    FH_HIP(hipStreamSynchronize(s));
    return h;
}

// ---------------------------------------------------------------- quadrature points
// quadrature.rs:83-219 (same literals as the oracle).
QuadTab quad_table(int dim, int degree) {
    QuadTab q{};
    if (dim == 2) {
        if (degree <= 1) {
            q.n = 1; q.pts[0][0] = 1.0 / 3.0; q.pts[0][1] = 1.0 / 3.0; q.w[0] = 0.5;
        } else if (degree == 2) {
            q.n = 3;
            const double P[3][2] = {{0.5, 0.0}, {0.5, 0.5}, {0.0, 0.5}};
            for (int i = 0; i < 3; ++i) { q.pts[i][0] = P[i][0]; q.pts[i][1] = P[i][1]; q.w[i] = 1.0 / 6.0; }
        } else {
            q.n = 4;
            const double P[4][2] = {{1.0 / 3.0, 1.0 / 3.0}, {0.6, 0.2}, {0.2, 0.6}, {0.2, 0.2}};
            for (int i = 0; i < 4; ++i) { q.pts[i][0] = P[i][0]; q.pts[i][1] = P[i][1]; }
            q.w[0] = -27.0 / 96.0; q.w[1] = q.w[2] = q.w[3] = 25.0 / 96.0;
        }
    } else {
        if (degree <= 1) {
            q.n = 1; q.pts[0][0] = q.pts[0][1] = q.pts[0][2] = 0.25; q.w[0] = 1.0 / 6.0;
        } else if (degree == 2) {
            const double s5 = std::sqrt(5.0);
            const double a = (5.0 - s5) / 20.0, b = (5.0 + 3.0 * s5) / 20.0;
            const double P[4][3] = {{a, a, a}, {b, a, a}, {a, b, a}, {a, a, b}};
            q.n = 4;
            for (int i = 0; i < 4; ++i) { for (int d = 0; d < 3; ++d) q.pts[i][d] = P[i][d]; q.w[i] = 1.0 / 24.0; }
        } else {
            const double P[5][3] = {{0.25, 0.25, 0.25}, {0.5, 1.0 / 6.0, 1.0 / 6.0},
                                    {1.0 / 6.0, 0.5, 1.0 / 6.0}, {1.0 / 6.0, 1.0 / 6.0, 0.5},
                                    {1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0}};
            q.n = 5;
            for (int i = 0; i < 5; ++i) for (int d = 0; d < 3; ++d) q.pts[i][d] = P[i][d];
            q.w[0] = -2.0 / 15.0; q.w[1] = q.w[2] = q.w[3] = q.w[4] = 3.0 / 40.0;
        }
    }
    return q;
}

template <int D>
__global__ void k_quad_points(const double* __restrict__ V, const uint32_t* __restrict__ C,
                              uint32_t nc, QuadTab q, double* __restrict__ xq) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= (uint64_t)nc * q.n) return;
    const uint32_t c = t / q.n, iq = t % q.n;
    const double* v0 = V + D * C[c * (D + 1)];
    double J[D][D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        const double* vj = V + D * C[c * (D + 1) + j + 1];
#pragma unroll
        for (int i = 0; i < D; ++i) J[i][j] = vj[i] - v0[i];
    }
#pragma unroll
    for (int i = 0; i < D; ++i) {
        double s = J[i][0] * q.pts[iq][0];
#pragma unroll
        for (int j = 1; j < D; ++j) s = s + J[i][j] * q.pts[iq][j];
        xq[t * D + i] = v0[i] + s;
    }
}

}  // namespace fh

using namespace fh;

static fh_mesh* new_mesh(int dev, int dim, uint64_t nv, uint64_t nc) {
    FH_CHECK(dim == 2 || dim == 3, "mesh dim must be 2 or 3");
    FH_CHECK(nv < 0xffffffffull && nc < 0x0fffffffull, "mesh too large for 32-bit indices");
    auto* m = new fh_mesh;
    m->dev = dev;
    m->dim = dim;
    m->k = dim + 1;
    m->nv = (uint32_t)nv;
    m->nc = (uint32_t)nc;
    m->verts.alloc(dev, nv * dim);
    m->cells.alloc(dev, nc * m->k);
    return m;
}

extern "C" {

int fh_mesh_create(int dim, uint64_t nv, const double* vertices, uint64_t nc,
                   const uint64_t* cells, int device_idx, fh_mesh** out) {
    return guard([&] {
        FH_CHECK(out && vertices && cells, "null argument");
        const int dev = resolve_device(device_idx);
        DeviceCtx& d = device(dev);
        std::unique_ptr<fh_mesh> m(new_mesh(dev, dim, nv, nc));
        std::vector<uint32_t> c32(nc * m->k);
        for (size_t i = 0; i < c32.size(); ++i) {
            FH_CHECK(cells[i] < nv, "cell vertex index out of range");
            c32[i] = (uint32_t)cells[i];
        }
        h2d(m->verts.p, vertices, nv * dim, d.stream);
        h2d(m->cells.p, c32.data(), c32.size(), d.stream);
        FH_HIP(hipStreamSynchronize(d.stream));
        *out = m.release();
    });
}

int fh_mesh_unit_square(uint64_t nx, uint64_t ny, int device_idx, fh_mesh** out) {
    return guard([&] {
        FH_CHECK(out, "null argument");
        FH_CHECK(nx > 0 && ny > 0, "nx and ny must be at least 1");
        const int dev = resolve_device(device_idx);
        DeviceCtx& d = device(dev);
        std::unique_ptr<fh_mesh> m(new_mesh(dev, 2, (nx + 1) * (ny + 1), 2 * nx * ny));
        const uint64_t nt = std::max<uint64_t>(m->nv, nx * ny);
        hipLaunchKernelGGL(k_unit_square, dim3(cdiv(nt, 256)), dim3(256), 0, d.stream,
                           (uint32_t)nx, (uint32_t)ny, m->verts.p, m->cells.p);
        FH_LAUNCH_CHECK("unit_square");
        FH_HIP(hipStreamSynchronize(d.stream));
        *out = m.release();
    });
}

int fh_mesh_box_beam(double L, double W, double H, uint64_t nx, uint64_t ny, uint64_t nz,
                     int device_idx, fh_mesh** out) {
    return guard([&] {
        FH_CHECK(out, "null argument");
        FH_CHECK(nx > 0 && ny > 0 && nz > 0, "nx, ny, and nz must be at least 1");
        const int dev = resolve_device(device_idx);
        DeviceCtx& d = device(dev);
        std::unique_ptr<fh_mesh> m(
            new_mesh(dev, 3, (nx + 1) * (ny + 1) * (nz + 1), 6 * nx * ny * nz));
        const uint64_t nt = std::max<uint64_t>(m->nv, nx * ny * nz);
        hipLaunchKernelGGL(k_box_beam, dim3(cdiv(nt, 256)), dim3(256), 0, d.stream, (uint32_t)nx,
                           (uint32_t)ny, (uint32_t)nz, L, W, H, m->verts.p, m->cells.p);
        FH_LAUNCH_CHECK("box_beam");
        FH_HIP(hipStreamSynchronize(d.stream));
        *out = m.release();
    });
}

int fh_mesh_sizes(const fh_mesh* m, int* dim, uint64_t* nv, uint64_t* nc) {
    return guard([&] {
        FH_CHECK(m, "null mesh");
        if (dim) *dim = m->dim;
        if (nv) *nv = m->nv;
        if (nc) *nc = m->nc;
    });
}

int fh_mesh_download(const fh_mesh* m, double* vertices, uint64_t* cells) {
    return guard([&] {
        FH_CHECK(m, "null mesh");
        DeviceCtx& d = device(m->dev);
        if (vertices) d2h(vertices, m->verts.p, (size_t)m->nv * m->dim, d.stream);
        if (cells) {
            std::vector<uint32_t> c32((size_t)m->nc * m->k);
            d2h(c32.data(), m->cells.p, c32.size(), d.stream);
            FH_HIP(hipStreamSynchronize(d.stream));
            for (size_t i = 0; i < c32.size(); ++i) cells[i] = c32[i];
        }
        FH_HIP(hipStreamSynchronize(d.stream));
    });
}

int fh_mesh_boundary_dofs(fh_mesh* m, uint64_t* out, uint64_t* count) {
    return guard([&] {
        FH_CHECK(m && count, "null argument");
        DeviceCtx& d = device(m->dev);
        DBuf<uint8_t> mask(m->dev, m->nv);
        boundary_vertex_mask(m, mask.p, d.stream, nullptr, nullptr);
        const std::vector<uint64_t> h = mask_to_sorted(m->dev, mask.p, m->nv, d.stream);
        if (out) {
            FH_CHECK(*count >= h.size() || *count == 0, "output buffer too small");
            std::copy(h.begin(), h.end(), out);
        }
        *count = h.size();
    });
}

// P2 boundary dofs: vertices and edges of the boundary facets
int fh_p2_boundary_dofs(fh_mesh* m, uint64_t* out, uint64_t* count) {
    return guard([&] {
        FH_CHECK(m && count, "null argument");
        DeviceCtx& d = device(m->dev);
        MeshP2& p2 = mesh_p2(m);
        const uint32_t nd = p2.adj->ndofs;
        DBuf<uint8_t> mask(m->dev, m->nv), mask2(m->dev, nd);
        fill_u8(mask2.p, nd, 0, d.stream);
        boundary_vertex_mask(m, mask.p, d.stream, p2.cell_dofs.p, mask2.p);
        const std::vector<uint64_t> h = mask_to_sorted(m->dev, mask2.p, nd, d.stream);
        if (out) {
            FH_CHECK(*count >= h.size() || *count == 0, "output buffer too small");
            std::copy(h.begin(), h.end(), out);
        }
        *count = h.size();
    });
}

int fh_quad_points(fh_mesh* m, int degree, double* xq_out) {
    return guard([&] {
        FH_CHECK(m && xq_out, "null argument");
        DeviceCtx& d = device(m->dev);
        const QuadTab q = quad_table(m->dim, degree);
        const size_t n = (size_t)m->nc * q.n * m->dim;
        DBuf<double> xq(m->dev, n);
        const uint64_t nt = (uint64_t)m->nc * q.n;
        if (m->dim == 2)
            hipLaunchKernelGGL(k_quad_points<2>, dim3(cdiv(nt, 256)), dim3(256), 0, d.stream,
                               m->verts.p, m->cells.p, m->nc, q, xq.p);
        else
            hipLaunchKernelGGL(k_quad_points<3>, dim3(cdiv(nt, 256)), dim3(256), 0, d.stream,
                               m->verts.p, m->cells.p, m->nc, q, xq.p);
        FH_LAUNCH_CHECK("quad_points");
        d2h(xq_out, xq.p, n, d.stream);
        FH_HIP(hipStreamSynchronize(d.stream));
    });
}

int fh_mesh_destroy(fh_mesh* m) {
    return guard([&] {
        if (!m) return;
        (void)hipSetDevice(m->dev);
        delete m;
    });
}

}  // extern "C"
