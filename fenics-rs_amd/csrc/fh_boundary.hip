// fh_boundary.hip — boundary facets and exterior boundary (ds) integrals on the
// GPU: extract_boundary_facets (mesh.rs:79-113, 145-210) and boundary.rs:8-139
// (assemble_neumann_2d, assemble_surface_traction_3d, assemble_robin_matrix_2d).
//
// Facets are found per (cell, local facet) from the vertex -> cell adjacency
// (a facet no other cell holds is on the boundary) and listed in the order of
// their sorted vertex tuples: each facet is counted under its smallest vertex,
// the per-vertex lists are filled, then sorted in place by the remaining
// vertices.  The reference lists them in HashMap order, which is random; no
// caller depends on it.
//
// The integrals take the caller's facet list.  Each vertex gathers the
// contributions of its facets in list order (a per-vertex incidence list,
// sorted), which is the order of the reference's sequential scatter, so the
// results are bitwise those of the same loop on the host.  Closures (flux,
// traction) are sampled by the caller at the points fh_boundary_points returns.
#include <algorithm>
#include <cmath>
#include <vector>

#include "fh_internal.hpp"

namespace fh {

// ------------------------------------------------------------------ facets
// K vertices per cell; the facet opposite local vertex a.  2D local facets
// follow mesh.rs:84-88: a = 0 -> (1,2), 1 -> (2,0), 2 -> (0,1).
template <int K>
__device__ __forceinline__ void facet_of(const uint32_t* __restrict__ C, uint32_t c, int a,
                                         uint32_t* f) {
    if (K == 3) {
        f[0] = C[c * 3 + (a + 1) % 3];
        f[1] = C[c * 3 + (a + 2) % 3];
    } else {
        int m = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (b != a) f[m++] = C[c * 4 + b];
    }
}

template <int K>
__device__ __forceinline__ void sort_small(uint32_t* f) {
#pragma unroll
    for (int i = 1; i < K - 1; ++i)
#pragma unroll
        for (int j = i; j > 0; --j)
            if (f[j] < f[j - 1]) {
                const uint32_t t = f[j];
                f[j] = f[j - 1];
                f[j - 1] = t;
            }
}

// flag[t] = 1 for a boundary facet t = c * K + a; counted under its smallest vertex
template <int K>
__global__ void k_facet_flag(const uint32_t* __restrict__ C, uint32_t nc,
                             const uint32_t* __restrict__ aptr, const uint32_t* __restrict__ aent,
                             uint8_t* __restrict__ flag, uint32_t* __restrict__ cnt) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= (uint64_t)nc * K) return;
    const uint32_t c = t / K;
    const int a = t % K;
    uint32_t f[K - 1];
    facet_of<K>(C, c, a, f);
    sort_small<K>(f);
    int shared = 0;
    for (uint32_t e = aptr[f[0]]; e < aptr[f[0] + 1] && !shared; ++e) {
        const uint32_t c2 = aent[e] / K;
        if (c2 == c) continue;
        int hits = 0;
#pragma unroll
        for (int b = 0; b < K; ++b) {
            const uint32_t v = C[c2 * K + b];
#pragma unroll
            for (int q = 1; q < K - 1; ++q) hits += (v == f[q]);
        }
        shared = hits == K - 2;
    }
    flag[t] = shared ? 0 : 1;
    if (!shared) atomicAdd(&cnt[f[0]], 1u);
}

template <int K>
__global__ void k_facet_fill(const uint32_t* __restrict__ C, uint32_t nc,
                             const uint8_t* __restrict__ flag, const uint32_t* __restrict__ off,
                             uint32_t* __restrict__ fill, uint32_t* __restrict__ rec) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= (uint64_t)nc * K || !flag[t]) return;
    uint32_t f[K - 1];
    facet_of<K>(C, t / K, t % K, f);
    sort_small<K>(f);
    rec[off[f[0]] + atomicAdd(&fill[f[0]], 1u)] = (uint32_t)t;
}

// per owner vertex: its facets by their remaining sorted vertices
template <int K>
__global__ void k_facet_sort(const uint32_t* __restrict__ C, uint32_t nv,
                             const uint32_t* __restrict__ off, uint32_t* __restrict__ rec) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const uint32_t b = off[v], e = off[v + 1];
    auto key = [&](uint32_t t) {
        uint32_t f[K - 1];
        facet_of<K>(C, t / K, t % K, f);
        sort_small<K>(f);
        return K == 3 ? (uint64_t)f[1] : ((uint64_t)f[1] << 32) | f[K - 2];
    };
    for (uint32_t i = b + 1; i < e; ++i) {
        const uint32_t t = rec[i];
        const uint64_t kt = key(t);
        uint32_t j = i;
        while (j > b && key(rec[j - 1]) > kt) {
            rec[j] = rec[j - 1];
            --j;
        }
        rec[j] = t;
    }
}

// BoundaryFacet records: 2D the sorted edge (mesh.rs:91-95); 3D the face in
// local order, [f0, f2, f1] when (f1 - f0) x (f2 - f0) points to the opposite
// vertex (mesh.rs:164-186)
template <int K>
__global__ void k_facet_emit(const double* __restrict__ V, const uint32_t* __restrict__ C,
                             uint32_t nf, const uint32_t* __restrict__ rec,
                             uint64_t* __restrict__ verts, uint64_t* __restrict__ cell,
                             uint64_t* __restrict__ local) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const uint32_t t = rec[i], c = t / K;
    const int a = t % K;
    uint32_t f[K - 1];
    facet_of<K>(C, c, a, f);
    if (K == 3) {
        sort_small<K>(f);
        verts[2 * (uint64_t)i] = f[0];
        verts[2 * (uint64_t)i + 1] = f[1];
    } else {
        const double* va = V + 3 * (uint64_t)f[0];
        const double* vb = V + 3 * (uint64_t)f[1];
        const double* vc = V + 3 * (uint64_t)f[2];
        const double* vo = V + 3 * (uint64_t)C[c * 4 + a];
        const double ab[3] = {vb[0] - va[0], vb[1] - va[1], vb[2] - va[2]};
        const double ac[3] = {vc[0] - va[0], vc[1] - va[1], vc[2] - va[2]};
        const double n[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2],
                             ab[0] * ac[1] - ab[1] * ac[0]};
        const double tf[3] = {va[0] - vo[0], va[1] - vo[1], va[2] - vo[2]};
        const double dot = n[0] * tf[0] + n[1] * tf[1] + n[2] * tf[2];
        verts[3 * (uint64_t)i] = f[0];
        verts[3 * (uint64_t)i + 1] = dot >= 0.0 ? f[1] : f[2];
        verts[3 * (uint64_t)i + 2] = dot >= 0.0 ? f[2] : f[1];
    }
    cell[i] = c;
    local[i] = (uint64_t)a;
}

// ------------------------------------------------ facet -> vertex incidence
// D vertices per facet; entries f * D + local, per vertex in ascending order
__global__ void k_inc_count(const uint32_t* __restrict__ F, uint64_t n, uint32_t* __restrict__ cnt) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t < n) atomicAdd(&cnt[F[t]], 1u);
}
__global__ void k_inc_fill(const uint32_t* __restrict__ F, uint64_t n, const uint32_t* __restrict__ ptr,
                           uint32_t* __restrict__ fill, uint32_t* __restrict__ ent) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t < n) ent[ptr[F[t]] + atomicAdd(&fill[F[t]], 1u)] = (uint32_t)t;
}
__global__ void k_inc_sort(uint32_t nv, const uint32_t* __restrict__ ptr, uint32_t* __restrict__ ent) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const uint32_t b = ptr[v], e = ptr[v + 1];
    for (uint32_t i = b + 1; i < e; ++i) {
        const uint32_t t = ent[i];
        uint32_t j = i;
        while (j > b && ent[j - 1] > t) {
            ent[j] = ent[j - 1];
            --j;
        }
        ent[j] = t;
    }
}

// -------------------------------------------------------------- integrals
__device__ __forceinline__ double edge_length(const double* __restrict__ V, uint32_t a, uint32_t b) {
    const double dx = V[2 * (uint64_t)b] - V[2 * (uint64_t)a];
    const double dy = V[2 * (uint64_t)b + 1] - V[2 * (uint64_t)a + 1];
    return sqrt(dx * dx + dy * dy);
}

// boundary.rs:29-40: x(s) = (1 - s) v0 + s v1 at the two Gauss points
__global__ void k_neumann_points(const double* __restrict__ V, const uint32_t* __restrict__ F,
                                 uint32_t nf, double s0, double s1, double* __restrict__ xq) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const double* v0 = V + 2 * (uint64_t)F[2 * f];
    const double* v1 = V + 2 * (uint64_t)F[2 * f + 1];
    for (int q = 0; q < 2; ++q) {
        const double s = q ? s1 : s0;
        xq[4 * (uint64_t)f + 2 * q] = (1.0 - s) * v0[0] + s * v1[0];
        xq[4 * (uint64_t)f + 2 * q + 1] = (1.0 - s) * v0[1] + s * v1[1];
    }
}

// boundary.rs:77-93: the three edge midpoints of each facet triangle
__global__ void k_traction_points(const double* __restrict__ V, const uint32_t* __restrict__ F,
                                  uint32_t nf, double* __restrict__ mids) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    for (int m = 0; m < 3; ++m) {
        const double* a = V + 3 * (uint64_t)F[3 * f + m];
        const double* b = V + 3 * (uint64_t)F[3 * f + (m + 1) % 3];
        for (int c = 0; c < 3; ++c) mids[9 * (uint64_t)f + 3 * m + c] = 0.5 * (a[c] + b[c]);
    }
}

// boundary.rs:19-43: rhs_v = sum over v's facets (list order) and Gauss points
// of (w_q L) g_q phi(s_q), phi = 1 - s for the facet's first vertex, s for its second
__global__ void k_neumann(const double* __restrict__ V, const uint32_t* __restrict__ F,
                          const uint32_t* __restrict__ ptr, const uint32_t* __restrict__ ent,
                          const double* __restrict__ g, uint32_t nv, double s0, double s1,
                          double* __restrict__ rhs) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double acc = 0.0;
    for (uint32_t e = ptr[v]; e < ptr[v + 1]; ++e) {
        const uint32_t f = ent[e] / 2, loc = ent[e] % 2;
        const double length = edge_length(V, F[2 * f], F[2 * f + 1]);
        for (int q = 0; q < 2; ++q) {
            const double s = q ? s1 : s0;
            const double w = 0.5 * length;
            const double gv = g[2 * (uint64_t)f + q];
            acc += loc == 0 ? w * gv * (1.0 - s) : w * gv * s;
        }
    }
    rhs[v] = acc;
}

// boundary.rs:60-104: every facet vertex receives (area / 9) t(m) for each
// edge midpoint m, in midpoint order
__global__ void k_traction(const double* __restrict__ V, const uint32_t* __restrict__ F,
                           const uint32_t* __restrict__ ptr, const uint32_t* __restrict__ ent,
                           const double* __restrict__ t, uint32_t nv, double* __restrict__ rhs) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double acc[3] = {0.0, 0.0, 0.0};
    for (uint32_t e = ptr[v]; e < ptr[v + 1]; ++e) {
        const uint32_t f = ent[e] / 3;
        const double* v0 = V + 3 * (uint64_t)F[3 * f];
        const double* v1 = V + 3 * (uint64_t)F[3 * f + 1];
        const double* v2 = V + 3 * (uint64_t)F[3 * f + 2];
        const double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
        const double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
        const double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                             e1[0] * e2[1] - e1[1] * e2[0]};
        const double area = 0.5 * sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        const double w = area / 3.0;
        for (int m = 0; m < 3; ++m)
            for (int a = 0; a < 3; ++a) acc[a] += (w / 3.0) * t[9 * (uint64_t)f + 3 * m + a];
    }
    for (int a = 0; a < 3; ++a) rhs[3 * (uint64_t)v + a] = acc[a];
}

// boundary.rs:111-139 as CSR rows.  Row v: v itself and the other vertex of
// each of its facets, ascending; the diagonal sums gamma L / 3, a coupling
// gamma L / 6, over v's facets in list order (the triplet summation order).
__device__ __forceinline__ uint32_t other_end(const uint32_t* __restrict__ F, uint32_t e) {
    return F[2 * (e / 2) + 1 - e % 2];
}
__global__ void k_robin_count(const uint32_t* __restrict__ F, const uint32_t* __restrict__ ptr,
                              const uint32_t* __restrict__ ent, uint32_t nv,
                              uint32_t* __restrict__ len) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const uint32_t b = ptr[v], e = ptr[v + 1];
    uint32_t n = b < e ? 1 : 0;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t w = other_end(F, ent[i]);
        bool seen = false;
        for (uint32_t j = b; j < i; ++j) seen = seen || other_end(F, ent[j]) == w;
        n += seen ? 0 : 1;
    }
    len[v] = n;
}
__global__ void k_robin_fill(const double* __restrict__ V, const uint32_t* __restrict__ F,
                             const uint32_t* __restrict__ ptr, const uint32_t* __restrict__ ent,
                             uint32_t nv, double gamma, const uint32_t* __restrict__ rp,
                             uint32_t* __restrict__ cols, double* __restrict__ vals) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const uint32_t b = ptr[v], e = ptr[v + 1];
    if (b == e) return;
    uint32_t o = rp[v];
    const uint32_t o_end = rp[v + 1];
    // columns: v and the distinct other ends, ascending (selection by minimum)
    uint32_t prev = 0;
    bool first = true;
    for (; o < o_end; ++o) {
        uint32_t best = 0xffffffffu;
        if ((first || v > prev) && v < best) best = v;
        for (uint32_t i = b; i < e; ++i) {
            const uint32_t w = other_end(F, ent[i]);
            if ((first || w > prev) && w < best) best = w;
        }
        double acc = -0.0;
        for (uint32_t i = b; i < e; ++i) {
            const uint32_t f = ent[i] / 2;
            const double length = edge_length(V, F[2 * f], F[2 * f + 1]);
            if (best == v) acc += gamma * length / 3.0;
            else if (other_end(F, ent[i]) == best) acc += gamma * length / 6.0;
        }
        cols[o] = best;
        vals[o] = acc;
        prev = best;
        first = false;
    }
}

namespace {

unsigned g256(uint64_t n) { return std::max(1u, cdiv(n, 256)); }

// host facet list -> device u32, checked against the vertex count
DBuf<uint32_t> upload_facets(const fh_mesh* m, uint64_t nf, const uint64_t* facets, int D,
                             bool distinct, hipStream_t s) {
    FH_CHECK(nf == 0 || facets, "null facet list");
    FH_CHECK(nf * D < 0xffffffffull, "too many facets");
    std::vector<uint32_t> h(nf * D);
    for (uint64_t i = 0; i < nf * D; ++i) {
        FH_CHECK(facets[i] < m->nv, "facet vertex out of range");
        h[i] = (uint32_t)facets[i];
    }
    if (distinct)
        for (uint64_t f = 0; f < nf; ++f)
            for (int a = 0; a < D; ++a)
                for (int b = a + 1; b < D; ++b)
                    FH_CHECK(h[f * D + a] != h[f * D + b], "a boundary facet repeats a vertex");
    DBuf<uint32_t> d(m->dev, std::max<uint64_t>(nf * D, 1));
    h2d(d.p, h.data(), nf * D, s);
    FH_HIP(hipStreamSynchronize(s));  // h is released on return
    return d;
}

struct Incidence {
    DBuf<uint32_t> ptr, ent;
};

Incidence incidence(const fh_mesh* m, const uint32_t* F, uint64_t nf, int D, hipStream_t s) {
    Incidence r;
    const uint64_t n = nf * D;
    DBuf<uint32_t> cnt(m->dev, m->nv + 1), fill(m->dev, m->nv);
    r.ptr.alloc(m->dev, m->nv + 1);
    r.ent.alloc(m->dev, std::max<uint64_t>(n, 1));
    fill_u32(cnt.p, m->nv + 1, 0, s);
    fill_u32(fill.p, m->nv, 0, s);
    if (n) hipLaunchKernelGGL(k_inc_count, dim3(g256(n)), dim3(256), 0, s, F, n, cnt.p);
    exclusive_scan_u32(cnt.p, r.ptr.p, m->nv + 1, m->dev, s);
    if (n) hipLaunchKernelGGL(k_inc_fill, dim3(g256(n)), dim3(256), 0, s, F, n, r.ptr.p, fill.p, r.ent.p);
    hipLaunchKernelGGL(k_inc_sort, dim3(g256(m->nv)), dim3(256), 0, s, m->nv, r.ptr.p, r.ent.p);
    FH_LAUNCH_CHECK("incidence");
    FH_HIP(hipStreamSynchronize(s));  // cnt and fill are released on return
    return r;
}

double gauss2(int q) {  // quadrature.rs:41-56, the 2-point rule on [0, 1]
    const double c = 1.0 / (2.0 * std::sqrt(3.0));
    return q == 0 ? 0.5 - c : 0.5 + c;
}

}  // namespace

}  // namespace fh

using namespace fh;

extern "C" {

int fh_mesh_boundary_facets(fh_mesh* m, uint64_t* vertices, uint64_t* cell_idx,
                            uint64_t* local_idx, uint64_t* count) {
    return guard([&] {
        FH_CHECK(m && count, "null argument");
        DeviceCtx& d = device(m->dev);
        hipStream_t s = d.stream;
        Adjacency& adj = mesh_adjacency(m);
        const int K = m->k, D = m->dim;
        const uint64_t nt = (uint64_t)m->nc * K;
        DBuf<uint8_t> flag(m->dev, std::max<uint64_t>(nt, 1));
        DBuf<uint32_t> cnt(m->dev, m->nv + 1), off(m->dev, m->nv + 1), fill(m->dev, m->nv);
        fill_u32(cnt.p, m->nv + 1, 0, s);
        fill_u32(fill.p, m->nv, 0, s);
        if (nt) {
            if (K == 3)
                hipLaunchKernelGGL(k_facet_flag<3>, dim3(g256(nt)), dim3(256), 0, s, m->cells.p,
                                   m->nc, adj.ptr.p, adj.ent.p, flag.p, cnt.p);
            else
                hipLaunchKernelGGL(k_facet_flag<4>, dim3(g256(nt)), dim3(256), 0, s, m->cells.p,
                                   m->nc, adj.ptr.p, adj.ent.p, flag.p, cnt.p);
            FH_LAUNCH_CHECK("facet_flag");
        }
        const uint64_t nf = exclusive_scan_u32(cnt.p, off.p, m->nv + 1, m->dev, s);
        if (vertices || cell_idx || local_idx) {
            FH_CHECK(*count >= nf || *count == 0, "output buffer too small");
            DBuf<uint32_t> rec(m->dev, std::max<uint64_t>(nf, 1));
            DBuf<uint64_t> dv(m->dev, std::max<uint64_t>(nf * D, 1)), dc(m->dev, std::max<uint64_t>(nf, 1)),
                dl(m->dev, std::max<uint64_t>(nf, 1));
            if (nf) {
                if (K == 3) {
                    hipLaunchKernelGGL(k_facet_fill<3>, dim3(g256(nt)), dim3(256), 0, s, m->cells.p,
                                       m->nc, flag.p, off.p, fill.p, rec.p);
                    hipLaunchKernelGGL(k_facet_sort<3>, dim3(g256(m->nv)), dim3(256), 0, s,
                                       m->cells.p, m->nv, off.p, rec.p);
                    hipLaunchKernelGGL(k_facet_emit<3>, dim3(g256(nf)), dim3(256), 0, s, m->verts.p,
                                       m->cells.p, (uint32_t)nf, rec.p, dv.p, dc.p, dl.p);
                } else {
                    hipLaunchKernelGGL(k_facet_fill<4>, dim3(g256(nt)), dim3(256), 0, s, m->cells.p,
                                       m->nc, flag.p, off.p, fill.p, rec.p);
                    hipLaunchKernelGGL(k_facet_sort<4>, dim3(g256(m->nv)), dim3(256), 0, s,
                                       m->cells.p, m->nv, off.p, rec.p);
                    hipLaunchKernelGGL(k_facet_emit<4>, dim3(g256(nf)), dim3(256), 0, s, m->verts.p,
                                       m->cells.p, (uint32_t)nf, rec.p, dv.p, dc.p, dl.p);
                }
                FH_LAUNCH_CHECK("facets");
                if (vertices) d2h(vertices, dv.p, nf * D, s);
                if (cell_idx) d2h(cell_idx, dc.p, nf, s);
                if (local_idx) d2h(local_idx, dl.p, nf, s);
            }
            FH_HIP(hipStreamSynchronize(s));
        }
        *count = nf;
    });
}

int fh_boundary_points(fh_mesh* m, uint64_t nf, const uint64_t* facets, double* points) {
    return guard([&] {
        FH_CHECK(m && points, "null argument");
        DeviceCtx& d = device(m->dev);
        hipStream_t s = d.stream;
        const int D = m->dim;
        DBuf<uint32_t> F = upload_facets(m, nf, facets, D, false, s);
        const uint64_t np = nf * (D == 2 ? 4 : 9);
        DBuf<double> out(m->dev, std::max<uint64_t>(np, 1));
        if (nf) {
            if (D == 2)
                hipLaunchKernelGGL(k_neumann_points, dim3(g256(nf)), dim3(256), 0, s, m->verts.p,
                                   F.p, (uint32_t)nf, gauss2(0), gauss2(1), out.p);
            else
                hipLaunchKernelGGL(k_traction_points, dim3(g256(nf)), dim3(256), 0, s, m->verts.p,
                                   F.p, (uint32_t)nf, out.p);
            FH_LAUNCH_CHECK("boundary_points");
            d2h(points, out.p, np, s);
        }
        FH_HIP(hipStreamSynchronize(s));
    });
}

int fh_assemble_neumann_2d(fh_mesh* m, uint64_t nf, const uint64_t* facets, const double* g,
                           double* rhs_out) {
    return guard([&] {
        FH_CHECK(m && rhs_out && (nf == 0 || g), "null argument");
        FH_CHECK(m->dim == 2, "assemble_neumann_2d needs a triangle mesh");
        DeviceCtx& d = device(m->dev);
        hipStream_t s = d.stream;
        DBuf<uint32_t> F = upload_facets(m, nf, facets, 2, false, s);
        Incidence inc = incidence(m, F.p, nf, 2, s);
        DBuf<double> dg(m->dev, std::max<uint64_t>(2 * nf, 1)), rhs(m->dev, m->nv);
        h2d(dg.p, g, 2 * nf, s);
        hipLaunchKernelGGL(k_neumann, dim3(g256(m->nv)), dim3(256), 0, s, m->verts.p, F.p,
                           inc.ptr.p, inc.ent.p, dg.p, m->nv, gauss2(0), gauss2(1), rhs.p);
        FH_LAUNCH_CHECK("neumann");
        d2h(rhs_out, rhs.p, m->nv, s);
        FH_HIP(hipStreamSynchronize(s));
    });
}

int fh_assemble_surface_traction_3d(fh_mesh* m, uint64_t nf, const uint64_t* facets,
                                    const double* t, double* rhs_out) {
    return guard([&] {
        FH_CHECK(m && rhs_out && (nf == 0 || t), "null argument");
        FH_CHECK(m->dim == 3, "assemble_surface_traction_3d needs a tetrahedral mesh");
        DeviceCtx& d = device(m->dev);
        hipStream_t s = d.stream;
        DBuf<uint32_t> F = upload_facets(m, nf, facets, 3, false, s);
        Incidence inc = incidence(m, F.p, nf, 3, s);
        DBuf<double> dt(m->dev, std::max<uint64_t>(9 * nf, 1)), rhs(m->dev, 3 * (uint64_t)m->nv);
        h2d(dt.p, t, 9 * nf, s);
        hipLaunchKernelGGL(k_traction, dim3(g256(m->nv)), dim3(256), 0, s, m->verts.p, F.p,
                           inc.ptr.p, inc.ent.p, dt.p, m->nv, rhs.p);
        FH_LAUNCH_CHECK("traction");
        d2h(rhs_out, rhs.p, 3 * (uint64_t)m->nv, s);
        FH_HIP(hipStreamSynchronize(s));
    });
}

int fh_assemble_robin_matrix_2d(fh_mesh* m, uint64_t nf, const uint64_t* facets, double gamma,
                                fh_matrix** out) {
    return guard([&] {
        FH_CHECK(m && out, "null argument");
        FH_CHECK(m->dim == 2, "assemble_robin_matrix_2d needs a triangle mesh");
        DeviceCtx& d = device(m->dev);
        hipStream_t s = d.stream;
        DBuf<uint32_t> F = upload_facets(m, nf, facets, 2, true, s);
        Incidence inc = incidence(m, F.p, nf, 2, s);
        DBuf<uint32_t> len(m->dev, m->nv + 1), rp(m->dev, m->nv + 1);
        fill_u32(len.p, m->nv + 1, 0, s);
        hipLaunchKernelGGL(k_robin_count, dim3(g256(m->nv)), dim3(256), 0, s, F.p, inc.ptr.p,
                           inc.ent.p, m->nv, len.p);
        FH_LAUNCH_CHECK("robin_count");
        const uint64_t nnz = exclusive_scan_u32(len.p, rp.p, m->nv + 1, m->dev, s);
        DBuf<uint32_t> cols(m->dev, std::max<uint64_t>(nnz, 1));
        DBuf<double> vals(m->dev, std::max<uint64_t>(nnz, 1));
        hipLaunchKernelGGL(k_robin_fill, dim3(g256(m->nv)), dim3(256), 0, s, m->verts.p, F.p,
                           inc.ptr.p, inc.ent.p, m->nv, gamma, rp.p, cols.p, vals.p);
        FH_LAUNCH_CHECK("robin_fill");
        std::vector<uint32_t> hrp(m->nv + 1), hc(nnz);
        std::vector<double> hv(nnz);
        d2h(hrp.data(), rp.p, m->nv + 1, s);
        d2h(hc.data(), cols.p, nnz, s);
        d2h(hv.data(), vals.p, nnz, s);
        FH_HIP(hipStreamSynchronize(s));
        *out = matrix_from_csr(m->dev, m->nv, 1, hrp.data(), hc.data(), hv.data());
    });
}

}  // extern "C"
