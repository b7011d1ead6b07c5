// fh_pattern.hip — symbolic phase of assembly.
//
// fenics-rs builds its sparsity by sorting every (row, col, val) triplet
// through faer's try_new_from_triplets (poisson.rs:68, elasticity.rs:148).
// Here the pattern is derived once per mesh from the connectivity: the row of
// dof v holds every dof of every cell incident to v, sorted and unique — the
// same canonical set faer produces (structural zeros included).  Built on the
// device in integer arithmetic, so it is deterministic and bit-exact:
//   1. dof -> cell adjacency by counting sort, segments sorted by cell;
//   2. row lengths and columns by repeated minimum selection over the
//      incident cells' dofs (no scratch arrays per thread);
//   3. a slot map: for every (dof, incident cell, local dof b) the position of
//      cell dof b inside row dof, so numeric assembly never searches;
//   4. the SELL-64 layout (64 consecutive rows per slice, column-major inside
//      the slice, slice width = longest row).
#include <algorithm>
#include <type_traits>

#include <hipcub/hipcub.hpp>

#include "fh_internal.hpp"

namespace fh {

__global__ void k_adj_count(const uint32_t* __restrict__ cd, uint64_t n, uint32_t* __restrict__ cnt) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t < n) atomicAdd(&cnt[cd[t]], 1u);
}
__global__ void k_adj_fill(const uint32_t* __restrict__ cd, uint64_t n, uint32_t* __restrict__ cur,
                           uint32_t* __restrict__ ent) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t < n) ent[atomicAdd(&cur[cd[t]], 1u)] = (uint32_t)t;  // t == cell*K + local
}
// insertion sort of each dof's segment (segments are short: the cells around one dof)
__global__ void k_adj_sort(const uint32_t* __restrict__ ptr, uint32_t* __restrict__ ent, uint32_t nd) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nd) return;
    const uint32_t a = ptr[v], b = ptr[v + 1];
    for (uint32_t i = a + 1; i < b; ++i) {
        const uint32_t x = ent[i];
        uint32_t j = i;
        while (j > a && ent[j - 1] > x) {
            ent[j] = ent[j - 1];
            --j;
        }
        ent[j] = x;
    }
}

std::unique_ptr<Adjacency> build_adjacency(int dev, const uint32_t* cd, uint32_t nc, int K,
                                           uint32_t ndofs, hipStream_t s) {
    auto a = std::make_unique<Adjacency>();
    a->ndofs = ndofs;
    a->K = K;
    const uint64_t n = (uint64_t)nc * K;
    DBuf<uint32_t> cnt(dev, std::max<uint32_t>(ndofs, 1));
    fill_u32(cnt.p, ndofs, 0, s);
    if (n) hipLaunchKernelGGL(k_adj_count, dim3(cdiv(n, 256)), dim3(256), 0, s, cd, n, cnt.p);
    FH_LAUNCH_CHECK("adj_count");
    a->ptr.alloc(dev, ndofs + 1);
    exclusive_scan_u32(cnt.p, a->ptr.p, ndofs, dev, s);
    a->ent.alloc(dev, std::max<uint64_t>(n, 1));
    d2d(cnt.p, a->ptr.p, ndofs, s);
    if (n) hipLaunchKernelGGL(k_adj_fill, dim3(cdiv(n, 256)), dim3(256), 0, s, cd, n, cnt.p, a->ent.p);
    FH_LAUNCH_CHECK("adj_fill");
    hipLaunchKernelGGL(k_adj_sort, dim3(cdiv(ndofs, 256)), dim3(256), 0, s, a->ptr.p, a->ent.p, ndofs);
    FH_LAUNCH_CHECK("adj_sort");
    return a;
}

// Row v = sorted unique dofs of the cells around v (v alone when it has none).
template <int K, bool FILL>
__global__ void k_row_pattern(const uint32_t* __restrict__ cd, const uint32_t* __restrict__ aptr,
                              const uint32_t* __restrict__ aent, uint32_t nd,
                              uint32_t* __restrict__ len, const uint32_t* __restrict__ rptr,
                              uint32_t* __restrict__ cols) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nd) return;
    const uint32_t e0 = aptr[v], e1 = aptr[v + 1];
    uint32_t cnt = 0;
    uint32_t out = FILL ? rptr[v] : 0;
    if (e0 == e1) {
        if (FILL) cols[out] = v;
        else len[v] = 1;
        return;
    }
    int64_t last = -1;
    for (;;) {
        uint32_t best = 0xffffffffu;
        for (uint32_t e = e0; e < e1; ++e) {
            const uint32_t c = aent[e] / K;
#pragma unroll
            for (int b = 0; b < K; ++b) {
                const uint32_t d = cd[(uint64_t)c * K + b];
                if ((int64_t)d > last && d < best) best = d;
            }
        }
        if (best == 0xffffffffu) break;
        if (FILL) cols[out + cnt] = best;
        ++cnt;
        last = best;
    }
    if (!FILL) len[v] = cnt;
}

template <int K, typename S>
__global__ void k_slots(const uint32_t* __restrict__ cd, const uint32_t* __restrict__ aptr,
                        const uint32_t* __restrict__ aent, uint32_t nd,
                        const uint32_t* __restrict__ rptr, const uint32_t* __restrict__ cols,
                        S* __restrict__ slots) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nd) return;
    const uint32_t r0 = rptr[v], r1 = rptr[v + 1];
    for (uint32_t e = aptr[v]; e < aptr[v + 1]; ++e) {
        const uint32_t c = aent[e] / K;
#pragma unroll
        for (int b = 0; b < K; ++b) {
            const uint32_t d = cd[(uint64_t)c * K + b];
            uint32_t lo = r0, hi = r1;  // lower_bound
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (cols[mid] < d) lo = mid + 1;
                else hi = mid;
            }
            slots[(uint64_t)e * K + b] = (S)(lo - r0);
        }
    }
}

__global__ void k_slice_width(const uint32_t* __restrict__ len, uint32_t nrows, uint32_t ns,
                              uint32_t* __restrict__ w) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    uint32_t m = 0;
    for (uint32_t l = 0; l < SLICE; ++l) {
        const uint32_t r = s * SLICE + l;
        if (r < nrows) m = max(m, len[r]);
    }
    w[s] = m * SLICE;
}

__global__ void k_sell_cols(const uint32_t* __restrict__ rptr, const uint32_t* __restrict__ cols,
                            const uint32_t* __restrict__ soff, uint32_t nrows, uint32_t nslices,
                            uint32_t* __restrict__ scols) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t s = r / SLICE, l = r % SLICE;
    if (s >= nslices) return;
    const uint32_t base = soff[s], w = (soff[s + 1] - base) / SLICE;
    const uint32_t len = r < nrows ? rptr[r + 1] - rptr[r] : 0;
    for (uint32_t k = 0; k < w; ++k)
        scols[base + k * SLICE + l] = k < len ? cols[rptr[r] + k] : (r < nrows ? r : 0);
}

__global__ void k_row_len(const uint32_t* __restrict__ rptr, uint32_t n, uint32_t* __restrict__ len) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) len[r] = rptr[r + 1] - rptr[r];
}

template <int K>
static void launch_rows(const uint32_t* cd, Adjacency& adj, Pattern& p, bool fill, uint32_t* len,
                        hipStream_t s) {
    const unsigned g = cdiv(adj.ndofs, 64);
    if (fill)
        hipLaunchKernelGGL((k_row_pattern<K, true>), dim3(g), dim3(64), 0, s, cd, adj.ptr.p,
                           adj.ent.p, adj.ndofs, len, p.row_ptr.p, p.cols.p);
    else
        hipLaunchKernelGGL((k_row_pattern<K, false>), dim3(g), dim3(64), 0, s, cd, adj.ptr.p,
                           adj.ent.p, adj.ndofs, len, p.row_ptr.p, p.cols.p);
    FH_LAUNCH_CHECK("row_pattern");
}

std::shared_ptr<Pattern> build_pattern(int dev, const uint32_t* cd, int K, Adjacency& adj,
                                       hipStream_t s) {
    auto p = std::make_shared<Pattern>();
    p->dev = dev;
    p->nrows = adj.ndofs;
    const uint32_t n = adj.ndofs;
    DBuf<uint32_t> len(dev, std::max<uint32_t>(n, 1));
    p->row_ptr.alloc(dev, n + 1);
    switch (K) {
        case 3: launch_rows<3>(cd, adj, *p, false, len.p, s); break;
        case 4: launch_rows<4>(cd, adj, *p, false, len.p, s); break;
        case 10: launch_rows<10>(cd, adj, *p, false, len.p, s); break;
        default: throw Error(FH_ERR_UNSUPPORTED, "unsupported cell size");
    }
    p->nnzb = exclusive_scan_u32(len.p, p->row_ptr.p, n, dev, s);
    p->cols.alloc(dev, std::max<uint64_t>(p->nnzb, 1));
    switch (K) {
        case 3: launch_rows<3>(cd, adj, *p, true, len.p, s); break;
        case 4: launch_rows<4>(cd, adj, *p, true, len.p, s); break;
        case 10: launch_rows<10>(cd, adj, *p, true, len.p, s); break;
    }
    // SELL-64 layout
    p->nslices = cdiv(n, SLICE);
    DBuf<uint32_t> w(dev, std::max<uint32_t>(p->nslices, 1));
    p->row_len.alloc(dev, std::max<uint32_t>(n, 1));
    hipLaunchKernelGGL(k_row_len, dim3(cdiv(n, 256)), dim3(256), 0, s, p->row_ptr.p, n, p->row_len.p);
    hipLaunchKernelGGL(k_slice_width, dim3(cdiv(p->nslices, 256)), dim3(256), 0, s, p->row_len.p, n,
                       p->nslices, w.p);
    FH_LAUNCH_CHECK("slice_width");
    p->slice_off.alloc(dev, p->nslices + 1);
    p->sell_entries = exclusive_scan_u32(w.p, p->slice_off.p, p->nslices, dev, s);
    p->sell_cols.alloc(dev, std::max<uint64_t>(p->sell_entries, 1));
    hipLaunchKernelGGL(k_sell_cols, dim3(cdiv((size_t)p->nslices * SLICE, 256)), dim3(256), 0, s,
                       p->row_ptr.p, p->cols.p, p->slice_off.p, n, p->nslices, p->sell_cols.p);
    FH_LAUNCH_CHECK("sell_cols");
    // widest row (slot indices are stored as u8, or as u16 past 255 blocks)
    std::vector<uint32_t> hw(p->nslices);
    d2h(hw.data(), w.p, p->nslices, s);
    FH_HIP(hipStreamSynchronize(s));
    uint32_t mw = 0;
    for (uint32_t x : hw) mw = std::max(mw, x / SLICE);
    p->max_width = mw;
    if (mw > 65535) throw Error(FH_ERR_UNSUPPORTED, "row longer than 65535 blocks");
    return p;
}

void build_slots(const uint32_t* cd, int K, Adjacency& adj, const Pattern& pat, hipStream_t s) {
    uint64_t nent = 0;
    FH_HIP(hipMemcpyAsync(&nent, adj.ptr.p + adj.ndofs, 4, hipMemcpyDeviceToHost, s));
    FH_HIP(hipStreamSynchronize(s));
    nent &= 0xffffffffull;
    const unsigned g = cdiv(adj.ndofs, 64);
    auto launch = [&](auto* out) {
        using S = std::remove_pointer_t<decltype(out)>;
        switch (K) {
            case 3:
                hipLaunchKernelGGL((k_slots<3, S>), dim3(g), dim3(64), 0, s, cd, adj.ptr.p,
                                   adj.ent.p, adj.ndofs, pat.row_ptr.p, pat.cols.p, out);
                break;
            case 4:
                hipLaunchKernelGGL((k_slots<4, S>), dim3(g), dim3(64), 0, s, cd, adj.ptr.p,
                                   adj.ent.p, adj.ndofs, pat.row_ptr.p, pat.cols.p, out);
                break;
            case 10:
                hipLaunchKernelGGL((k_slots<10, S>), dim3(g), dim3(64), 0, s, cd, adj.ptr.p,
                                   adj.ent.p, adj.ndofs, pat.row_ptr.p, pat.cols.p, out);
                break;
        }
    };
    // u8 slots while every row fits (half the bytes the assembly kernels read)
    if (pat.max_width > 255) {
        adj.slots.reset();
        adj.slots16.alloc(pat.dev, std::max<uint64_t>(nent * K, 1));
        launch(adj.slots16.p);
    } else {
        adj.slots16.reset();
        adj.slots.alloc(pat.dev, std::max<uint64_t>(nent * K, 1));
        launch(adj.slots.p);
    }
    FH_LAUNCH_CHECK("slots");
}

Adjacency& mesh_adjacency(fh_mesh* m) {
    if (!m->adj) {
        DeviceCtx& d = device(m->dev);
        m->adj = build_adjacency(m->dev, m->cells.p, m->nc, m->k, m->nv, d.stream);
    }
    return *m->adj;
}

std::shared_ptr<Pattern> mesh_pattern(fh_mesh* m) {
    if (!m->pat) {
        DeviceCtx& d = device(m->dev);
        Adjacency& adj = mesh_adjacency(m);
        m->pat = build_pattern(m->dev, m->cells.p, m->k, adj, d.stream);
        build_slots(m->cells.p, m->k, adj, *m->pat, d.stream);
    }
    return m->pat;
}

// ---------------------------------------------------------------- P2 dofmap
// Vertices keep their numbers; edge dofs follow, ordered by the number of
// cells around the edge (most first; counts above 255 count as 255), then by
// the vertex difference max - min, then by min.  Edges with equal counts
// have rows of nearly equal length, so SELL slices of edge rows carry little
// padding (unit_cube(64^3): 1.02 stored entries per matrix entry, against
// 1.15 in plain lexicographic order; the C2 product 15% faster).  On a
// structured mesh max - min is the edge's direction, so a slice holds edges
// of one direction at consecutive vertices, whose neighbours sit at the same
// offsets: the product's gathers are contiguous runs (C2 product 185 ->
// 177 us against ties in (min, max) order).
// The lexicographic position comes first: the upper triangle (col > row) of
// the vertex graph in CSR order is exactly that order, position = upper
// entries before it; a radix sort of (count class, max - min, min) keys then
// gives the numbering.
__global__ void k_upper_count(const uint32_t* __restrict__ rptr, const uint32_t* __restrict__ cols,
                              uint32_t n, uint32_t* __restrict__ cnt) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    uint32_t c = 0;
    for (uint32_t p = rptr[v]; p < rptr[v + 1]; ++p) c += cols[p] > v;
    cnt[v] = c;
}

__device__ __forceinline__ uint32_t edge_id(const uint32_t* rptr, const uint32_t* cols,
                                            const uint32_t* uptr, uint32_t nv, uint32_t a,
                                            uint32_t b) {
    const uint32_t lo = min(a, b), hi = max(a, b);
    // position of hi among the entries of row lo that exceed lo
    uint32_t p0 = rptr[lo], p1 = rptr[lo + 1];
    uint32_t first = p0;
    {  // first entry > lo
        uint32_t l = p0, h = p1;
        while (l < h) {
            const uint32_t m = (l + h) >> 1;
            if (cols[m] <= lo) l = m + 1;
            else h = m;
        }
        first = l;
    }
    uint32_t l = first, h = p1;
    while (l < h) {
        const uint32_t m = (l + h) >> 1;
        if (cols[m] < hi) l = m + 1;
        else h = m;
    }
    return nv + uptr[lo] + (l - first);
}

// P2Tetrahedron local edges (lagrange.rs:291-296): 01, 12, 02, 03, 13, 23.
__global__ void k_p2_cell_dofs(const uint32_t* __restrict__ C, uint32_t nc,
                               const uint32_t* __restrict__ rptr, const uint32_t* __restrict__ cols,
                               const uint32_t* __restrict__ uptr, uint32_t nv,
                               uint32_t* __restrict__ cd) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const uint32_t v[4] = {C[4 * c], C[4 * c + 1], C[4 * c + 2], C[4 * c + 3]};
    const int E[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
    uint32_t* o = cd + 10 * (uint64_t)c;
#pragma unroll
    for (int a = 0; a < 4; ++a) o[a] = v[a];
#pragma unroll
    for (int e = 0; e < 6; ++e) o[4 + e] = edge_id(rptr, cols, uptr, nv, v[E[e][0]], v[E[e][1]]);
}

// cells around each edge (lexicographic position e): cnt[e]
__global__ void k_edge_cells(const uint32_t* __restrict__ cd, uint32_t nc, uint32_t nv,
                             uint32_t* __restrict__ cnt) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= 6 * (uint64_t)nc) return;
    const uint64_t c = i / 6, k = i % 6;
    atomicAdd(&cnt[cd[10 * c + 4 + k] - nv], 1u);
}
// sort key of the edge at lexicographic position e = (lo, hi):
// (255 - capped count, hi - lo, lo), most cells first; (hi - lo, lo) is unique
__global__ void k_edge_keys(const uint32_t* __restrict__ rptr, const uint32_t* __restrict__ cols,
                            const uint32_t* __restrict__ uptr, uint32_t nv,
                            const uint32_t* __restrict__ cnt, uint64_t* __restrict__ key,
                            uint32_t* __restrict__ pos) {
    const uint32_t lo = blockIdx.x * blockDim.x + threadIdx.x;
    if (lo >= nv) return;
    uint32_t e = uptr[lo];
    for (uint32_t p = rptr[lo]; p < rptr[lo + 1]; ++p) {
        const uint32_t hi = cols[p];
        if (hi <= lo) continue;
        key[e] = (uint64_t)(255u - min(cnt[e], 255u)) << 56 | (uint64_t)(hi - lo) << 28 | lo;
        pos[e] = e;
        ++e;
    }
}
__global__ void k_edge_rank(const uint32_t* __restrict__ sorted_pos, uint32_t ne,
                            uint32_t* __restrict__ rank) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ne) rank[sorted_pos[i]] = i;
}
__global__ void k_edge_renumber(uint32_t* __restrict__ cd, uint32_t nc, uint32_t nv,
                                const uint32_t* __restrict__ rank) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= 6 * (uint64_t)nc) return;
    uint32_t& d = cd[10 * (i / 6) + 4 + i % 6];
    d = nv + rank[d - nv];
}

// lexicographic edge positions of the cell dofs -> the edge dof numbers above
static void order_edges(uint32_t* cd, uint32_t nc, const Pattern& vp, const uint32_t* uptr,
                        uint32_t nv, uint32_t ne, int dev, hipStream_t s) {
    if (!ne) return;
    FH_CHECK(nv < (1u << 28), "P2 edge numbering needs fewer than 2^28 vertices");
    FH_CHECK(ne < (1u << 31), "P2 edge numbering needs fewer than 2^31 edges");
    DBuf<uint32_t> cnt(dev, ne), pos(dev, ne), spos(dev, ne), rank(dev, ne);
    DBuf<uint64_t> key(dev, ne), skey(dev, ne);
    fill_u32(cnt.p, ne, 0u, s);
    const unsigned g6 = std::max(1u, cdiv(6 * (uint64_t)nc, 256)), ge = cdiv(ne, 256);
    hipLaunchKernelGGL(k_edge_cells, dim3(g6), dim3(256), 0, s, cd, nc, nv, cnt.p);
    hipLaunchKernelGGL(k_edge_keys, dim3(cdiv(nv, 256)), dim3(256), 0, s, vp.row_ptr.p,
                       vp.cols.p, uptr, nv, cnt.p, key.p, pos.p);
    size_t tb = 0;
    FH_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key.p, skey.p, pos.p, spos.p, (int)ne,
                                              0, 64, s));
    DBuf<uint8_t> tmp(dev, std::max<size_t>(tb, 1));
    FH_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, key.p, skey.p, pos.p, spos.p, (int)ne, 0,
                                              64, s));
    hipLaunchKernelGGL(k_edge_rank, dim3(ge), dim3(256), 0, s, spos.p, ne, rank.p);
    hipLaunchKernelGGL(k_edge_renumber, dim3(g6), dim3(256), 0, s, cd, nc, nv, rank.p);
    FH_LAUNCH_CHECK("order_edges");
}

MeshP2& mesh_p2(fh_mesh* m) {
    if (!m->p2) {
        FH_CHECK(m->dim == 3, "P2 dofs need a tetrahedral mesh");
        DeviceCtx& d = device(m->dev);
        auto pat = mesh_pattern(m);
        auto p2 = std::make_unique<MeshP2>();
        DBuf<uint32_t> cnt(m->dev, m->nv), uptr(m->dev, m->nv + 1);
        hipLaunchKernelGGL(k_upper_count, dim3(cdiv(m->nv, 256)), dim3(256), 0, d.stream,
                           pat->row_ptr.p, pat->cols.p, m->nv, cnt.p);
        p2->nedges = (uint32_t)exclusive_scan_u32(cnt.p, uptr.p, m->nv, m->dev, d.stream);
        p2->cell_dofs.alloc(m->dev, (size_t)m->nc * 10);
        hipLaunchKernelGGL(k_p2_cell_dofs, dim3(cdiv(m->nc, 256)), dim3(256), 0, d.stream,
                           m->cells.p, m->nc, pat->row_ptr.p, pat->cols.p, uptr.p, m->nv,
                           p2->cell_dofs.p);
        FH_LAUNCH_CHECK("p2_cell_dofs");
        order_edges(p2->cell_dofs.p, m->nc, *pat, uptr.p, m->nv, p2->nedges, m->dev, d.stream);
        const uint32_t nd = m->nv + p2->nedges;
        p2->adj = build_adjacency(m->dev, p2->cell_dofs.p, m->nc, 10, nd, d.stream);
        p2->pat = build_pattern(m->dev, p2->cell_dofs.p, 10, *p2->adj, d.stream);
        build_slots(p2->cell_dofs.p, 10, *p2->adj, *p2->pat, d.stream);
        FH_HIP(hipStreamSynchronize(d.stream));
        m->p2 = std::move(p2);
    }
    return *m->p2;
}

}  // namespace fh
