// fh_assemble.hip — numeric assembly on the device.
//
// Row-gather formulation of the reference's cell loops (poisson.rs:23-218,
// elasticity.rs:73-227, mass.rs:10-82): one thread owns one block row (a
// vertex, or a P2 dof), walks the cells incident to it in ascending cell
// order and adds that cell's contributions for its row into the row's slots.
// Consequences:
//   * no atomics and no colouring: every value slot has exactly one writer;
//   * each slot's duplicates are summed in ascending cell order, the order in
//     which rayon's collect() lays the triplets out and the order the oracle
//     sums them in, so FP64 values are bit-identical to the oracle;
//   * load vectors use the same per-dof order as the reference's sequential
//     scatter (poisson.rs:116-119, elasticity.rs:157-182), so they are
//     bit-identical too.
// Values are written straight into the SELL-64 layout the CG kernels read.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <type_traits>
#include <vector>

#include "fh_geom.hpp"
#include "fh_internal.hpp"

namespace fh {

using namespace fhg;

// value index of (block row r, entry k, block component j) in SELL-64
__device__ __forceinline__ uint64_t vidx(uint32_t soff_s, uint32_t k, int j, uint32_t lane,
                                         int bs2) {
    return vpos((uint64_t)soff_s + (uint64_t)k * SLICE, j, lane, bs2);
}

struct RowCtx {
    uint32_t soff, len, lane;
};
__device__ __forceinline__ RowCtx row_ctx(const uint32_t* soff, const uint32_t* rlen, uint32_t r) {
    return {soff[r / SLICE], rlen[r], r % SLICE};
}

template <int BS2>
__device__ __forceinline__ void init_row(double* vals, const RowCtx& rc) {
    for (uint32_t k = 0; k < rc.len; ++k)
#pragma unroll
        for (int j = 0; j < BS2; ++j) vals[vidx(rc.soff, k, j, rc.lane, BS2)] = -0.0;
}

// poisson.rs:33-66 (kind 0: stiffness) and mass.rs:16-41 (kind 1: mass), P1 triangles
template <int KIND, typename S>
__global__ __launch_bounds__(256) void k_asm_p1_tri(
    const double* __restrict__ V, const uint32_t* __restrict__ C, const uint32_t* __restrict__ aptr,
    const uint32_t* __restrict__ aent, const S* __restrict__ slots,
    const uint32_t* __restrict__ soff, const uint32_t* __restrict__ rlen, uint32_t nv,
    double* __restrict__ vals) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const RowCtx rc = row_ctx(soff, rlen, v);
    init_row<1>(vals, rc);
    for (uint32_t e = aptr[v]; e < aptr[v + 1]; ++e) {
        const uint32_t c = aent[e] / 3, a = aent[e] % 3;
        double x[3][2];
        uint32_t t[3];
        load_cell<2, 3>(V, C, c, x, t);
        Map2 m;
        if (!affine2(x[0], x[1], x[2], m)) continue;
        const double area = 0.5 * fabs(m.det);
        if (KIND == 0) {
            double g[3][2];
            p1_tri_grads(m, g);
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const double val = (g[a][0] * g[b][0] + g[a][1] * g[b][1]) * area;
                vals[vidx(rc.soff, slots[(uint64_t)e * 3 + b], 0, rc.lane, 1)] += val;
            }
        } else {
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const double factor = ((int)a == b) ? 1.0 / 6.0 : 1.0 / 12.0;
                vals[vidx(rc.soff, slots[(uint64_t)e * 3 + b], 0, rc.lane, 1)] += area * factor;
            }
        }
    }
}

// poisson.rs:129-162 (kind 0) and mass.rs:53-79 (kind 1), P1 tetrahedra
template <int KIND, typename S>
__global__ __launch_bounds__(256) void k_asm_p1_tet(
    const double* __restrict__ V, const uint32_t* __restrict__ C, const uint32_t* __restrict__ aptr,
    const uint32_t* __restrict__ aent, const S* __restrict__ slots,
    const uint32_t* __restrict__ soff, const uint32_t* __restrict__ rlen, uint32_t nv,
    double* __restrict__ vals) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const RowCtx rc = row_ctx(soff, rlen, v);
    init_row<1>(vals, rc);
    for (uint32_t e = aptr[v]; e < aptr[v + 1]; ++e) {
        const uint32_t c = aent[e] / 4, a = aent[e] % 4;
        double x[4][3];
        uint32_t t[4];
        load_cell<3, 4>(V, C, c, x, t);
        Map3 m;
        if (!affine3(x[0], x[1], x[2], x[3], m)) continue;
        const double vol = (1.0 / 6.0) * fabs(m.det);
        if (KIND == 0) {
            double g[4][3];
            p1_tet_grads(m, g);
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const double val = (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]) * vol;
                vals[vidx(rc.soff, slots[(uint64_t)e * 4 + b], 0, rc.lane, 1)] += val;
            }
        } else {
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const double factor = ((int)a == b) ? 1.0 / 10.0 : 1.0 / 20.0;
                vals[vidx(rc.soff, slots[(uint64_t)e * 4 + b], 0, rc.lane, 1)] += vol * factor;
            }
        }
    }
}

// elasticity.rs:86-146 in two kernels.
//
// 1. k_elas_cell_geom, one thread per cell: the cell's P1 gradients
//    g[0..3][0..2], its volume and, with a per-cell E, its Lame constants,
//    into a 128-byte record of four 32-byte quarters, quarter b = (g[b], X_b)
//    with X = (vol, lambda, mu, 0) (vol < 0 marks a cell the reference skips,
//    |det J| < 1e-14).  The geometry is computed once per cell instead of
//    once per (row, cell) pair.
// 2. k_asm_elasticity_rows: the 3x3 blocks of each row vertex v (local index
//    a in the cell) against every vertex b of each incident cell.  One wave
//    takes 16 rows of a SELL slice, four lanes per row, lane q the block of
//    cell vertex b = q, so a wave's accumulator is 16 rows wide.  The blocks
//    are summed in LDS (elas_acc_index: each half-wave's 8 rows own four
//    consecutive doubles of every 32-double line, picked by k mod 4, so two
//    lanes of a half-wave meet in a bank only when two blocks of one cell are
//    congruent mod 4), and written to the SELL values once, with 16-byte
//    stores of the component pairs.  Every row still walks its incident cells
//    in ascending order, so every value is the oracle's sum, bit for bit.
//    The accumulator holds `wp` blocks per row; a slice wider than that is
//    done in passes over ranges of wp blocks.  Work unit u = 4 * slice +
//    quarter; each block loops over units of its XCD (row_unit).
constexpr int ELAS_GS = 16;  // doubles per cell record

template <bool EC>
__global__ __launch_bounds__(256) void k_elas_cell_geom(const double* __restrict__ V,
                                                        const uint32_t* __restrict__ C,
                                                        uint32_t nc, double nu,
                                                        const double* __restrict__ Ecell,
                                                        double* __restrict__ G) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    double o[ELAS_GS];
#pragma unroll
    for (int i = 0; i < ELAS_GS; ++i) o[i] = 0.0;
    double x[4][3];
    uint32_t t[4];
    Map3 m;
    if (c >= nc) {
        // no cell: the record is never stored
    } else if (load_cell<3, 4>(V, C, c, x, t), affine3(x[0], x[1], x[2], x[3], m)) {
        double g[4][3];
        p1_tet_grads(m, g);
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int i = 0; i < 3; ++i) o[4 * b + i] = g[b][i];
        o[3] = (1.0 / 6.0) * fabs(m.det);
        if (EC) {
            const double ec = Ecell[c];
            o[7] = lame_lambda(ec, nu);
            o[11] = lame_mu(ec, nu);
        }
    } else {
        o[3] = -1.0;
    }
    // the block's records are one contiguous range: stage them in LDS and
    // store them as consecutive 16-byte pieces
    // (a thread's 8 pieces at stride 9, so a wave's stores spread over the banks)
    constexpr uint32_t NP = ELAS_GS / 2, SP = NP + 1;
    __shared__ d2 stage[256 * SP];
#pragma unroll
    for (uint32_t i = 0; i < NP; ++i) {
        d2 v;
        v.x = o[2 * i];
        v.y = o[2 * i + 1];
        stage[threadIdx.x * SP + i] = v;
    }
    __syncthreads();
    const uint32_t c0 = blockIdx.x * 256;
    const uint32_t npieces = min(256u, nc - c0) * NP;
    d2* out = reinterpret_cast<d2*>(G + (uint64_t)c0 * ELAS_GS);
    for (uint32_t i = threadIdx.x; i < npieces; i += 256) out[i] = stage[(i / NP) * SP + i % NP];
}

// Work units of the row kernels (16 rows of a slice) are dealt to the 8 XCDs
// in chunks of ROW_CHUNK consecutive units, round robin, so rows that share
// cells run on one XCD (its L2) and a range of heavy rows (the vertex dofs of
// a P2 numbering) spreads over all XCDs.  XCD x's j-th unit:
constexpr uint32_t ROW_CHUNK = 16;
__device__ __forceinline__ uint32_t row_unit(uint32_t x, uint32_t j) {
    return ((j / ROW_CHUNK) * 8 + x) * ROW_CHUNK + j % ROW_CHUNK;
}
// units each XCD takes (its last chunk may run past the end: units >= nunits
// are not processed)
inline uint32_t row_units_per_xcd(uint32_t nunits) {
    return (uint32_t)cdiv(nunits, 8 * ROW_CHUNK) * ROW_CHUNK;
}

constexpr int ELAS_RPW = 16;  // rows per wave of k_asm_elasticity_rows
// LDS position of component j of block kk (counted from the pass's first
// block) of row r (0-15): line j + 9 (kk / 4) holds, for each half of the
// rows, 8 rows x 4 blocks
__device__ __forceinline__ uint32_t elas_acc_index(uint32_t r, uint32_t kk, uint32_t j) {
    return ((j + 9 * (kk / 4)) * 2 + r / 8) * 32 + (r % 8) * 4 + kk % 4;
}
__host__ __device__ inline uint32_t elas_acc_size(uint32_t nk) { return 9 * ((nk + 3) / 4) * 64; }

// Lane q of a row's four loads quarter q of an incident cell's record (two
// 16-byte loads: g[q] and X_q), so the quad reads the whole 128-byte record
// as one line; the row's own gradient g[a] then comes from lane a of the
// quad (ds_bpermute) and vol, lambda, mu from lanes 0-2 (DPP broadcasts).
// The record also holds the column's block slot and the index entry and slot
// of the cell it takes after this one.
struct ElasRec {
    double w[4];
    uint32_t a, k, nen, nk;
};
template <bool EC>
__device__ __forceinline__ void rec_load(const double* __restrict__ G, uint32_t en, uint32_t kq,
                                         uint32_t q, ElasRec& r) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2* gp = reinterpret_cast<const d2*>(G + (uint64_t)(en / 4) * ELAS_GS + 4 * q);
    const d2 u = gp[0], v = gp[1];
    r.w[0] = u.x;
    r.w[1] = u.y;
    r.w[2] = v.x;
    r.w[3] = v.y;
    r.a = en % 4;
    r.k = kq;
}
// value v of lane src (an absolute lane index) of the wave
__device__ __forceinline__ double lane_read(double v, uint32_t src) {
    const int lo = __builtin_amdgcn_ds_bpermute((int)(src << 2), __double2loint(v));
    const int hi = __builtin_amdgcn_ds_bpermute((int)(src << 2), __double2hiint(v));
    return __hiloint2double(hi, lo);
}
// value v of lane P of each quad, to every lane of the quad (DPP quad_perm)
template <int P>
__device__ __forceinline__ double quad_bcast(double v) {
    constexpr int ctrl = P | P << 2 | P << 4 | P << 6;
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), ctrl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), ctrl, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
// records in flight per lane: LDS holds the kernel to two waves per SIMD, so
// registers are spent on depth instead
constexpr int ELAS_RING = 6;
// first fill: data of cell e, index of cell e + ELAS_RING (positions past
// the row's last cell re-read the last one and are never used)
template <bool EC, typename S>
__device__ __forceinline__ void rec_start(const double* __restrict__ G,
                                          const uint32_t* __restrict__ aent,
                                          const S* __restrict__ slots, uint32_t e,
                                          uint32_t last, uint32_t q, ElasRec& r) {
    const uint32_t ec = min(e, last), en = min(e + ELAS_RING, last);
    rec_load<EC>(G, aent[ec], slots[(uint64_t)ec * 4 + q], q, r);
    r.nen = aent[en];
    r.nk = slots[(uint64_t)en * 4 + q];
}
// add cell e's block into the accumulator (the nine old values are read
// first, so the LDS round trip overlaps the arithmetic), then load the data
// of cell e + ELAS_RING and the index of cell e + 2 ELAS_RING.  The four
// lanes of a row run the same steps, so the quad exchanges below read active
// lanes only.
template <bool EC, typename S>
__device__ __forceinline__ void rec_step(const double* __restrict__ G,
                                         const uint32_t* __restrict__ aent,
                                         const S* __restrict__ slots, uint32_t e,
                                         uint32_t last, uint32_t q, uint32_t lane, double lambda0,
                                         double mu0, uint32_t k0, uint32_t kmax, double* acc,
                                         uint32_t racc, ElasRec& r) {
    const uint32_t src = (lane & ~3u) | r.a;
    double ga[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) ga[i] = lane_read(r.w[i], src);
    const double vol = quad_bcast<0>(r.w[3]);
    const double lam = EC ? quad_bcast<1>(r.w[3]) : lambda0;
    const double mu = EC ? quad_bcast<2>(r.w[3]) : mu0;
    if (!(vol < 0.0) && r.k >= k0 && r.k < kmax) {
        // components j of one block are 64 doubles apart
        double* ab = acc + elas_acc_index(racc, r.k - k0, 0);
        double old[9];
#pragma unroll
        for (int j = 0; j < 9; ++j) old[j] = ab[j * 64];
#pragma unroll
        for (int al = 0; al < 3; ++al)
#pragma unroll
            for (int be = 0; be < 3; ++be)
                ab[(al * 3 + be) * 64] =
                    old[al * 3 + be] + elas_entry(ga, r.w, al, be, lam, mu, vol);
    }
    const uint32_t en = min(e + 2 * ELAS_RING, last);
    rec_load<EC>(G, r.nen, r.nk, q, r);
    r.nen = aent[en];
    r.nk = slots[(uint64_t)en * 4 + q];
}

template <bool EC, typename S>
__global__ __launch_bounds__(64) void k_asm_elasticity_rows(
    const double* __restrict__ G, const uint32_t* __restrict__ aptr,
    const uint32_t* __restrict__ aent, const S* __restrict__ slots,
    const uint32_t* __restrict__ soff, const uint32_t* __restrict__ rlen, uint32_t nv,
    double lambda0, double mu0, uint32_t wp, uint32_t nunits, uint32_t per,
    double* __restrict__ vals) {
    extern __shared__ double acc[];
    typedef double d2 __attribute__((ext_vector_type(2)));
    // this XCD's units, strided by the XCD's blocks
    for (uint32_t j = blockIdx.x / 8; j < per; j += gridDim.x / 8) {
        const uint32_t u = row_unit(blockIdx.x % 8, j);
        if (u >= nunits) break;
        const uint32_t sl = u / 4, h = u % 4, lane = threadIdx.x;
        const uint32_t q = lane & 3, row = sl * SLICE + h * ELAS_RPW + (lane >> 2);
        const uint32_t base = soff[sl], w = (soff[sl + 1] - base) / SLICE;
        const bool live = row < nv;
        const uint32_t e0 = live ? aptr[row] : 0, e1 = live ? aptr[row + 1] : 0;
        // the row and value pair this lane writes out
        const uint32_t wr = lane % ELAS_RPW, wpp = lane / ELAS_RPW;
        const uint32_t wrow = sl * SLICE + h * ELAS_RPW + wr;
        const uint32_t wlen = wrow < nv ? rlen[wrow] : 0;
        const uint32_t racc = lane >> 2;
        for (uint32_t k0 = 0; k0 < w; k0 += wp) {
            const uint32_t nk = min(wp, w - k0);
            for (uint32_t i = lane; i < elas_acc_size(nk); i += SLICE) acc[i] = -0.0;
            __syncthreads();
            if (e0 < e1) {
                const uint32_t last = e1 - 1;
                // ELAS_RING records in rotation: record i holds the data of the
                // cell it processes next and the index of the cell ELAS_RING
                // after that, so data arrive ELAS_RING steps and indices twice
                // that ahead of use
                ElasRec r[ELAS_RING];
#pragma unroll
                for (int i = 0; i < ELAS_RING; ++i)
                    rec_start<EC, S>(G, aent, slots, e0 + i, last, q, r[i]);
                const uint32_t kmax = nk + k0;
                uint32_t e = e0;
                for (; e + ELAS_RING - 1 < e1; e += ELAS_RING) {
#pragma unroll
                    for (int i = 0; i < ELAS_RING; ++i)
                        rec_step<EC, S>(G, aent, slots, e + i, last, q, lane, lambda0, mu0, k0, kmax,
                                     acc, racc, r[i]);
                }
#pragma unroll
                for (int i = 0; i < ELAS_RING - 1; ++i)
                    if (e + i < e1)
                        rec_step<EC, S>(G, aent, slots, e + i, last, q, lane, lambda0, mu0, k0, kmax,
                                     acc, racc, r[i]);
            }
            __syncthreads();
            // this pass's blocks: lane (wpp, wr) stores the component pair wpp of
            // row wr, lanes 0-15 also component 8; zeros past the row's length
            for (uint32_t kk = 0; kk < nk; ++kk) {
                const uint32_t k = k0 + kk;
                double* vk = vals + ((uint64_t)base + (uint64_t)k * SLICE) * 9;
                const bool on = k < wlen;
                const double* ak = acc + elas_acc_index(wr, kk, 0);
                d2 v;
                v.x = on ? ak[(2 * wpp) * 64] : 0.0;
                v.y = on ? ak[(2 * wpp + 1) * 64] : 0.0;
                *reinterpret_cast<d2*>(vk + 128 * wpp + 2 * (h * ELAS_RPW + wr)) = v;
                if (wpp == 0) vk[512 + h * ELAS_RPW + wr] = on ? ak[8 * 64] : 0.0;
            }
            __syncthreads();
        }
    }
}

// Row gather of stored FP32 elasticity element blocks (the MFMA path,
// k_elas_elem_mfma): block (a, b) of cell c is components 0-7 at
// Ke8[((4c + a) 4 + b) 8 ..] and component 8 at Ke9[(4c + a) 4 + b].  The
// row walk, the LDS sums (elas_acc_index, in floats) and the write-out are
// those of k_asm_elasticity_rows: lane q of a row's four adds the block of
// cell vertex b = q, in ascending cell order, and the slice's values are
// written once, as FP32 and (vals64 given) as their FP64 widening.
constexpr int ELAS_GRING = 4;  // element blocks in flight per lane
typedef float f4v_t __attribute__((ext_vector_type(4)));
struct ElasBlk {
    f4v_t lo, hi;
    float v8;
    uint32_t k;
};
template <typename S>
__device__ __forceinline__ void elas_blk_load(const float* __restrict__ Ke8,
                                              const float* __restrict__ Ke9,
                                              const uint32_t* __restrict__ aent,
                                              const S* __restrict__ slots, uint32_t e, uint32_t q,
                                              ElasBlk& x) {
    const uint64_t blk = (uint64_t)aent[e] * 4 + q;  // aent = 4 cell + a
    const f4v_t* p = reinterpret_cast<const f4v_t*>(Ke8 + blk * 8);
    x.lo = p[0];
    x.hi = p[1];
    x.v8 = Ke9[blk];
    x.k = slots[(uint64_t)e * 4 + q];
}

template <typename S>
__global__ __launch_bounds__(64) void k_gather_elas_rows(
    const float* __restrict__ Ke8, const float* __restrict__ Ke9,
    const uint32_t* __restrict__ aptr, const uint32_t* __restrict__ aent,
    const S* __restrict__ slots, const uint32_t* __restrict__ soff,
    const uint32_t* __restrict__ rlen, uint32_t nv, uint32_t wp, uint32_t nunits, uint32_t per,
    float* __restrict__ vals32, double* __restrict__ vals64) {
    extern __shared__ float accf[];
    typedef float f2 __attribute__((ext_vector_type(2)));
    typedef double d2 __attribute__((ext_vector_type(2)));
    for (uint32_t j = blockIdx.x / 8; j < per; j += gridDim.x / 8) {
        const uint32_t u = row_unit(blockIdx.x % 8, j);
        if (u >= nunits) break;
        const uint32_t sl = u / 4, h = u % 4, lane = threadIdx.x;
        const uint32_t q = lane & 3, racc = lane >> 2, row = sl * SLICE + h * ELAS_RPW + racc;
        const uint32_t base = soff[sl], w = (soff[sl + 1] - base) / SLICE;
        const bool live = row < nv;
        const uint32_t e0 = live ? aptr[row] : 0, e1 = live ? aptr[row + 1] : 0;
        const uint32_t wr = lane % ELAS_RPW, wpp = lane / ELAS_RPW;
        const uint32_t wrow = sl * SLICE + h * ELAS_RPW + wr;
        const uint32_t wlen = wrow < nv ? rlen[wrow] : 0;
        for (uint32_t k0 = 0; k0 < w; k0 += wp) {
            const uint32_t nk = min(wp, w - k0);
            for (uint32_t i = lane; i < elas_acc_size(nk); i += SLICE) accf[i] = -0.0f;
            __syncthreads();
            if (e0 < e1) {
                const uint32_t last = e1 - 1;
                ElasBlk x[ELAS_GRING];
#pragma unroll
                for (int i = 0; i < ELAS_GRING; ++i)
                    elas_blk_load(Ke8, Ke9, aent, slots, min(e0 + i, last), q, x[i]);
                for (uint32_t e = e0; e < e1; e += ELAS_GRING) {
#pragma unroll
                    for (int i = 0; i < ELAS_GRING; ++i) {
                        if (e + i < e1 && x[i].k - k0 < nk) {
                            float* ab = accf + elas_acc_index(racc, x[i].k - k0, 0);
                            const float v[9] = {x[i].lo.x, x[i].lo.y, x[i].lo.z, x[i].lo.w,
                                                x[i].hi.x, x[i].hi.y, x[i].hi.z, x[i].hi.w,
                                                x[i].v8};
                            float old[9];
#pragma unroll
                            for (int c = 0; c < 9; ++c) old[c] = ab[c * 64];
#pragma unroll
                            for (int c = 0; c < 9; ++c) ab[c * 64] = old[c] + v[c];
                        }
                        // the entry ELAS_GRING ahead (past the row's last: the last again, unused)
                        elas_blk_load(Ke8, Ke9, aent, slots, min(e + i + ELAS_GRING, last), q, x[i]);
                    }
                }
            }
            __syncthreads();
            for (uint32_t kk = 0; kk < nk; ++kk) {
                const uint32_t k = k0 + kk;
                const uint64_t o = ((uint64_t)base + (uint64_t)k * SLICE) * 9;
                const bool on = k < wlen;
                const float* ak = accf + elas_acc_index(wr, kk, 0);
                f2 v;
                v.x = on ? ak[(2 * wpp) * 64] : 0.0f;
                v.y = on ? ak[(2 * wpp + 1) * 64] : 0.0f;
                const uint64_t pi = o + 128 * wpp + 2 * (h * ELAS_RPW + wr);
                *reinterpret_cast<f2*>(vals32 + pi) = v;
                const float v8 = on ? ak[8 * 64] : 0.0f;
                if (wpp == 0) vals32[o + 512 + h * ELAS_RPW + wr] = v8;
                if (vals64) {
                    d2 dv;
                    dv.x = (double)v.x;
                    dv.y = (double)v.y;
                    *reinterpret_cast<d2*>(vals64 + pi) = dv;
                    if (wpp == 0) vals64[o + 512 + h * ELAS_RPW + wr] = (double)v8;
                }
            }
            __syncthreads();
        }
    }
}

// mass.rs:86-126: M_{(v,al),(b,al)} += density * vol * (1/10 if a == b else
// 1/20) for every incident cell, the same scalar mass on each component.
template <typename S>
__global__ __launch_bounds__(256) void k_asm_elasticity_mass(
    const double* __restrict__ V, const uint32_t* __restrict__ C, const uint32_t* __restrict__ aptr,
    const uint32_t* __restrict__ aent, const S* __restrict__ slots,
    const uint32_t* __restrict__ soff, const uint32_t* __restrict__ rlen, uint32_t nv,
    double density, double* __restrict__ vals) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const RowCtx rc = row_ctx(soff, rlen, v);
    init_row<9>(vals, rc);
    for (uint32_t e = aptr[v]; e < aptr[v + 1]; ++e) {
        const uint32_t c = aent[e] / 4, a = aent[e] % 4;
        double x[4][3];
        uint32_t t[4];
        load_cell<3, 4>(V, C, c, x, t);
        Map3 m;
        if (!affine3(x[0], x[1], x[2], x[3], m)) continue;
        const double vol = (1.0 / 6.0) * fabs(m.det);
        for (int b = 0; b < 4; ++b) {
            const uint32_t k = slots[(uint64_t)e * 4 + b];
            const double factor = (a == b) ? 1.0 / 10.0 : 1.0 / 20.0;
            const double mv = density * vol * factor;
#pragma unroll
            for (int al = 0; al < 3; ++al) vals[vidx(rc.soff, k, al * 3 + al, rc.lane, 9)] += mv;
        }
    }
}

// mass.rs:136-180: the P2 vector mass, M_{(i,al),(j,al)} += density |det J|
// I_ab per cell (I_ab the reference-tetrahedron product of P2 basis
// functions a and b), the same scalar on each component.  Nodes [n][3],
// cells [nc][10] (4 corners, then edges 01, 12, 02, 03, 13, 23); the map
// uses the corners only.
struct P2MassTab {
    double m[100];
};
template <typename S>
__global__ __launch_bounds__(256) void k_asm_p2_vec_mass(
    const double* __restrict__ V, const uint32_t* __restrict__ C, const uint32_t* __restrict__ aptr,
    const uint32_t* __restrict__ aent, const S* __restrict__ slots,
    const uint32_t* __restrict__ soff, const uint32_t* __restrict__ rlen, uint32_t nd,
    double density, P2MassTab tab, double* __restrict__ vals) {
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= nd) return;
    const RowCtx rc = row_ctx(soff, rlen, d);
    init_row<9>(vals, rc);
    for (uint32_t e = aptr[d]; e < aptr[d + 1]; ++e) {
        const uint32_t c = aent[e] / 10, a = aent[e] % 10;
        double x[4][3];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int q = 0; q < 3; ++q) x[k][q] = V[3 * (uint64_t)C[(uint64_t)c * 10 + k] + q];
        Map3 m;
        if (!affine3(x[0], x[1], x[2], x[3], m)) continue;
        const double det = fabs(m.det);
        for (int b = 0; b < 10; ++b) {
            const uint32_t k = slots[(uint64_t)e * 10 + b];
            const double mv = density * det * tab.m[a * 10 + b];
#pragma unroll
            for (int al = 0; al < 3; ++al) vals[vidx(rc.soff, k, al * 3 + al, rc.lane, 9)] += mv;
        }
    }
}

// poisson_solver.rs:62-96 (triangles) / 135-172 (tetrahedra): per cell,
// sum_q w_q |det J| (u_h(x_q) - u_exact(x_q))^2 with the degree-3 rule, the P1
// basis at the reference points (1 - sum xi, xi...); block partial sums (the
// host adds them in block order).  A degenerate cell raises *bad (the
// reference's unwrap panics there).
template <int D>
__global__ __launch_bounds__(256) void k_l2_err(const double* __restrict__ V,
                                                const uint32_t* __restrict__ C, uint32_t nc,
                                                const double* __restrict__ u,
                                                const double* __restrict__ uq, QuadTab q,
                                                double* __restrict__ part, uint32_t* bad) {
    __shared__ double sh[4];
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    double tot = 0.0;
    if (c < nc) {
        constexpr int K = D + 1;
        double x[K][D];
        uint32_t t[K];
        load_cell<D, K>(V, C, c, x, t);
        double detj;
        bool ok;
        if constexpr (D == 2) {
            Map2 m;
            ok = affine2(x[0], x[1], x[2], m);
            detj = fabs(m.det);
        } else {
            Map3 m;
            ok = affine3(x[0], x[1], x[2], x[3], m);
            detj = fabs(m.det);
        }
        if (!ok) atomicOr(bad, 1u);
        for (int k = 0; k < q.n; ++k) {
            double phi[K];
            phi[0] = D == 2 ? 1.0 - q.pts[k][0] - q.pts[k][1]
                            : 1.0 - q.pts[k][0] - q.pts[k][1] - q.pts[k][2];
#pragma unroll
            for (int a = 1; a < K; ++a) phi[a] = q.pts[k][a - 1];
            double ap = u[t[0]] * phi[0];
#pragma unroll
            for (int a = 1; a < K; ++a) ap = ap + u[t[a]] * phi[a];
            const double err = ap - uq[(uint64_t)c * q.n + k];
            tot += q.w[k] * detj * err * err;
        }
    }
    tot = fhd::block_sum256(tot, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// Dirichlet fold (eigen.rs:51-78): the diagonal value of a constrained dof is
// `diag` (1 for the stiffness, 0 for the mass), whatever its multiplicity.
__global__ void k_fold_count(const uint32_t* __restrict__ count, uint32_t n, uint32_t diag,
                             uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = count[i] ? diag : 0u;
}

// P2 tetrahedron stiffness, FP64 VALU: K(a,b) = sum_q (w_q |det J|) (g_a^q . g_b^q)
// with g^q = J^{-T} grad phi(xi_q) (lagrange.rs:324-346 at the
// quadrature.rs:157-184 points).  G = reference gradients [q][dof][3].
//
// k_p2_elem_f64 computes each cell's 10 x 10 matrix once (one thread per
// cell; the 55 entries a <= b, since g_a . g_b and g_b . g_a are the same
// products summed in the same order) into Ke [nc][10][10]; k_gather_p2_rows
// then sums the rows into the SELL values.  A cell the reference skips
// (|det J| < 1e-14) has every entry set to P2_SKIP_BITS, a signalling NaN
// that no arithmetic produces.
struct P2Tab {
    double G[4][10][3];
    double w[4];
};
constexpr uint64_t P2_SKIP_BITS = 0x7ff0000000000badull;

__global__ __launch_bounds__(64) void k_p2_elem_f64(const double* __restrict__ V,
                                                    const uint32_t* __restrict__ C, uint32_t nc,
                                                    P2Tab tab, double* __restrict__ Ke) {
    __shared__ double stage[16 * 100];
    typedef double d2 __attribute__((ext_vector_type(2)));
    const uint32_t lane = threadIdx.x, c0 = blockIdx.x * 64, c = c0 + lane;
    double K[55];
#pragma unroll
    for (int i = 0; i < 55; ++i) K[i] = __longlong_as_double((long long)P2_SKIP_BITS);
    Map3 m;
    bool ok = false;
    if (c < nc) {
        double x[4][3];
        uint32_t t[4];
        load_cell<3, 4>(V, C, c, x, t);
        ok = affine3(x[0], x[1], x[2], x[3], m);
    }
    if (ok) {
        const double detj = fabs(m.det);
#pragma unroll
        for (int i = 0; i < 55; ++i) K[i] = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double sq = tab.w[q] * detj;
            double g[10][3];
#pragma unroll
            for (int b = 0; b < 10; ++b) grad3(m, tab.G[q][b], g[b]);
            int i = 0;
#pragma unroll
            for (int a = 0; a < 10; ++a)
#pragma unroll
                for (int b = a; b < 10; ++b, ++i)
                    K[i] += sq * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]);
        }
    }
    // four rounds of 16 cells: the round's lanes expand their matrices into
    // LDS, then the wave stores the 16 matrices as one contiguous range
#pragma unroll 1
    for (uint32_t rr = 0; rr < 4; ++rr) {
        if (lane / 16 == rr) {
            double* o = stage + (lane % 16) * 100;
            int i = 0;
#pragma unroll
            for (int a = 0; a < 10; ++a)
#pragma unroll
                for (int b = a; b < 10; ++b, ++i) {
                    o[a * 10 + b] = K[i];
                    o[b * 10 + a] = K[i];
                }
        }
        __syncthreads();
        const uint32_t cr = c0 + rr * 16;
        if (cr < nc) {
            const uint32_t np = min(16u, nc - cr) * 50;
            d2* out = reinterpret_cast<d2*>(Ke + (uint64_t)cr * 100);
            const d2* st = reinterpret_cast<const d2*>(stage);
            for (uint32_t i = lane; i < np; i += 64) out[i] = st[i];
        }
        __syncthreads();
    }
}

// Row gather of element matrices Ke [nc][10][10] (T = double: the FP64
// matrices, skipped cells marked; T = float: the MFMA matrices) into the
// SELL-64 values of T.  One wave takes 16 rows of a slice, four lanes per row:
// lane q adds entries (2q, 2q+1) of the row's element-matrix row, lanes 0 and
// 1 also entry 8 + q, so a row's 10 entries arrive as four 16-byte (8-byte for
// float) loads and two small ones.  Rows are summed in LDS (p2_acc_index,
// banks as in the elasticity kernel) in ascending cell order per slot, and
// written once (every SELL entry, padding included; with vals64 also as
// doubles).  The accumulator holds wp blocks per row (512: 64 KB of LDS for
// doubles); a wider row is done in passes over ranges of wp blocks.
__device__ __forceinline__ uint32_t p2_acc_index(uint32_t r, uint32_t k) {
    return ((k / 4) * 2 + r / 8) * 32 + (r % 8) * 4 + k % 4;
}
__device__ __forceinline__ bool p2_skipped(double v) {
    return (uint64_t)__double_as_longlong(v) == P2_SKIP_BITS;
}
__device__ __forceinline__ bool p2_skipped(float) { return false; }

// two consecutive slots in one load, as (first | second << 16)
__device__ __forceinline__ uint32_t slot_pair(const uint8_t* p) {
    const uint32_t v = *reinterpret_cast<const uint16_t*>(p);
    return (v & 0xffu) | (v >> 8) << 16;
}
__device__ __forceinline__ uint32_t slot_pair(const uint16_t* p) {
    return *reinterpret_cast<const uint32_t*>(p);
}

// lane q's part of one incident cell's element-matrix row: entries (2q,
// 2q+1) and, for q < 2, 8 + q, their slots, and the index entry of the cell
// this record takes after it
template <typename T>
struct P2Rec {
    T v0, v1, v2;
    uint32_t s01, s2, nen, ne;
};
template <typename T, typename S>
__device__ __forceinline__ void p2_load(const T* __restrict__ Ke, const S* __restrict__ slots,
                                        uint32_t en, uint32_t e, uint32_t q, P2Rec<T>& r) {
    typedef T t2 __attribute__((ext_vector_type(2)));
    const T* kr = Ke + (uint64_t)(en / 10) * 100 + (en % 10) * 10;
    const t2 v = *reinterpret_cast<const t2*>(kr + 2 * q);
    r.v0 = v.x;
    r.v1 = v.y;
    r.s01 = slot_pair(slots + (uint64_t)e * 10 + 2 * q);
    if (q < 2) {
        r.v2 = kr[8 + q];
        r.s2 = slots[(uint64_t)e * 10 + 8 + q];
    }
}
template <typename T, typename S>
__device__ __forceinline__ void p2_start(const T* __restrict__ Ke, const uint32_t* __restrict__ aent,
                                         const S* __restrict__ slots, uint32_t e,
                                         uint32_t last, uint32_t q, P2Rec<T>& r) {
    const uint32_t ec = min(e, last);
    p2_load(Ke, slots, aent[ec], ec, q, r);
    r.ne = min(e + 3, last);
    r.nen = aent[r.ne];
}
template <typename T, typename S>
__device__ __forceinline__ void p2_step(const T* __restrict__ Ke, const uint32_t* __restrict__ aent,
                                        const S* __restrict__ slots, uint32_t e,
                                        uint32_t last, uint32_t q, uint32_t r, uint32_t k0,
                                        uint32_t nk, T* acc, P2Rec<T>& x) {
    if (!p2_skipped(x.v0)) {
        // slots of this pass's blocks [k0, k0 + nk) only
        const uint32_t ka = (x.s01 & 0xffffu) - k0, kb = (x.s01 >> 16) - k0;  // slot_pair
        if (ka < nk) {
            T* a0 = acc + p2_acc_index(r, ka);
            *a0 = *a0 + x.v0;
        }
        if (kb < nk) {
            T* a1 = acc + p2_acc_index(r, kb);
            *a1 = *a1 + x.v1;
        }
        if (q < 2 && x.s2 - k0 < nk) {
            T* a2 = acc + p2_acc_index(r, x.s2 - k0);
            *a2 = *a2 + x.v2;
        }
    }
    p2_load(Ke, slots, x.nen, x.ne, q, x);
    x.ne = min(e + 6, last);
    x.nen = aent[x.ne];
}

template <typename T, typename S>
__global__ __launch_bounds__(64) void k_gather_p2_rows(
    const T* __restrict__ Ke, const uint32_t* __restrict__ aptr, const uint32_t* __restrict__ aent,
    const S* __restrict__ slots, const uint32_t* __restrict__ soff,
    const uint32_t* __restrict__ rlen, uint32_t nd, uint32_t wp, uint32_t nunits, uint32_t per,
    T* __restrict__ vals, double* __restrict__ vals64) {
    extern __shared__ __align__(16) unsigned char smem_p2[];
    T* acc = reinterpret_cast<T*>(smem_p2);
    // this XCD's units, strided by the XCD's blocks
    for (uint32_t j = blockIdx.x / 8; j < per; j += gridDim.x / 8) {
        const uint32_t u = row_unit(blockIdx.x % 8, j);
        if (u >= nunits) break;
        const uint32_t sl = u / 4, h = u % 4, lane = threadIdx.x;
        const uint32_t q = lane & 3, r = lane >> 2, row = sl * SLICE + h * 16 + r;
        const uint32_t base = soff[sl], w = (soff[sl + 1] - base) / SLICE;
        const bool live = row < nd;
        const uint32_t e0 = live ? aptr[row] : 0, e1 = live ? aptr[row + 1] : 0;
        // the row's writer lane (kq, wr) stores blocks kq, kq + 4, ... of row wr
        const uint32_t wr = lane % 16, wrow = sl * SLICE + h * 16 + wr;
        const uint32_t wlen = wrow < nd ? rlen[wrow] : 0;
        for (uint32_t k0 = 0; k0 < w; k0 += wp) {
            const uint32_t nk = min(wp, w - k0), accn = ((nk + 3) / 4) * 64;
            for (uint32_t i = lane; i < accn; i += SLICE) acc[i] = (T)-0.0;
            __syncthreads();
            if (e0 < e1) {
                // three records in rotation, as in k_asm_elasticity_rows: data
                // three cells and indices six cells ahead of use
                const uint32_t last = e1 - 1;
                P2Rec<T> r0, r1, r2;
                p2_start(Ke, aent, slots, e0, last, q, r0);
                p2_start(Ke, aent, slots, e0 + 1, last, q, r1);
                p2_start(Ke, aent, slots, e0 + 2, last, q, r2);
                uint32_t e = e0;
                for (; e + 2 < e1; e += 3) {
                    p2_step(Ke, aent, slots, e, last, q, r, k0, nk, acc, r0);
                    p2_step(Ke, aent, slots, e + 1, last, q, r, k0, nk, acc, r1);
                    p2_step(Ke, aent, slots, e + 2, last, q, r, k0, nk, acc, r2);
                }
                if (e < e1) p2_step(Ke, aent, slots, e, last, q, r, k0, nk, acc, r0);
                if (e + 1 < e1) p2_step(Ke, aent, slots, e + 1, last, q, r, k0, nk, acc, r1);
            }
            __syncthreads();
            // this pass's blocks; zeros past the row's length
            for (uint32_t k = k0 + lane / 16; k < k0 + nk; k += 4) {
                const uint64_t ix = (uint64_t)base + (uint64_t)k * SLICE + h * 16 + wr;
                const T v = k < wlen ? acc[p2_acc_index(wr, k - k0)] : (T)0;
                vals[ix] = v;
                if (vals64) vals64[ix] = (double)v;
            }
            __syncthreads();  // the next pass or unit clears the accumulator
        }
    }
}

// poisson.rs:80-119 (2D, K=3) / 176-216 (3D, K=4): b_v = sum over incident cells
// (ascending) of sum_q (w_q |det J|) f(x_q) phi_a(xi_q).  phi = basis at points.
template <int D>
__global__ __launch_bounds__(256) void k_rhs_p1(const double* __restrict__ V,
                                                const uint32_t* __restrict__ C,
                                                const uint32_t* __restrict__ aptr,
                                                const uint32_t* __restrict__ aent, uint32_t nv,
                                                const double* __restrict__ fq, int nq,
                                                const double* __restrict__ wq,
                                                const double* __restrict__ phi,
                                                double* __restrict__ b) {
    constexpr int K = D + 1;
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double acc = 0.0;
    for (uint32_t e = aptr[v]; e < aptr[v + 1]; ++e) {
        const uint32_t c = aent[e] / K, a = aent[e] % K;
        double x[K][D];
        uint32_t t[K];
        load_cell<D, K>(V, C, c, x, t);
        double det;
        if constexpr (D == 2) {
            Map2 m;
            if (!affine2(x[0], x[1], x[2], m)) continue;
            det = m.det;
        } else {
            Map3 m;
            if (!affine3(x[0], x[1], x[2], x[K - 1], m)) continue;
            det = m.det;
        }
        const double detj = fabs(det);
        double lb = 0.0;
        for (int q = 0; q < nq; ++q) lb += wq[q] * detj * fq[(uint64_t)c * nq + q] * phi[q * K + a];
        acc += lb;
    }
    b[v] = acc;
}

// P2 load with a per-cell constant source: the same formula with 10 basis functions.
__global__ __launch_bounds__(256) void k_rhs_p2(const double* __restrict__ V,
                                                const uint32_t* __restrict__ C,
                                                const uint32_t* __restrict__ aptr,
                                                const uint32_t* __restrict__ aent, uint32_t nd,
                                                const double* __restrict__ fcell, P2Tab tab,
                                                const double* __restrict__ phi,
                                                double* __restrict__ b) {
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= nd) return;
    double acc = 0.0;
    for (uint32_t e = aptr[d]; e < aptr[d + 1]; ++e) {
        const uint32_t c = aent[e] / 10, a = aent[e] % 10;
        double x[4][3];
        uint32_t t[4];
        load_cell<3, 4>(V, C, c, x, t);
        Map3 m;
        if (!affine3(x[0], x[1], x[2], x[3], m)) continue;
        const double detj = fabs(m.det);
        double lb = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) lb += tab.w[q] * detj * fcell[c] * phi[q * 10 + a];
        acc += lb;
    }
    b[d] = acc;
}

// elasticity.rs:157-182: b(3v+alpha) = sum over incident cells of f_alpha (vol/4).
__global__ __launch_bounds__(256) void k_body_force(const double* __restrict__ V,
                                                    const uint32_t* __restrict__ C,
                                                    const uint32_t* __restrict__ aptr,
                                                    const uint32_t* __restrict__ aent, uint32_t nv,
                                                    double f0, double f1, double f2,
                                                    double* __restrict__ b) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    for (uint32_t e = aptr[v]; e < aptr[v + 1]; ++e) {
        const uint32_t c = aent[e] / 4;
        double x[4][3];
        uint32_t t[4];
        load_cell<3, 4>(V, C, c, x, t);
        Map3 m;
        if (!affine3(x[0], x[1], x[2], x[3], m)) continue;
        const double vol = (1.0 / 6.0) * fabs(m.det);
        acc0 += f0 * (vol * 0.25);
        acc1 += f1 * (vol * 0.25);
        acc2 += f2 * (vol * 0.25);
    }
    b[3 * (uint64_t)v] = acc0;
    b[3 * (uint64_t)v + 1] = acc1;
    b[3 * (uint64_t)v + 2] = acc2;
}

// ---------------------------------------------------------------- Dirichlet
// poisson.rs:231-234 / elasticity.rs:197-200: the last listing of a dof
// decides its value; every listing appends one (d, d, 1.0) triplet.
__global__ void k_bc_mark(const uint32_t* __restrict__ dofs, uint64_t nfix,
                          uint32_t* __restrict__ winner, uint32_t* __restrict__ count) {
    const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (k >= nfix) return;
    atomicMax(&winner[dofs[k]], (uint32_t)(k + 1));
    atomicAdd(&count[dofs[k]], 1u);
}
__global__ void k_bc_values(const uint32_t* __restrict__ winner, const double* __restrict__ vals,
                            uint32_t n, uint8_t* __restrict__ fixed, double* __restrict__ g) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t w = winner[i];
    fixed[i] = w ? 1 : 0;
    g[i] = w ? vals[w - 1] : 0.0;
}

// poisson.rs:242-259 row by row: for a free row r, b_r -= A_rc g_c over the
// fixed columns c in ascending order and those couplings become explicit
// zeros; a fixed row becomes (count) on the diagonal and b_d = g_d.
template <int BS>
__global__ __launch_bounds__(256) void k_condense(const uint32_t* __restrict__ soff,
                                                  const uint32_t* __restrict__ rlen,
                                                  const uint32_t* __restrict__ scols,
                                                  uint32_t nrows, const uint8_t* __restrict__ fixed,
                                                  const double* __restrict__ g,
                                                  const uint32_t* __restrict__ count,
                                                  const double* __restrict__ b_in,
                                                  double* __restrict__ vals,
                                                  double* __restrict__ b_out) {
    constexpr int BS2 = BS * BS;
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= (uint64_t)nrows * BS) return;
    const uint32_t r = t / BS, al = t % BS;
    const uint64_t i = t;  // scalar row
    const RowCtx rc = row_ctx(soff, rlen, r);
    if (fixed[i]) {
        for (uint32_t k = 0; k < rc.len; ++k) {
            const uint32_t cb = scols[(uint64_t)rc.soff + (uint64_t)k * SLICE + rc.lane];
#pragma unroll
            for (int be = 0; be < BS; ++be) {
                const uint64_t col = (uint64_t)cb * BS + be;
                vals[vidx(rc.soff, k, al * BS + be, rc.lane, BS2)] =
                    (col == i) ? (double)count[i] : 0.0;
            }
        }
        b_out[i] = g[i];
        return;
    }
    double acc = b_in[i];
    for (uint32_t k = 0; k < rc.len; ++k) {
        const uint32_t cb = scols[(uint64_t)rc.soff + (uint64_t)k * SLICE + rc.lane];
#pragma unroll
        for (int be = 0; be < BS; ++be) {
            const uint64_t col = (uint64_t)cb * BS + be;
            if (fixed[col]) {
                const uint64_t ix = vidx(rc.soff, k, al * BS + be, rc.lane, BS2);
                acc -= vals[ix] * g[col];
                vals[ix] = 0.0;
            }
        }
    }
    b_out[i] = acc;
}

template <class T>
__global__ void k_to_f32(const double* __restrict__ a, float* __restrict__ b, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

// ---------------------------------------------------------------- host side

static uint64_t scalar_nnz(const Pattern& p, int bs) { return p.nnzb * (uint64_t)bs * bs; }

uint64_t sell_value_count(const fh_matrix* a) {
    return a->pat->sell_entries * (uint64_t)a->bs * a->bs;
}

void ensure_vals32(fh_matrix* a) {
    if (a->vals32.p) return;
    DeviceCtx& d = device(a->dev);
    const uint64_t n = sell_value_count(a);
    a->vals32.alloc(a->dev, std::max<uint64_t>(n, 1));
    const unsigned g = std::min<uint64_t>(cdiv(n, 256), 16384);
    hipLaunchKernelGGL(k_to_f32<float>, dim3(std::max(g, 1u)), dim3(256), 0, d.stream, a->vals.p,
                       a->vals32.p, n);
    FH_LAUNCH_CHECK("to_f32");
}

static fh_matrix* new_matrix(int dev, std::shared_ptr<Pattern> pat, int bs) {
    auto* a = new fh_matrix;
    a->dev = dev;
    a->bs = bs;
    a->pat = std::move(pat);
    a->n = a->pat->nrows * bs;
    a->nnz_scalar = scalar_nnz(*a->pat, bs);
    a->vals.alloc(dev, std::max<uint64_t>(a->pat->sell_entries * bs * bs, 1));
    // SELL padding must hold zeros: the CG product reads it without a length test
    fill_f64(a->vals.p, a->vals.n, 0.0, device(dev).stream);
    return a;
}

static P2Tab p2_table() {
    // lagrange.rs:324-346 gradients at the degree-2 points (host arithmetic,
    // identical to the oracle's table)
    P2Tab t{};
    const double s5 = std::sqrt(5.0);
    const double a = (5.0 - s5) / 20.0, b = (5.0 + 3.0 * s5) / 20.0;
    const double P[4][3] = {{a, a, a}, {b, a, a}, {a, b, a}, {a, a, b}};
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
        for (int i = 0; i < 10; ++i)
            for (int k = 0; k < 3; ++k) t.G[q][i][k] = G[i][k];
        t.w[q] = 1.0 / 24.0;
    }
    return t;
}

// calls f with the adjacency's slot map (u8, or u16 for rows past 255 blocks)
template <class F>
static void with_slots(const Adjacency& adj, F&& f) {
    if (adj.slots16.p) f(adj.slots16.p);
    else f(adj.slots.p);
}
#define FH_SLOT_T(sl) std::remove_const_t<std::remove_pointer_t<decltype(sl)>>

// Blocks of the row kernels (one wave each, looping over the units of
// their XCD): FH_ASM_WAVES_PER_CU (default 8) per compute unit, at most one
// per unit.  A multiple of 8, so block b stays on XCD b % 8.
static uint32_t rows_grid(int dev, uint32_t per) {
    static int wpc = [] {
        const char* e = std::getenv("FH_ASM_WAVES_PER_CU");
        return e ? std::max(1, std::atoi(e)) : 8;
    }();
    const uint32_t cus = (uint32_t)std::max(8, device(dev).prop.multiProcessorCount);
    return 8 * std::max(1u, std::min(per, cus / 8 * (uint32_t)wpc));
}

void p2_gather(fh_mesh* m, const void* Ke, bool f64, void* vals, double* vals64, hipStream_t s) {
    MeshP2& p2 = mesh_p2(m);
    const Pattern& pat = *p2.pat;
    const uint32_t nunits = pat.nslices * 4, per = row_units_per_xcd(nunits);
    if (!nunits) return;
    const uint32_t wp = std::max(1u, std::min(pat.max_width, 512u));
    const size_t accn = (size_t)((wp + 3) / 4) * 64;
    with_slots(*p2.adj, [&](auto* sl) {
        using S = FH_SLOT_T(sl);
        if (f64)
            hipLaunchKernelGGL((k_gather_p2_rows<double, S>), dim3(rows_grid(m->dev, per)),
                               dim3(SLICE), accn * sizeof(double), s, (const double*)Ke,
                               p2.adj->ptr.p, p2.adj->ent.p, sl, pat.slice_off.p, pat.row_len.p,
                               p2.adj->ndofs, wp, nunits, per, (double*)vals, nullptr);
        else
            hipLaunchKernelGGL((k_gather_p2_rows<float, S>), dim3(rows_grid(m->dev, per)),
                               dim3(SLICE), accn * sizeof(float), s, (const float*)Ke,
                               p2.adj->ptr.p, p2.adj->ent.p, sl, pat.slice_off.p, pat.row_len.p,
                               p2.adj->ndofs, wp, nunits, per, (float*)vals, vals64);
    });
    FH_LAUNCH_CHECK("gather_p2_rows");
}

// P2 basis values at the degree-2 points, [q][10] (lagrange.rs:301-320)
static std::vector<double> p2_basis_table() {
    const double s5 = std::sqrt(5.0);
    const double a = (5.0 - s5) / 20.0, b = (5.0 + 3.0 * s5) / 20.0;
    const double P[4][3] = {{a, a, a}, {b, a, a}, {a, b, a}, {a, a, b}};
    std::vector<double> phi(40);
    for (int q = 0; q < 4; ++q) {
        const double x = P[q][0], y = P[q][1], z = P[q][2];
        const double l0 = 1.0 - x - y - z, l1 = x, l2 = y, l3 = z;
        const double v[10] = {l0 * (2.0 * l0 - 1.0), l1 * (2.0 * l1 - 1.0), l2 * (2.0 * l2 - 1.0),
                              l3 * (2.0 * l3 - 1.0), 4.0 * l0 * l1, 4.0 * l1 * l2,
                              4.0 * l0 * l2, 4.0 * l0 * l3, 4.0 * l1 * l3, 4.0 * l2 * l3};
        for (int i = 0; i < 10; ++i) phi[q * 10 + i] = v[i];
    }
    return phi;
}

// P1 basis values [q][K] at the degree-2 points and the weights.
struct P1Tab {
    std::vector<double> w, phi;
    int nq;
};
static P1Tab p1_table(int dim) {
    const QuadTab q = quad_table(dim, 2);
    P1Tab t;
    t.nq = q.n;
    const int K = dim + 1;
    for (int i = 0; i < q.n; ++i) {
        t.w.push_back(q.w[i]);
        const double* x = q.pts[i];
        if (dim == 2) {
            t.phi.push_back(1.0 - x[0] - x[1]);
            t.phi.push_back(x[0]);
            t.phi.push_back(x[1]);
        } else {
            t.phi.push_back(1.0 - x[0] - x[1] - x[2]);
            t.phi.push_back(x[0]);
            t.phi.push_back(x[1]);
            t.phi.push_back(x[2]);
        }
        (void)K;
    }
    return t;
}

// per-cell records of 16 doubles (one 128-byte line): g (12), vol, and
// lambda, mu with a per-cell E (k_elas_cell_geom)
void elas_cell_records(fh_mesh* m, double nu, const double* Ecell_dev, double* G, hipStream_t s) {
    if (!m->nc) return;
    if (Ecell_dev)
        hipLaunchKernelGGL((k_elas_cell_geom<true>), dim3(cdiv(m->nc, 256)), dim3(256), 0, s,
                           m->verts.p, m->cells.p, m->nc, nu, Ecell_dev, G);
    else
        hipLaunchKernelGGL((k_elas_cell_geom<false>), dim3(cdiv(m->nc, 256)), dim3(256), 0, s,
                           m->verts.p, m->cells.p, m->nc, nu, Ecell_dev, G);
    FH_LAUNCH_CHECK("elas_cell_geom");
}

// blocks per row held in LDS at once by the elasticity row kernels (48 blocks:
// 58.8 KB of doubles); wider rows take passes
static uint32_t elas_pass_blocks(const Pattern& p) { return std::max(1u, std::min(p.max_width, 48u)); }

void elas_gather_f32(fh_mesh* m, const Pattern& p, const float* Ke8, const float* Ke9,
                     float* vals32, double* vals64, hipStream_t s) {
    Adjacency& adj = mesh_adjacency(m);
    const uint32_t wp = elas_pass_blocks(p);
    const uint32_t nunits = p.nslices * (SLICE / ELAS_RPW);
    const uint32_t per = row_units_per_xcd(nunits);
    if (!nunits) return;
    with_slots(adj, [&](auto* sl) {
        hipLaunchKernelGGL((k_gather_elas_rows<FH_SLOT_T(sl)>), dim3(rows_grid(m->dev, per)),
                           dim3(SLICE), (size_t)elas_acc_size(wp) * sizeof(float), s, Ke8, Ke9,
                           adj.ptr.p, adj.ent.p, sl, p.slice_off.p, p.row_len.p, m->nv, wp,
                           nunits, per, vals32, vals64);
    });
    FH_LAUNCH_CHECK("gather_elas_rows");
}

void assemble_elasticity_into(fh_mesh* m, double E, double nu, const double* Ecell_dev,
                              fh_matrix* a, hipStream_t s) {
    Adjacency& adj = mesh_adjacency(m);
    const Pattern& p = *a->pat;
    const bool ec = Ecell_dev != nullptr;
    DBuf<double> geom(a->dev, (size_t)m->nc * ELAS_GS);
    elas_cell_records(m, nu, Ecell_dev, geom.p, s);
    const uint32_t wp = elas_pass_blocks(p);
    const size_t lds = (size_t)elas_acc_size(wp) * sizeof(double);
    const uint32_t nunits = p.nslices * (SLICE / ELAS_RPW);
    const uint32_t per = row_units_per_xcd(nunits);
    if (!nunits) return;
    const dim3 grid(rows_grid(a->dev, per));
    with_slots(adj, [&](auto* sl) {
        using S = FH_SLOT_T(sl);
        if (ec)
            hipLaunchKernelGGL((k_asm_elasticity_rows<true, S>), grid, dim3(SLICE), lds, s, geom.p,
                               adj.ptr.p, adj.ent.p, sl, p.slice_off.p, p.row_len.p, m->nv, 0.0,
                               0.0, wp, nunits, per, a->vals.p);
        else
            hipLaunchKernelGGL((k_asm_elasticity_rows<false, S>), grid, dim3(SLICE), lds, s,
                               geom.p, adj.ptr.p, adj.ent.p, sl, p.slice_off.p, p.row_len.p, m->nv,
                               lame_lambda(E, nu), lame_mu(E, nu), wp, nunits, per, a->vals.p);
    });
    FH_LAUNCH_CHECK("asm_elasticity");
    // geom goes back to the allocator's cache; the next user of that memory is
    // ordered after these kernels on the same stream
}

fh_matrix* assemble_elasticity(fh_mesh* m, double E, double nu, const double* Ecell_dev) {
    FH_CHECK(m->dim == 3, "elasticity needs a tetrahedral mesh");
    FH_CHECK(nu > -1.0 && nu < 0.5, "Poisson's ratio must be in (-1.0, 0.5) for physical stability");
    DeviceCtx& d = device(m->dev);
    auto pat = mesh_pattern(m);
    std::unique_ptr<fh_matrix> a(new_matrix(m->dev, pat, 3));
    assemble_elasticity_into(m, E, nu, Ecell_dev, a.get(), d.stream);
    a->skipped_cells = mesh_degenerate_cells(m);
    return a.release();
}

fh_matrix* assemble_elasticity_mass(fh_mesh* m, double density) {
    FH_CHECK(m->dim == 3, "the elasticity mass needs a tetrahedral mesh");
    DeviceCtx& d = device(m->dev);
    auto pat = mesh_pattern(m);
    Adjacency& adj = mesh_adjacency(m);
    std::unique_ptr<fh_matrix> a(new_matrix(m->dev, pat, 3));
    with_slots(adj, [&](auto* sl) {
        hipLaunchKernelGGL((k_asm_elasticity_mass<FH_SLOT_T(sl)>), dim3(cdiv(m->nv, 256)),
                           dim3(256), 0, d.stream, m->verts.p, m->cells.p, adj.ptr.p, adj.ent.p,
                           sl, pat->slice_off.p, pat->row_len.p, m->nv, density, a->vals.p);
    });
    FH_LAUNCH_CHECK("asm_elasticity_mass");
    FH_HIP(hipStreamSynchronize(d.stream));
    a->comp_diag = true;
    a->nnz_scalar = a->pat->nnzb * 3;  // one entry per component of each block
    a->skipped_cells = mesh_degenerate_cells(m);
    return a.release();
}

// mass.rs:190-250: int phi_a phi_b over the reference tetrahedron, from the
// barycentric moments prod m_i! / (sum m_i + 3)!; corner node i is
// lambda_i (2 lambda_i - 1), edge node (i, j) is 4 lambda_i lambda_j.  Host
// arithmetic in the reference's term order.
static P2MassTab p2_mass_table() {
    static const double F[12] = {1.0,   1.0,    2.0,     6.0,      24.0,      120.0,
                                 720.0, 5040.0, 40320.0, 362880.0, 3628800.0, 39916800.0};
    struct Term {
        int e[4];
        double c;
    };
    std::vector<std::vector<Term>> T(10);
    for (int i = 0; i < 4; ++i) {
        Term sq{{0, 0, 0, 0}, 2.0}, lin{{0, 0, 0, 0}, -1.0};
        sq.e[i] = 2;
        lin.e[i] = 1;
        T[i] = {sq, lin};
    }
    const int E[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
    for (int k = 0; k < 6; ++k) {
        Term t{{0, 0, 0, 0}, 4.0};
        t.e[E[k][0]] = 1;
        t.e[E[k][1]] = 1;
        T[4 + k] = {t};
    }
    P2MassTab tab{};
    for (int a = 0; a < 10; ++a)
        for (int b = 0; b < 10; ++b) {
            double acc = 0.0;
            for (const Term& ta : T[a])
                for (const Term& tb : T[b]) {
                    int m[4], deg = 0;
                    for (int q = 0; q < 4; ++q) {
                        m[q] = ta.e[q] + tb.e[q];
                        deg += m[q];
                    }
                    acc += ta.c * tb.c * F[m[0]] * F[m[1]] * F[m[2]] * F[m[3]] / F[deg + 3];
                }
            tab.m[a * 10 + b] = acc;
        }
    return tab;
}

fh_matrix* assemble_elasticity_mass_p2(int dev, uint64_t n_nodes, const double* nodes,
                                       uint64_t nc, const uint64_t* cells, double density) {
    FH_CHECK(n_nodes < 0x3fffffffull && nc * 10 < 0xffffffffull, "mesh too large");
    FH_CHECK(nc == 0 || (nodes && cells), "null mesh arrays");
    std::vector<uint32_t> c32(nc * 10);
    for (uint64_t i = 0; i < nc * 10; ++i) {
        FH_CHECK(cells[i] < n_nodes, "cell node index out of range");
        c32[i] = (uint32_t)cells[i];
    }
    DeviceCtx& d = device(dev);
    hipStream_t s = d.stream;
    DBuf<double> V(dev, std::max<uint64_t>(3 * n_nodes, 1));
    DBuf<uint32_t> C(dev, std::max<uint64_t>(nc * 10, 1));
    h2d(V.p, nodes, 3 * n_nodes, s);
    h2d(C.p, c32.data(), nc * 10, s);
    auto adj = build_adjacency(dev, C.p, (uint32_t)nc, 10, (uint32_t)n_nodes, s);
    auto pat = build_pattern(dev, C.p, 10, *adj, s);
    build_slots(C.p, 10, *adj, *pat, s);
    std::unique_ptr<fh_matrix> a(new_matrix(dev, pat, 3));
    if (n_nodes)
        with_slots(*adj, [&](auto* sl) {
            hipLaunchKernelGGL((k_asm_p2_vec_mass<FH_SLOT_T(sl)>), dim3(cdiv(n_nodes, 256)),
                               dim3(256), 0, s, V.p, C.p, adj->ptr.p, adj->ent.p, sl,
                               pat->slice_off.p, pat->row_len.p, (uint32_t)n_nodes, density,
                               p2_mass_table(), a->vals.p);
        });
    FH_LAUNCH_CHECK("asm_p2_vec_mass");
    FH_HIP(hipStreamSynchronize(s));
    a->comp_diag = true;
    a->nnz_scalar = pat->nnzb * 3;
    return a.release();
}

// Condense a copy of `a` (values shared pattern) and lift b (device).
fh_matrix* apply_dirichlet(const fh_matrix* a, const double* b_in, uint64_t nfix,
                           const uint64_t* dofs, const double* vals, double* b_out,
                           hipStream_t s, int fold_diag) {
    const int dev = a->dev;
    for (uint64_t k = 0; k < nfix; ++k)
        FH_CHECK(dofs[k] < a->n, "Dirichlet dof out of range");
    std::vector<uint32_t> d32(nfix);
    for (uint64_t k = 0; k < nfix; ++k) d32[k] = (uint32_t)dofs[k];
    DBuf<uint32_t> ddofs(dev, std::max<uint64_t>(nfix, 1)), winner(dev, a->n);
    DBuf<double> dvals(dev, std::max<uint64_t>(nfix, 1)), g(dev, a->n);
    h2d(ddofs.p, d32.data(), nfix, s);
    h2d(dvals.p, vals, nfix, s);
    std::unique_ptr<fh_matrix> o(new fh_matrix);
    o->dev = dev;
    o->bs = a->bs;
    o->pat = a->pat;
    o->n = a->n;
    o->nnz_scalar = 0;
    o->condensed = true;
    o->comp_diag = a->comp_diag;
    o->skipped_cells = a->skipped_cells;
    o->vals.alloc(dev, a->vals.n);
    d2d(o->vals.p, a->vals.p, a->vals.n, s);
    o->fixed.alloc(dev, a->n);
    o->fixed_count.alloc(dev, a->n);
    fill_u32(winner.p, a->n, 0, s);
    fill_u32(o->fixed_count.p, a->n, 0, s);
    if (nfix)
        hipLaunchKernelGGL(k_bc_mark, dim3(cdiv(nfix, 256)), dim3(256), 0, s, ddofs.p, nfix,
                           winner.p, o->fixed_count.p);
    hipLaunchKernelGGL(k_bc_values, dim3(cdiv(a->n, 256)), dim3(256), 0, s, winner.p, dvals.p,
                       a->n, o->fixed.p, g.p);
    if (fold_diag >= 0) {  // Dirichlet fold: fixed diagonal = fold_diag, no multiplicity
        hipLaunchKernelGGL(k_fold_count, dim3(cdiv(a->n, 256)), dim3(256), 0, s,
                           o->fixed_count.p, a->n, (uint32_t)fold_diag, o->fixed_count.p);
        o->folded = true;
    }
    const Pattern& p = *a->pat;
    const uint64_t nt = (uint64_t)p.nrows * a->bs;
    if (a->bs == 1)
        hipLaunchKernelGGL(k_condense<1>, dim3(cdiv(nt, 256)), dim3(256), 0, s, p.slice_off.p,
                           p.row_len.p, p.sell_cols.p, p.nrows, o->fixed.p, g.p,
                           o->fixed_count.p, b_in, o->vals.p, b_out);
    else
        hipLaunchKernelGGL(k_condense<3>, dim3(cdiv(nt, 256)), dim3(256), 0, s, p.slice_off.p,
                           p.row_len.p, p.sell_cols.p, p.nrows, o->fixed.p, g.p,
                           o->fixed_count.p, b_in, o->vals.p, b_out);
    FH_LAUNCH_CHECK("condense");
    FH_HIP(hipStreamSynchronize(s));
    // scalar nnz of the condensed CSC: both-free couplings + one per fixed dof
    return o.release();
}

// Host export to faer's canonical CSC raw parts (optionally dropping the
// condensed couplings).  Not on the hot path.
HostCsr export_scalar_csr(const fh_matrix* a) {
    DeviceCtx& d = device(a->dev);
    const Pattern& p = *a->pat;
    const int bs = a->bs, bs2 = bs * bs;
    std::vector<uint32_t> rptr(p.nrows + 1), cols(p.nnzb), soff(p.nslices + 1);
    std::vector<double> vals(sell_value_count(a));
    std::vector<uint8_t> fixed;
    std::vector<uint32_t> cnt;
    d2h(rptr.data(), p.row_ptr.p, rptr.size(), d.stream);
    d2h(cols.data(), p.cols.p, cols.size(), d.stream);
    d2h(soff.data(), p.slice_off.p, soff.size(), d.stream);
    d2h(vals.data(), a->vals.p, vals.size(), d.stream);
    if (a->condensed) {
        fixed.resize(a->n);
        d2h(fixed.data(), a->fixed.p, a->n, d.stream);
    }
    FH_HIP(hipStreamSynchronize(d.stream));
    HostCsr h;
    h.rp.assign((size_t)a->n + 1, 0);
    h.ci.reserve(p.nnzb * bs2);
    h.v.reserve(p.nnzb * bs2);
    for (uint32_t r = 0; r < p.nrows; ++r) {
        const uint32_t s = r / SLICE, l = r % SLICE, len = rptr[r + 1] - rptr[r];
        for (int al = 0; al < bs; ++al) {
            const uint64_t i = (uint64_t)r * bs + al;
            for (uint32_t k = 0; k < len; ++k) {
                const uint32_t cb = cols[rptr[r] + k];
                for (int be = 0; be < bs; ++be) {
                    if (a->comp_diag && be != al) continue;  // mass.rs:115-118
                    const uint64_t c = (uint64_t)cb * bs + be;
                    if (a->condensed && (fixed[i] || fixed[c]) && i != c) continue;
                    const uint64_t ix =
                        vpos((uint64_t)soff[s] + (uint64_t)k * SLICE, al * bs + be, l, bs2);
                    if (a->folded && vals[ix] == 0.0) continue;  // eigen.rs:68
                    h.ci.push_back((uint32_t)c);
                    h.v.push_back(vals[ix]);
                }
            }
            h.rp[i + 1] = h.ci.size();
        }
    }
    return h;
}

// cells whose affine map the assembly rejects (jacobian.rs:29-31), counted
// with the same geometry routines the assembly kernels use
template <int D>
__global__ __launch_bounds__(256) void k_count_degenerate(const double* __restrict__ V,
                                                          const uint32_t* __restrict__ C,
                                                          uint32_t nc, int K,
                                                          uint32_t* __restrict__ count) {
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    bool bad = false;
    if (c < nc) {
        double x[4][3];
#pragma unroll
        for (int a = 0; a < D + 1; ++a)
#pragma unroll
            for (int q = 0; q < D; ++q) x[a][q] = V[(uint64_t)C[(uint64_t)c * K + a] * D + q];
        if constexpr (D == 2) {
            Map2 m;
            bad = !affine2(x[0], x[1], x[2], m);
        } else {
            Map3 m;
            bad = !affine3(x[0], x[1], x[2], x[3], m);
        }
    }
    const uint64_t mask = __ballot(bad);
    if ((threadIdx.x & 63) == 0 && mask) atomicAdd(count, (uint32_t)__popcll(mask));
}

uint64_t mesh_degenerate_cells(fh_mesh* m) {
    if (m->n_degenerate >= 0) return (uint64_t)m->n_degenerate;
    hipStream_t s = device(m->dev).stream;
    DBuf<uint32_t> cnt(m->dev, 1);
    fill_u32(cnt.p, 1, 0u, s);
    if (m->nc) {
        if (m->dim == 2)
            hipLaunchKernelGGL(k_count_degenerate<2>, dim3(cdiv(m->nc, 256)), dim3(256), 0, s,
                               m->verts.p, m->cells.p, m->nc, m->k, cnt.p);
        else
            hipLaunchKernelGGL(k_count_degenerate<3>, dim3(cdiv(m->nc, 256)), dim3(256), 0, s,
                               m->verts.p, m->cells.p, m->nc, m->k, cnt.p);
        FH_LAUNCH_CHECK("count_degenerate");
    }
    uint32_t h = 0;
    d2h(&h, cnt.p, 1, s);
    FH_HIP(hipStreamSynchronize(s));
    m->n_degenerate = h;
    return h;
}

// kind 0: stiffness (poisson.rs:23-69 / 124-165), kind 1: mass (mass.rs:10-82)
fh_matrix* assemble_p1(fh_mesh* m, int kind) {
    hipStream_t s = device(m->dev).stream;
    auto pat = mesh_pattern(m);
    Adjacency& adj = mesh_adjacency(m);
    std::unique_ptr<fh_matrix> a(new_matrix(m->dev, pat, 1));
    const dim3 g(cdiv(m->nv, 256)), b(256);
    with_slots(adj, [&](auto* sl) {
        using S = FH_SLOT_T(sl);
        auto go = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, g, b, 0, s, m->verts.p, m->cells.p, adj.ptr.p, adj.ent.p,
                               sl, pat->slice_off.p, pat->row_len.p, m->nv, a->vals.p);
        };
        if (m->dim == 2) {
            if (kind == 0) go(k_asm_p1_tri<0, S>);
            else go(k_asm_p1_tri<1, S>);
        } else {
            if (kind == 0) go(k_asm_p1_tet<0, S>);
            else go(k_asm_p1_tet<1, S>);
        }
    });
    FH_LAUNCH_CHECK("asm_p1");
    a->skipped_cells = mesh_degenerate_cells(m);
    return a.release();
}

// f sampled at the degree-2 points (device, [nc][nq]) -> b (device, [nv])
void rhs_p1_dev(fh_mesh* m, const double* fq, double* b, hipStream_t s) {
    Adjacency& adj = mesh_adjacency(m);
    const P1Tab t = p1_table(m->dim);
    DBuf<double> dw(m->dev, t.w.size()), dphi(m->dev, t.phi.size());
    h2d(dw.p, t.w.data(), t.w.size(), s);
    h2d(dphi.p, t.phi.data(), t.phi.size(), s);
    if (m->dim == 2)
        hipLaunchKernelGGL(k_rhs_p1<2>, dim3(cdiv(m->nv, 256)), dim3(256), 0, s, m->verts.p,
                           m->cells.p, adj.ptr.p, adj.ent.p, m->nv, fq, t.nq, dw.p, dphi.p, b);
    else
        hipLaunchKernelGGL(k_rhs_p1<3>, dim3(cdiv(m->nv, 256)), dim3(256), 0, s, m->verts.p,
                           m->cells.p, adj.ptr.p, adj.ent.p, m->nv, fq, t.nq, dw.p, dphi.p, b);
    FH_LAUNCH_CHECK("rhs_p1");
    FH_HIP(hipStreamSynchronize(s));  // the small tables above are freed on return
}

void body_force_dev(fh_mesh* m, const double f[3], double* b, hipStream_t s) {
    FH_CHECK(m->dim == 3, "body force needs a tetrahedral mesh");
    Adjacency& adj = mesh_adjacency(m);
    hipLaunchKernelGGL(k_body_force, dim3(cdiv(m->nv, 256)), dim3(256), 0, s, m->verts.p,
                       m->cells.p, adj.ptr.p, adj.ent.p, m->nv, f[0], f[1], f[2], b);
    FH_LAUNCH_CHECK("body_force");
}

}  // namespace fh

using namespace fh;

static hipStream_t stream_of(int dev) { return device(dev).stream; }

extern "C" {

int fh_assemble_stiffness(fh_mesh* m, fh_matrix** out) {
    return guard([&] {
        FH_CHECK(m && out, "null argument");
        *out = assemble_p1(m, 0);
        FH_HIP(hipStreamSynchronize(stream_of(m->dev)));
    });
}

int fh_assemble_mass(fh_mesh* m, fh_matrix** out) {
    return guard([&] {
        FH_CHECK(m && out, "null argument");
        *out = assemble_p1(m, 1);
        FH_HIP(hipStreamSynchronize(stream_of(m->dev)));
    });
}

int fh_assemble_elasticity_mass_p2(uint64_t n_nodes, const double* nodes, uint64_t nc,
                                   const uint64_t* cells, double density, int device_idx,
                                   fh_matrix** out) {
    return guard([&] {
        FH_CHECK(out, "null argument");
        *out = assemble_elasticity_mass_p2(resolve_device(device_idx), n_nodes, nodes, nc, cells,
                                           density);
    });
}

int fh_l2_error(fh_mesh* m, const double* u, const double* u_exact_q, double* out) {
    return guard([&] {
        FH_CHECK(m && u && u_exact_q && out, "null argument");
        DeviceCtx& d = device(m->dev);
        hipStream_t s = d.stream;
        const QuadTab q = quad_table(m->dim, 3);
        const unsigned g = std::max(1u, cdiv(m->nc, 256));
        DBuf<double> du(m->dev, std::max<uint32_t>(m->nv, 1)),
            dq(m->dev, std::max<uint64_t>((uint64_t)m->nc * q.n, 1)), part(m->dev, g);
        DBuf<uint32_t> bad(m->dev, 1);
        fill_u32(bad.p, 1, 0, s);
        h2d(du.p, u, m->nv, s);
        h2d(dq.p, u_exact_q, (uint64_t)m->nc * q.n, s);
        if (m->dim == 2)
            hipLaunchKernelGGL(k_l2_err<2>, dim3(g), dim3(256), 0, s, m->verts.p, m->cells.p, m->nc,
                               du.p, dq.p, q, part.p, bad.p);
        else
            hipLaunchKernelGGL(k_l2_err<3>, dim3(g), dim3(256), 0, s, m->verts.p, m->cells.p, m->nc,
                               du.p, dq.p, q, part.p, bad.p);
        FH_LAUNCH_CHECK("l2_err");
        std::vector<double> h(g);
        uint32_t hb = 0;
        d2h(h.data(), part.p, g, s);
        d2h(&hb, bad.p, 1, s);
        FH_HIP(hipStreamSynchronize(s));
        FH_CHECK(!hb, "degenerate cell in the L2 error integral");
        double tot = 0.0;
        for (double v : h) tot += v;
        *out = std::sqrt(tot);
    });
}

int fh_assemble_elasticity_dev(fh_mesh* m, double E, double nu, const double* E_cell_dev,
                               fh_matrix** out) {
    return guard([&] {
        FH_CHECK(m && out, "null argument");
        *out = assemble_elasticity(m, E, nu, E_cell_dev);
        FH_HIP(hipStreamSynchronize(stream_of(m->dev)));
    });
}

int fh_assemble_elasticity(fh_mesh* m, double E, double nu, const double* E_cell,
                           fh_matrix** out) {
    return guard([&] {
        FH_CHECK(m && out, "null argument");
        hipStream_t s = stream_of(m->dev);
        DBuf<double> ec;
        if (E_cell) {
            ec.alloc(m->dev, m->nc);
            h2d(ec.p, E_cell, m->nc, s);
        }
        *out = assemble_elasticity(m, E, nu, ec.p);
        FH_HIP(hipStreamSynchronize(s));
    });
}

int fh_assemble_elasticity_method(fh_mesh* m, double E, double nu, const double* E_cell,
                                  int method, fh_matrix** out) {
    return guard([&] {
        FH_CHECK(m && out, "null argument");
        FH_CHECK(method >= 0 && method <= 2,
                 "elasticity method must be 0 (VALU FP64), 1 (MFMA FP32) "
                 "or 2 (both: FP64 values + MFMA FP32 copy)");
        FH_CHECK(m->dim == 3, "elasticity needs a tetrahedral mesh");
        FH_CHECK(nu > -1.0 && nu < 0.5,
                 "Poisson's ratio must be in (-1.0, 0.5) for physical stability");
        hipStream_t s = stream_of(m->dev);
        DBuf<double> ec;
        if (E_cell) {
            ec.alloc(m->dev, m->nc);
            h2d(ec.p, E_cell, m->nc, s);
        }
        std::unique_ptr<fh_matrix> a;
        if (method == 1) {
            a.reset(new_matrix(m->dev, mesh_pattern(m), 3));
            a->skipped_cells = mesh_degenerate_cells(m);
        } else {
            a.reset(assemble_elasticity(m, E, nu, ec.p));
        }
        if (method != 0) assemble_elasticity_mfma(m, E, nu, ec.p, a.get(), method == 1, s);
        FH_HIP(hipStreamSynchronize(s));
        *out = a.release();
    });
}

int fh_assemble_rhs(fh_mesh* m, const double* fq, double* b_out) {
    return guard([&] {
        FH_CHECK(m && fq && b_out, "null argument");
        hipStream_t s = stream_of(m->dev);
        const int nq = m->dim == 2 ? 3 : 4;
        DBuf<double> dfq(m->dev, (size_t)m->nc * nq), db(m->dev, m->nv);
        h2d(dfq.p, fq, dfq.n, s);
        rhs_p1_dev(m, dfq.p, db.p, s);
        d2h(b_out, db.p, m->nv, s);
        FH_HIP(hipStreamSynchronize(s));
    });
}

int fh_assemble_body_force_dev(fh_mesh* m, const double f[3], double* b_dev) {
    return guard([&] {
        FH_CHECK(m && f && b_dev, "null argument");
        hipStream_t s = stream_of(m->dev);
        body_force_dev(m, f, b_dev, s);
        FH_HIP(hipStreamSynchronize(s));
    });
}

int fh_assemble_body_force(fh_mesh* m, const double f[3], double* b_out) {
    return guard([&] {
        FH_CHECK(m && f && b_out, "null argument");
        hipStream_t s = stream_of(m->dev);
        DBuf<double> db(m->dev, 3 * (size_t)m->nv);
        if (fh_assemble_body_force_dev(m, f, db.p) != FH_OK) throw Error(FH_ERR_HIP, fh_last_error());
        d2h(b_out, db.p, db.n, s);
        FH_HIP(hipStreamSynchronize(s));
    });
}

int fh_assemble_p2_stiffness(fh_mesh* m, int method, fh_matrix** out) {
    return guard([&] {
        FH_CHECK(m && out, "null argument");
        hipStream_t s = stream_of(m->dev);
        MeshP2& p2 = mesh_p2(m);
        std::unique_ptr<fh_matrix> a(new_matrix(m->dev, p2.pat, 1));
        FH_CHECK(method >= 0 && method <= 2, "P2 method must be 0 (VALU FP64), 1 (MFMA FP32) "
                                             "or 2 (both: FP64 values + MFMA FP32 copy)");
        if (method != 1) {
            DBuf<double> Ke(m->dev, (size_t)m->nc * 100);
            if (m->nc) {
                hipLaunchKernelGGL(k_p2_elem_f64, dim3(cdiv(m->nc, 64)), dim3(64), 0, s,
                                   m->verts.p, m->cells.p, m->nc, p2_table(), Ke.p);
                FH_LAUNCH_CHECK("p2_elem_f64");
            }
            p2_gather(m, Ke.p, true, a->vals.p, nullptr, s);
        }
        if (method != 0) assemble_p2_mfma(m, a.get(), method == 1, s);
        FH_HIP(hipStreamSynchronize(s));
        a->skipped_cells = mesh_degenerate_cells(m);
        *out = a.release();
    });
}

int fh_p2_info(fh_mesh* m, uint64_t* ndofs, uint64_t* nedges) {
    return guard([&] {
        FH_CHECK(m, "null mesh");
        MeshP2& p2 = mesh_p2(m);
        if (ndofs) *ndofs = p2.adj->ndofs;
        if (nedges) *nedges = p2.nedges;
    });
}

int fh_p2_cell_dofs(fh_mesh* m, uint64_t* out) {
    return guard([&] {
        FH_CHECK(m && out, "null argument");
        hipStream_t s = stream_of(m->dev);
        MeshP2& p2 = mesh_p2(m);
        std::vector<uint32_t> h((size_t)m->nc * 10);
        d2h(h.data(), p2.cell_dofs.p, h.size(), s);
        FH_HIP(hipStreamSynchronize(s));
        for (size_t i = 0; i < h.size(); ++i) out[i] = h[i];
    });
}

int fh_assemble_p2_rhs_cellconst(fh_mesh* m, const double* f_cell, double* b_out) {
    return guard([&] {
        FH_CHECK(m && f_cell && b_out, "null argument");
        hipStream_t s = stream_of(m->dev);
        MeshP2& p2 = mesh_p2(m);
        const std::vector<double> phi = p2_basis_table();
        DBuf<double> df(m->dev, m->nc), dphi(m->dev, 40), db(m->dev, p2.adj->ndofs);
        h2d(df.p, f_cell, m->nc, s);
        h2d(dphi.p, phi.data(), 40, s);
        hipLaunchKernelGGL(k_rhs_p2, dim3(cdiv(p2.adj->ndofs, 256)), dim3(256), 0, s, m->verts.p,
                           m->cells.p, p2.adj->ptr.p, p2.adj->ent.p, p2.adj->ndofs, df.p,
                           p2_table(), dphi.p, db.p);
        FH_LAUNCH_CHECK("rhs_p2");
        d2h(b_out, db.p, db.n, s);
        FH_HIP(hipStreamSynchronize(s));
    });
}

int fh_apply_dirichlet_dev(const fh_matrix* a, const double* b_in_dev, uint64_t nfix,
                           const uint64_t* dofs, const double* vals, fh_matrix** out,
                           double* b_out_dev) {
    return guard([&] {
        FH_CHECK(a && b_in_dev && out && b_out_dev, "null argument");
        FH_CHECK(nfix == 0 || (dofs && vals), "null Dirichlet list");
        *out = apply_dirichlet(a, b_in_dev, nfix, dofs, vals, b_out_dev, stream_of(a->dev));
    });
}

int fh_apply_dirichlet(const fh_matrix* a, const double* b_in, uint64_t nfix,
                       const uint64_t* dofs, const double* vals, fh_matrix** out,
                       double* b_out) {
    return guard([&] {
        FH_CHECK(a && b_in && out && b_out, "null argument");
        FH_CHECK(nfix == 0 || (dofs && vals), "null Dirichlet list");
        hipStream_t s = stream_of(a->dev);
        DBuf<double> bi(a->dev, a->n), bo(a->dev, a->n);
        h2d(bi.p, b_in, a->n, s);
        *out = apply_dirichlet(a, bi.p, nfix, dofs, vals, bo.p, s);
        d2h(b_out, bo.p, a->n, s);
        FH_HIP(hipStreamSynchronize(s));
    });
}

int fh_matrix_skipped_cells(const fh_matrix* a, uint64_t* count) {
    return guard([&] {
        FH_CHECK(a && count, "null argument");
        *count = a->skipped_cells;
    });
}

int fh_matrix_info(const fh_matrix* a, uint64_t* n, uint64_t* nnz, int* block) {
    return guard([&] {
        FH_CHECK(a, "null matrix");
        if (n) *n = a->n;
        if (block) *block = a->bs;
        if (nnz) {
            if (!a->condensed) {
                *nnz = a->nnz_scalar;
            } else {
                const HostCsr h = export_scalar_csr(a);
                *nnz = h.ci.size();
                const_cast<fh_matrix*>(a)->nnz_scalar = h.ci.size();
            }
        }
    });
}

int fh_matrix_export_csc(const fh_matrix* a, uint64_t* col_ptr, uint64_t* row_idx, double* vals) {
    return guard([&] {
        FH_CHECK(a && col_ptr, "null argument");
        const HostCsr h = export_scalar_csr(a);
        const uint64_t n = a->n, nnz = h.ci.size();
        // transpose the scalar CSR (rows ascending within each column)
        std::vector<uint64_t> cp(n + 1, 0);
        for (uint64_t i = 0; i < nnz; ++i) cp[h.ci[i] + 1]++;
        for (uint64_t c = 0; c < n; ++c) cp[c + 1] += cp[c];
        std::copy(cp.begin(), cp.end(), col_ptr);
        if (row_idx || vals) {
            std::vector<uint64_t> cur(cp.begin(), cp.end() - 1);
            for (uint64_t r = 0; r < n; ++r)
                for (uint64_t p = h.rp[r]; p < h.rp[r + 1]; ++p) {
                    const uint64_t q = cur[h.ci[p]]++;
                    if (row_idx) row_idx[q] = r;
                    if (vals) vals[q] = h.v[p];
                }
        }
    });
}

int fh_matrix_export_bsr3(const fh_matrix* a, uint32_t* row_ptr, uint32_t* cols,
                          double* vals_flat, double* dinv_flat) {
    return guard([&] {
        FH_CHECK(a && row_ptr, "null argument");
        FH_CHECK(a->bs == 3, "matrix is not 3x3-blocked (assemble an elasticity operator)");
        DeviceCtx& d = device(a->dev);
        const Pattern& p = *a->pat;
        std::vector<uint32_t> rp(p.nrows + 1);
        d2h(rp.data(), p.row_ptr.p, rp.size(), d.stream);
        FH_HIP(hipStreamSynchronize(d.stream));
        std::copy(rp.begin(), rp.end(), row_ptr);
        if (!cols && !vals_flat && !dinv_flat) return;
        std::vector<uint32_t> cc(p.nnzb), soff(p.nslices + 1);
        std::vector<double> vals(sell_value_count(a));
        d2h(cc.data(), p.cols.p, cc.size(), d.stream);
        d2h(soff.data(), p.slice_off.p, soff.size(), d.stream);
        d2h(vals.data(), a->vals.p, vals.size(), d.stream);
        FH_HIP(hipStreamSynchronize(d.stream));
        if (cols) std::copy(cc.begin(), cc.end(), cols);
        // Stored values are the operator itself: condensed couplings are
        // explicit zeros, a constrained diagonal holds its value.
        for (uint32_t r = 0; r < p.nrows; ++r) {
            const uint32_t s = r / SLICE, l = r % SLICE;
            bool have_diag = false;
            for (uint32_t k = 0; k < rp[r + 1] - rp[r]; ++k) {
                double blk[9];
                for (int e = 0; e < 9; ++e)
                    blk[e] = vals[vpos((uint64_t)soff[s] + (uint64_t)k * SLICE, e, l, 9)];
                if (vals_flat) std::copy(blk, blk + 9, vals_flat + 9 * ((uint64_t)rp[r] + k));
                if (dinv_flat && cc[rp[r] + k] == r) {
                    // adjugate over determinant, cofactors in row-major order
                    const double c00 = blk[4] * blk[8] - blk[5] * blk[7];
                    const double c01 = blk[5] * blk[6] - blk[3] * blk[8];
                    const double c02 = blk[3] * blk[7] - blk[4] * blk[6];
                    const double det = blk[0] * c00 + blk[1] * c01 + blk[2] * c02;
                    FH_CHECK(det != 0.0 && std::isfinite(det), "singular diagonal block at node " +
                                                                   std::to_string(r));
                    const double id = 1.0 / det;
                    double* o = dinv_flat + 9 * (uint64_t)r;
                    o[0] = c00 * id;
                    o[1] = (blk[2] * blk[7] - blk[1] * blk[8]) * id;
                    o[2] = (blk[1] * blk[5] - blk[2] * blk[4]) * id;
                    o[3] = c01 * id;
                    o[4] = (blk[0] * blk[8] - blk[2] * blk[6]) * id;
                    o[5] = (blk[2] * blk[3] - blk[0] * blk[5]) * id;
                    o[6] = c02 * id;
                    o[7] = (blk[1] * blk[6] - blk[0] * blk[7]) * id;
                    o[8] = (blk[0] * blk[4] - blk[1] * blk[3]) * id;
                    have_diag = true;
                }
            }
            FH_CHECK(!dinv_flat || have_diag, "node " + std::to_string(r) +
                                                  " has no diagonal block");
        }
    });
}

int fh_matrix_destroy(fh_matrix* a) {
    return guard([&] {
        if (!a) return;
        (void)hipSetDevice(a->dev);
        delete a;
    });
}

}  // extern "C"
