// fh_cg.hip — Jacobi-preconditioned CG on the device (replaces cg.rs:40-149
// and the host-driven loops of gpu.rs:775-1141 / 1145-1442).
//
// One iteration is two kernels, and no scalar ever visits the host:
//   k_cg_spmv   : reduces the previous update's (r.r, r.z) partials (every
//                 block, same fixed order -> same beta and the same stopping
//                 decision everywhere), forms p_new = z + beta p_old on the fly
//                 for the gathered columns, Ap = A p_new (SELL-64, one wave per
//                 64-row slice) and the block partials of p.Ap;
//   k_cg_update : reduces the p.Ap partials, alpha = rz/pAp, x += alpha p,
//                 r -= alpha Ap, z = D^{-1} r and the (r.r, r.z) partials.
// Vectors inside the solve use the lane-interleaved layout of the matrix values:
// component a of block row r sits at ((r / 64) * bs + a) * 64 + r % 64, so
// every load and store of a wave is one aligned, contiguous 512-byte access per
// component (the natural [row][bs] layout gives 24-byte records per lane).
// Inputs and the solution are converted at entry and exit.
// Iteration bookkeeping lives in a device CgState that only block 0 writes
// (fields read in the same kernel are double-buffered by iteration parity) and
// the next kernel reads, so a chunk of iterations has fixed arguments and is
// captured once into a HIP graph and replayed; the host reads the state once
// per chunk.
//
// Difference form.  Stiffness matrices annihilate rigid motions, so for the
// smooth fields CG works with A x is a sum of large, nearly cancelling terms
// and its rounding error (eps * sum_j |a_ij| |x_j|) swamps small residuals:
// for the 1M-dof cantilever the plain FP64 product cannot resolve a relative
// residual of 1e-10, and an FP32 copy of the matrix is a 1e-2-accurate
// operator.  The difference form
//       (A x)_i = sum_j a_ij (x_j - x_i) + s_i x_i,   s_i = sum_j a_ij
// (s_i the block row sum, tiny except next to Dirichlet rows) is the same
// operator with rounding errors proportional to the local variation
// |x_j - x_i| ~ h |grad x|, about L/h times smaller.
//
// Modes:
//   FH_PREC_FP64      : the reference's arithmetic (plain SpMV, stop on the
//                       recursive residual); FH_ORDER_ORDERED makes SpMV and
//                       axpys bitwise equal to cg.rs, only the dots differ;
//   FH_PREC_FP64_TRUE : FP64 SpMV (plain form) and residual replacement:
//                       whenever the recursive residual has fallen by
//                       delta64 (0.1) since the last replacement, and when it
//                       reaches tol, r is recomputed as the true residual
//                       b - A x (A x summed in double-double).  Replacing often keeps
//                       the gap between the two small against the residual, so
//                       beta can be kept (van der Vorst & Ye 2000); a gap above
//                       10x restarts the recurrence instead.  Convergence is
//                       declared on the true residual only (evaluated in
//                       double-double): below tol, or at the residual that
//                       rounding x to FP64 alone leaves (on ill-conditioned
//                       systems such as a 1M-dof slender cantilever that floor
//                       is ~1e-10 relative, so no FP64 vector reaches below it);
//   FH_PREC_MIXED     : as FP64_TRUE, but once ||r||/||b|| < sw (1e-3) the
//                       SpMV reads the FP32 copy of the matrix (difference
//                       form, FP64 vectors and accumulation), from the next
//                       chunk of iterations on (a separate kernel
//                       instantiation, so each has its own profile).  An inexact
//                       product is only harmless when its error is small
//                       against the residual it acts on (inexact Krylov
//                       theory: eta_k <~ tol / ||r_k||), so the FP32 operator
//                       is used only where it cannot slow CG down; on it the
//                       residual is replaced every delta (0.1) drop, and a
//                       cycle whose true residual fails to halve, or that
//                       takes over 1.25x the iterations of the last FP64
//                       cycle, sends the rest of the solve back to FP64 values;
//   FH_PREC_MIXED_IR  : FP32 operator from the start, restarting the
//                       recurrences at each replacement: classic FP64
//                       iterative refinement around FP32 CG (for comparison).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include <hip/hip_ext.h>

#include "fh_internal.hpp"

namespace fh {

enum : uint32_t { ST_RUNNING = 0, ST_CONVERGED = 1, ST_NOT_CONVERGED = 2, ST_BREAKDOWN = 3 };
enum : uint32_t { M_REF = 0, M_TRUE = 1, M_MIXED = 2, M_MIXED_IR = 3 };

struct CgState {
    uint32_t status;
    uint32_t need_rep;
    uint64_t it_spmv;
    uint64_t it_upd;
    uint64_t rep_mark;  // spmv index that consumes fresh true-residual partials
    uint64_t n_rep;
    uint64_t iters;
    uint64_t it_low;    // first iteration on the FP32 operator (0: none)
    uint64_t n_low;     // iterations whose product read the FP32 values
    uint32_t want_low;  // relaxation threshold crossed: switch to the FP32 operator
    uint32_t stalled;   // FP32-operator cycle failed to reduce the true residual
    double rz[2];
    double beta[2];     // beta of iteration j, by parity (the update rebuilds p with it)
    double alpha[2];    // alpha of iteration j, by parity (the deferred x update reads it)
    uint64_t x_base;    // last iteration at which x is exact (deferred x updates; k_x_base)
    uint32_t low[2];    // operator precision of iteration j, by parity
    double bnorm;
    double res;
    // Fields written by block 0 of a kernel are never what other blocks of the
    // same kernel decide on: per-replacement values are double-buffered by the
    // parity of n_rep (advanced between kernels, by k_mark), per-iteration
    // values by the parity of the iteration.
    double r_rep[2];    // true residual at replacement n (or the recursive one
                        // where the FP32 operator took over), by n_rep parity
    uint32_t n_stag[2]; // consecutive replacement cycles that did not halve it
    uint32_t restore;   // 1 + slot of the saved iterate to return (0: the current x)
    double r_trig;      // recursive residual that triggered the pending replacement
    uint64_t n_restart;
    double rel;
    double rel_last;    // relative recursive residual seen by the last spmv
    double floor_rel;   // FP64 rounding floor of the residual at the last true residual
    uint64_t rep_it;    // iterations done at the last replacement
    uint64_t cyc_last;  // length of the last replacement cycle (0: none yet)
    uint64_t cyc_hi;    // length of the last replacement cycle on FP64 values (0: none)
    // configuration
    double tol;
    double delta;       // replacement factor on the FP32 operator
    double delta64;     // replacement factor on the FP64 operator (true-residual modes)
    double sw;
    uint64_t max_iter;
    uint32_t mode;
    uint32_t restart;
};

struct CgMat {
    const uint32_t* soff;
    const uint32_t* rlen;
    const uint32_t* cols;
    const double* v64;
    const float* v32;
    const uint32_t* t64;  // (last value, column) records, see fh_matrix::tail64
    const uint32_t* q32;  // FP32 3x3 blocks in 16-byte groups, see fh_matrix::quad32
    const double* S64;  // block row sums, lane-interleaved [slice][bs*bs][64]
    const float* S32;
    uint32_t nrows;
    uint32_t nslices;
};

// lane-interleaved index of component a of block row r
template <int BS>
__device__ __forceinline__ uint64_t li(uint32_t r, int a) {
    return ((uint64_t)(r / SLICE) * BS + a) * SLICE + r % SLICE;
}

__device__ __forceinline__ double madd(double acc, double a, double b, bool ordered) {
    return ordered ? acc + a * b : fma(a, b, acc);
}

// acc[a] = sum_k sum_q v_k[a][q] * (x_{c_k}[q] (- xr[q] when DIFF)) for the
// lane's row, stored entries only (k < len), in column order; `gather(c, xv)`
// produces the gathered vector entries.
template <int BS, class TM, bool ORD, bool DIFF, class G>
__device__ __forceinline__ void row_product(const CgMat& A, const TM* __restrict__ vals,
                                            uint32_t s, uint32_t lane, uint32_t len,
                                            const double* xr, G&& gather, double* acc) {
    const uint32_t base = A.soff[s], w = (A.soff[s + 1] - base) / SLICE;
#pragma unroll
    for (int a = 0; a < BS; ++a) acc[a] = 0.0;
    for (uint32_t k = 0; k < w; ++k) {
        if (k < len) {
            const uint64_t e = (uint64_t)base + (uint64_t)k * SLICE;
            const uint32_t c = A.cols[e + lane];
            double xv[BS];
            gather(c, xv);
            if (DIFF)
#pragma unroll
                for (int q = 0; q < BS; ++q) xv[q] = xv[q] - xr[q];
#pragma unroll
            for (int a = 0; a < BS; ++a)
#pragma unroll
                for (int q = 0; q < BS; ++q)
                    acc[a] = madd(acc[a], (double)vals[vpos(e, a * BS + q, lane, BS * BS)], xv[q], ORD);
        }
    }
}

// acc += s_i xr (block row sum, lane-interleaved)
template <int BS, class TS>
__device__ __forceinline__ void add_rowsum(const TS* __restrict__ S, uint32_t s, uint32_t lane,
                                           const double* xr, double* acc) {
    const TS* Sr = S + (uint64_t)s * BS * BS * SLICE + lane;
#pragma unroll
    for (int a = 0; a < BS; ++a)
#pragma unroll
        for (int q = 0; q < BS; ++q) acc[a] = fma((double)Sr[(a * BS + q) * SLICE], xr[q], acc[a]);
}

// ---------------------------------------------------------------- plain SpMV
// y = A x on natural-layout vectors, one wave per slice (the reference's
// product; ORD: cg.rs order).
template <int BS, bool ORD>
__global__ __launch_bounds__(256) void k_spmv(CgMat A, const double* __restrict__ x,
                                              double* __restrict__ y) {
    const unsigned blk = fhd::xcd_block(blockIdx.x, gridDim.x);
    const unsigned s = blk * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (s >= A.nslices) return;
    const uint32_t row = s * SLICE + lane;
    const uint32_t len = row < A.nrows ? A.rlen[row] : 0;
    double acc[BS], xr[BS];
    row_product<BS, double, ORD, false>(
        A, A.v64, s, lane, len, xr,
        [&](uint32_t c, double* xv) {
#pragma unroll
            for (int q = 0; q < BS; ++q) xv[q] = x[(uint64_t)c * BS + q];
        },
        acc);
    if (row < A.nrows)
#pragma unroll
        for (int a = 0; a < BS; ++a) y[(uint64_t)row * BS + a] = acc[a];
}

// Block row sums s_i = sum_j a_ij (FP64, column order), lane-interleaved:
// element t = (slice * bs*bs + j) * 64 + lane; rows past the last are zero.
template <int BS>
__global__ void k_rowsum(CgMat A, double* __restrict__ S) {
    constexpr int B2 = BS * BS;
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= (uint64_t)A.nslices * SLICE * B2) return;
    const uint32_t s = t / (SLICE * B2), j = (t / SLICE) % B2, lane = t % SLICE;
    const uint32_t row = s * SLICE + lane;
    const uint32_t len = row < A.nrows ? A.rlen[row] : 0, base = A.soff[s];
    double acc = 0.0;
    for (uint32_t k = 0; k < len; ++k)
        acc += A.v64[vpos((uint64_t)base + (uint64_t)k * SLICE, j, lane, B2)];
    S[t] = acc;
}

// Error-free transformations (Knuth TwoSum, FMA TwoProduct) for residuals
// evaluated in double-double: the sum of the stored products is carried as
// hi + lo, so b - A x is accurate to ~eps |b - A x| + eps^2 sum|a_ij||x_j|.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void dd_add_prod(double& hi, double& lo, double a, double b) {
    const double p = a * b;
    const double ep = fma(a, b, -p);
    double sum, e;
    two_sum(hi, p, sum, e);
    hi = sum;
    lo += e + ep;
}

// t = b - A x for the lane's block row, in double-double (plain form, all
// stored entries in column order; x lane-interleaved).  `sq` accumulates
// sum_j (a_ij x_j)^2 over the row: eps^2 times it is the expected square of
// (A dx)_i when x is rounded to FP64 (dx_j ~ eps x_j), the residual floor no
// FP64 solution vector can go below.
template <int BS>
__device__ __forceinline__ void row_residual_dd(const CgMat& A, uint32_t s, uint32_t lane,
                                                uint32_t len, const double* __restrict__ x,
                                                const double* bi, double* t, double& sq) {
    const uint32_t base = A.soff[s];
    double hi[BS], lo[BS];
#pragma unroll
    for (int a = 0; a < BS; ++a) {
        hi[a] = 0.0;
        lo[a] = 0.0;
    }
    for (uint32_t k = 0; k < len; ++k) {
        const uint64_t e = (uint64_t)base + (uint64_t)k * SLICE;
        const uint32_t c = A.cols[e + lane];
#pragma unroll
        for (int q = 0; q < BS; ++q) {
            const double xv = x[li<BS>(c, q)];
#pragma unroll
            for (int a = 0; a < BS; ++a) {
                const double av = A.v64[vpos(e, a * BS + q, lane, BS * BS)];
                dd_add_prod(hi[a], lo[a], av, xv);
                sq += (av * xv) * (av * xv);
            }
        }
    }
#pragma unroll
    for (int a = 0; a < BS; ++a) {
        double d, e;
        two_sum(bi[a], -hi[a], d, e);
        t[a] = d + (e - lo[a]);
    }
}

// zv = D^{-1} r for one block row (lane-interleaved D); returns r.z
template <int BS, int PRE>
__device__ __forceinline__ double precond_row(const double* __restrict__ dinv, uint32_t s,
                                              uint32_t lane, const double* rv, double* zv) {
    double rz = 0.0;
    if constexpr (PRE == 0) {
#pragma unroll
        for (int a = 0; a < BS; ++a) {
            zv[a] = rv[a] * dinv[((uint64_t)s * BS + a) * SLICE + lane];
            rz += rv[a] * zv[a];
        }
    } else {  // 3x3 block Jacobi, D lane-interleaved [slice][9][64], row-major blocks
        const double* D = dinv + (uint64_t)s * 9 * SLICE + lane;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            zv[a] = D[(3 * a) * SLICE] * rv[0] + D[(3 * a + 1) * SLICE] * rv[1] +
                    D[(3 * a + 2) * SLICE] * rv[2];
            rz += rv[a] * zv[a];
        }
    }
    return rz;
}

// t = b - A x; r = t, z = D^{-1} r; partials (t.t, r.z) and optionally b.b.
// EXACT: t in double-double (true-residual modes: initial residual and
// replacement) plus the rounding-floor partials; otherwise FP64 in the
// reference's product order (ORD).  z goes to the z slots of the (z, p) pair
// vector, p is left as it is.  Every lane writes r (unless r is null: the
// update then keeps no r) and z (zeros past the last row), so the padded
// vector entries stay zero.
template <int BS, bool ORD, bool EXACT, int PRE>
__global__ __launch_bounds__(256) void k_residual(CgMat A, const double* __restrict__ b,
                                                  const double* __restrict__ x,
                                                  const double* __restrict__ dinv,
                                                  double* __restrict__ r, double* __restrict__ zp,
                                                  double2* __restrict__ red,
                                                  double* __restrict__ bb,
                                                  double* __restrict__ fl) {
    __shared__ double sh[4];
    const unsigned blk = fhd::xcd_block(blockIdx.x, gridDim.x);
    const unsigned s = blk * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    double tt = 0.0, rz = 0.0, bsq = 0.0, sq = 0.0;
    if (s < A.nslices) {
        const uint32_t row = s * SLICE + lane;
        const bool live = row < A.nrows;
        const uint32_t len = live ? A.rlen[row] : 0;
        double rv[BS], bi[BS];
#pragma unroll
        for (int a = 0; a < BS; ++a) bi[a] = b[((uint64_t)s * BS + a) * SLICE + lane];
        if constexpr (EXACT) {
            row_residual_dd<BS>(A, s, lane, len, x, bi, rv, sq);
        } else {
            double xr[BS], acc[BS];
            row_product<BS, double, ORD, false>(
                A, A.v64, s, lane, len, xr,
                [&](uint32_t c, double* xv) {
#pragma unroll
                    for (int q = 0; q < BS; ++q) xv[q] = x[li<BS>(c, q)];
                },
                acc);
#pragma unroll
            for (int a = 0; a < BS; ++a) rv[a] = bi[a] - acc[a];
        }
#pragma unroll
        for (int a = 0; a < BS; ++a) {
            if (!live) rv[a] = 0.0;
            if (r) r[((uint64_t)s * BS + a) * SLICE + lane] = rv[a];
            tt += rv[a] * rv[a];
            bsq += bi[a] * bi[a];
        }
        double zv[BS];
        rz = precond_row<BS, PRE>(dinv, s, lane, rv, zv);
#pragma unroll
        for (int a = 0; a < BS; ++a) zp[2 * (((uint64_t)s * BS + a) * SLICE + lane)] = zv[a];
    }
    tt = fhd::block_sum256(tt, sh);
    rz = fhd::block_sum256(rz, sh);
    if (bb) bsq = fhd::block_sum256(bsq, sh);
    if (EXACT) sq = fhd::block_sum256(sq, sh);
    if (threadIdx.x == 0) {
        red[blockIdx.x] = make_double2(tt, rz);
        if (bb) bb[blockIdx.x] = bsq;
        if (EXACT) fl[blockIdx.x] = sq;
    }
}

// ---------------------------------------------------------------- CG kernels
template <int NT = 256>
__device__ __forceinline__ double2 reduce_pairs(const double2* __restrict__ p, int m, double* sh) {
    double a = 0.0, b = 0.0;
    int i = threadIdx.x;
    for (; i + 3 * NT < m; i += 4 * NT) {  // four independent loads in flight
        const double2 u = p[i], v = p[i + NT], w = p[i + 2 * NT], t = p[i + 3 * NT];
        a += (u.x + v.x) + (w.x + t.x);
        b += (u.y + v.y) + (w.y + t.y);
    }
    for (; i < m; i += NT) {
        const double2 v = p[i];
        a += v.x;
        b += v.y;
    }
    a = fhd::block_sum<NT>(a, sh);
    b = fhd::block_sum<NT>(b, sh);
    return make_double2(a, b);
}

// matrix stream loads: non-temporal for the FP64 values (measured faster),
// ordinary for the FP32 ones (measured faster)
template <bool NT, class T>
__device__ __forceinline__ T ld_s(const T* p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

// two consecutive values (16 bytes for FP64, 8 for FP32) in one load
template <bool NT, class T>
__device__ __forceinline__ void ld_pair(const T* p, T& a, T& b) {
    typedef T pair_t __attribute__((ext_vector_type(2)));
    const pair_t v = ld_s<NT>(reinterpret_cast<const pair_t*>(p));
    a = v.x;
    b = v.y;
}

// words per (last value, column) tail record: 3 (12 bytes, one dwordx3 load)
constexpr int TAILW = 3;

template <bool NT>
__device__ __forceinline__ void ld_tail(const uint32_t* p, double& v, uint32_t& c) {
    typedef uint32_t wv __attribute__((ext_vector_type(TAILW)));
    const wv t = ld_s<NT>(reinterpret_cast<const wv*>(p));
    v = __hiloint2double((int)t.y, (int)t.x);
    c = t.z;
}

// Software-pipelined form of row_product for the CG loop (fused-multiply-add
// order): U stored blocks per step, the column indices of the next step
// loaded before this step's gathers.  No per-lane length test: padded SELL
// entries hold zero values and the row's own column (or column 0 past the
// last row), so they add exact zeros.
template <int BS, int U, class TM, bool DIFF, bool NT, class G>
__device__ __forceinline__ void row_product_pipe(const CgMat& A, const TM* __restrict__ vals,
                                                 uint32_t s, uint32_t lane, const double* xr,
                                                 G&& gather, double* acc) {
    constexpr int B2 = BS * BS;
    const uint32_t base = A.soff[s], w = (A.soff[s + 1] - base) / SLICE;
#pragma unroll
    for (int a = 0; a < BS; ++a) acc[a] = 0.0;
    if (w == 0) return;
    const uint32_t* cp = A.cols + (uint64_t)base + lane;
    const TM* vb = vals + (uint64_t)base * B2;  // step k: vb + k * SLICE * B2, see vpos
    const uint32_t last = w - 1;
    uint32_t c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = ld_s<NT>(cp + (uint64_t)min((uint32_t)u, last) * SLICE);
    for (uint32_t k = 0; k < w; k += U) {
        uint32_t cn[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            cn[u] = ld_s<NT>(cp + (uint64_t)min(k + U + u, last) * SLICE);
        TM v[U][B2];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const TM* vk = vb + (uint64_t)min(k + u, last) * SLICE * B2;
            if constexpr (B2 == 9) {
#pragma unroll
                for (int pp = 0; pp < 4; ++pp)
                    ld_pair<NT>(vk + 128 * pp + 2 * lane, v[u][2 * pp], v[u][2 * pp + 1]);
                v[u][8] = ld_s<NT>(vk + 512 + lane);
            } else {
#pragma unroll
                for (int j = 0; j < B2; ++j) v[u][j] = ld_s<NT>(vk + j * SLICE + lane);
            }
        }
        double xg[U][BS];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            gather(c[u], xg[u]);
            // steps past the slice width re-read its last entry: add zero
            const bool ok = u == 0 || k + u < w;
#pragma unroll
            for (int q = 0; q < BS; ++q) {
                const double d = DIFF ? xg[u][q] - xr[q] : xg[u][q];
                xg[u][q] = ok ? d : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int a = 0; a < BS; ++a)
#pragma unroll
                for (int q = 0; q < BS; ++q)
                    acc[a] = fma((double)v[u][a * BS + q], xg[u][q], acc[a]);
#pragma unroll
        for (int u = 0; u < U; ++u) c[u] = cn[u];
    }
}

// The same for FP64 3x3 blocks, one block per step, with the block's last
// value and column index in one 12-byte tail record loaded a step ahead: one
// memory instruction fewer per block (measured -1.5% at C1; the scalar
// product measured 7-8% slower with tails and keeps the form above; the FP32
// 3x3 product has a layout of its own, row_product_quad32).
// The first step's loads (tail record and the eight paired values), which
// depend on nothing the kernel computes: issued before the kernel's prologue.
struct Tail3Head {
    uint32_t base, w;  // slice start (entries) and width (steps)
    uint32_t c;        // column of step 0
    double vl;         // last value of step 0
    double v[8];       // values 0..7 of step 0
};

__device__ __forceinline__ void tail3_head(const CgMat& A, uint32_t s, uint32_t lane,
                                           Tail3Head& h) {
    h.base = A.soff[s];
    h.w = (A.soff[s + 1] - h.base) / SLICE;
    if (h.w == 0) return;
    ld_tail<true>(A.t64 + ((uint64_t)h.base + lane) * TAILW, h.vl, h.c);
    const double* vk = A.v64 + (uint64_t)h.base * 9;
#pragma unroll
    for (int pp = 0; pp < 4; ++pp) ld_pair<true>(vk + 128 * pp + 2 * lane, h.v[2 * pp], h.v[2 * pp + 1]);
}

template <bool DIFF, class G>
__device__ __forceinline__ void row_product_tail3(const CgMat& A, const Tail3Head& h, uint32_t lane,
                                                  const double* xr, G&& gather, double* acc) {
    const uint32_t base = h.base, w = h.w;
#pragma unroll
    for (int a = 0; a < 3; ++a) acc[a] = 0.0;
    if (w == 0) return;
    const uint32_t* tp = A.t64 + ((uint64_t)base + lane) * TAILW;  // step k: + k * SLICE * TAILW
    const double* vb = A.v64 + (uint64_t)base * 9;              // step k: vb + k * SLICE * 9
    const uint32_t last = w - 1;
    auto step = [&](const double* v, uint32_t c) {
        double xg[3];
        gather(c, xg);
#pragma unroll
        for (int q = 0; q < 3; ++q)
            if (DIFF) xg[q] -= xr[q];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int q = 0; q < 3; ++q) acc[a] = fma(v[a * 3 + q], xg[q], acc[a]);
    };
    double vl, vn;
    uint32_t c, cn;
    // step 0 from the head, the tail record of step 1 loaded ahead
    ld_tail<true>(tp + (uint64_t)min(1u, last) * SLICE * TAILW, vl, c);
    {
        double v[9];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = h.v[e];
        v[8] = h.vl;
        step(v, h.c);
    }
    for (uint32_t k = 1; k < w; ++k) {
        ld_tail<true>(tp + (uint64_t)min(k + 1, last) * SLICE * TAILW, vn, cn);
        double v[9];
        const double* vk = vb + (uint64_t)k * SLICE * 9;
#pragma unroll
        for (int pp = 0; pp < 4; ++pp) ld_pair<true>(vk + 128 * pp + 2 * lane, v[2 * pp], v[2 * pp + 1]);
        v[8] = vl;
        step(v, c);
        c = cn;
        vl = vn;
    }
}

// FP32 3x3 blocks in the grouped layout (fh_matrix::quad32): a block is two
// 16-byte loads (values 0-7) and one 8-byte load (value 8 and the column),
// 3 memory instructions instead of the 5 value loads and 1 column load of the
// lane-pair layout.  The first step is loaded before the kernel's prologue,
// the (value 8, column) record one step ahead.
struct Quad32Head {
    uint32_t base, w;  // slice start (entries) and width (steps)
    uint32_t c;        // column of step 0
    float v[9];        // values of step 0
};

typedef float f4v __attribute__((ext_vector_type(4)));
typedef uint32_t u2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void ld_quads(const uint32_t* qs, uint32_t lane, float* v) {
    const f4v g0 = *reinterpret_cast<const f4v*>(qs + 4 * lane);
    const f4v g1 = *reinterpret_cast<const f4v*>(qs + 256 + 4 * lane);
    v[0] = g0.x; v[1] = g0.y; v[2] = g0.z; v[3] = g0.w;
    v[4] = g1.x; v[5] = g1.y; v[6] = g1.z; v[7] = g1.w;
}
__device__ __forceinline__ void ld_q32_tail(const uint32_t* qs, uint32_t lane, float& v8,
                                            uint32_t& c) {
    const u2v t = *reinterpret_cast<const u2v*>(qs + 512 + 2 * lane);
    v8 = __uint_as_float(t.x);
    c = t.y;
}

__device__ __forceinline__ void quad32_head(const CgMat& A, uint32_t s, uint32_t lane,
                                            Quad32Head& h) {
    h.base = A.soff[s];
    h.w = (A.soff[s + 1] - h.base) / SLICE;
    if (h.w == 0) return;
    const uint32_t* qs = A.q32 + (uint64_t)h.base * 10;
    ld_q32_tail(qs, lane, h.v[8], h.c);
    ld_quads(qs, lane, h.v);
}

template <class G>
__device__ __forceinline__ void row_product_quad32(const CgMat& A, const Quad32Head& h,
                                                   uint32_t lane, const double* xr, G&& gather,
                                                   double* acc) {
    const uint32_t w = h.w;
#pragma unroll
    for (int a = 0; a < 3; ++a) acc[a] = 0.0;
    if (w == 0) return;
    const uint32_t* qb = A.q32 + (uint64_t)h.base * 10;  // step k: qb + k * 640
    const uint32_t last = w - 1;
    auto step = [&](const float* v, uint32_t c) {
        double xg[3];
        gather(c, xg);
#pragma unroll
        for (int q = 0; q < 3; ++q) xg[q] -= xr[q];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int q = 0; q < 3; ++q) acc[a] = fma((double)v[a * 3 + q], xg[q], acc[a]);
    };
    float vl, vn;
    uint32_t c, cn;
    ld_q32_tail(qb + (uint64_t)min(1u, last) * 640, lane, vl, c);
    step(h.v, h.c);
    for (uint32_t k = 1; k < w; ++k) {
        ld_q32_tail(qb + (uint64_t)min(k + 1, last) * 640, lane, vn, cn);
        float v[9];
        ld_quads(qb + (uint64_t)k * 640, lane, v);
        v[8] = vl;
        step(v, c);
        c = cn;
        vl = vn;
    }
}

// Block shapes of the two CG kernels.  Every block of each kernel first
// reduces all block partials of the other one, so the partial counts are kept
// small: CG_SPB slices (waves) per k_cg_spmv block, UPD_T threads per
// k_cg_update block and at most UPD_MAX update blocks (measured against 4 /
// 256 / 1024: -1% per iteration at the 1M-dof cantilever, -6.5% on a 4.2M-row
// scalar Poisson matrix, where the partial reads had grown to ~400 MB per
// iteration).
constexpr int CG_SPB = 8;
constexpr int CG_T = 64 * CG_SPB;
constexpr int UPD_T = 1024;
constexpr int UPD_MAX = 256;

// LOW: the product reads the FP32 copy of the matrix (difference form).  The
// host picks the instantiation per chunk of iterations; the FP64 one raises
// st->want_low when the relaxation threshold is crossed.
// REC: A p_new = A z + beta A p_old (Chronopoulos & Gear 1989): only z is
// gathered, A p_old is the previous Ap of the lane's own row.  The recurrence
// drifts from the exact product by rounding, so it is used only in the
// true-residual modes, whose replacements follow an exact product.
//
// zp holds (z_i, p_i) pairs at lane-interleaved index i (one 16-byte gather per
// component).  The kernel forms p_new = z + beta p_old on the fly for every
// gathered column and for its own rows; it writes only Ap: the update kernel
// rebuilds the same p_new (same beta, same arithmetic) and stores the pairs.
// Slices of the 8 XCDs for k_cg_spmv (the hardware deals block b to XCD
// b mod 8).  Uniform slice widths: XCD x owns slices [b[x], b[x+1]), one
// contiguous range with 1/8 of the stored entries, for L2 reuse of the
// gathered vector.  Widths that vary by more than 1.5x their mean (P2:
// vertex dofs first, 3x wider than edge dofs): interleaved, block i takes
// slices i * CG_SPB.., so every XCD gets the same mix of wide and narrow
// slices (C2: 472 us contiguous by slices, 267 us contiguous by entries,
// 213 us interleaved; C1: interleaved 2% slower than contiguous).
struct XcdMap {
    uint32_t b[9];
    uint32_t interleave;
};

template <int BS, bool ORD, bool DIFF, bool LOW, bool REC = false>
__global__ __launch_bounds__(CG_T) void k_cg_spmv(CgMat A, XcdMap xm, const double* __restrict__ zp,
                                                 double* __restrict__ Ap, CgState* st,
                                                 const double2* __restrict__ red_in, int n_in,
                                                 const double* __restrict__ bb_in,
                                                 const double* __restrict__ fl_in,
                                                 double* __restrict__ red_out) {
    static_assert(DIFF || !LOW, "the FP32 operator is only used in difference form");
    __shared__ double sh[CG_SPB];
    // wave-uniform slice index (lets the compiler keep slice bounds in scalars)
    const unsigned xcd = xm.interleave ? 7 : blockIdx.x % 8;
    const unsigned s = __builtin_amdgcn_readfirstlane(
        xm.interleave ? blockIdx.x * CG_SPB + (threadIdx.x >> 6)
                      : xm.b[xcd] + (blockIdx.x / 8) * CG_SPB + (threadIdx.x >> 6));
    const unsigned lane = threadIdx.x & 63;
    const bool active = s < xm.b[xcd + 1];
    // 3x3 products (EARLY): loads that depend on nothing computed below
    // (the lane's own (z, p) pairs, the first stored block of its row) are
    // issued before the state is read and the partials reduced, and the state
    // is read before the reduction, so these memory round trips overlap.  A
    // launch that turns out to be a no-op (the rest of a chunk after
    // convergence or a replacement request) then pays for those loads: the
    // other products keep the plain order, where no-ops are a larger share
    // of the launches (short solves, scalar operators) and the gain smaller.
    constexpr bool EARLY64 = BS == 3 && !ORD && !LOW;
    constexpr bool EARLY32 = BS == 3 && LOW;  // the FP32 3x3 product: its first step
    constexpr bool EARLY = EARLY64 || EARLY32;
    double xz[BS], po[BS];
    Tail3Head head;
    Quad32Head head32;
    auto own_pairs = [&] {
#pragma unroll
        for (int q = 0; q < BS; ++q)
            ld_pair<false>(zp + 2 * (((uint64_t)s * BS + q) * SLICE + lane), xz[q], po[q]);
    };
    if (EARLY && active) {
        own_pairs();
        if constexpr (EARLY64) tail3_head(A, s, lane, head);
        if constexpr (EARLY32) quad32_head(A, s, lane, head32);
    }
    const uint32_t status = st->status, need_rep = st->need_rep;
    if (!EARLY && (status != ST_RUNNING || need_rep)) return;
    const uint64_t j = st->it_upd + 1;
    const uint32_t mode = st->mode;
    const double tol = st->tol;
    const double2 rr_rz = reduce_pairs<CG_T>(red_in, n_in, sh);
    if (EARLY && (status != ST_RUNNING || need_rep)) return;
    const double rr = rr_rz.x, rz = rr_rz.y;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const bool low_prev = j > 1 ? st->low[(j - 1) & 1] != 0 : false;
    double beta;
    double rel;
    if (j == 1) {
        // cg.rs:80-84 initial test on ||r0||
        const double bnorm = sqrt(fhd::reduce_partials<CG_T>(bb_in, n_in, sh));
        const double r0 = sqrt(rr);
        rel = bnorm > 0.0 ? r0 / bnorm : r0;
        // true-residual modes: also stop at the FP64 rounding floor of x
        const double floor0 =
            mode != M_REF ? 2.0 * 1.1102230246251565e-16 * sqrt(fhd::reduce_partials<CG_T>(fl_in, n_in, sh))
                          : 0.0;
        if (lead) {
            st->bnorm = bnorm;
            st->r_rep[0] = r0;
            st->floor_rel = bnorm > 0.0 ? floor0 / bnorm : floor0;
        }
        if (r0 < tol || (bnorm > 0.0 && rel < tol) || (mode != M_REF && r0 <= floor0)) {
            if (lead) {
                st->status = ST_CONVERGED;
                st->iters = 0;
                st->res = r0;
                st->rel = rel;
            }
            return;
        }
        beta = 0.0;
    } else {
        const double bnorm = st->bnorm;
        const double res = sqrt(rr);
        rel = bnorm > 0.0 ? res / bnorm : res;  // cg.rs:118-119
        const bool fresh = mode != M_REF && st->rep_mark == j;
        const uint64_t nrep = st->n_rep;
        double floorr = 0.0;
        bool halved = true;
        uint32_t nstag = 0;
        bool done = rel < tol;  // cg.rs:120-122
        if (fresh) {
            // r is the true residual (double-double).  Converged when it is
            // below tol, or at the floor that rounding x to FP64 alone leaves
            // (no FP64 vector does better), or close to that floor once
            // replacement cycles stop halving it (the attainable accuracy)
            floorr = 2.0 * 1.1102230246251565e-16 * sqrt(fhd::reduce_partials<CG_T>(fl_in, n_in, sh));
            const double prev = st->r_rep[(nrep - 1) & 1];
            halved = res <= 0.5 * prev;
            nstag = halved ? 0u : st->n_stag[(nrep - 1) & 1] + 1u;
            const bool stop_stag = (!halved && res <= 2.0 * floorr) || nstag >= 2;
            done = done || res <= floorr || stop_stag;
            if (lead) {
                st->floor_rel = bnorm > 0.0 ? floorr / bnorm : floorr;
                // near the floor, further cycles can make the true residual
                // worse: return the iterate of the previous replacement (saved
                // by the host in slot nrep & 1) when it was better
                if (done && rel >= tol && res > floorr && nrep >= 2 && res > prev)
                    st->restore = 1u + (uint32_t)(nrep & 1);
            }
        }
        if (mode == M_REF || fresh) {
            if (done) {
                if (lead) {
                    // stagnation well above the floor is reported, not hidden
                    const bool ok = rel < tol || res <= 8.0 * floorr || mode == M_REF;
                    st->status = ok ? ST_CONVERGED : ST_NOT_CONVERGED;
                    st->iters = j - 1;
                    st->res = res;
                    st->rel = rel;
                }
                return;
            }
        } else {
            // replacement: the recursive residual reached tol or fell by delta
            // since the last one, or the cycle runs long (a stagnating
            // recursion is checked against the true residual)
            const uint64_t cyc = (j - 1) - st->rep_it;
            if (rel < tol || res < (low_prev ? st->delta : st->delta64) * st->r_rep[nrep & 1] ||
                (st->cyc_last > 0 && cyc > 2 * st->cyc_last + 16)) {
                if (lead) {
                    st->need_rep = 1;  // recompute r as the true residual first
                    st->r_trig = res;
                }
                return;
            }
        }
        if (j > st->max_iter) {
            if (lead) {
                st->status = ST_NOT_CONVERGED;
                st->iters = j - 1;
                st->res = res;
                st->rel = rel;
            }
            return;
        }
        if (lead) st->rel_last = rel;
        if (lead && fresh) {
            // a replacement cycle on the FP32 operator that did not halve the
            // true residual, or took over 1.25x the iterations of the last
            // FP64 cycle: the FP32 operator's error slows CG down here; the
            // host finishes the solve on FP64 values
            const uint64_t cyc = (j - 1) - st->rep_it;
            if (low_prev && (!halved || (st->cyc_hi > 0 && 4 * cyc > 5 * st->cyc_hi + 8)))
                st->stalled = 1;
            if (!low_prev) st->cyc_hi = cyc;
            st->cyc_last = cyc;
            st->rep_it = j - 1;
            st->r_rep[nrep & 1] = res;
            st->n_stag[nrep & 1] = nstag;
        }
        // the FP32 operator's cycles are measured from the recursive residual
        // where it took over.  Blocks of this kernel that read the slot after
        // this write see a replacement threshold below res: the same "no
        // replacement" that block 0 decided before writing it.
        if (lead && LOW && !low_prev && !fresh) st->r_rep[nrep & 1] = res;
        // keep the recurrence across a replacement (van der Vorst & Ye) unless
        // the true residual is far from the recursive one it replaces: then
        // the Krylov information is stale and CG restarts from the true one
        const bool restart = fresh && (st->restart || res > 10.0 * st->r_trig);
        if (lead && restart) st->n_restart = st->n_restart + 1;
        beta = restart ? 0.0 : rz / st->rz[(j - 1) & 1];
    }
    if (lead) {
        st->rz[j & 1] = rz;
        st->beta[j & 1] = beta;
        st->low[j & 1] = LOW ? 1u : 0u;
        st->it_spmv = j;
        if (LOW && !low_prev) st->it_low = j;
        if (LOW) st->n_low = st->n_low + 1;
        if (!LOW && mode == M_MIXED && rel < st->sw) st->want_low = 1;
    }

    double dot = 0.0;
    if (active) {
        const uint32_t row = s * SLICE + lane;
        const uint32_t len = row < A.nrows ? A.rlen[row] : 0;
        double xr[BS], acc[BS];
        if constexpr (!EARLY) own_pairs();
#pragma unroll
        for (int q = 0; q < BS; ++q) xr[q] = madd(xz[q], beta, po[q], ORD);
        auto gather = [&](uint32_t c, double* xv) {
#pragma unroll
            for (int q = 0; q < BS; ++q) {
                double zc, pc;
                ld_pair<false>(zp + 2 * li<BS>(c, q), zc, pc);
                xv[q] = REC ? zc : madd(zc, beta, pc, ORD);
            }
        };
        // own-row entry of the gathered vector (difference form)
        const double* xo = REC ? xz : xr;
        // blocks per pipelined step: one for FP64 3x3 blocks (fewer registers,
        // measured 2-4% faster than two), two for FP32 ones, four scalars
        constexpr int U = BS == 1 ? 4 : (LOW ? 2 : 1);
        if constexpr (LOW) {
            if constexpr (EARLY32)
                row_product_quad32(A, head32, lane, xo, gather, acc);
            else
                row_product_pipe<BS, U, float, true, false>(A, A.v32, s, lane, xo, gather, acc);
            add_rowsum<BS>(A.S32, s, lane, xo, acc);
        } else if constexpr (ORD) {
            row_product<BS, double, true, DIFF>(A, A.v64, s, lane, len, xr, gather, acc);
        } else if constexpr (BS == 3) {
            row_product_tail3<DIFF>(A, head, lane, xo, gather, acc);
        } else {
            row_product_pipe<BS, U, double, DIFF, true>(A, A.v64, s, lane, xo, gather, acc);
            if (DIFF) add_rowsum<BS>(A.S64, s, lane, xo, acc);
        }
        // every lane stores (rows past the last hold zeros): full 512-byte stores
#pragma unroll
        for (int a = 0; a < BS; ++a) {
            const uint64_t i = ((uint64_t)s * BS + a) * SLICE + lane;
            if constexpr (REC) acc[a] = fma(beta, Ap[i], acc[a]);
            Ap[i] = acc[a];
            dot += xr[a] * acc[a];
        }
    }
    dot = fhd::block_sum<CG_T>(dot, sh);
    if (threadIdx.x == 0) red_out[blockIdx.x] = dot;
}

// PRE 0: scalar Jacobi (cg.rs:52-64), elementwise over the padded vectors;
// PRE 1: 3x3 block Jacobi (gpu.rs bsr3_apply_dinv), one thread per block row.
// p_new = z + beta p_old is rebuilt from the (z, p) pairs exactly as the spmv
// kernel formed it; the pairs are overwritten with (z_new, p_new).
// r null (PRE 0, a D^{-1} with no zero entry): r is not stored; its entries
// come back from the pairs as z_i / D^{-1}_i (within two roundings of the r
// that z was formed from), which saves 16 of the 80 bytes per unknown this
// kernel streams.
// lazy (PRE 0): x is updated every second iteration.  An update an odd number
// of iterations after x_base leaves x as it is; the next one, which still
// finds p_{j-1} in the pairs, adds alpha_{j-1} p_{j-1} and then alpha_j p_j:
// the same two FMAs in the same order as updating every iteration, so x is
// bitwise the same, for half of the x traffic.  k_x_flush completes x before
// anything reads it (a replacement, the end of the solve).
template <bool ORD, int PRE, bool EARLY>
__global__ __launch_bounds__(UPD_T) void k_cg_update(uint32_t npad, double* __restrict__ x,
                                                   double* __restrict__ r,
                                                   const double* __restrict__ Ap,
                                                   double* __restrict__ zp,
                                                   const double* __restrict__ dinv, CgState* st,
                                                   const double* __restrict__ red_in, int n_in,
                                                   double2* __restrict__ red_out, int lazy) {
    typedef double pair_t __attribute__((ext_vector_type(2)));
    __shared__ double sh[UPD_T / 64];
    pair_t* zpv = reinterpret_cast<pair_t*>(zp);
    const uint64_t stride = (uint64_t)gridDim.x * UPD_T;
    const uint64_t i0 = (uint64_t)blockIdx.x * UPD_T + threadIdx.x;
    // EARLY (with the FP64 3x3 product, as in k_cg_spmv): the operands of the
    // thread's first element do not depend on the reduced scalars, so their
    // loads are issued before the state is read and the partials reduced, and
    // the state is read before the reduction.  One element per thread covers
    // that prologue; more only add to what a no-op launch streams.
    constexpr bool PF = EARLY && PRE == 0;
    const bool zr = r == nullptr;
    pair_t f_zp = {0.0, 0.0};
    double f_x = 0.0, f_r = 0.0, f_ap = 0.0, f_d = 0.0;
    if (PF && i0 < npad) {
        f_zp = zpv[i0];
        f_x = x[i0];
        if (!zr) f_r = r[i0];
        f_ap = Ap[i0];
        f_d = dinv[i0];
    }
    // Whether this launch updates is decided on fields that only earlier
    // kernels write (status, need_rep, it_spmv): every k_cg_spmv launch that
    // does not advance it_spmv sets status or need_rep, so the two tests
    // below are exactly "the preceding product was a full one".  it_upd, which
    // block 0 of this kernel advances, is never read here: blocks need not be
    // co-resident, and one that starts after block 0 has finished would see
    // the new value.  The one field this kernel may write that it also reads,
    // status (breakdown), is written only when every block decides the same.
    const uint32_t status = st->status, need_rep = st->need_rep;
    if (!EARLY && (status != ST_RUNNING || need_rep)) return;
    const uint64_t j = st->it_spmv;
    const double pap = fhd::reduce_partials<UPD_T>(red_in, n_in, sh);
    if (EARLY && (status != ST_RUNNING || need_rep)) return;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (fabs(pap) < 1e-30) {  // cg.rs:102-107
        if (lead) {
            st->status = ST_BREAKDOWN;
            st->iters = j;
        }
        return;
    }
    const double al = st->rz[j & 1] / pap;
    const double beta = st->beta[j & 1];
    // deferred x: skip (x then lags by alpha_j p_j), or catch up the lag
    const bool skip_x = lazy && ((j - st->x_base) & 1);
    const bool two = lazy && !skip_x;
    const double al_prev = two ? st->alpha[(j - 1) & 1] : 0.0;
    if (lead) st->alpha[j & 1] = al;
    double rr = 0.0, rz = 0.0;
    if (PRE == 0) {
        // cg.rs:112-115 with p of cg.rs:133-135, z of cg.rs:124-126
        auto step = [&](uint64_t i, pair_t o, double xo, double ro, double ap, double di) {
            const double pi = madd(o.x, beta, o.y, ORD);
            const double xl = two ? fma(al_prev, o.y, xo) : xo;
            const double xi = ORD ? xo + al * pi : fma(al, pi, xl);
            const double rp = zr ? o.x / di : ro;
            const double ri = ORD ? rp - al * ap : fma(-al, ap, rp);
            const double zi = ri * di;
            if (!skip_x) x[i] = xi;
            if (!zr) r[i] = ri;
            pair_t o2;
            o2.x = zi;
            o2.y = pi;
            zpv[i] = o2;
            rr += ri * ri;
            rz += ri * zi;
        };
        uint64_t i = i0;
        if (PF) {
            if (i < npad) step(i, f_zp, f_x, f_r, f_ap, f_d);
            i += stride;
        }
        for (; i < npad; i += stride)
            step(i, zpv[i], skip_x ? 0.0 : x[i], zr ? 0.0 : r[i], Ap[i], dinv[i]);
    } else {
        const uint32_t nn = npad / 3;
        for (uint32_t v = blockIdx.x * UPD_T + threadIdx.x; v < nn; v += stride) {
            const uint32_t s = v / SLICE, lane = v % SLICE;
            double rv[3], pv[3], zv[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const uint64_t i = ((uint64_t)s * 3 + a) * SLICE + lane;
                const pair_t o = zpv[i];
                pv[a] = madd(o.x, beta, o.y, ORD);
                x[i] = ORD ? x[i] + al * pv[a] : fma(al, pv[a], x[i]);
                rv[a] = ORD ? r[i] - al * Ap[i] : fma(-al, Ap[i], r[i]);
                r[i] = rv[a];
                rr += rv[a] * rv[a];
            }
            rz += precond_row<3, 1>(dinv, s, lane, rv, zv);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                pair_t o2;
                o2.x = zv[a];
                o2.y = pv[a];
                zpv[((uint64_t)s * 3 + a) * SLICE + lane] = o2;
            }
        }
    }
    rr = fhd::block_sum<UPD_T>(rr, sh);
    rz = fhd::block_sum<UPD_T>(rz, sh);
    if (threadIdx.x == 0) red_out[blockIdx.x] = make_double2(rr, rz);
    if (lead) st->it_upd = j;
}

// Deferred x: when the last update (it_upd) left x lagging by alpha p (an odd
// number of iterations after x_base), add it; p is the pairs' second entry.
// x_base is then moved by k_x_base, a launch of its own, so no block of this
// kernel can see it change.
__global__ __launch_bounds__(256) void k_x_flush(uint32_t npad, double* __restrict__ x,
                                                 const double* __restrict__ zp,
                                                 const CgState* st) {
    const uint64_t j = st->it_upd;
    if (((j - st->x_base) & 1) == 0) return;
    const double al = st->alpha[j & 1];
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < npad;
         i += (uint64_t)gridDim.x * blockDim.x)
        x[i] = fma(al, zp[2 * i + 1], x[i]);
}
__global__ void k_x_base(CgState* st) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->x_base = st->it_upd;
}

// After a replacement: the next spmv consumes fresh true-residual partials.
__global__ void k_mark(CgState* st) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->rep_mark = st->it_upd + 1;
        st->need_rep = 0;
        st->n_rep = st->n_rep + 1;
    }
}

// diagonal entry a of block row r (0 when the row has no diagonal block)
template <int BS>
__device__ __forceinline__ double diag_entry(const CgMat& A, uint32_t r, int a, bool& found) {
    const uint32_t s = r / SLICE, lane = r % SLICE, base = A.soff[s], len = A.rlen[r];
    found = false;
    for (uint32_t k = 0; k < len; ++k) {
        const uint64_t e = (uint64_t)base + (uint64_t)k * SLICE;
        if (A.cols[e + lane] == r) {
            found = true;
            return A.v64[vpos(e, a * BS + a, lane, BS * BS)];
        }
    }
    return 0.0;
}

// Inverse diagonal, natural layout: 1/d when |d| > 1e-15, else 1 (cg.rs:52-64).
template <int BS>
__global__ void k_inv_diag(CgMat A, double* __restrict__ out) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= (uint64_t)A.nrows * BS) return;
    bool found;
    const double d = diag_entry<BS>(A, t / BS, t % BS, found);
    out[t] = (found && fabs(d) > 1e-15) ? 1.0 / d : 1.0;
}

// The same, lane-interleaved over the padded rows (1 past the last row).
template <int BS>
__global__ void k_inv_diag_li(CgMat A, double* __restrict__ out) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= (uint64_t)A.nslices * SLICE * BS) return;
    const uint32_t s = t / (SLICE * BS), a = (t / SLICE) % BS, lane = t % SLICE;
    const uint32_t r = s * SLICE + lane;
    double inv = 1.0;
    if (r < A.nrows) {
        bool found;
        const double d = diag_entry<BS>(A, r, a, found);
        if (found && fabs(d) > 1e-15) inv = 1.0 / d;
    }
    out[t] = inv;
}

template <class TI, class TO>
__global__ void k_convert(const TI* __restrict__ a, TO* __restrict__ b, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        b[i] = (TO)a[i];
}

// natural [row][W] -> lane-interleaved [slice][W][64]; `pad` past the last row
template <int W>
__global__ void k_to_li(const double* __restrict__ nat, double* __restrict__ out, uint32_t nrows,
                        uint32_t nslices, double pad) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= (uint64_t)nslices * SLICE * W) return;
    const uint32_t s = t / (SLICE * W), a = (t / SLICE) % W, lane = t % SLICE;
    const uint32_t r = s * SLICE + lane;
    out[t] = r < nrows ? nat[(uint64_t)r * W + a] : pad;
}

// lane-interleaved -> natural [row][W]
template <int W>
__global__ void k_from_li(const double* __restrict__ in, double* __restrict__ nat, uint32_t nrows) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= (uint64_t)nrows * W) return;
    const uint32_t r = t / W, a = t % W;
    nat[t] = in[((uint64_t)(r / SLICE) * W + a) * SLICE + r % SLICE];
}

// ||b - A x||^2 (double-double evaluation) and ||b||^2 block partials
// (lane-interleaved b and x)
template <int BS>
__global__ __launch_bounds__(256) void k_true_res(CgMat A, const double* __restrict__ b,
                                                  const double* __restrict__ x,
                                                  double2* __restrict__ red) {
    __shared__ double sh[4];
    const unsigned blk = fhd::xcd_block(blockIdx.x, gridDim.x);
    const unsigned s = blk * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    double tt = 0.0, bsq = 0.0;
    if (s < A.nslices) {
        const uint32_t row = s * SLICE + lane;
        if (row < A.nrows) {
            double bi[BS], t[BS];
#pragma unroll
            for (int a = 0; a < BS; ++a) bi[a] = b[((uint64_t)s * BS + a) * SLICE + lane];
            double sq = 0.0;
            row_residual_dd<BS>(A, s, lane, A.rlen[row], x, bi, t, sq);
#pragma unroll
            for (int a = 0; a < BS; ++a) {
                tt += t[a] * t[a];
                bsq += bi[a] * bi[a];
            }
        }
    }
    tt = fhd::block_sum256(tt, sh);
    bsq = fhd::block_sum256(bsq, sh);
    if (threadIdx.x == 0) red[blockIdx.x] = make_double2(tt, bsq);
}

// t = b - A (xh + xl) in double-double for natural-layout [row][bs] vectors
// (xl may be null: a plain FP64 vector), t rounded to FP64 into r (may be
// null), block partials (t.t, b.b).  Every product a_ij xh_j and a_ij xl_j
// enters the double-double sum with its rounding error (TwoProduct), so t is
// accurate to about eps |t| + eps^2 sum_j |a_ij| |x_j|.
template <int BS>
__global__ __launch_bounds__(256) void k_residual_pair(CgMat A, const double* __restrict__ b,
                                                       const double* __restrict__ xh,
                                                       const double* __restrict__ xl,
                                                       double* __restrict__ r,
                                                       double2* __restrict__ red) {
    __shared__ double sh[4];
    const unsigned blk = fhd::xcd_block(blockIdx.x, gridDim.x);
    const unsigned s = blk * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    double tt = 0.0, bsq = 0.0;
    if (s < A.nslices) {
        const uint32_t row = s * SLICE + lane;
        if (row < A.nrows) {
            const uint32_t base = A.soff[s], len = A.rlen[row];
            double hi[BS], lo[BS];
#pragma unroll
            for (int a = 0; a < BS; ++a) hi[a] = lo[a] = 0.0;
            for (uint32_t k = 0; k < len; ++k) {
                const uint64_t e = (uint64_t)base + (uint64_t)k * SLICE;
                const uint32_t c = A.cols[e + lane];
#pragma unroll
                for (int q = 0; q < BS; ++q) {
                    const double vh = xh[(uint64_t)c * BS + q];
                    const double vl = xl ? xl[(uint64_t)c * BS + q] : 0.0;
#pragma unroll
                    for (int a = 0; a < BS; ++a) {
                        const double av = A.v64[vpos(e, a * BS + q, lane, BS * BS)];
                        dd_add_prod(hi[a], lo[a], av, vh);
                        dd_add_prod(hi[a], lo[a], av, vl);
                    }
                }
            }
#pragma unroll
            for (int a = 0; a < BS; ++a) {
                const double bi = b[(uint64_t)row * BS + a];
                double d, e;
                two_sum(bi, -hi[a], d, e);
                const double t = d + (e - lo[a]);
                if (r) r[(uint64_t)row * BS + a] = t;
                tt += t * t;
                bsq += bi * bi;
            }
        }
    }
    tt = fhd::block_sum256(tt, sh);
    bsq = fhd::block_sum256(bsq, sh);
    if (threadIdx.x == 0) red[blockIdx.x] = make_double2(tt, bsq);
}

// (hi, lo) += sc d elementwise (sc a power of two, so sc d is exact),
// renormalised so that hi is the FP64 value nearest to hi + lo (TwoSum, then
// Fast-TwoSum)
__global__ void k_scale(double* __restrict__ v, double sc, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        v[i] *= sc;
}

__global__ void k_dd_add(double* __restrict__ hi, double* __restrict__ lo,
                         const double* __restrict__ d, double sc, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        double s, e;
        two_sum(hi[i], sc * d[i], s, e);
        e += lo[i];
        const double h = s + e;
        lo[i] = e - (h - s);
        hi[i] = h;
    }
}

// ---------------------------------------------------------------- host side

static CgMat cgmat(const fh_matrix* a) {
    const Pattern& p = *a->pat;
    return {p.slice_off.p, p.row_len.p, p.sell_cols.p, a->vals.p, a->vals32.p, a->tail64.p,
            a->quad32.p, a->rowsum64.p, a->rowsum32.p, p.nrows, p.nslices};
}

static unsigned spmv_grid(const fh_matrix* a) { return std::max(1u, cdiv(a->pat->nslices, 4)); }

double spmv_algorithmic_bytes(const fh_matrix* a, int vbytes, int mbytes) {
    // matrix values + one 4-byte index per stored block + row pointers, one
    // read of x and one write of y (the CSR / BSR algorithmic byte count)
    const double nnzb = (double)a->pat->nnzb;
    const int bs2 = a->bs * a->bs;
    return nnzb * (bs2 * (double)mbytes + 4.0) + 4.0 * (a->pat->nrows + 1) +
           2.0 * (double)a->n * vbytes;
}

// the difference form also reads one bs*bs block row sum per block row
static double diff_spmv_bytes(const fh_matrix* a, int mbytes) {
    return spmv_algorithmic_bytes(a, 8, mbytes) +
           (double)a->pat->nrows * a->bs * a->bs * (double)mbytes;
}

void spmv_f64(const fh_matrix* a, const double* x, double* y, int order, hipStream_t s) {
    const CgMat A = cgmat(a);
    const unsigned g = spmv_grid(a);
    if (a->bs == 1) {
        if (order == FH_ORDER_ORDERED) hipLaunchKernelGGL((k_spmv<1, true>), dim3(g), dim3(256), 0, s, A, x, y);
        else hipLaunchKernelGGL((k_spmv<1, false>), dim3(g), dim3(256), 0, s, A, x, y);
    } else {
        if (order == FH_ORDER_ORDERED) hipLaunchKernelGGL((k_spmv<3, true>), dim3(g), dim3(256), 0, s, A, x, y);
        else hipLaunchKernelGGL((k_spmv<3, false>), dim3(g), dim3(256), 0, s, A, x, y);
    }
    FH_LAUNCH_CHECK("spmv");
}

void inv_diag_f64(const fh_matrix* a, double* out, hipStream_t s) {
    const CgMat A = cgmat(a);
    if (a->bs == 1) hipLaunchKernelGGL(k_inv_diag<1>, dim3(cdiv(a->n, 256)), dim3(256), 0, s, A, out);
    else hipLaunchKernelGGL(k_inv_diag<3>, dim3(cdiv(a->n, 256)), dim3(256), 0, s, A, out);
    FH_LAUNCH_CHECK("inv_diag");
}

// block row sums for the difference form (FP64, and an FP32 copy when asked),
// lane-interleaved over the padded rows
void ensure_rowsum(fh_matrix* a, bool f32) {
    DeviceCtx& d = device(a->dev);
    const uint64_t m = (uint64_t)a->pat->nslices * SLICE * a->bs * a->bs;
    if (!a->rowsum64.p) {
        a->rowsum64.alloc(a->dev, std::max<uint64_t>(m, 1));
        const CgMat A = cgmat(a);
        if (a->bs == 1) hipLaunchKernelGGL(k_rowsum<1>, dim3(cdiv(m, 256)), dim3(256), 0, d.stream, A, a->rowsum64.p);
        else hipLaunchKernelGGL(k_rowsum<3>, dim3(cdiv(m, 256)), dim3(256), 0, d.stream, A, a->rowsum64.p);
        FH_LAUNCH_CHECK("rowsum");
    }
    if (f32 && !a->rowsum32.p) {
        a->rowsum32.alloc(a->dev, std::max<uint64_t>(m, 1));
        hipLaunchKernelGGL((k_convert<double, float>), dim3(std::max(1u, std::min(cdiv(m, 256), 4096u))),
                           dim3(256), 0, d.stream, a->rowsum64.p, a->rowsum32.p, m);
        FH_LAUNCH_CHECK("rowsum32");
    }
}

// (last value, column) record of every SELL entry e of a 3x3-block matrix
// (slices start at multiples of 64, so e's step base is e rounded down to 64
// and its lane e mod 64)
__global__ void k_make_tail(const double* __restrict__ vals, const uint32_t* __restrict__ cols,
                            uint64_t n, uint32_t* __restrict__ tail) {
    const uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const uint64_t b =
        (uint64_t)__double_as_longlong(vals[vpos(e & ~63ull, 8, (uint32_t)(e & 63), 9)]);
    tail[TAILW * e] = (uint32_t)b;
    tail[TAILW * e + 1] = (uint32_t)(b >> 32);
    tail[TAILW * e + 2] = cols[e];
    if constexpr (TAILW == 4) tail[TAILW * e + 3] = 0u;
}

// the FP32 3x3 blocks of every SELL entry e in the grouped layout of
// fh_matrix::quad32, from the FP32 values (same rounding as the values)
__global__ void k_make_quad32(const float* __restrict__ v32, const uint32_t* __restrict__ cols,
                              uint64_t n, uint32_t* __restrict__ q) {
    const uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const uint64_t e0 = e & ~(uint64_t)(SLICE - 1);
    const uint32_t lane = (uint32_t)(e & (SLICE - 1));
    uint32_t* qs = q + e0 * 10;
#pragma unroll
    for (int j = 0; j < 8; ++j)
        qs[(j < 4 ? 0 : 256) + 4 * lane + (j & 3)] = __float_as_uint(v32[vpos(e0, j, lane, 9)]);
    qs[512 + 2 * lane] = __float_as_uint(v32[vpos(e0, 8, lane, 9)]);
    qs[512 + 2 * lane + 1] = cols[e];
}

void ensure_quad32(fh_matrix* a) {
    if (a->quad32.p || a->bs != 3) return;
    ensure_vals32(a);
    DeviceCtx& d = device(a->dev);
    const uint64_t n = a->pat->sell_entries;
    a->quad32.alloc(a->dev, std::max<uint64_t>(10 * n, 1));
    if (n) {
        hipLaunchKernelGGL(k_make_quad32, dim3(std::max(1u, cdiv(n, 256))), dim3(256), 0,
                           d.stream, a->vals32.p, a->pat->sell_cols.p, n, a->quad32.p);
        FH_LAUNCH_CHECK("make_quad32");
    }
}

void ensure_tail(fh_matrix* a) {
    if (a->tail64.p || a->bs != 3) return;
    DeviceCtx& d = device(a->dev);
    const uint64_t n = a->pat->sell_entries;
    a->tail64.alloc(a->dev, std::max<uint64_t>(TAILW * n, 1));
    if (n) {
        hipLaunchKernelGGL(k_make_tail, dim3(std::max(1u, cdiv(n, 256))), dim3(256), 0, d.stream,
                           a->vals.p, a->pat->sell_cols.p, n, a->tail64.p);
        FH_LAUNCH_CHECK("make_tail");
    }
}

// Pinned host mirror of the state, one per device (exported calls hold the
// library's call lock, so one solve uses it at a time)
static CgState* host_state(int dev) {
    static std::mutex mu;
    static std::map<int, CgState*> m;
    std::lock_guard<std::mutex> lk(mu);
    auto it = m.find(dev);
    if (it != m.end()) return it->second;
    CgState* p = nullptr;
    FH_HIP(hipHostMalloc((void**)&p, sizeof(CgState), hipHostMallocDefault));
    m[dev] = p;
    return p;
}

static int env_int(const char* name, int dflt) {
    const char* v = std::getenv(name);
    return v ? std::atoi(v) : dflt;
}
static double env_dbl(const char* name, double dflt) {
    const char* v = std::getenv(name);
    return v ? std::atof(v) : dflt;
}

template <int W>
static void to_li(const double* nat, double* out, uint32_t nrows, uint32_t nslices, double pad,
                  hipStream_t s) {
    const uint64_t m = (uint64_t)nslices * SLICE * W;
    hipLaunchKernelGGL(k_to_li<W>, dim3(std::max(1u, cdiv(m, 256))), dim3(256), 0, s, nat, out, nrows,
                       nslices, pad);
    FH_LAUNCH_CHECK("to_li");
}

// XCD ranges of k_cg_spmv for a pattern (computed once and kept on it; the
// slice offsets are copied to the host for that): slice boundaries at each
// eighth of the stored entries, or interleaved blocks when slice widths vary.
static void xcd_map(const Pattern& pc, hipStream_t s, XcdMap& xm, unsigned& gc) {
    Pattern& p = const_cast<Pattern&>(pc);
    if (!p.xcd_ready) {
        const uint32_t nsl = p.nslices;
        std::vector<uint32_t> soff(nsl + 1);
        d2h(soff.data(), p.slice_off.p, nsl + 1, s);
        FH_HIP(hipStreamSynchronize(s));
        const uint64_t tot = soff[nsl];
        uint32_t per = 0;
        p.xcd_b[0] = 0;
        p.xcd_b[8] = nsl;
        for (int xx = 1; xx < 8; ++xx) {
            const uint64_t want = tot * (uint64_t)xx / 8;
            p.xcd_b[xx] = (uint32_t)(std::lower_bound(soff.begin(), soff.end() - 1, (uint32_t)want) -
                                     soff.begin());
            p.xcd_b[xx] = std::max(p.xcd_b[xx], p.xcd_b[xx - 1]);
        }
        for (int xx = 0; xx < 8; ++xx) per = std::max(per, cdiv(p.xcd_b[xx + 1] - p.xcd_b[xx], CG_SPB));
        p.xcd_blocks = 8 * std::max(per, 1u);
        p.xcd_interleave = 0;
        uint32_t wmax = 0;
        for (uint32_t i = 0; i < nsl; ++i) wmax = std::max(wmax, soff[i + 1] - soff[i]);
        if (nsl && (double)wmax > 1.5 * (double)tot / nsl) {
            p.xcd_interleave = 1;  // b[8] = nsl bounds every block
            p.xcd_blocks = std::max(1u, cdiv(nsl, CG_SPB));
        }
        p.xcd_ready = true;
    }
    for (int xx = 0; xx < 9; ++xx) xm.b[xx] = p.xcd_b[xx];
    xm.interleave = p.xcd_interleave;
    gc = p.xcd_blocks;
}

CgCache::~CgCache() {
    for (int g = 0; g < 2; ++g) {
        if (exec[g]) (void)hipGraphExecDestroy(exec[g]);
        if (graph[g]) (void)hipGraphDestroy(graph[g]);
    }
}

// Device buffers of one solve, carved from one workspace.
struct CgWork {
    double *r, *zp, *Ap, *xl, *bl, *xs0, *xs1, *dinv, *redA, *bb, *fl;
    double2 *redB, *tr, *sum;
    CgState* st;
};

template <class T>
static T* carve(uint8_t* base, size_t& off, size_t count) {
    off = (off + 255) & ~(size_t)255;
    T* q = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += std::max<size_t>(count, 1) * sizeof(T);
    return q;
}

// Lays the buffers of a solve out from `base` (null: only sizes them);
// returns the workspace bytes.  gc: k_cg_spmv blocks (xcd_map).
static size_t cg_work(const fh_matrix* a, int pre, bool diff, unsigned gc, uint8_t* base,
                      CgWork* w) {
    const uint32_t nsl = a->pat->nslices;
    const size_t npad = (size_t)nsl * SLICE * a->bs;
    const unsigned gs = spmv_grid(a);
    const unsigned gu = std::max(1u, std::min<unsigned>(cdiv(npad, UPD_T * 4), UPD_MAX));
    size_t off = 0;
    CgWork t{};
    t.r = carve<double>(base, off, npad);
    t.zp = carve<double>(base, off, 2 * npad);  // (z, p) pairs at each lane-interleaved index
    t.Ap = carve<double>(base, off, npad);
    t.xl = carve<double>(base, off, npad);
    t.bl = carve<double>(base, off, npad);
    if (diff) {  // the iterate at each of the last two replacements
        t.xs0 = carve<double>(base, off, npad);
        t.xs1 = carve<double>(base, off, npad);
    }
    t.dinv = carve<double>(base, off, pre == 0 ? npad : (size_t)nsl * SLICE * 9);
    t.redA = carve<double>(base, off, gc);
    t.bb = carve<double>(base, off, gs);
    t.fl = carve<double>(base, off, gs);
    t.redB = carve<double2>(base, off, std::max(gs, gu));
    t.tr = carve<double2>(base, off, gs);
    t.sum = carve<double2>(base, off, 1);
    t.st = carve<CgState>(base, off, 1);
    if (w) *w = t;
    return (off + 255) & ~(size_t)255;
}

// fixed-order sum of m (a, b) partials into out[0] (one block)
__global__ __launch_bounds__(256) void k_sum_pairs(const double2* __restrict__ p, int m,
                                                   double2* __restrict__ out) {
    __shared__ double sh[4];
    const double2 v = reduce_pairs<256>(p, m, sh);
    if (threadIdx.x == 0) out[0] = v;
}

// The solve proper.  ws: the workspace (cg_work bytes) when the caller
// brings one; the call then allocates nothing (no HIP graphs, no events, no
// pinned host state), and the matrix must carry what its precision reads
// (FP32 copy, row sums, tail records, XCD map), built by an earlier solve.
template <int BS, bool ORD, bool DIFF, int PRE>
static int run_cg(const fh_matrix* a, const SolveArgs& args, fh_solve_stats* stats,
                  uint8_t* ws = nullptr) {
    const int dev = a->dev;
    DeviceCtx& d = device(dev);
    hipStream_t s = d.stream;
    const bool noalloc = ws != nullptr;
    const uint32_t n = a->n, nrows = a->pat->nrows, nsl = a->pat->nslices;
    const uint32_t npad = nsl * SLICE * BS;  // lane-interleaved vector length
    const unsigned gs = spmv_grid(a);
    const unsigned gu = std::max(1u, std::min<unsigned>(cdiv(npad, UPD_T * 4), UPD_MAX));
    const uint32_t mode = args.precision == FH_PREC_FP64 ? M_REF
                        : args.precision == FH_PREC_FP64_TRUE ? M_TRUE
                        : args.precision == FH_PREC_MIXED ? M_MIXED
                                                          : M_MIXED_IR;
    const bool uses32 = mode == M_MIXED || mode == M_MIXED_IR;
    // FP64-value phase of the true-residual modes: the plain product by default.
    // With double-double replacements it converges in fewer iterations than the
    // difference form (3237 vs 3822 on the 1M-dof cantilever) and reads no row
    // sums; FH_CG_DIFF64=1 selects the difference form.
    const bool diff64 = DIFF && env_int("FH_CG_DIFF64", 0) != 0;
    fh_matrix* am = const_cast<fh_matrix*>(a);
    if (noalloc) {
        FH_CHECK(a->pat->xcd_ready && (!uses32 || a->vals32.p) &&
                     (!(uses32 || diff64) || a->rowsum64.p) && (!uses32 || a->rowsum32.p) &&
                     (ORD || BS != 3 || a->tail64.p) && (!uses32 || BS != 3 || a->quad32.p),
                 "the allocation-free solve needs the matrix prepared for this precision: "
                 "solve it once with fh_solve_dev (or fh_solve) in the same precision first");
    } else {
        if (uses32) ensure_vals32(am);
        if (uses32 || diff64) ensure_rowsum(am, uses32);
        if (!ORD && BS == 3) ensure_tail(am);
        if (uses32 && BS == 3) ensure_quad32(am);
    }
    XcdMap xm{};
    unsigned gc = 8;  // k_cg_spmv blocks
    xcd_map(*a->pat, s, xm, gc);
    const auto t_wall0 = std::chrono::steady_clock::now();
    const CgMat A = cgmat(a);

    const int M = std::max(1, env_int("FH_CG_CHUNK", 16));
    const bool use_graph = !noalloc && env_int("FH_CG_GRAPH", 1) != 0;
    // optional per-launch timing of the SpMV kernel: a start and a stop event
    // carried by the kernel's own dispatch (hipExtLaunchKernelGGL), so the pair
    // brackets its execution on this stream and nothing else
    const bool time_spmv = !noalloc && env_int("FH_TIME_SPMV", 0) != 0;
    // true-residual modes: Ap by recurrence (FH_CG_APREC=1) except right after
    // the initial residual or a replacement, where the exact product restarts it
    const bool aprec = DIFF && env_int("FH_CG_APREC", 0) != 0;
    // workspace and graphs: kept on the matrix per kernel configuration (the
    // graphs reference only the matrix and the workspace), so a repeated solve
    // of this operator allocates and captures nothing
    // scalar Jacobi with the library's own D^{-1} (never zero): no stored r,
    // its entries are recovered from z (k_cg_update); the reference's ordered
    // arithmetic and a caller's D^{-1} keep r
    const bool keep_r = PRE != 0 || ORD || args.inv_diag_override;
    // x updated every second iteration (bitwise the same x): scalar Jacobi,
    // outside the reference's ordered arithmetic
    const bool lazy_x = PRE == 0 && !ORD;
    CgCache* cc = nullptr;
    if (!noalloc) {
        const uint64_t key = (uint64_t)BS | (uint64_t)ORD << 4 | (uint64_t)DIFF << 5 |
                             (uint64_t)PRE << 6 | (uint64_t)diff64 << 8 | (uint64_t)aprec << 9 |
                             (uint64_t)keep_r << 10 | (uint64_t)M << 16 | (uint64_t)gc << 40;
        std::unique_ptr<CgCache>& slot = am->cg_cache[key];
        if (!slot) {
            slot = std::make_unique<CgCache>();
            slot->ws.alloc(dev, cg_work(a, PRE, DIFF, gc, nullptr, nullptr));
        }
        cc = slot.get();
        ws = cc->ws.p;
    }
    CgWork w{};
    cg_work(a, PRE, DIFF, gc, ws, &w);
    double* const x = w.xl;
    double* const rvec = keep_r ? w.r : nullptr;
    if (PRE == 1) to_li<9>(args.block_dinv, w.dinv, nrows, nsl, 0.0, s);
    else if (args.inv_diag_override) to_li<BS>(args.inv_diag_override, w.dinv, nrows, nsl, 1.0, s);
    else {
        hipLaunchKernelGGL(k_inv_diag_li<BS>, dim3(std::max(1u, cdiv(npad, 256))), dim3(256), 0, s, A,
                           w.dinv);
        FH_LAUNCH_CHECK("inv_diag_li");
    }
    to_li<BS>(args.b, w.bl, nrows, nsl, 0.0, s);
    if (args.x0) to_li<BS>(args.x0, x, nrows, nsl, 0.0, s);
    else fill_f64(x, npad, 0.0, s);
    fill_f64(w.zp, 2 * (size_t)npad, 0.0, s);

    // host mirror of the state: pinned, or (allocation-free calls) a static
    // pageable copy per device
    static CgState hs_static[64];
    FH_CHECK(!noalloc || (dev >= 0 && dev < 64), "device index out of range");
    CgState* hs = noalloc ? &hs_static[dev] : host_state(dev);
    std::memset(hs, 0, sizeof *hs);
    hs->tol = args.tol;
    hs->max_iter = args.max_iter;
    hs->delta = env_dbl("FH_RR_DELTA", 1e-1);
    hs->delta64 = env_dbl("FH_RR_DELTA64", 1e-1);
    hs->sw = env_dbl("FH_MIXED_SWITCH", 1e-3);
    hs->mode = mode;
    hs->restart = mode == M_MIXED_IR ? 1 : 0;
    hs->rep_mark = 1;
    FH_HIP(hipMemcpyAsync(w.st, hs, sizeof(CgState), hipMemcpyHostToDevice, s));

    std::unique_ptr<Timer> tk;
    if (!noalloc) tk = std::make_unique<Timer>(s);
    // r0 = b - A x0 (FP64; double-double in the true-residual modes),
    // z0 = D^{-1} r0, partials (r.r, r.z), b.b
    hipLaunchKernelGGL((k_residual<BS, ORD, DIFF, PRE>), dim3(gs), dim3(256), 0, s, A, w.bl, x,
                       w.dinv, rvec, w.zp, w.redB, w.bb, w.fl);
    FH_LAUNCH_CHECK("residual");
    // the next spmv follows k_residual (gs partials) rather than an update (gu)
    bool after_res = true;

    const bool trace = env_int("FH_CG_TRACE", 0) != 0;  // per-chunk state on stderr
    std::vector<hipEvent_t> ev;
    double ms_hi = 0.0, ms_lo = 0.0;
    uint64_t n_hi = 0, n_lo = 0;
    if (time_spmv) {
        ev.resize(2 * M);
        for (auto& e : ev) FH_HIP(hipEventCreate(&e));
    }
    auto launch_chunk = [&](bool res_first, bool lowk) {
        for (int m = 0; m < M; ++m) {
            const int nin = (m == 0 && res_first) ? (int)gs : (int)gu;
            const bool rec = aprec && !(m == 0 && res_first);
            hipEvent_t e0 = time_spmv ? ev[2 * m] : nullptr, e1 = time_spmv ? ev[2 * m + 1] : nullptr;
#define FH_SPMV(D, L, R)                                                                         \
    do {                                                                                         \
        if (time_spmv)                                                                           \
            hipExtLaunchKernelGGL((k_cg_spmv<BS, ORD, D, L, R>), dim3(gc), dim3(CG_T), 0, s, e0, \
                                  e1, 0, A, xm, (const double*)w.zp, w.Ap, w.st,                 \
                                  (const double2*)w.redB, nin, (const double*)w.bb,              \
                                  (const double*)w.fl, w.redA);                                  \
        else                                                                                     \
            hipLaunchKernelGGL((k_cg_spmv<BS, ORD, D, L, R>), dim3(gc), dim3(CG_T), 0, s, A, xm, \
                               w.zp, w.Ap, w.st, w.redB, nin, w.bb, w.fl, w.redA);               \
    } while (0)
            if constexpr (DIFF) {
                if (rec) {
                    if (lowk) FH_SPMV(true, true, true);
                    else if (diff64) FH_SPMV(true, false, true);
                    else FH_SPMV(false, false, true);
                } else {
                    if (lowk) FH_SPMV(true, true, false);
                    else if (diff64) FH_SPMV(true, false, false);
                    else FH_SPMV(false, false, false);
                }
            } else {
                FH_SPMV(false, false, false);
            }
#undef FH_SPMV
            hipLaunchKernelGGL((k_cg_update<ORD, PRE, BS == 3 && !ORD>), dim3(gu), dim3(UPD_T), 0, s, npad, x,
                               rvec, w.Ap, w.zp, w.dinv, w.st, w.redA, (int)gc, w.redB,
                               lazy_x ? 1 : 0);
        }
    };
    // complete a deferred x before it is read (replacement, end of the solve)
    auto flush_x = [&] {
        if (!lazy_x) return;
        hipLaunchKernelGGL(k_x_flush, dim3(std::max(1u, std::min<unsigned>(cdiv(npad, 256), 2048))),
                           dim3(256), 0, s, npad, x, (const double*)w.zp, (const CgState*)w.st);
        hipLaunchKernelGGL(k_x_base, dim3(1), dim3(64), 0, s, w.st);
        FH_LAUNCH_CHECK("x_flush");
    };
    uint64_t launched = 0;
    bool first_chunk = true;
    bool lowk = mode == M_MIXED_IR;  // operator precision of the next chunk
    for (;;) {
        const uint64_t j0 = hs->it_upd;  // iterations done before this chunk
        // the first spmv after init / replacement reads gs residual partials;
        // later ones read gu update partials
        if (use_graph && !time_spmv && !first_chunk && !after_res) {
            const int g = lowk ? 1 : 0;
            if (!cc->exec[g]) {
                FH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                launch_chunk(false, lowk);
                FH_HIP(hipStreamEndCapture(s, &cc->graph[g]));
                FH_HIP(hipGraphInstantiate(&cc->exec[g], cc->graph[g], nullptr, nullptr, 0));
            }
            FH_HIP(hipGraphLaunch(cc->exec[g], s));
        } else {
            launch_chunk(after_res, lowk);
            FH_LAUNCH_CHECK("cg_chunk");
        }
        first_chunk = false;
        after_res = false;
        launched += M;
        FH_HIP(hipMemcpyAsync(hs, w.st, sizeof(CgState), hipMemcpyDeviceToHost, s));
        FH_HIP(hipStreamSynchronize(s));
        if (time_spmv && hs->status == ST_RUNNING && !hs->need_rep && hs->it_upd == j0 + M) {
            // every SpMV of this chunk did a full product: attribute it by precision
            for (int m = 0; m < M; ++m) {
                float ms = 0.f;
                FH_HIP(hipEventElapsedTime(&ms, ev[2 * m], ev[2 * m + 1]));
                if (lowk) {
                    ms_lo += ms;
                    ++n_lo;
                } else {
                    ms_hi += ms;
                    ++n_hi;
                }
            }
        }
        if (hs->status != ST_RUNNING) break;
        if (hs->want_low && !hs->stalled) lowk = true;
        if (hs->stalled) lowk = false;
        if (trace)
            std::fprintf(stderr,
                         "cg: it %llu rel %.3e rep %llu restarts %llu low %d need_rep %d stalled %d\n",
                         (unsigned long long)hs->it_upd, hs->rel_last, (unsigned long long)hs->n_rep,
                         (unsigned long long)hs->n_restart, (int)lowk, (int)hs->need_rep,
                         (int)hs->stalled);
        if (hs->need_rep) {
            flush_x();
            // keep the iterate this replacement evaluates (slot = count before k_mark)
            d2d((hs->n_rep & 1) ? w.xs1 : w.xs0, x, npad, s);
            // residual replacement: r = b - A x (double-double evaluation)
            hipLaunchKernelGGL((k_residual<BS, false, true, PRE>), dim3(gs), dim3(256), 0, s, A,
                               w.bl, x, w.dinv, rvec, w.zp, w.redB, (double*)nullptr, w.fl);
            hipLaunchKernelGGL(k_mark, dim3(1), dim3(64), 0, s, w.st);
            FH_LAUNCH_CHECK("replacement");
            after_res = true;
        }
        if (launched > args.max_iter + (hs->n_rep + 2) * (uint64_t)M + 64) break;  // safety net
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    flush_x();

    FH_HIP(hipMemcpyAsync(hs, w.st, sizeof(CgState), hipMemcpyDeviceToHost, s));
    FH_HIP(hipStreamSynchronize(s));
    if (DIFF && hs->restore) d2d(x, hs->restore == 1 ? w.xs0 : w.xs1, npad, s);
    // the solution in the caller's layout, and its true residual (double-double)
    hipLaunchKernelGGL(k_from_li<BS>, dim3(std::max(1u, cdiv(n, 256))), dim3(256), 0, s, x, args.x,
                       nrows);
    hipLaunchKernelGGL(k_true_res<BS>, dim3(gs), dim3(256), 0, s, A, w.bl, x, w.tr);
    hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(256), 0, s, (const double2*)w.tr, (int)gs, w.sum);
    FH_LAUNCH_CHECK("true_res");
    double2 trs{0.0, 0.0};
    FH_HIP(hipMemcpyAsync(&trs, w.sum, sizeof trs, hipMemcpyDeviceToHost, s));
    FH_HIP(hipMemcpyAsync(hs, w.st, sizeof(CgState), hipMemcpyDeviceToHost, s));
    FH_HIP(hipStreamSynchronize(s));
    const double kms = tk ? tk->stop_ms() : 0.0;
    const double tt = trs.x, bsq = trs.y;
    const double wall =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wall0).count();

    int code = FH_OK;
    if (hs->status == ST_BREAKDOWN) code = FH_ERR_BREAKDOWN;
    else if (hs->status != ST_CONVERGED) code = FH_ERR_NOT_CONVERGED;
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->iterations = hs->status == ST_RUNNING ? hs->it_upd : hs->iters;
        stats->iterations_low = hs->n_low;
        stats->replacements = hs->n_rep;
        stats->final_residual = hs->res;
        stats->rel_residual = hs->rel;
        stats->true_rel_residual = bsq > 0 ? std::sqrt(tt) / std::sqrt(bsq) : std::sqrt(tt);
        stats->residual_floor = hs->floor_rel;
        stats->wall_ms = wall;
        stats->kernel_ms = tk ? kms : wall;
        stats->spmv_bytes = diff64 ? diff_spmv_bytes(a, 8) : spmv_algorithmic_bytes(a, 8, 8);
        stats->spmv32_bytes = diff_spmv_bytes(a, 4);
        // SpMV + the vector passes of the fused loop: the spmv kernel gathers
        // (z, p_old) pairs (one vector beyond the SpMV's x) and writes Ap, the
        // update reads Ap, the pairs and D^{-1}, reads and writes the pairs, x
        // (every second iteration when deferred) and r (when kept)
        const double dinv_b = PRE == 0 ? 8.0 * n : 24.0 * n;
        const double upd_vectors = 1.0 + 4.0 + (lazy_x ? 1.0 : 2.0) + (keep_r ? 2.0 : 0.0);
        stats->iter_bytes = stats->spmv_bytes + 8.0 * n * (1.0 + upd_vectors) + dinv_b;
        const double it = (double)std::max<uint64_t>(stats->iterations, 1);
        const double tms = std::max(stats->kernel_ms, 1e-9);
        stats->effective_bw_gbps = stats->iter_bytes * it / (tms * 1e-3) / 1e9;
        const double flops = 2.0 * (double)a->nnz_scalar + 12.0 * n;
        stats->gflops = flops * it / (tms * 1e-3) / 1e9;
        stats->spmv_avg_ms = n_hi ? ms_hi / (double)n_hi : 0.0;
        stats->spmv_timed = n_hi;
        stats->spmv32_avg_ms = n_lo ? ms_lo / (double)n_lo : 0.0;
        stats->spmv32_timed = n_lo;
        stats->device = dev;
        stats->precision = args.precision;
        stats->status = code;
    }
    return code;
}

int solve_device(const fh_matrix* a, const SolveArgs& args, fh_solve_stats* st, uint8_t* ws) {
    const bool ref = args.precision == FH_PREC_FP64;
    const bool ord = ref && args.order == FH_ORDER_ORDERED;
    const int pre = args.block_dinv ? 1 : 0;
    FH_CHECK(!(pre && a->bs != 3), "block Jacobi needs 3x3 blocks");
    FH_CHECK(args.precision >= 0 && args.precision <= 3, "unknown precision");
#define FH_RUN(BS, PRE)                                                              \
    return ref ? (ord ? run_cg<BS, true, false, PRE>(a, args, st, ws)               \
                      : run_cg<BS, false, false, PRE>(a, args, st, ws))             \
               : run_cg<BS, false, true, PRE>(a, args, st, ws)
    if (a->bs == 1) FH_RUN(1, 0);
    if (pre) FH_RUN(3, 1);
    FH_RUN(3, 0);
#undef FH_RUN
}

// workspace bytes of a solve of `a` (the allocation-free entry point's size check)
size_t solve_workspace_bytes(const fh_matrix* a, const SolveArgs& args) {
    XcdMap xm{};
    unsigned gc = 8;
    xcd_map(*a->pat, device(a->dev).stream, xm, gc);
    const bool diff = args.precision != FH_PREC_FP64;
    return cg_work(a, args.block_dinv ? 1 : 0, diff, gc, nullptr, nullptr);
}

// ||b - A (xh + xl)|| / ||b|| in double-double for natural-layout device
// vectors (xl may be null); r (may be null) receives the residual rounded to
// FP64.  Synchronises the stream.
static void residual_pair(const fh_matrix* a, const double* b, const double* xh,
                          const double* xl, double* r, hipStream_t s, double& rnorm,
                          double& bnorm) {
    const CgMat A = cgmat(a);
    const unsigned g = spmv_grid(a);
    DBuf<double2> red(a->dev, g);
    if (a->bs == 1)
        hipLaunchKernelGGL(k_residual_pair<1>, dim3(g), dim3(256), 0, s, A, b, xh, xl, r, red.p);
    else
        hipLaunchKernelGGL(k_residual_pair<3>, dim3(g), dim3(256), 0, s, A, b, xh, xl, r, red.p);
    FH_LAUNCH_CHECK("residual_pair");
    std::vector<double2> h(g);
    d2h(h.data(), red.p, g, s);
    FH_HIP(hipStreamSynchronize(s));
    double tt = 0.0, bsq = 0.0;
    for (const auto& v : h) {
        tt += v.x;
        bsq += v.y;
    }
    rnorm = std::sqrt(tt);
    bnorm = std::sqrt(bsq);
}

double true_rel_residual_pair(const fh_matrix* a, const double* b, const double* xh,
                              const double* xl, double* r, hipStream_t s) {
    double rn, bn;
    residual_pair(a, b, xh, xl, r, s, rn, bn);
    return bn > 0.0 ? rn / bn : rn;
}

// Iterative refinement with the solution carried in double-double: r = b -
// A (xh + xl) evaluated in double-double, A d = r solved by the FP64
// true-residual CG to inner_tol, (xh, xl) += d, until the double-double
// residual is below `target`.  xh is then hi + lo rounded to FP64, so its own
// true residual is the residual of the FP64 vector nearest to the solution
// found: a measured version of the floor that the FP64 solves estimate.
static int refine_dd(const fh_matrix* a, const double* b, const double* x0, double target,
                     uint64_t max_outer, double inner_tol, uint64_t max_iter, double* xh,
                     double* xl, fh_refine_stats* out) {
    DeviceCtx& d = device(a->dev);
    hipStream_t s = d.stream;
    const uint64_t n = a->n;
    if (x0) d2d(xh, x0, n, s);
    else fill_f64(xh, n, 0.0, s);
    fill_f64(xl, n, 0.0, s);
    DBuf<double> r(a->dev, n), dx(a->dev, n);
    uint64_t k = 0, iters = 0;
    double rn, bn;
    residual_pair(a, b, xh, xl, r.p, s, rn, bn);
    double rel = bn > 0.0 ? rn / bn : rn;
    int code = FH_OK;
    const unsigned g = std::max(1u, std::min(cdiv(n, 256), 4096u));
    for (; k < max_outer && rel >= target && rn > 0.0; ++k) {
        // the correction system is scaled to ||r|| ~ 1 by a power of two (exact):
        // cg.rs's initial test on the absolute ||r0|| must not end it early
        const double sc = std::ldexp(1.0, -std::ilogb(rn));
        hipLaunchKernelGGL(k_scale, dim3(g), dim3(256), 0, s, r.p, sc, n);
        SolveArgs sa{r.p, nullptr, dx.p, inner_tol, max_iter, FH_PREC_FP64_TRUE, FH_ORDER_FAST};
        fh_solve_stats st{};
        code = solve_device(a, sa, &st);
        iters += st.iterations;
        if (code == FH_ERR_BREAKDOWN) break;
        hipLaunchKernelGGL(k_dd_add, dim3(g), dim3(256), 0, s, xh, xl, dx.p, 1.0 / sc, n);
        FH_LAUNCH_CHECK("dd_add");
        residual_pair(a, b, xh, xl, r.p, s, rn, bn);
        rel = bn > 0.0 ? rn / bn : rn;
    }
    if (out) {
        out->outer_steps = k;
        out->cg_iterations = iters;
        out->rel_residual_pair = rel;
        out->rel_residual_rounded = true_rel_residual_pair(a, b, xh, nullptr, nullptr, s);
    }
    if (code == FH_ERR_BREAKDOWN) return code;
    return rel < target ? FH_OK : FH_ERR_NOT_CONVERGED;
}

// Scratch memory of the allocation-free solve (fh_solve_dev_noalloc): one
// copy of this array per GPU, held until the process ends.  In host code
// `scratch` is a host-side stand-in, not a device address: host code only
// uses the address hipGetSymbolAddress returns for the current device
// (scratch_ptr), in kernel arguments and copies.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)FH_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

// Device address of `scratch` on device dev, looked up once per device with
// that device current (the caller has made it so through device(dev)).
// Exported calls hold the library's call lock, so the table needs no lock of
// its own.
static uint8_t* scratch_ptr(int dev) {
    static uint8_t* ptr[64] = {};
    FH_CHECK(dev >= 0 && dev < 64, "device index out of range");
    if (!ptr[dev]) {
        void* p = nullptr;
        FH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        FH_CHECK(p != nullptr, "no device address for the solver scratch");
        ptr[dev] = static_cast<uint8_t*>(p);
    }
    return ptr[dev];
}

}  // namespace fh

using namespace fh;

extern "C" {

int fh_solve_dev_noalloc(const fh_matrix* a, const double* b_dev, const double* x0_dev,
                         double tol, uint64_t max_iter, int precision, double* x_dev,
                         fh_solve_stats* stats) {
    int code = FH_OK;
    const int g = guard([&] {
        FH_CHECK(a && b_dev && x_dev, "null argument");
        FH_CHECK(precision >= 0 && precision <= 3, "unknown precision");
        device(a->dev);
        uint8_t* ws = scratch_ptr(a->dev);
        FH_CHECK(a->pat->xcd_ready,
                 "the allocation-free solve needs the matrix prepared for this precision: "
                 "solve it once with fh_solve_dev (or fh_solve) in the same precision first");
        SolveArgs sa{b_dev, x0_dev, x_dev, tol, max_iter, precision,
                     env_int("FH_ORDER", FH_ORDER_FAST)};
        const size_t need = solve_workspace_bytes(a, sa);
        if (need > NOALLOC_SCRATCH_BYTES) {
            char m[256];
            std::snprintf(m, sizeof m,
                          "the allocation-free solve holds %d MiB of scratch per GPU; this "
                          "system needs %.1f MiB (use fh_solve_dev)",
                          (int)FH_NOALLOC_SCRATCH_MIB, need / 1048576.0);
            throw Error(FH_ERR_BAD_INPUT, m);
        }
        code = solve_device(a, sa, stats, ws);
        if (code == FH_ERR_NOT_CONVERGED)
            set_last_error("CG failed to converge within the iteration limit");
        else if (code == FH_ERR_BREAKDOWN)
            set_last_error("Breakdown in CG: curvature p^T A p is near zero");
    });
    return g != FH_OK ? g : code;
}

int fh_true_residual_dev(const fh_matrix* a, const double* b_dev, const double* x_hi_dev,
                         const double* x_lo_dev, double* rel_out) {
    return guard([&] {
        FH_CHECK(a && b_dev && x_hi_dev && rel_out, "null argument");
        DeviceCtx& d = device(a->dev);
        *rel_out = true_rel_residual_pair(a, b_dev, x_hi_dev, x_lo_dev, nullptr, d.stream);
    });
}

int fh_solve_refined_dev(const fh_matrix* a, const double* b_dev, const double* x0_dev,
                         double target, uint64_t max_outer, double inner_tol, uint64_t max_iter,
                         double* x_hi_dev, double* x_lo_dev, fh_refine_stats* stats) {
    int code = FH_OK;
    const int g = guard([&] {
        FH_CHECK(a && b_dev && x_hi_dev && x_lo_dev, "null argument");
        FH_CHECK(target > 0.0 && inner_tol > 0.0 && max_outer > 0, "bad refinement parameters");
        device(a->dev);
        code = refine_dd(a, b_dev, x0_dev, target, max_outer, inner_tol, max_iter, x_hi_dev,
                         x_lo_dev, stats);
        if (code == FH_ERR_NOT_CONVERGED)
            set_last_error("refinement did not reach the target residual");
        else if (code == FH_ERR_BREAKDOWN)
            set_last_error("Breakdown in CG: curvature p^T A p is near zero");
    });
    return g != FH_OK ? g : code;
}

int fh_matrix_spmv_dev(const fh_matrix* a, const double* x, double* y, int order) {
    return guard([&] {
        FH_CHECK(a && x && y, "null argument");
        DeviceCtx& d = device(a->dev);
        spmv_f64(a, x, y, order, d.stream);
        FH_HIP(hipStreamSynchronize(d.stream));
    });
}

int fh_matrix_spmv(const fh_matrix* a, const double* x, double* y, int order) {
    return guard([&] {
        FH_CHECK(a && x && y, "null argument");
        DeviceCtx& d = device(a->dev);
        DBuf<double> dx(a->dev, a->n), dy(a->dev, a->n);
        h2d(dx.p, x, a->n, d.stream);
        spmv_f64(a, dx.p, dy.p, order, d.stream);
        d2h(y, dy.p, a->n, d.stream);
        FH_HIP(hipStreamSynchronize(d.stream));
    });
}

int fh_solve_dev(const fh_matrix* a, const double* b_dev, const double* x0_dev, double tol,
                 uint64_t max_iter, int precision, double* x_dev, fh_solve_stats* stats) {
    int code = FH_OK;
    const int g = guard([&] {
        FH_CHECK(a && b_dev && x_dev, "null argument");
        device(a->dev);
        SolveArgs sa{b_dev, x0_dev, x_dev, tol, max_iter, precision,
                     env_int("FH_ORDER", FH_ORDER_FAST)};
        code = solve_device(a, sa, stats);
        if (code == FH_ERR_NOT_CONVERGED)
            set_last_error("CG failed to converge within the iteration limit");
        else if (code == FH_ERR_BREAKDOWN)
            set_last_error("Breakdown in CG: curvature p^T A p is near zero");
    });
    return g != FH_OK ? g : code;
}

int fh_solve(const fh_matrix* a, const double* b, const double* x0, double tol, uint64_t max_iter,
             int precision, double* x_out, fh_solve_stats* stats) {
    int code = FH_OK;
    const int g = guard([&] {
        FH_CHECK(a && b && x_out, "null argument");
        DeviceCtx& d = device(a->dev);
        DBuf<double> db(a->dev, a->n), dx(a->dev, a->n), dx0;
        h2d(db.p, b, a->n, d.stream);
        if (x0) {
            dx0.alloc(a->dev, a->n);
            h2d(dx0.p, x0, a->n, d.stream);
        }
        code = fh_solve_dev(a, db.p, dx0.p, tol, max_iter, precision, dx.p, stats);
        d2h(x_out, dx.p, a->n, d.stream);
        FH_HIP(hipStreamSynchronize(d.stream));
    });
    return g != FH_OK ? g : code;
}

}  // extern "C"
