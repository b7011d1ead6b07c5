// fh_batched.hip — many independent Poisson P1 problems with up to four
// right-hand sides each, solved together by block CG (new; no counterpart in
// the reference, which solves one system per solve_poisson_2d call,
// poisson_solver.rs:26).
//
// Layout.  The instances form one block-diagonal system over the disjoint
// union of their unit_square meshes (generators.rs:24-55), assembled and
// condensed by the single-system kernels.  Each instance's rows are padded to
// a multiple of 256 (one row tile, four SELL slices) with isolated, fixed
// rows, so every tile belongs to one instance.  Vectors hold the four
// right-hand-side columns interleaved per row ([row][4], 32 bytes): one
// product pass reads each matrix entry once for all of them.
//
// Block CG per instance with A-orthonormal directions (O'Leary 1980 in the
// breakdown-free form of Dubrulle 2001 / Ji & Li 2017), Jacobi D^{-1}:
//     W = Z + P beta,  Q_W = A W,  W^T A W = L L^T,  P = W L^{-T},  Q = Q_W L^{-T},
//     alpha = P^T R,  X += P alpha,  R -= Q alpha,  Z = D^{-1} R,  beta = -Q^T Z.
// The columns share one Krylov space, so each converges in fewer iterations
// than an independent CG on it; with one column this is cg.rs:40-149.  The
// Cholesky factor drops a direction whose A-norm is not above 1e-12 of its
// own outside the span of the others (rank-revealing deflation: the block
// shrinks instead of breaking down), and a column whose residual reaches tol
// leaves the block (its X and R stay; its W column is zero from then on).
//
// One block iteration is five launches: W = Z + P beta over the active tiles
// (so the product gathers one vector, not two), the product Q_W = A W, a
// one-workgroup-per-instance reduction (L, alpha), the update over the
// active tiles, and a second reduction (convergence per column, beta).  The
// s x s algebra runs once per instance, in fixed order, so runs are
// bitwise reproducible.  The tile kernels have a fixed grid that strides over
// a device-side work list of the tiles of unfinished instances, rebuilt by a
// compaction kernel after every chunk of iterations: a chunk is captured once
// into a HIP graph and replayed with fixed arguments, and the host queues
// the next chunk before it looks at the previous chunk's count of unfinished
// instances (so it never waits with the device idle).  Kernel boundaries are
// the device-wide synchronisation: a single persistent launch whose
// workgroups meet at barriers inside the kernel was measured 1.2-2.2x slower
// here, because each agent-scope barrier costs an L2 write-back and
// invalidate (buffer_wbl2 / buffer_inv) on this 8-XCD part.
//
// Precision: FH_PREC_FP64 reads FP64 matrix values; the mixed modes read the
// FP32 copy (the unit-square P1 stiffness entries are small integers and
// halves, so the copy is exact here).  Vectors, reductions and the s x s
// algebra are FP64; convergence is on the recursive residual of each column
// (cg.rs:118-122), and the true FP64 residual b - A x of every column is
// evaluated at the end and reported.
#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <string>
#include <vector>

#include "fh_internal.hpp"

namespace fh {

constexpr int NR = 4;    // columns stored per row (block size s <= NR)
constexpr int BT = 256;  // threads per workgroup (4 waves)
constexpr int TS = 16;   // SELL slices per row tile: each wave takes 4 of them
constexpr int PU = 8;    // stored entries per step of the product (4: 729 ms, 8: 671 ms of
                         // product time in the 256-instance mixed solve)
constexpr int TR = TS * 64;  // rows per tile (each lane accumulates 4 rows before the reduction)
constexpr int NSYM = 10; // entries of a symmetric 4 x 4 matrix

template <class T>
struct SellB {
    const uint32_t* soff;
    const uint32_t* rlen;
    const uint32_t* cols;
    const T* vals;
    uint32_t nrows;
    uint32_t nslices;
};

__device__ __forceinline__ int sym(int a, int b) {  // packed index of (a, b) in a symmetric 4 x 4
    const int i = a < b ? a : b, j = a < b ? b : a;
    return i * NR - (i * (i - 1)) / 2 + (j - i);
}

// Sum over the workgroup of n values per thread (fixed butterfly + wave
// order); every thread gets the sums.
template <int N>
__device__ __forceinline__ void wg_sum(double* v, double* sh /* [4][N] */) {
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = fhd::wave_sum(v[k]);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < N; ++k) sh[w * N + k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = (sh[k] + sh[N + k]) + (sh[2 * N + k] + sh[3 * N + k]);
    __syncthreads();
}

// Sum of the partial records [slot0, slot0 + w) of width N (fixed order).
template <int N>
__device__ __forceinline__ void reduce_slots(const double* part, uint32_t slot0, uint32_t w,
                                             double* out, double* sh) {
    double v[N];
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = 0.0;
    for (uint32_t t = threadIdx.x; t < w; t += BT)
#pragma unroll
        for (int k = 0; k < N; ++k) v[k] += part[(uint64_t)(slot0 + t) * N + k];
    wg_sum<N>(v, sh);
#pragma unroll
    for (int k = 0; k < N; ++k) out[k] = v[k];
}

// Cholesky factor of a symmetric 4 x 4 (packed) that drops dependent
// directions: column j is kept when the part of it not in the span of the
// kept earlier ones, d_j = G_jj - sum_k L_jk^2, exceeds 1e-12 G_jj (and is
// positive); a dropped direction has a zero column and row in L.  Returns the
// kept mask.  Only directions in `cand` are considered.
__device__ uint32_t chol_drop4(const double* g, uint32_t cand, double L[NR][NR]) {
    uint32_t kept = 0;
    for (int i = 0; i < NR; ++i)
        for (int j = 0; j < NR; ++j) L[i][j] = 0.0;
    for (int j = 0; j < NR; ++j) {
        if (!((cand >> j) & 1)) continue;
        const double gjj = g[sym(j, j)];
        double d = gjj;
        for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
        if (!(gjj > 0.0) || !(d > 1e-12 * gjj)) continue;
        const double l = sqrt(d);
        L[j][j] = l;
        kept |= 1u << j;
        for (int i = j + 1; i < NR; ++i) {
            if (!((cand >> i) & 1)) continue;
            double t = g[sym(i, j)];
            for (int k = 0; k < j; ++k) t -= L[i][k] * L[j][k];
            L[i][j] = t / l;
        }
    }
    return kept;
}

// v = L^{-1} u for a row vector u of the W basis (the P = W L^{-T} change of
// basis, one row): forward substitution over the kept directions.
__device__ __forceinline__ void lsolve_row(const double L[NR][NR], uint32_t kept, const double* u,
                                           double* v) {
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        double t = u[j];
#pragma unroll
        for (int k = 0; k < j; ++k) t = fma(-L[j][k], v[k], t);
        v[j] = ((kept >> j) & 1) ? t / L[j][j] : 0.0;
    }
}

__device__ __forceinline__ void ld4(const double* p, double* v) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 a = *reinterpret_cast<const d2*>(p), b = *reinterpret_cast<const d2*>(p + 2);
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
    v[3] = b.y;
}
__device__ __forceinline__ void st4(double* p, const double* v) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    d2 a, b;
    a.x = v[0];
    a.y = v[1];
    b.x = v[2];
    b.y = v[3];
    *reinterpret_cast<d2*>(p) = a;
    *reinterpret_cast<d2*>(p + 2) = b;
}

// Position of this workgroup in a stride loop over a grid that is a multiple
// of 8: the hardware deals workgroup b to XCD b mod 8, so XCD x gets the
// contiguous positions [x G/8, (x+1) G/8) of every stride: consecutive tiles
// (grid neighbours whose (z, p) rows the product gathers) share one L2.
__device__ __forceinline__ uint32_t xcd_pos() {
    const uint32_t per = gridDim.x / 8;
    return (blockIdx.x % 8) * per + blockIdx.x / 8;
}

// four values of a vector row as doubles, and back (rounded to TV)
__device__ __forceinline__ void ld4(const float* p, double* v) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 a = *reinterpret_cast<const f4*>(p);
    v[0] = a.x;
    v[1] = a.y;
    v[2] = a.z;
    v[3] = a.w;
}
__device__ __forceinline__ void st4(float* p, const double* v) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    f4 a;
    a.x = (float)v[0];
    a.y = (float)v[1];
    a.z = (float)v[2];
    a.w = (float)v[3];
    *reinterpret_cast<f4*>(p) = a;
}
// v rounded to the storage type TV (identity for double)
template <class TV>
__device__ __forceinline__ double rnd(double v) {
    return (double)(TV)v;
}

// w = z + p beta on the active right-hand-side columns (zero on the others)
__device__ __forceinline__ void form_w(const double* zc, const double* pc, const double beta[NR][NR],
                                       uint32_t act, double* w) {
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        double t = zc[k];
#pragma unroll
        for (int m = 0; m < NR; ++m) t = fma(pc[m], beta[m][k], t);
        w[k] = ((act >> k) & 1) ? t : 0.0;
    }
}

// ---------------------------------------------------------------- kernels
// Per-instance state, written by the one-workgroup reduction kernels and read
// by the tile kernels of the next launch.
struct BbState {
    uint32_t* act;     // [n_inst] right-hand-side columns still iterating
    uint32_t* kept;    // [n_inst] directions kept by the last orthonormalisation
    uint32_t* it;      // [n_inst] block iterations done
    uint32_t* done;    // [n_inst] 1 when no column iterates any more
    double* bnorm;     // [n_inst][4]
    double* beta;      // [n_inst][16]  W = Z + P beta
    double* L;         // [n_inst][16]  W^T A W = L L^T (kept directions)
    double* alpha;     // [n_inst][16]  X += P alpha
    uint32_t* iters;   // [n_inst][4]   iteration at which each column converged
    uint8_t* status;   // [n_inst][4]   0 iterating, 1 converged, 2 not converged, 3 breakdown
    const uint32_t* tile_inst;  // [n_tiles]
    const uint32_t* inst_t0;    // [n_inst] first tile
    const uint32_t* inst_t1;    // one past the last
    uint32_t* active;  // [n_tiles] tiles of unfinished instances, then [n_inst] instances
    uint32_t* nact;    // [2] active tiles, active instances
    uint32_t n_tiles, n_inst;
    int s;
    double tol;
    uint32_t max_iter;
};

constexpr int NA = NSYM + NR * NR;  // W^T A W (packed, from the product), W^T R (from W = Z + P beta)
constexpr int NB = NR * NR + NR;    // update partials: Q^T Z, R^T R (diagonal)

// R = B (X0 = 0), Z = D^{-1} R (FP64 rows of four), P = 0 (rows of four in
// the storage type TV); tile partials R^T R and B^T B (diagonals) in the
// first 2 NR entries of an NB record.  R itself is not stored: the kernels
// that need it take R = Z / D^{-1} (D^{-1} has no zero entry)
template <class TV>
__global__ __launch_bounds__(BT) void k_bb_init(const double* __restrict__ b,
                                                const double* __restrict__ dinv,
                                                double* __restrict__ z, TV* __restrict__ p,
                                                double* __restrict__ partB, int s) {
    __shared__ double sh[4 * 2 * NR];
    const uint32_t t = blockIdx.x;
    double acc[2 * NR];
#pragma unroll
    for (int k = 0; k < 2 * NR; ++k) acc[k] = 0.0;
    for (int sub = 0; sub < TR / BT; ++sub) {
        const uint64_t row = (uint64_t)t * TR + sub * BT + threadIdx.x;
        double bv[NR], zv[NR];
        ld4(b + row * NR, bv);
        const double di = dinv[row];
        const double zero[NR] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            if (k >= s) bv[k] = 0.0;
            zv[k] = bv[k] * di;
            acc[k] += bv[k] * bv[k];
        }
        st4(z + row * NR, zv);
        st4(p + row * NR, zero);
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) acc[NR + k] = acc[k];
    wg_sum<2 * NR>(acc, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 2 * NR; ++k) partB[(uint64_t)t * NB + k] = acc[k];
}

// one workgroup per instance: the initial test per column (cg.rs:80-84; X0 = 0
// so r0 = b and only ||b|| < tol stops it), beta = 0
__global__ __launch_bounds__(BT) void k_bb_start(BbState S, const double* __restrict__ partB) {
    __shared__ double sh[4 * NB];
    const uint32_t i = blockIdx.x;
    double red[NB];
    reduce_slots<NB>(partB, S.inst_t0[i], S.inst_t1[i] - S.inst_t0[i], red, sh);
    if (threadIdx.x != 0) return;
    uint32_t act = 0;
    for (int k = 0; k < NR; ++k) {
        const double bn = sqrt(red[NR + k]);
        S.bnorm[i * NR + k] = bn;
        const bool on = k < S.s && !(bn < S.tol);
        if (on) act |= 1u << k;
        S.status[i * NR + k] = on ? 0 : 1;
        S.iters[i * NR + k] = 0;
    }
    for (int k = 0; k < NR * NR; ++k) S.beta[i * NR * NR + k] = 0.0;
    S.act[i] = act;
    S.kept[i] = 0;
    S.it[i] = 0;
    S.done[i] = act ? 0u : 1u;
}

// W = Z + P beta (rounded to the storage type TV) for the rows of the active
// tiles; tile partials W^T R (R = Z / D^{-1}) into entries NSYM.. of the
// tile's NA record
template <class TV>
__global__ __launch_bounds__(BT) void k_bb_formw(const double* __restrict__ z,
                                                 const TV* __restrict__ p,
                                                 const double* __restrict__ dinv,
                                                 TV* __restrict__ w, BbState S,
                                                 double* __restrict__ partA) {
    __shared__ double sh[4 * NR * NR];
    const uint32_t na = S.nact[0];
    for (uint32_t a = xcd_pos(); a < na; a += gridDim.x) {
        const uint32_t t = S.active[a], inst = S.tile_inst[t];
        if (S.done[inst]) continue;
        const uint32_t act = S.act[inst];
        double beta[NR][NR];
#pragma unroll
        for (int m = 0; m < NR; ++m)
#pragma unroll
            for (int k = 0; k < NR; ++k) beta[m][k] = S.beta[(uint64_t)inst * NR * NR + m * NR + k];
        double acc[NR * NR];
#pragma unroll
        for (int k = 0; k < NR * NR; ++k) acc[k] = 0.0;
        for (int sub = 0; sub < TR / BT; ++sub) {
            const uint64_t row = (uint64_t)t * TR + sub * BT + threadIdx.x;
            double zo[NR], po[NR], wo[NR], rv[NR];
            ld4(z + row * NR, zo);
            ld4(p + row * NR, po);
            const double di = dinv[row];
#pragma unroll
            for (int k = 0; k < NR; ++k) rv[k] = zo[k] / di;
            form_w(zo, po, beta, act, wo);
#pragma unroll
            for (int k = 0; k < NR; ++k) wo[k] = rnd<TV>(wo[k]);
            st4(w + row * NR, wo);
#pragma unroll
            for (int a2 = 0; a2 < NR; ++a2)
#pragma unroll
                for (int c = 0; c < NR; ++c) acc[a2 * NR + c] += wo[a2] * rv[c];
        }
        wg_sum<NR * NR>(acc, sh);
        if (threadIdx.x == 0)
#pragma unroll
            for (int k = 0; k < NR * NR; ++k) partA[(uint64_t)t * NA + NSYM + k] = acc[k];
    }
}

// Q_W = A W (FP64 accumulation of the stored W); tile partials W^T Q_W into
// entries 0..NSYM of the tile's NA record.  Grid-stride over the active tiles.
template <class TM, class TV>
__global__ __launch_bounds__(BT) void k_bb_prod(SellB<TM> A, const TV* __restrict__ w,
                                                double* __restrict__ q, BbState S,
                                                double* __restrict__ partA) {
    __shared__ double sh[4 * NSYM];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t na = S.nact[0];
    for (uint32_t a = xcd_pos(); a < na; a += gridDim.x) {
        const uint32_t t = S.active[a], inst = S.tile_inst[t];
        if (S.done[inst]) continue;  // uniform per workgroup
        double acc[NSYM];
#pragma unroll
        for (int k = 0; k < NSYM; ++k) acc[k] = 0.0;
        for (int sub = 0; sub < TS / 4; ++sub) {
            const uint32_t sl = t * TS + sub * 4 + wv;
            const uint64_t row = (uint64_t)sl * SLICE + lane;
            const uint32_t base = A.soff[sl], wd = (A.soff[sl + 1] - base) / SLICE;
            const uint32_t* cp = A.cols + (uint64_t)base + lane;
            const TM* vp = A.vals + (uint64_t)base + lane;
            double qv[NR] = {0.0, 0.0, 0.0, 0.0};
            // PU stored entries per step (a P1 row of a structured 2D mesh has
            // 7: one step), every load of a step issued before the
            // arithmetic; SELL padding holds zero values and the row's own
            // column, so it needs no length test (a step past the width
            // re-reads the last entry with a zero value)
            for (uint32_t e0 = 0; e0 < wd; e0 += PU) {
                uint32_t c[PU];
                double v[PU];
#pragma unroll
                for (int u = 0; u < PU; ++u) {
                    const uint32_t e = min(e0 + u, wd - 1);
                    c[u] = cp[(uint64_t)e * SLICE];
                    v[u] = e0 + u < wd ? (double)vp[(uint64_t)e * SLICE] : 0.0;
                }
                double wc[PU][NR];
#pragma unroll
                for (int u = 0; u < PU; ++u) ld4(w + (uint64_t)c[u] * NR, wc[u]);
#pragma unroll
                for (int u = 0; u < PU; ++u)
#pragma unroll
                    for (int k = 0; k < NR; ++k) qv[k] = fma(v[u], wc[u][k], qv[k]);
            }
            double wo[NR];
            ld4(w + row * NR, wo);
            st4(q + row * NR, qv);
#pragma unroll
            for (int a2 = 0; a2 < NR; ++a2)
#pragma unroll
                for (int c = a2; c < NR; ++c) acc[sym(a2, c)] += wo[a2] * qv[c];
        }
        wg_sum<NSYM>(acc, sh);
        if (threadIdx.x == 0)
#pragma unroll
            for (int k = 0; k < NSYM; ++k) partA[(uint64_t)t * NA + k] = acc[k];
    }
}

// one workgroup per active instance: A-orthonormal directions P = W L^{-T}
// (dependent ones dropped) and alpha = P^T R = L^{-1} (W^T R)
__global__ __launch_bounds__(BT) void k_bb_red1(BbState S, const double* __restrict__ partA) {
    __shared__ double sh[4 * NA];
    const uint32_t ni = S.nact[1];
    for (uint32_t a = blockIdx.x; a < ni; a += gridDim.x) {
        const uint32_t i = S.active[S.n_tiles + a];
        if (S.done[i]) continue;
        double red[NA];
        reduce_slots<NA>(partA, S.inst_t0[i], S.inst_t1[i] - S.inst_t0[i], red, sh);
        if (threadIdx.x != 0) continue;
        const uint32_t act = S.act[i];
        double L[NR][NR];
        const uint32_t kept = chol_drop4(red, act, L);
        S.kept[i] = kept;
        for (int m = 0; m < NR; ++m)
            for (int k = 0; k < NR; ++k) S.L[(uint64_t)i * NR * NR + m * NR + k] = L[m][k];
        for (int c = 0; c < NR; ++c) {
            double u[NR], v[NR];
            for (int m = 0; m < NR; ++m) u[m] = red[NSYM + m * NR + c];
            lsolve_row(L, kept, u, v);
            for (int m = 0; m < NR; ++m)
                S.alpha[(uint64_t)i * NR * NR + m * NR + c] = ((act >> c) & 1) ? v[m] : 0.0;
        }
        if (!kept) {
            // every direction vanished while columns are left: their residuals
            // lie outside what W can reach
            for (int k = 0; k < NR; ++k)
                if ((act >> k) & 1) {
                    S.status[i * NR + k] = 3;
                    S.iters[i * NR + k] = S.it[i] + 1;
                }
            S.act[i] = 0;
            S.done[i] = 1;
        }
    }
}

// P = W L^{-T}, Q = Q_W L^{-T} (FP64, from the stored W and Q_W = A W, so X
// and R move by a consistent pair and the recursive residual tracks b - A x
// to FP64 rounding whatever TV is); R = Z / D^{-1} (R is not stored: 64
// bytes per row less to stream), X += P alpha, R -= Q alpha, Z = D^{-1} R;
// Z stored in FP64, P as TV; tile partials Q^T Z and R^T R
template <class TV>
__global__ __launch_bounds__(BT) void k_bb_upd(const TV* __restrict__ w,
                                               const double* __restrict__ qw,
                                               const double* __restrict__ dinv,
                                               double* __restrict__ x, double* __restrict__ z,
                                               TV* __restrict__ p, BbState S,
                                               double* __restrict__ partB) {
    __shared__ double sh[4 * NB];
    const uint32_t na = S.nact[0];
    for (uint32_t a = xcd_pos(); a < na; a += gridDim.x) {
        const uint32_t t = S.active[a], inst = S.tile_inst[t];
        if (S.done[inst]) continue;
        const uint32_t kept = S.kept[inst];
        double L[NR][NR], alpha[NR][NR];
#pragma unroll
        for (int m = 0; m < NR; ++m)
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                L[m][k] = S.L[(uint64_t)inst * NR * NR + m * NR + k];
                alpha[m][k] = S.alpha[(uint64_t)inst * NR * NR + m * NR + k];
            }
        double acc[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) acc[k] = 0.0;
        for (int sub = 0; sub < TR / BT; ++sub) {
        const uint64_t row = (uint64_t)t * TR + sub * BT + threadIdx.x;
        double wv4[NR], qw4[NR], pv[NR], qv[NR], xv[NR], rv[NR], zv[NR];
        ld4(w + row * NR, wv4);
        ld4(qw + row * NR, qw4);
        ld4(x + row * NR, xv);
        ld4(z + row * NR, zv);
        lsolve_row(L, kept, wv4, pv);
        lsolve_row(L, kept, qw4, qv);
        const double di = dinv[row];
#pragma unroll
        for (int k = 0; k < NR; ++k) rv[k] = zv[k] / di;
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            double dx = 0.0, dr = 0.0;
#pragma unroll
            for (int m = 0; m < NR; ++m) {
                dx = fma(pv[m], alpha[m][k], dx);
                dr = fma(qv[m], alpha[m][k], dr);
            }
            xv[k] += dx;
            rv[k] -= dr;
            zv[k] = rv[k] * di;
        }
        st4(x + row * NR, xv);
        st4(z + row * NR, zv);
        st4(p + row * NR, pv);
#pragma unroll
        for (int a2 = 0; a2 < NR; ++a2)
#pragma unroll
            for (int c = 0; c < NR; ++c) acc[a2 * NR + c] += qv[a2] * zv[c];
#pragma unroll
        for (int k = 0; k < NR; ++k) acc[NR * NR + k] += rv[k] * rv[k];
        }
        wg_sum<NB>(acc, sh);
        if (threadIdx.x == 0)
#pragma unroll
            for (int k = 0; k < NB; ++k) partB[(uint64_t)t * NB + k] = acc[k];
    }
}

// one workgroup per active instance: per-column test on the recursive
// residual (cg.rs:118-122; a converged column leaves the block, its X and R
// stay), beta = -(P^T A P)^{-1} Q^T Z = -Q^T Z (P is A-orthonormal), the
// iteration limit (cg.rs:139-148)
__global__ __launch_bounds__(BT) void k_bb_red2(BbState S, const double* __restrict__ partB) {
    __shared__ double sh[4 * NB];
    const uint32_t ni = S.nact[1];
    for (uint32_t a = blockIdx.x; a < ni; a += gridDim.x) {
        const uint32_t i = S.active[S.n_tiles + a];
        if (S.done[i]) continue;
        double red[NB];
        reduce_slots<NB>(partB, S.inst_t0[i], S.inst_t1[i] - S.inst_t0[i], red, sh);
        if (threadIdx.x != 0) continue;
        const uint32_t j = S.it[i] + 1;
        uint32_t act = S.act[i];
        const uint32_t kept = S.kept[i];
        for (int k = 0; k < NR; ++k)
            if ((act >> k) & 1) {
                const double bn = S.bnorm[i * NR + k];
                const double res = sqrt(red[NR * NR + k]);
                const double rel = bn > 0.0 ? res / bn : res;
                if (rel < S.tol) {
                    S.status[i * NR + k] = 1;
                    S.iters[i * NR + k] = j;
                    act &= ~(1u << k);
                }
            }
        if (act && j >= S.max_iter) {
            for (int k = 0; k < NR; ++k)
                if ((act >> k) & 1) {
                    S.status[i * NR + k] = 2;
                    S.iters[i * NR + k] = j;
                }
            act = 0;
        }
        for (int m = 0; m < NR; ++m)
            for (int c = 0; c < NR; ++c)
                S.beta[(uint64_t)i * NR * NR + m * NR + c] =
                    (((kept >> m) & 1) && ((act >> c) & 1)) ? -red[m * NR + c] : 0.0;
        S.act[i] = act;
        S.it[i] = j;
        if (!act) S.done[i] = 1;
    }
}

// the work list: tiles and instances not yet done, in order (one workgroup)
__global__ __launch_bounds__(1024) void k_bb_compact(BbState S) {
    __shared__ uint32_t cnt[1024];
    __shared__ uint32_t tot;
    for (int pass = 0; pass < 2; ++pass) {
        const uint32_t n = pass == 0 ? S.n_tiles : S.n_inst;
        const uint32_t per = (n + 1023) / 1024;
        const uint32_t b0 = threadIdx.x * per, b1 = min(n, b0 + per);
        uint32_t c = 0;
        for (uint32_t k = b0; k < b1; ++k) {
            const uint32_t inst = pass == 0 ? S.tile_inst[k] : k;
            c += S.done[inst] ? 0u : 1u;
        }
        cnt[threadIdx.x] = c;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t run = 0;
            for (int k = 0; k < 1024; ++k) {
                const uint32_t v = cnt[k];
                cnt[k] = run;
                run += v;
            }
            tot = run;
        }
        __syncthreads();
        uint32_t o = cnt[threadIdx.x];
        uint32_t* out = S.active + (pass == 0 ? 0 : S.n_tiles);
        for (uint32_t k = b0; k < b1; ++k) {
            const uint32_t inst = pass == 0 ? S.tile_inst[k] : k;
            if (!S.done[inst]) out[o++] = k;
        }
        if (threadIdx.x == 0) S.nact[pass] = tot;
        __syncthreads();
    }
}

// tile partials of ||b - A64 x||^2 per column
__global__ __launch_bounds__(BT) void k_bb_true_res(SellB<double> A, const double* __restrict__ b,
                                                    const double* __restrict__ x,
                                                    double* __restrict__ part, int s) {
    __shared__ double sh[4 * NR];
    const uint32_t t = blockIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double acc[NR] = {0.0, 0.0, 0.0, 0.0};
    for (int sub = 0; sub < TS / 4; ++sub) {
        const uint32_t sl = t * TS + sub * 4 + wv;
        const uint64_t row = (uint64_t)sl * SLICE + lane;
        const uint32_t base = A.soff[sl], wd = (A.soff[sl + 1] - base) / SLICE;
        const uint32_t len = A.rlen[row];
        double ax[NR] = {0.0, 0.0, 0.0, 0.0};
        for (uint32_t e = 0; e < wd && e < len; ++e) {
            const uint64_t ix = (uint64_t)base + (uint64_t)e * SLICE + lane;
            double xc[NR];
            ld4(x + (uint64_t)A.cols[ix] * NR, xc);
#pragma unroll
            for (int k = 0; k < NR; ++k) ax[k] = fma(A.vals[ix], xc[k], ax[k]);
        }
        double bv[NR];
        ld4(b + row * NR, bv);
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const double tk = k < s ? bv[k] - ax[k] : 0.0;
            acc[k] += tk * tk;
        }
    }
    wg_sum<NR>(acc, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NR; ++k) part[(uint64_t)t * NR + k] = acc[k];
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

// Launch parameters of the tile kernels: workgroups per CU of their
// grid-stride grid and block iterations per graph chunk, from
// FH_BATCH_WGS_PER_CU / FH_BATCH_CHUNK, else the tuning file written by
// fenics-rs_amd/tune.py (fh_tune.json beside the library, or the path in
// FH_TUNE_FILE), else 16 and 16.
struct BatchTune {
    int wgs_per_cu = 16;
    int chunk = 16;
};

// the integer after "key": in a small JSON text (-1 when absent)
static int json_int(const std::string& t, const char* key) {
    const std::string k = std::string("\"") + key + "\"";
    const size_t p = t.find(k);
    if (p == std::string::npos) return -1;
    const size_t c = t.find(':', p + k.size());
    if (c == std::string::npos) return -1;
    return std::atoi(t.c_str() + c + 1);
}

static BatchTune batch_tune() {
    BatchTune t;
    std::string path;
    if (const char* f = std::getenv("FH_TUNE_FILE")) {
        path = f;
    } else {
        Dl_info info{};
        if (dladdr((void*)&json_int, &info) && info.dli_fname) {
            path = info.dli_fname;
            const size_t sl = path.rfind('/');
            path = (sl == std::string::npos ? std::string(".") : path.substr(0, sl)) + "/fh_tune.json";
        }
    }
    if (!path.empty()) {
        if (FILE* fp = std::fopen(path.c_str(), "r")) {
            std::string txt;
            char buf[4096];
            size_t n;
            while ((n = std::fread(buf, 1, sizeof buf, fp)) > 0) txt.append(buf, n);
            std::fclose(fp);
            const int a = json_int(txt, "batched_wgs_per_cu"), b = json_int(txt, "batched_chunk");
            if (a >= 1 && a <= 64) t.wgs_per_cu = a;
            if (b >= 1 && b <= 256) t.chunk = b;
        }
    }
    if (const char* e = std::getenv("FH_BATCH_WGS_PER_CU"))
        t.wgs_per_cu = std::max(1, std::min(64, std::atoi(e)));
    if (const char* e = std::getenv("FH_BATCH_CHUNK")) t.chunk = std::max(1, std::min(256, std::atoi(e)));
    return t;
}

template <class T>
static int run_batched(uint64_t n_inst, const uint64_t* sizes, uint64_t nrhs,
                       const double* f_nodal, double tol, uint64_t max_iter, int precision,
                       int dev, double* u_out, uint64_t* iters_out, fh_solve_stats* stats) {
    DeviceCtx& d = device(dev);
    hipStream_t s = d.stream;
    const auto tw0 = std::chrono::steady_clock::now();
    // ---- union mesh with per-instance padding to TR rows (one tile)
    std::vector<uint32_t> voff(n_inst + 1), coff(n_inst + 1), nvi(n_inst);
    uint64_t V = 0, Cn = 0;
    for (uint64_t i = 0; i < n_inst; ++i) {
        FH_CHECK(sizes[i] >= 1 && sizes[i] < 20000, "instance size out of range");
        voff[i] = (uint32_t)V;
        coff[i] = (uint32_t)Cn;
        nvi[i] = (uint32_t)((sizes[i] + 1) * (sizes[i] + 1));
        V += (nvi[i] + TR - 1) / TR * TR;
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
    // ---- right-hand sides b = M f (f given per real vertex, nrhs columns)
    const uint64_t nvec = V * NR;
    DBuf<double> f(dev, nvec), b(dev, nvec);
    {
        fill_f64(f.p, nvec, 0.0, s);
        std::vector<double> hf;
        uint64_t src = 0;
        for (uint64_t i = 0; i < n_inst; ++i) {
            if (nrhs == (uint64_t)NR) {
                h2d(f.p + (uint64_t)voff[i] * NR, f_nodal + src * NR, (size_t)nvi[i] * NR, s);
            } else {
                hf.assign((size_t)nvi[i] * NR, 0.0);
                for (uint32_t r = 0; r < nvi[i]; ++r)
                    for (uint64_t k = 0; k < nrhs; ++k)
                        hf[(size_t)r * NR + k] = f_nodal[(src + r) * nrhs + k];
                h2d(f.p + (uint64_t)voff[i] * NR, hf.data(), hf.size(), s);
                FH_HIP(hipStreamSynchronize(s));
            }
            src += nvi[i];
        }
        FH_HIP(hipStreamSynchronize(s));
    }
    const unsigned gslices = cdiv(Ac->pat->nslices, 4);
    hipLaunchKernelGGL(k_spmm_f64, dim3(gslices), dim3(256), 0, s, sellb(M.get(), M->vals.p), f.p, b.p);
    hipLaunchKernelGGL(k_zero_fixed_rows, dim3(cdiv(V, 256)), dim3(256), 0, s, Ac->fixed.p, b.p,
                       (uint32_t)V);
    FH_LAUNCH_CHECK("batched_rhs");
    M.reset();
    f.reset();
    DBuf<double> dinv(dev, V);
    inv_diag_f64(Ac.get(), dinv.p, s);
    constexpr bool MIX = std::is_same<T, float>::value;
    if (MIX) ensure_vals32(Ac.get());

    // ---- tiles (TR rows) and the per-instance state
    const uint32_t n_tiles = (uint32_t)(V / TR);
    std::vector<uint32_t> tile_inst(n_tiles), t0(n_inst), t1(n_inst);
    for (uint64_t i = 0; i < n_inst; ++i) {
        t0[i] = voff[i] / TR;
        t1[i] = voff[i + 1] / TR;
        for (uint32_t t = t0[i]; t < t1[i]; ++t) tile_inst[t] = (uint32_t)i;
    }
    DBuf<uint32_t> dti(dev, n_tiles), dt0(dev, n_inst), dt1(dev, n_inst), act(dev, n_inst),
        kept(dev, n_inst), itc(dev, n_inst), done(dev, n_inst), iters(dev, n_inst * NR),
        active(dev, (size_t)n_tiles + n_inst), nact(dev, 2);
    h2d(dti.p, tile_inst.data(), n_tiles, s);
    h2d(dt0.p, t0.data(), n_inst, s);
    h2d(dt1.p, t1.data(), n_inst, s);
    DBuf<uint8_t> status(dev, n_inst * NR);
    DBuf<double> bnorm(dev, n_inst * NR), beta(dev, n_inst * NR * NR), Lm(dev, n_inst * NR * NR),
        alpha(dev, n_inst * NR * NR);
    // Krylov vectors: x, z and q = A w in FP64 (r is not stored: R = Z / D^{-1});
    // p and the gathered w in TV (FP32 in the mixed modes)
    using TV = T;
    DBuf<double> x(dev, nvec), z(dev, nvec), q(dev, nvec), partA(dev, (size_t)n_tiles * NA),
        partB(dev, (size_t)n_tiles * NB), part_t(dev, (size_t)n_tiles * NR);
    DBuf<TV> pb(dev, nvec), wb(dev, nvec);
    // padding rows and the rows of the unused columns stay zero in every vector
    for (double* v : {x.p, z.p, q.p}) fill_f64(v, nvec, 0.0, s);
    if constexpr (MIX) {
        fill_f32((float*)pb.p, nvec, 0.0f, s);
        fill_f32((float*)wb.p, nvec, 0.0f, s);
    } else {
        fill_f64((double*)pb.p, nvec, 0.0, s);
        fill_f64((double*)wb.p, nvec, 0.0, s);
    }
    BbState S{act.p, kept.p, itc.p, done.p, bnorm.p, beta.p, Lm.p, alpha.p, iters.p, status.p,
              dti.p, dt0.p, dt1.p, active.p, nact.p, n_tiles, (uint32_t)n_inst, (int)nrhs, tol,
              (uint32_t)std::min<uint64_t>(std::max<uint64_t>(max_iter, 1), 0xffffffffu)};
    const SellB<T> AK = MIX ? sellb<T>(Ac.get(), (const T*)(const void*)Ac->vals32.p)
                            : sellb<T>(Ac.get(), (const T*)(const void*)Ac->vals.p);
    const SellB<double> A64 = sellb(Ac.get(), Ac->vals.p);
    const BatchTune tn = batch_tune();
    // a multiple of 8 (xcd_pos), at most the tile count rounded up
    const unsigned gt = std::max(8u, std::min<unsigned>(cdiv(n_tiles, 8) * 8,
                                                        (unsigned)(d.n_cu * tn.wgs_per_cu)) / 8 * 8);
    const unsigned gi = std::max(1u, std::min<unsigned>((unsigned)n_inst, (unsigned)(d.n_cu * 4)));
    FH_HIP(hipStreamSynchronize(s));

    Timer tk(s);
    hipLaunchKernelGGL(k_bb_init<TV>, dim3(n_tiles), dim3(BT), 0, s, b.p, dinv.p, z.p, pb.p,
                       partB.p, (int)nrhs);
    hipLaunchKernelGGL(k_bb_start, dim3((unsigned)n_inst), dim3(BT), 0, s, S, partB.p);
    hipLaunchKernelGGL(k_bb_compact, dim3(1), dim3(1024), 0, s, S);
    FH_LAUNCH_CHECK("batched_init");
    const int Mch = tn.chunk;
    auto chunk = [&]() {
        for (int m = 0; m < Mch; ++m) {
            hipLaunchKernelGGL(k_bb_formw<TV>, dim3(gt), dim3(BT), 0, s, (const double*)z.p,
                               (const TV*)pb.p, (const double*)dinv.p, wb.p, S, partA.p);
            hipLaunchKernelGGL((k_bb_prod<T, TV>), dim3(gt), dim3(BT), 0, s, AK, (const TV*)wb.p,
                               q.p, S, partA.p);
            hipLaunchKernelGGL(k_bb_red1, dim3(gi), dim3(BT), 0, s, S, partA.p);
            hipLaunchKernelGGL(k_bb_upd<TV>, dim3(gt), dim3(BT), 0, s, (const TV*)wb.p,
                               (const double*)q.p, (const double*)dinv.p, x.p, z.p, pb.p, S,
                               partB.p);
            hipLaunchKernelGGL(k_bb_red2, dim3(gi), dim3(BT), 0, s, S, partB.p);
        }
        hipLaunchKernelGGL(k_bb_compact, dim3(1), dim3(1024), 0, s, S);
    };
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    FH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    chunk();
    FH_HIP(hipStreamEndCapture(s, &graph));
    FH_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    // chunks are queued two ahead of the host's look at the work count: the
    // host never waits on a chunk it has not already queued the next one behind
    uint32_t* hn = nullptr;
    FH_HIP(hipHostMalloc((void**)&hn, 2 * 2 * sizeof(uint32_t), hipHostMallocDefault));
    hipEvent_t evs[2];
    FH_HIP(hipEventCreateWithFlags(&evs[0], hipEventDisableTiming));
    FH_HIP(hipEventCreateWithFlags(&evs[1], hipEventDisableTiming));
    uint64_t chunks = 0;
    const uint64_t max_chunks = (uint64_t)S.max_iter / (uint64_t)Mch + 2;
    bool finished = false;
    for (uint64_t c = 0; c < max_chunks + 1 && !finished; ++c) {
        if (c < max_chunks) {
            FH_HIP(hipGraphLaunch(gexec, s));
            FH_HIP(hipMemcpyAsync(hn + 2 * (c & 1), nact.p, 2 * sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, s));
            FH_HIP(hipEventRecord(evs[c & 1], s));
            ++chunks;
        }
        if (c >= 1) {  // the chunk before this one
            FH_HIP(hipEventSynchronize(evs[(c - 1) & 1]));
            if (hn[2 * ((c - 1) & 1) + 1] == 0) finished = true;
        }
    }
    FH_HIP(hipStreamSynchronize(s));
    const double kms = tk.stop_ms();
    (void)hipEventDestroy(evs[0]);
    (void)hipEventDestroy(evs[1]);
    (void)hipHostFree(hn);
    (void)hipGraphExecDestroy(gexec);
    (void)hipGraphDestroy(graph);
    // ---- true FP64 residuals
    hipLaunchKernelGGL(k_bb_true_res, dim3(n_tiles), dim3(BT), 0, s, A64, b.p, x.p, part_t.p,
                       (int)nrhs);
    FH_LAUNCH_CHECK("batched_true_res");
    std::vector<double> hpt((size_t)n_tiles * NR), hbn(n_inst * NR), htr(n_inst * NR, 0.0);
    d2h(hpt.data(), part_t.p, hpt.size(), s);
    d2h(hbn.data(), bnorm.p, hbn.size(), s);
    FH_HIP(hipStreamSynchronize(s));
    for (uint64_t i = 0; i < n_inst; ++i)
        for (int k = 0; k < NR; ++k) {
            double tt = 0.0;
            for (uint32_t t = t0[i]; t < t1[i]; ++t) tt += hpt[(size_t)t * NR + k];
            const double bn = hbn[i * NR + k];
            htr[i * NR + k] = bn > 0.0 ? std::sqrt(tt) / bn : std::sqrt(tt);
        }
    // ---- outputs
    std::vector<uint32_t> hit(n_inst * NR);
    std::vector<uint8_t> hd(n_inst * NR);
    {
        std::vector<double> hx;
        uint64_t dst = 0;
        for (uint64_t i = 0; i < n_inst; ++i) {  // the caller's rows of each instance
            if (nrhs == (uint64_t)NR) {
                d2h(u_out + dst * NR, x.p + (uint64_t)voff[i] * NR, (size_t)nvi[i] * NR, s);
            } else {
                hx.resize((size_t)nvi[i] * NR);
                d2h(hx.data(), x.p + (uint64_t)voff[i] * NR, hx.size(), s);
                FH_HIP(hipStreamSynchronize(s));
                for (uint32_t rr = 0; rr < nvi[i]; ++rr)
                    for (uint64_t k = 0; k < nrhs; ++k)
                        u_out[(dst + rr) * nrhs + k] = hx[(size_t)rr * NR + k];
            }
            dst += nvi[i];
        }
    }
    d2h(hit.data(), iters.p, n_inst * NR, s);
    d2h(hd.data(), status.p, n_inst * NR, s);
    FH_HIP(hipStreamSynchronize(s));
    double dof_iters = 0.0, tr_max = 0.0;
    uint64_t max_it = 0, inst_iters = 0;
    int code = FH_OK;
    for (uint64_t i = 0; i < n_inst; ++i) {
        uint64_t imax = 0;
        for (uint64_t k = 0; k < nrhs; ++k) {
            const uint64_t col = i * NR + k;
            if (iters_out) iters_out[i * nrhs + k] = hit[col];
            imax = std::max<uint64_t>(imax, hit[col]);
            tr_max = std::max(tr_max, htr[col]);
            if (hd[col] == 3) code = FH_ERR_BREAKDOWN;
            else if (hd[col] != 1 && code == FH_OK) code = FH_ERR_NOT_CONVERGED;
        }
        // every block iteration of the instance advances all nrhs columns
        dof_iters += (double)nvi[i] * (double)nrhs * (double)imax;
        inst_iters += imax;
        max_it = std::max(max_it, imax);
    }
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->iterations = max_it;
        stats->replacements = inst_iters;  // block iterations summed over instances
        stats->true_rel_residual = tr_max;
        stats->kernel_ms = kms;
        stats->wall_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tw0).count();
        // DOF (unknowns x columns) * iterations per second
        stats->gflops = dof_iters / (kms * 1e-3) / 1e9;
        // one product pass: the matrix once for all columns, one read of the
        // gathered w, the q store; vector rows are 4 columns of 8 bytes (x, z,
        // q) or of sizeof(T) (p, w)
        const double vb = 4.0 * sizeof(T);
        stats->spmv_bytes = (double)Ac->pat->nnzb * (sizeof(T) + 4.0) + 4.0 * (V + 1) +
                            (vb + 32.0) * (double)V;
        // one block iteration: W = Z + P beta (reads z, p and D^{-1}, writes
        // w), the product pass, and the update (reads w, q, x, z and D^{-1};
        // writes x, z and p)
        stats->iter_bytes = stats->spmv_bytes + (2.0 * vb + 40.0) * (double)V +
                            (2.0 * vb + 32.0 * 3.0 + 8.0 + 32.0 * 2.0) * (double)V;
        stats->effective_bw_gbps = 0.0;
        if (kms > 0.0) {
            double bytes = 0.0;
            for (uint64_t i = 0; i < n_inst; ++i) {
                uint64_t imax = 0;
                for (uint64_t k = 0; k < nrhs; ++k) imax = std::max<uint64_t>(imax, hit[i * NR + k]);
                bytes += (double)imax * stats->iter_bytes * (double)(voff[i + 1] - voff[i]) / (double)V;
            }
            stats->effective_bw_gbps = bytes / (kms * 1e-3) / 1e9;
        }
        stats->spmv_timed = chunks;
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
        FH_CHECK(nrhs >= 1 && nrhs <= (uint64_t)NR, "the block size must be 1 to 4 right-hand sides");
        FH_CHECK(precision >= 0 && precision <= 3, "unknown precision");
        const int dev = resolve_device(device_idx);
        if (precision == FH_PREC_FP64 || precision == FH_PREC_FP64_TRUE)
            code = run_batched<double>(n_inst, sizes, nrhs, f_nodal, tol, max_iter, precision, dev,
                                       u_out, iters_out, stats);
        else
            code = run_batched<float>(n_inst, sizes, nrhs, f_nodal, tol, max_iter, precision, dev,
                                      u_out, iters_out, stats);
        if (code == FH_ERR_NOT_CONVERGED) set_last_error("a batched column failed to converge");
        if (code == FH_ERR_BREAKDOWN) set_last_error("breakdown in a batched block");
    });
    return g != FH_OK ? g : code;
}
