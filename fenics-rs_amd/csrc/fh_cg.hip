// fh_cg.hip — Jacobi-preconditioned CG on the device (replaces cg.rs:40-149
// and the host-driven loops of gpu.rs:775-1141 / 1145-1442).
//
// One iteration is two kernels, and no scalar ever visits the host:
//   k_cg_spmv   : reduces the previous update's (r.r, r.z) partials (every
//                 block, same fixed order -> same beta everywhere), applies the
//                 stopping rule, then computes p_new = z + beta p_old on the fly
//                 for the gathered columns, Ap = A p_new (SELL-64, one wave per
//                 64-row slice) and the block partials of p.Ap;
//   k_cg_update : reduces the p.Ap partials, alpha = rz/pAp, x += alpha p,
//                 r -= alpha Ap, z = D^{-1} r and the (r.r, r.z) partials.
// Iteration bookkeeping lives in a device CgState written only by block 0 and
// read by the next kernel, so a chunk of iterations has fixed arguments and is
// captured once into a HIP graph and replayed.  The host looks at the state
// once per chunk.
//
// Precision modes:
//   FP64      : everything FP64 (FH_ORDER_ORDERED also reproduces the
//               reference's SpMV/axpy arithmetic bit for bit; only the dot
//               products are tree-summed);
//   MIXED     : FP32 matrix/vectors in the Krylov loop, dots accumulated in
//               FP64; when the recursive residual has dropped by `delta` since
//               the last replacement (or reached tol), the residual is replaced
//               by the true FP64 residual b - A64 (x64 + x32), x32 is folded
//               into the FP64 solution and the Krylov recurrences continue
//               (residual replacement, van der Vorst & Ye 2000).  Convergence
//               is declared only on the true residual;
//   MIXED_IR  : the same but restarting the recurrences (p = z) at each
//               replacement: classic FP64 iterative refinement of FP32 CG.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "fh_internal.hpp"

namespace fh {

enum : uint32_t { ST_RUNNING = 0, ST_CONVERGED = 1, ST_NOT_CONVERGED = 2, ST_BREAKDOWN = 3 };

struct CgState {
    uint32_t status;
    uint32_t need_rep;
    uint64_t it_spmv;
    uint64_t it_upd;
    uint64_t rep_mark;  // spmv index that consumes fresh true-residual partials
    uint64_t n_rep;
    uint64_t iters;
    double rz[2];
    double bnorm;
    double res;
    double r_rep;
    double rel;
    // configuration
    double tol;
    double delta;
    uint64_t max_iter;
    uint32_t mixed;
    uint32_t restart;
};

template <class T>
struct SellA {
    const uint32_t* soff;
    const uint32_t* rlen;
    const uint32_t* cols;
    const T* vals;
    uint32_t nrows;
    uint32_t nslices;
};

template <class T>
__device__ __forceinline__ T madd(T acc, T a, T b, bool ordered) {
    return ordered ? acc + a * b : fma(a, b, acc);
}

// ---------------------------------------------------------------- SpMV kernels
// y = A x (plain), one wave per slice.
template <int BS, class T, bool ORD>
__global__ __launch_bounds__(256) void k_spmv(SellA<T> A, const T* __restrict__ x,
                                              T* __restrict__ y) {
    const unsigned blk = fhd::xcd_block(blockIdx.x, gridDim.x);
    const unsigned s = blk * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (s >= A.nslices) return;
    const uint32_t r = s * SLICE + lane;
    const uint32_t base = A.soff[s], w = (A.soff[s + 1] - base) / SLICE;
    const uint32_t len = r < A.nrows ? A.rlen[r] : 0;
    T acc[BS];
#pragma unroll
    for (int a = 0; a < BS; ++a) acc[a] = T(0);
    for (uint32_t k = 0; k < w; ++k) {
        if (k < len) {
            const uint64_t e = (uint64_t)base + (uint64_t)k * SLICE;
            const uint32_t c = A.cols[e + lane];
            T xv[BS];
#pragma unroll
            for (int j = 0; j < BS; ++j) xv[j] = x[(uint64_t)c * BS + j];
            const T* v = A.vals + e * BS * BS + lane;
#pragma unroll
            for (int a = 0; a < BS; ++a)
#pragma unroll
                for (int j = 0; j < BS; ++j) acc[a] = madd(acc[a], v[(a * BS + j) * SLICE], xv[j], ORD);
        }
    }
    if (r < A.nrows)
#pragma unroll
        for (int a = 0; a < BS; ++a) y[(uint64_t)r * BS + a] = acc[a];
}

// Residual in FP64: t = b - A64 (x + x32), r = (TV) t, z = D^{-1} r; partials
// (t.t, r.z) and optionally b.b.  Used for the initial residual and for the
// residual replacement of the mixed modes.
template <int BS, class TV, bool ORD, int PRE>
__global__ __launch_bounds__(256) void k_residual(SellA<double> A, const double* __restrict__ b,
                                                  const double* __restrict__ x,
                                                  const float* __restrict__ x32,
                                                  const TV* __restrict__ dinv, TV* __restrict__ r,
                                                  TV* __restrict__ z, double2* __restrict__ red,
                                                  double* __restrict__ bb) {
    __shared__ double sh[4];
    const unsigned blk = fhd::xcd_block(blockIdx.x, gridDim.x);
    const unsigned s = blk * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    double tt = 0.0, rz = 0.0, bsq = 0.0;
    if (s < A.nslices) {
        const uint32_t row = s * SLICE + lane;
        const uint32_t base = A.soff[s], w = (A.soff[s + 1] - base) / SLICE;
        const uint32_t len = row < A.nrows ? A.rlen[row] : 0;
        double acc[BS];
#pragma unroll
        for (int a = 0; a < BS; ++a) acc[a] = 0.0;
        for (uint32_t k = 0; k < w; ++k) {
            if (k < len) {
                const uint64_t e = (uint64_t)base + (uint64_t)k * SLICE;
                const uint32_t c = A.cols[e + lane];
                double xv[BS];
#pragma unroll
                for (int j = 0; j < BS; ++j) {
                    xv[j] = x[(uint64_t)c * BS + j];
                    if (x32) xv[j] = xv[j] + (double)x32[(uint64_t)c * BS + j];
                }
                const double* v = A.vals + e * BS * BS + lane;
#pragma unroll
                for (int a = 0; a < BS; ++a)
#pragma unroll
                    for (int j = 0; j < BS; ++j)
                        acc[a] = madd(acc[a], v[(a * BS + j) * SLICE], xv[j], ORD);
            }
        }
        if (row < A.nrows) {
            TV rv[BS];
#pragma unroll
            for (int a = 0; a < BS; ++a) {
                const uint64_t i = (uint64_t)row * BS + a;
                const double t = b[i] - acc[a];
                tt += t * t;
                rv[a] = (TV)t;
                r[i] = rv[a];
                if (bb) bsq += b[i] * b[i];
            }
            if constexpr (PRE == 0) {
#pragma unroll
                for (int a = 0; a < BS; ++a) {
                    const uint64_t i = (uint64_t)row * BS + a;
                    const TV zv = rv[a] * dinv[i];
                    z[i] = zv;
                    rz += (double)rv[a] * (double)zv;
                }
            } else {  // 3x3 block Jacobi, dinv row-major per node
                const TV* D = dinv + (uint64_t)row * 9;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const TV zv = D[3 * a] * rv[0] + D[3 * a + 1] * rv[1] + D[3 * a + 2] * rv[2];
                    z[(uint64_t)row * 3 + a] = zv;
                    rz += (double)rv[a] * (double)zv;
                }
            }
        }
    }
    tt = fhd::block_sum256(tt, sh);
    rz = fhd::block_sum256(rz, sh);
    if (bb) bsq = fhd::block_sum256(bsq, sh);
    if (threadIdx.x == 0) {
        red[blockIdx.x] = make_double2(tt, rz);
        if (bb) bb[blockIdx.x] = bsq;
    }
}

// ---------------------------------------------------------------- CG kernels
__device__ __forceinline__ double2 reduce_pairs(const double2* __restrict__ p, int m, double* sh) {
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < m; i += 256) {
        const double2 v = p[i];
        a += v.x;
        b += v.y;
    }
    a = fhd::block_sum256(a, sh);
    b = fhd::block_sum256(b, sh);
    return make_double2(a, b);
}

template <int BS, class T, bool ORD>
__global__ __launch_bounds__(256) void k_cg_spmv(SellA<T> A, const T* __restrict__ z, T* p0, T* p1,
                                                 T* __restrict__ Ap, CgState* st,
                                                 const double2* __restrict__ red_in, int n_in,
                                                 const double* __restrict__ bb_in,
                                                 double* __restrict__ red_out) {
    __shared__ double sh[4];
    const uint32_t status = st->status, need_rep = st->need_rep;
    if (status != ST_RUNNING || need_rep) return;
    const uint64_t j = st->it_upd + 1;
    const double tol = st->tol;
    const bool mixed = st->mixed != 0;
    const bool fresh = mixed && st->rep_mark == j;
    const double2 rr_rz = reduce_pairs(red_in, n_in, sh);
    const double rr = rr_rz.x, rz = rr_rz.y;
    double bnorm, beta;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (j == 1) {
        // cg.rs:80-84 initial test on ||r0||
        bnorm = sqrt(fhd::reduce_partials256(bb_in, n_in, sh));
        const double r0 = sqrt(rr);
        if (lead) {
            st->bnorm = bnorm;
            st->r_rep = r0;
        }
        if (r0 < tol || (bnorm > 0.0 && r0 / bnorm < tol)) {
            if (lead) {
                st->status = ST_CONVERGED;
                st->iters = 0;
                st->res = r0;
                st->rel = bnorm > 0.0 ? r0 / bnorm : r0;
            }
            return;
        }
        beta = 0.0;
    } else {
        bnorm = st->bnorm;
        const double res = sqrt(rr);
        const double rel = bnorm > 0.0 ? res / bnorm : res;  // cg.rs:118-119
        if (!mixed || fresh) {
            if (rel < tol) {  // cg.rs:120-122
                if (lead) {
                    st->status = ST_CONVERGED;
                    st->iters = j - 1;
                    st->res = res;
                    st->rel = rel;
                }
                return;
            }
        } else if (rel < tol || res < st->delta * st->r_rep) {
            if (lead) st->need_rep = 1;  // replace r by the FP64 true residual first
            return;
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
        if (fresh && lead) st->r_rep = res;
        beta = (fresh && st->restart) ? 0.0 : rz / st->rz[(j - 1) & 1];
    }
    if (lead) {
        st->rz[j & 1] = rz;
        st->it_spmv = j;
    }
    const T* pold = (j & 1) ? p0 : p1;
    T* pnew = (j & 1) ? p1 : p0;
    const T bt = (T)beta;

    const unsigned blk = fhd::xcd_block(blockIdx.x, gridDim.x);
    const unsigned s = blk * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    double dot = 0.0;
    if (s < A.nslices) {
        const uint32_t row = s * SLICE + lane;
        const uint32_t base = A.soff[s], w = (A.soff[s + 1] - base) / SLICE;
        const uint32_t len = row < A.nrows ? A.rlen[row] : 0;
        T acc[BS];
#pragma unroll
        for (int a = 0; a < BS; ++a) acc[a] = T(0);
        for (uint32_t k = 0; k < w; ++k) {
            if (k < len) {
                const uint64_t e = (uint64_t)base + (uint64_t)k * SLICE;
                const uint32_t c = A.cols[e + lane];
                T xv[BS];
#pragma unroll
                for (int q = 0; q < BS; ++q) {
                    const uint64_t ci = (uint64_t)c * BS + q;
                    xv[q] = madd(z[ci], bt, pold[ci], ORD);
                }
                const T* v = A.vals + e * BS * BS + lane;
#pragma unroll
                for (int a = 0; a < BS; ++a)
#pragma unroll
                    for (int q = 0; q < BS; ++q)
                        acc[a] = madd(acc[a], v[(a * BS + q) * SLICE], xv[q], ORD);
            }
        }
        if (row < A.nrows) {
#pragma unroll
            for (int a = 0; a < BS; ++a) {
                const uint64_t i = (uint64_t)row * BS + a;
                const T pn = madd(z[i], bt, pold[i], ORD);
                pnew[i] = pn;
                Ap[i] = acc[a];
                dot += (double)pn * (double)acc[a];
            }
        }
    }
    dot = fhd::block_sum256(dot, sh);
    if (threadIdx.x == 0) red_out[blockIdx.x] = dot;
}

// PRE 0: scalar Jacobi (cg.rs:52-64); PRE 1: 3x3 block Jacobi (gpu.rs bsr3_apply_dinv)
template <class T, bool ORD, int PRE>
__global__ __launch_bounds__(256) void k_cg_update(uint32_t n, T* __restrict__ x, T* __restrict__ r,
                                                   const T* __restrict__ Ap, const T* p0,
                                                   const T* p1, const T* __restrict__ dinv,
                                                   T* __restrict__ z, CgState* st,
                                                   const double* __restrict__ red_in, int n_in,
                                                   double2* __restrict__ red_out) {
    __shared__ double sh[4];
    if (st->status != ST_RUNNING || st->need_rep) return;
    const uint64_t j = st->it_spmv;
    if (j != st->it_upd + 1) return;
    const double pap = fhd::reduce_partials256(red_in, n_in, sh);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (fabs(pap) < 1e-30) {  // cg.rs:102-107
        if (lead) {
            st->status = ST_BREAKDOWN;
            st->iters = j;
        }
        return;
    }
    const double alpha = st->rz[j & 1] / pap;
    const T al = (T)alpha;
    const T* p = (j & 1) ? p1 : p0;
    double rr = 0.0, rz = 0.0;
    const uint32_t stride = gridDim.x * 256;
    if (PRE == 0) {
        for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
            // cg.rs:113-116, 125-127
            const T xi = ORD ? x[i] + al * p[i] : fma(al, p[i], x[i]);
            const T ri = ORD ? r[i] - al * Ap[i] : fma(-al, Ap[i], r[i]);
            const T zi = ri * dinv[i];
            x[i] = xi;
            r[i] = ri;
            z[i] = zi;
            rr += (double)ri * (double)ri;
            rz += (double)ri * (double)zi;
        }
    } else {
        const uint32_t nn = n / 3;
        for (uint32_t v = blockIdx.x * 256 + threadIdx.x; v < nn; v += stride) {
            T rv[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const uint32_t i = 3 * v + a;
                x[i] = ORD ? x[i] + al * p[i] : fma(al, p[i], x[i]);
                rv[a] = ORD ? r[i] - al * Ap[i] : fma(-al, Ap[i], r[i]);
                r[i] = rv[a];
                rr += (double)rv[a] * (double)rv[a];
            }
            const T* D = dinv + (uint64_t)v * 9;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const T zv = D[3 * a] * rv[0] + D[3 * a + 1] * rv[1] + D[3 * a + 2] * rv[2];
                z[3 * v + a] = zv;
                rz += (double)rv[a] * (double)zv;
            }
        }
    }
    rr = fhd::block_sum256(rr, sh);
    rz = fhd::block_sum256(rz, sh);
    if (threadIdx.x == 0) red_out[blockIdx.x] = make_double2(rr, rz);
    if (lead) st->it_upd = j;
}

// Fold the FP32 correction into the FP64 solution after a replacement.
__global__ void k_commit(uint32_t n, double* __restrict__ x, float* __restrict__ x32, CgState* st,
                         int mark) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x[i] = x[i] + (double)x32[i];
        x32[i] = 0.0f;
    }
    if (mark && blockIdx.x == 0 && threadIdx.x == 0) {
        st->rep_mark = st->it_upd + 1;
        st->need_rep = 0;
        st->n_rep = st->n_rep + 1;
    }
}

// Inverse diagonal: 1/d when |d| > 1e-15, else 1 (cg.rs:52-64).
template <int BS, class TO>
__global__ void k_inv_diag(SellA<double> A, TO* __restrict__ out) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= (uint64_t)A.nrows * BS) return;
    const uint32_t r = t / BS, a = t % BS;
    const uint32_t s = r / SLICE, lane = r % SLICE, base = A.soff[s], len = A.rlen[r];
    double inv = 1.0;
    for (uint32_t k = 0; k < len; ++k) {
        const uint64_t e = (uint64_t)base + (uint64_t)k * SLICE;
        if (A.cols[e + lane] == r) {
            const double d = A.vals[e * BS * BS + (uint64_t)(a * BS + a) * SLICE + lane];
            if (fabs(d) > 1e-15) inv = 1.0 / d;
            break;
        }
    }
    out[t] = (TO)inv;
}

template <class TI, class TO>
__global__ void k_convert(const TI* __restrict__ a, TO* __restrict__ b, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        b[i] = (TO)a[i];
}

// ||b - A x|| and ||b|| in FP64 (true residual, for the statistics)
template <int BS>
__global__ __launch_bounds__(256) void k_true_res(SellA<double> A, const double* __restrict__ b,
                                                  const double* __restrict__ x,
                                                  double2* __restrict__ red) {
    __shared__ double sh[4];
    const unsigned blk = fhd::xcd_block(blockIdx.x, gridDim.x);
    const unsigned s = blk * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    double tt = 0.0, bsq = 0.0;
    if (s < A.nslices) {
        const uint32_t row = s * SLICE + lane;
        const uint32_t base = A.soff[s], w = (A.soff[s + 1] - base) / SLICE;
        const uint32_t len = row < A.nrows ? A.rlen[row] : 0;
        double acc[BS];
#pragma unroll
        for (int a = 0; a < BS; ++a) acc[a] = 0.0;
        for (uint32_t k = 0; k < w; ++k)
            if (k < len) {
                const uint64_t e = (uint64_t)base + (uint64_t)k * SLICE;
                const uint32_t c = A.cols[e + lane];
                const double* v = A.vals + e * BS * BS + lane;
#pragma unroll
                for (int a = 0; a < BS; ++a)
#pragma unroll
                    for (int q = 0; q < BS; ++q)
                        acc[a] = fma(v[(a * BS + q) * SLICE], x[(uint64_t)c * BS + q], acc[a]);
            }
        if (row < A.nrows)
#pragma unroll
            for (int a = 0; a < BS; ++a) {
                const uint64_t i = (uint64_t)row * BS + a;
                const double t = b[i] - acc[a];
                tt += t * t;
                bsq += b[i] * b[i];
            }
    }
    tt = fhd::block_sum256(tt, sh);
    bsq = fhd::block_sum256(bsq, sh);
    if (threadIdx.x == 0) red[blockIdx.x] = make_double2(tt, bsq);
}

// ---------------------------------------------------------------- host side

template <class T>
static SellA<T> sell_of(const fh_matrix* a, const T* vals) {
    const Pattern& p = *a->pat;
    return {p.slice_off.p, p.row_len.p, p.sell_cols.p, vals, p.nrows, p.nslices};
}

static unsigned spmv_grid(const fh_matrix* a) { return std::max(1u, cdiv(a->pat->nslices, 4)); }

double spmv_algorithmic_bytes(const fh_matrix* a, int vbytes, int mbytes) {
    // matrix values + one 4-byte index per stored block + row pointers, one
    // read of x and one write of y (the SURVEY-style CSR/BSR formula)
    const double nnzb = (double)a->pat->nnzb;
    const int bs2 = a->bs * a->bs;
    return nnzb * (bs2 * (double)mbytes + 4.0) + 4.0 * (a->pat->nrows + 1) +
           2.0 * (double)a->n * vbytes;
}

void spmv_f64(const fh_matrix* a, const double* x, double* y, int order, hipStream_t s) {
    const SellA<double> A = sell_of(a, a->vals.p);
    const unsigned g = spmv_grid(a);
    if (a->bs == 1) {
        if (order == FH_ORDER_ORDERED)
            hipLaunchKernelGGL((k_spmv<1, double, true>), dim3(g), dim3(256), 0, s, A, x, y);
        else
            hipLaunchKernelGGL((k_spmv<1, double, false>), dim3(g), dim3(256), 0, s, A, x, y);
    } else {
        if (order == FH_ORDER_ORDERED)
            hipLaunchKernelGGL((k_spmv<3, double, true>), dim3(g), dim3(256), 0, s, A, x, y);
        else
            hipLaunchKernelGGL((k_spmv<3, double, false>), dim3(g), dim3(256), 0, s, A, x, y);
    }
    FH_LAUNCH_CHECK("spmv");
}

void inv_diag_f64(const fh_matrix* a, double* out, hipStream_t s) {
    const SellA<double> A = sell_of(a, a->vals.p);
    if (a->bs == 1)
        hipLaunchKernelGGL((k_inv_diag<1, double>), dim3(cdiv(a->n, 256)), dim3(256), 0, s, A, out);
    else
        hipLaunchKernelGGL((k_inv_diag<3, double>), dim3(cdiv(a->n, 256)), dim3(256), 0, s, A, out);
    FH_LAUNCH_CHECK("inv_diag");
}

void spmv_f32(const fh_matrix* a, const float* x, float* y, hipStream_t s) {
    const SellA<float> A = sell_of(a, a->vals32.p);
    const unsigned g = spmv_grid(a);
    if (a->bs == 1) hipLaunchKernelGGL((k_spmv<1, float, false>), dim3(g), dim3(256), 0, s, A, x, y);
    else hipLaunchKernelGGL((k_spmv<3, float, false>), dim3(g), dim3(256), 0, s, A, x, y);
    FH_LAUNCH_CHECK("spmv32");
}

// pinned host mirror of the state, one per device
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

template <int BS, class T, bool ORD, int PRE>
struct Launcher {
    static void spmv(const SellA<T>& A, unsigned g, hipStream_t s, const T* z, T* p0, T* p1, T* Ap,
                     CgState* st, const double2* red_in, int n_in, const double* bb,
                     double* red_out) {
        hipLaunchKernelGGL((k_cg_spmv<BS, T, ORD>), dim3(g), dim3(256), 0, s, A, z, p0, p1, Ap, st,
                           red_in, n_in, bb, red_out);
    }
    static void update(unsigned g, hipStream_t s, uint32_t n, T* x, T* r, const T* Ap, const T* p0,
                       const T* p1, const T* dinv, T* z, CgState* st, const double* red_in,
                       int n_in, double2* red_out) {
        hipLaunchKernelGGL((k_cg_update<T, ORD, PRE>), dim3(g), dim3(256), 0, s, n, x, r, Ap, p0,
                           p1, dinv, z, st, red_in, n_in, red_out);
    }
    static void residual(const SellA<double>& A, unsigned g, hipStream_t s, const double* b,
                         const double* x, const float* x32, const T* dinv, T* r, T* z,
                         double2* red, double* bb) {
        hipLaunchKernelGGL((k_residual<BS, T, ORD, PRE>), dim3(g), dim3(256), 0, s, A, b, x, x32,
                           dinv, r, z, red, bb);
    }
};

template <class T>
static void fill_vec(T* p, size_t n, hipStream_t s);
template <>
void fill_vec<double>(double* p, size_t n, hipStream_t s) { fill_f64(p, n, 0.0, s); }
template <>
void fill_vec<float>(float* p, size_t n, hipStream_t s) { fill_f32(p, n, 0.0f, s); }


// The solve proper.  T = Krylov vector/matrix type (double for FP64, float for
// the mixed modes).
template <int BS, class T, bool ORD, int PRE>
static int run_cg(const fh_matrix* a, const SolveArgs& args, fh_solve_stats* stats) {
    using L = Launcher<BS, T, ORD, PRE>;
    constexpr bool MIX = std::is_same<T, float>::value;
    const int dev = a->dev;
    DeviceCtx& d = device(dev);
    hipStream_t s = d.stream;
    const uint32_t n = a->n;
    const unsigned gs = spmv_grid(a);
    const unsigned gu = std::max(1u, std::min<unsigned>(cdiv(n, 256 * 4), 1024));
    const auto t_wall0 = std::chrono::steady_clock::now();

    const SellA<double> A64 = sell_of(a, a->vals.p);
    SellA<T> AK;
    if constexpr (MIX) {
        ensure_vals32(const_cast<fh_matrix*>(a));
        AK = sell_of(a, a->vals32.p);
    } else {
        AK = A64;
    }

    // Krylov vectors
    DBuf<T> r(dev, n), z(dev, n), p0(dev, n), p1(dev, n), Ap(dev, n);
    DBuf<float> x32;
    DBuf<T> dinv;
    const size_t ndinv = PRE == 0 ? n : (size_t)(n / 3) * 9;
    dinv.alloc(dev, ndinv);
    if (PRE == 1) {
        if constexpr (MIX) {
            hipLaunchKernelGGL((k_convert<double, float>), dim3(256), dim3(256), 0, s,
                               args.block_dinv, (float*)dinv.p, (uint64_t)ndinv);
        } else {
            d2d((double*)dinv.p, args.block_dinv, ndinv, s);
        }
    } else if (args.inv_diag_override) {
        hipLaunchKernelGGL((k_convert<double, T>), dim3(256), dim3(256), 0, s,
                           args.inv_diag_override, dinv.p, (uint64_t)n);
    } else {
        hipLaunchKernelGGL((k_inv_diag<BS, T>), dim3(cdiv(n, 256)), dim3(256), 0, s, A64, dinv.p);
    }
    FH_LAUNCH_CHECK("dinv");
    fill_vec(p0.p, n, s);
    fill_vec(p1.p, n, s);
    // x64 = x0 or 0
    double* x = args.x;
    if (args.x0) d2d(x, args.x0, n, s);
    else fill_f64(x, n, 0.0, s);
    T* xk;  // the vector the Krylov updates accumulate into
    if constexpr (MIX) {
        x32.alloc(dev, n);
        fill_f32(x32.p, n, 0.0f, s);
        xk = (T*)x32.p;
    } else {
        xk = (T*)x;
    }
    DBuf<double> redA(dev, gs), bb(dev, gs);
    DBuf<double2> redB(dev, std::max(gs, gu));

    CgState* hs = host_state(dev);
    std::memset(hs, 0, sizeof *hs);
    hs->tol = args.tol;
    hs->max_iter = args.max_iter;
    hs->delta = MIX ? std::atof(std::getenv("FH_RR_DELTA") ? std::getenv("FH_RR_DELTA") : "0.1") : 0.0;
    hs->mixed = MIX ? 1 : 0;
    hs->restart = args.precision == FH_PREC_MIXED_IR ? 1 : 0;
    hs->rep_mark = 1;
    DBuf<CgState> st(dev, 1);
    FH_HIP(hipMemcpyAsync(st.p, hs, sizeof(CgState), hipMemcpyHostToDevice, s));

    Timer tk(s);
    // r0 = b - A x0 (FP64), z0 = D^{-1} r0, partials (r.r, r.z), b.b
    L::residual(A64, gs, s, args.b, x, nullptr, dinv.p, r.p, z.p, redB.p, bb.p);
    FH_LAUNCH_CHECK("residual");
    int n_in_spmv = (int)gs;  // partial count feeding the next spmv kernel

    const int M = std::max(1, env_int("FH_CG_CHUNK", 16));
    const bool use_graph = env_int("FH_CG_GRAPH", 1) != 0;
    // optional per-launch timing of the SpMV kernel with HIP events on this stream
    const bool time_spmv = env_int("FH_TIME_SPMV", 0) != 0;
    std::vector<hipEvent_t> ev;
    double spmv_ms = 0.0;
    uint64_t spmv_n = 0;
    if (time_spmv) {
        ev.resize(2 * M);
        for (auto& e : ev) FH_HIP(hipEventCreate(&e));
    }
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    auto launch_chunk = [&](int nin_first) {
        for (int m = 0; m < M; ++m) {
            if (time_spmv) FH_HIP(hipEventRecord(ev[2 * m], s));
            L::spmv(AK, gs, s, z.p, p0.p, p1.p, Ap.p, st.p, redB.p, m == 0 ? nin_first : (int)gu,
                    bb.p, redA.p);
            if (time_spmv) FH_HIP(hipEventRecord(ev[2 * m + 1], s));
            L::update(gu, s, n, xk, r.p, Ap.p, p0.p, p1.p, dinv.p, z.p, st.p, redA.p, (int)gs,
                      redB.p);
        }
    };
    uint64_t launched = 0;
    bool first_chunk = true;
    for (;;) {
        // the first spmv after init / replacement reads gs residual partials;
        // later ones read gu update partials
        if (use_graph && !time_spmv && !first_chunk && n_in_spmv == (int)gu) {
            if (!gexec) {
                FH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                launch_chunk((int)gu);
                FH_HIP(hipStreamEndCapture(s, &graph));
                FH_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
            }
            FH_HIP(hipGraphLaunch(gexec, s));
        } else {
            launch_chunk(n_in_spmv);
            FH_LAUNCH_CHECK("cg_chunk");
        }
        first_chunk = false;
        n_in_spmv = (int)gu;
        launched += M;
        FH_HIP(hipMemcpyAsync(hs, st.p, sizeof(CgState), hipMemcpyDeviceToHost, s));
        FH_HIP(hipStreamSynchronize(s));
        if (time_spmv && hs->status == ST_RUNNING && !hs->need_rep) {
            // every SpMV of this chunk did a full product: include it
            for (int m = 0; m < M; ++m) {
                float ms = 0.f;
                FH_HIP(hipEventElapsedTime(&ms, ev[2 * m], ev[2 * m + 1]));
                spmv_ms += ms;
                ++spmv_n;
            }
        }
        if (hs->status != ST_RUNNING) break;
        if (hs->need_rep) {
            // residual replacement: t = b - A64 (x64 + x32) in FP64
            L::residual(A64, gs, s, args.b, x, x32.p, dinv.p, r.p, z.p, redB.p, nullptr);
            hipLaunchKernelGGL(k_commit, dim3(std::min(cdiv(n, 256), 1024u)), dim3(256), 0, s, n,
                               x, x32.p, st.p, 1);
            FH_LAUNCH_CHECK("replacement");
            n_in_spmv = (int)gs;
        }
        if (launched > args.max_iter + (hs->n_rep + 2) * (uint64_t)M + 64) {
            // safety net: the device state machine always terminates before this
            break;
        }
    }
    if constexpr (MIX) {
        if (hs->status != ST_CONVERGED)
            hipLaunchKernelGGL(k_commit, dim3(std::min(cdiv(n, 256), 1024u)), dim3(256), 0, s, n, x,
                               x32.p, st.p, 0);
    }
    const double kms = tk.stop_ms();
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);

    // true FP64 residual of the returned x
    DBuf<double2> tr(dev, gs);
    if (a->bs == 1)
        hipLaunchKernelGGL(k_true_res<1>, dim3(gs), dim3(256), 0, s, A64, args.b, x, tr.p);
    else
        hipLaunchKernelGGL(k_true_res<3>, dim3(gs), dim3(256), 0, s, A64, args.b, x, tr.p);
    std::vector<double2> htr(gs);
    d2h(htr.data(), tr.p, gs, s);
    FH_HIP(hipMemcpyAsync(hs, st.p, sizeof(CgState), hipMemcpyDeviceToHost, s));
    FH_HIP(hipStreamSynchronize(s));
    double tt = 0, bsq = 0;
    for (auto& v : htr) {
        tt += v.x;
        bsq += v.y;
    }
    const double wall =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wall0).count();

    int code = FH_OK;
    if (hs->status == ST_BREAKDOWN) code = FH_ERR_BREAKDOWN;
    else if (hs->status != ST_CONVERGED) code = FH_ERR_NOT_CONVERGED;
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->iterations = hs->status == ST_RUNNING ? hs->it_upd : hs->iters;
        stats->replacements = hs->n_rep;
        stats->final_residual = hs->res;
        stats->rel_residual = hs->rel;
        stats->true_rel_residual = bsq > 0 ? std::sqrt(tt) / std::sqrt(bsq) : std::sqrt(tt);
        stats->wall_ms = wall;
        stats->kernel_ms = kms;
        const int vb = sizeof(T), mb = sizeof(T);
        stats->spmv_bytes = spmv_algorithmic_bytes(a, vb, mb);
        // SpMV + the vector passes of the fused loop: spmv gathers z, p_old and
        // writes p_new, Ap; update reads x, r, Ap, p, dinv and writes x, r, z
        const double dinv_b = PRE == 0 ? (double)n * sizeof(T) : (double)n * 3 * sizeof(T);
        stats->iter_bytes = stats->spmv_bytes + (double)n * sizeof(T) * 11.0 + dinv_b -
                            2.0 * (double)n * vb;
        const double it = (double)std::max<uint64_t>(stats->iterations, 1);
        stats->effective_bw_gbps = stats->iter_bytes * it / (kms * 1e-3) / 1e9;
        const double flops = 2.0 * (double)a->nnz_scalar + 12.0 * n;
        stats->gflops = flops * it / (kms * 1e-3) / 1e9;
        stats->spmv_avg_ms = spmv_n ? spmv_ms / (double)spmv_n : 0.0;
        stats->spmv_timed = spmv_n;
        stats->device = dev;
        stats->precision = args.precision;
        stats->status = code;
    }
    return code;
}

int solve_device(const fh_matrix* a, const SolveArgs& args, fh_solve_stats* st) {
    const bool mixed = args.precision == FH_PREC_MIXED || args.precision == FH_PREC_MIXED_IR;
    const bool ord = args.order == FH_ORDER_ORDERED && !mixed;
    const int pre = args.block_dinv ? 1 : 0;
    FH_CHECK(!(pre && a->bs != 3), "block Jacobi needs 3x3 blocks");
    if (a->bs == 1) {
        if (mixed) return run_cg<1, float, false, 0>(a, args, st);
        return ord ? run_cg<1, double, true, 0>(a, args, st) : run_cg<1, double, false, 0>(a, args, st);
    }
    if (pre) {
        if (mixed) return run_cg<3, float, false, 1>(a, args, st);
        return ord ? run_cg<3, double, true, 1>(a, args, st) : run_cg<3, double, false, 1>(a, args, st);
    }
    if (mixed) return run_cg<3, float, false, 0>(a, args, st);
    return ord ? run_cg<3, double, true, 0>(a, args, st) : run_cg<3, double, false, 0>(a, args, st);
}

}  // namespace fh

using namespace fh;

extern "C" {

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
        FH_CHECK(precision >= 0 && precision <= 2, "unknown precision");
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
