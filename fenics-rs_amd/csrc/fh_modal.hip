// fh_modal.hip — structural modal analysis on the GPU: the consumer of the CG
// named by fenics-rs's solve_vibration_modes / solve_vibration_modes_with_inner
// (eigen.rs:94-300), with the elasticity mass (mass.rs:86-126, assembled in
// fh_assemble.hip) and the Dirichlet fold (eigen.rs:51-78).
//
// Inverse power iteration with M-deflation against the converged modes, in the
// Dirichlet frame on full-length vectors, exactly the reference's sequence:
// mass rescaled by s = mean|diag K| / mean|diag M|, deterministic start vector
// per mode, inner solve K w = s M v from x0 = v by the device Jacobi-CG
// (eigen.rs:288-300 uses tol 1e-9, 1000 iterations), stop when the Rayleigh
// quotient moves by less than tol relative to itself.  Vectors stay on the
// device; dot products are fixed-order block reductions summed on the host.
#include <algorithm>
#include <cmath>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "fh_internal.hpp"

namespace fh {

__global__ __launch_bounds__(256) void k_dot_part(const double* __restrict__ a,
                                                  const double* __restrict__ b, uint64_t n,
                                                  double* __restrict__ part) {
    __shared__ double sh[4];
    double s = 0.0;
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
        s += a[i] * b[i];
    s = fhd::block_sum256(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// y -= c x (the reference's v[i] -= coeff * p[i])
__global__ void k_sub_scaled(double c, const double* __restrict__ x, double* __restrict__ y,
                             uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        y[i] -= c * x[i];
}

// x /= d (the reference's normalisations divide)
__global__ void k_div(double d, double* __restrict__ x, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        x[i] /= d;
}

// x *= a (the reference's mv *= m_scale); entries of constrained dofs set to
// 0 when `fixed` is given
__global__ void k_scale_mask(double a, double* __restrict__ x, const uint8_t* __restrict__ fixed,
                             uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        x[i] = (fixed && fixed[i]) ? 0.0 : x[i] * a;
}

// eigen.rs:144-151: deterministic start vector, zero on constrained entries
__global__ void k_init_mode(const uint8_t* __restrict__ fixed, uint64_t mode, double* __restrict__ v,
                            uint64_t n) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t q = ((i + 1) * 31 + (mode + 1) * 97) % 1000;
    v[i] = fixed[i] ? 0.0 : (double)q / 1000.0 - 0.5;
}

// eigen.rs:22-45 over the stored (non-zero after the fold) diagonal entries:
// block partials of sum |d| and of their count
template <int BS>
__global__ __launch_bounds__(256) void k_diag_stats(const uint32_t* __restrict__ soff,
                                                    const uint32_t* __restrict__ rlen,
                                                    const uint32_t* __restrict__ cols,
                                                    const double* __restrict__ vals, uint32_t nrows,
                                                    double2* __restrict__ part) {
    __shared__ double sh[4];
    double sum = 0.0, cnt = 0.0;
    for (uint64_t t = blockIdx.x * 256ull + threadIdx.x; t < (uint64_t)nrows * BS;
         t += (uint64_t)gridDim.x * 256) {
        const uint32_t r = t / BS, a = t % BS, s = r / SLICE, lane = r % SLICE;
        for (uint32_t k = 0; k < rlen[r]; ++k) {
            const uint64_t e0 = (uint64_t)soff[s] + (uint64_t)k * SLICE;
            if (cols[e0 + lane] == r) {
                const double d = vals[vpos(e0, a * BS + a, lane, BS * BS)];
                if (d != 0.0) {
                    sum += fabs(d);
                    cnt += 1.0;
                }
                break;
            }
        }
    }
    sum = fhd::block_sum256(sum, sh);
    cnt = fhd::block_sum256(cnt, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = make_double2(sum, cnt);
}

namespace {

struct Reducer {
    int dev;
    hipStream_t s;
    unsigned g = 512;
    DBuf<double> part;
    DBuf<double2> part2;
    std::vector<double> h;
    std::vector<double2> h2;
    Reducer(int d, hipStream_t st) : dev(d), s(st), part(d, 512), part2(d, 512), h(512), h2(512) {}
    double dot(const double* a, const double* b, uint64_t n) {
        hipLaunchKernelGGL(k_dot_part, dim3(g), dim3(256), 0, s, a, b, n, part.p);
        FH_LAUNCH_CHECK("dot");
        d2h(h.data(), part.p, g, s);
        FH_HIP(hipStreamSynchronize(s));
        double acc = 0.0;
        for (unsigned i = 0; i < g; ++i) acc += h[i];
        return acc;
    }
    double mean_abs_diag(const fh_matrix* a) {
        const Pattern& p = *a->pat;
        if (a->bs == 1)
            hipLaunchKernelGGL(k_diag_stats<1>, dim3(g), dim3(256), 0, s, p.slice_off.p,
                               p.row_len.p, p.sell_cols.p, a->vals.p, p.nrows, part2.p);
        else
            hipLaunchKernelGGL(k_diag_stats<3>, dim3(g), dim3(256), 0, s, p.slice_off.p,
                               p.row_len.p, p.sell_cols.p, a->vals.p, p.nrows, part2.p);
        FH_LAUNCH_CHECK("diag_stats");
        d2h(h2.data(), part2.p, g, s);
        FH_HIP(hipStreamSynchronize(s));
        double sum = 0.0, cnt = 0.0;
        for (unsigned i = 0; i < g; ++i) {
            sum += h2[i].x;
            cnt += h2[i].y;
        }
        return cnt > 0.0 ? sum / cnt : 0.0;
    }
};

unsigned vgrid(uint64_t n) { return std::max(1u, std::min(cdiv(n, 256), 4096u)); }

// Solves kd w = b (device vectors, x0 the previous iterate); returns the
// inner iteration count it spent, or 0 when it does not count them.
using InnerSolve = std::function<uint64_t(const fh_matrix* kd, const double* b_dev,
                                          const double* x0_dev, double* w_dev)>;

// eigen.rs:94-266: the iteration shared by both entry points
void vibration_modes(const fh_matrix* K, const fh_matrix* M, uint64_t nfix, const uint64_t* fixed,
                     uint64_t num_modes, uint64_t max_power_iters, double tol,
                     const InnerSolve& inner, double* freq_hz, double* omega, double* eigenvalue,
                     double* shapes, uint64_t* inner_iterations) {
    FH_CHECK(K && M && freq_hz && omega && eigenvalue, "null argument");
    FH_CHECK(K->n == M->n && K->dev == M->dev, "K and M must match in size and device");
    FH_CHECK(nfix == 0 || fixed, "null Dirichlet list");
    const int dev = K->dev;
    DeviceCtx& d = device(dev);
    hipStream_t s = d.stream;
    const uint64_t n = K->n;
    std::vector<uint64_t> keep;
    std::vector<uint8_t> isf(n, 0);
    for (uint64_t i = 0; i < nfix; ++i)
        if (fixed[i] < n) {
            keep.push_back(fixed[i]);
            isf[fixed[i]] = 1;
        }
    uint64_t nfree = 0;
    for (uint64_t i = 0; i < n; ++i) nfree += isf[i] ? 0 : 1;
    if (nfree == 0) throw Error(FH_ERR_BAD_INPUT, "No free degrees of freedom in system");
    const std::vector<double> zeros(keep.size(), 0.0);
    DBuf<double> b0(dev, n), b1(dev, n);
    fill_f64(b0.p, n, 0.0, s);
    std::unique_ptr<fh_matrix> kd(
        apply_dirichlet(K, b0.p, keep.size(), keep.data(), zeros.data(), b1.p, s, 1));
    std::unique_ptr<fh_matrix> md(
        apply_dirichlet(M, b0.p, keep.size(), keep.data(), zeros.data(), b1.p, s, 0));
    const uint8_t* dfix = kd->fixed.p;  // device mask of the constrained dofs

    Reducer red(dev, s);
    const double kdd = red.mean_abs_diag(kd.get()), mdd = red.mean_abs_diag(md.get());
    const double sc = (kdd > 0.0 && mdd > 0.0) ? kdd / mdd : 1.0;  // eigen.rs:122-131
    const unsigned g = vgrid(n);
    DBuf<double> v(dev, n), mv(dev, n), w(dev, n), kw(dev, n), mw(dev, n);
    DBuf<double> modes(dev, std::max<uint64_t>(num_modes, 1) * n);
    DBuf<double> mprev(dev, std::max<uint64_t>(num_modes, 1) * n);
    uint64_t inner_total = 0;
    const double pi = 3.14159265358979323846;
    for (uint64_t mi = 0; mi < num_modes; ++mi) {
        hipLaunchKernelGGL(k_init_mode, dim3(cdiv(n, 256)), dim3(256), 0, s, dfix, mi, v.p, n);
        // s M p for the converged modes (eigen.rs:155-160)
        if (mi > 0) {
            double* mp = mprev.p + (mi - 1) * n;
            spmv_f64(md.get(), modes.p + (mi - 1) * n, mp, FH_ORDER_FAST, s);
            hipLaunchKernelGGL(k_scale_mask, dim3(g), dim3(256), 0, s, sc, mp,
                               (const uint8_t*)nullptr, n);
        }
        auto deflate = [&](double* x) {  // x -= (x . sMp) p over the converged modes
            for (uint64_t q = 0; q < mi; ++q) {
                const double coeff = red.dot(x, mprev.p + q * n, n);
                hipLaunchKernelGGL(k_sub_scaled, dim3(g), dim3(256), 0, s, coeff,
                                   modes.p + q * n, x, n);
            }
        };
        deflate(v.p);
        spmv_f64(md.get(), v.p, mv.p, FH_ORDER_FAST, s);
        hipLaunchKernelGGL(k_scale_mask, dim3(g), dim3(256), 0, s, sc, mv.p,
                           (const uint8_t*)nullptr, n);
        const double norm_m = std::sqrt(red.dot(v.p, mv.p, n));
        if (norm_m > 1e-14) hipLaunchKernelGGL(k_div, dim3(g), dim3(256), 0, s, norm_m, v.p, n);
        double prev = 0.0;
        for (uint64_t it = 0; it < max_power_iters; ++it) {
            spmv_f64(md.get(), v.p, mv.p, FH_ORDER_FAST, s);
            hipLaunchKernelGGL(k_scale_mask, dim3(g), dim3(256), 0, s, sc, mv.p,
                               (const uint8_t*)nullptr, n);
            // inner solve K w = s M v from x0 = v (eigen.rs:184-189)
            inner_total += inner(kd.get(), mv.p, v.p, w.p);
            hipLaunchKernelGGL(k_scale_mask, dim3(g), dim3(256), 0, s, 1.0, w.p, dfix, n);
            deflate(w.p);
            spmv_f64(md.get(), w.p, mw.p, FH_ORDER_FAST, s);
            const double mwn = std::sqrt(red.dot(w.p, mw.p, n)) * std::sqrt(sc);
            if (mwn < 1e-15) break;
            hipLaunchKernelGGL(k_div, dim3(g), dim3(256), 0, s, mwn, w.p, n);
            spmv_f64(kd.get(), w.p, kw.p, FH_ORDER_FAST, s);
            const double ray = red.dot(w.p, kw.p, n);
            std::swap(v, w);
            const double denom = std::max(std::fabs(ray), std::fabs(prev));
            if (denom == 0.0 || std::fabs(ray - prev) / denom < tol) break;
            prev = ray;
        }
        // eigen.rs:242-252: final quotient with the unscaled mass
        spmv_f64(kd.get(), v.p, kw.p, FH_ORDER_FAST, s);
        spmv_f64(md.get(), v.p, mw.p, FH_ORDER_FAST, s);
        const double lam = std::max(red.dot(v.p, kw.p, n) / red.dot(v.p, mw.p, n), 0.0);
        const double om = std::sqrt(lam);
        eigenvalue[mi] = lam;
        omega[mi] = om;
        freq_hz[mi] = om / (2.0 * pi);
        d2d(modes.p + mi * n, v.p, n, s);
        if (shapes) d2h(shapes + mi * n, v.p, n, s);
    }
    FH_HIP(hipStreamSynchronize(s));
    if (inner_iterations) *inner_iterations = inner_total;
}

}  // namespace

}  // namespace fh

using namespace fh;

extern "C" {

int fh_assemble_elasticity_mass(fh_mesh* m, double density, fh_matrix** out) {
    return guard([&] {
        FH_CHECK(m && out, "null argument");
        *out = assemble_elasticity_mass(m, density);
    });
}

int fh_fold_dirichlet(const fh_matrix* a, uint64_t nfix, const uint64_t* dofs, int unit_diag,
                      fh_matrix** out) {
    return guard([&] {
        FH_CHECK(a && out, "null argument");
        FH_CHECK(nfix == 0 || dofs, "null Dirichlet list");
        hipStream_t s = device(a->dev).stream;
        // eigen.rs:57-62 ignores indices past the end
        std::vector<uint64_t> keep;
        for (uint64_t i = 0; i < nfix; ++i)
            if (dofs[i] < a->n) keep.push_back(dofs[i]);
        const std::vector<double> zeros(keep.size(), 0.0);
        DBuf<double> b0(a->dev, a->n), b1(a->dev, a->n);
        fill_f64(b0.p, a->n, 0.0, s);
        *out = apply_dirichlet(a, b0.p, keep.size(), keep.data(), zeros.data(), b1.p, s,
                               unit_diag ? 1 : 0);
    });
}

int fh_vibration_modes(const fh_matrix* K, const fh_matrix* M, uint64_t nfix,
                       const uint64_t* fixed, uint64_t num_modes, uint64_t max_power_iters,
                       double tol, double inner_tol, uint64_t inner_max_iter, int precision,
                       double* freq_hz, double* omega, double* eigenvalue, double* shapes,
                       uint64_t* inner_iterations) {
    return guard([&] {
        // eigen.rs:288-300: the device Jacobi-CG on the folded K
        const InnerSolve cg = [&](const fh_matrix* kd, const double* b, const double* x0,
                                  double* w) -> uint64_t {
            SolveArgs sa{b, x0, w, inner_tol, inner_max_iter, precision, FH_ORDER_FAST};
            fh_solve_stats st;
            const int code = solve_device(kd, sa, &st);
            if (code != FH_OK)
                throw Error(code, code == FH_ERR_BREAKDOWN
                                      ? "Breakdown in CG: curvature p^T A p is near zero"
                                      : "CG failed to converge within the iteration limit");
            return st.iterations;
        };
        vibration_modes(K, M, nfix, fixed, num_modes, max_power_iters, tol, cg, freq_hz, omega,
                        eigenvalue, shapes, inner_iterations);
    });
}

int fh_vibration_modes_with_inner(const fh_matrix* K, const fh_matrix* M, uint64_t nfix,
                                  const uint64_t* fixed, uint64_t num_modes,
                                  uint64_t max_power_iters, double tol, fh_inner_solve_fn inner,
                                  void* inner_ctx, double* freq_hz, double* omega,
                                  double* eigenvalue, double* shapes) {
    return guard([&] {
        FH_CHECK(inner, "null inner solve");
        FH_CHECK(K, "null argument");
        const uint64_t n = K->n;
        std::vector<double> hb(n), hx0(n), hw(n);
        const InnerSolve host = [&](const fh_matrix* kd, const double* b, const double* x0,
                                    double* w) -> uint64_t {
            hipStream_t s = device(kd->dev).stream;
            d2h(hb.data(), b, n, s);
            d2h(hx0.data(), x0, n, s);
            FH_HIP(hipStreamSynchronize(s));
            const int code = inner(n, hb.data(), hx0.data(), hw.data(), inner_ctx);
            if (code != 0)
                throw Error(FH_ERR_BAD_INPUT, "inner solve failed with code " +
                                                  std::to_string(code));
            h2d(w, hw.data(), n, s);
            return 0;
        };
        vibration_modes(K, M, nfix, fixed, num_modes, max_power_iters, tol, host, freq_hz, omega,
                        eigenvalue, shapes, nullptr);
    });
}

}  // extern "C"
