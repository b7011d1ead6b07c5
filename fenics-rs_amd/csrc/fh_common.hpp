// fh_common.hpp — error handling, device contexts, device buffers and small
// device-side helpers shared by every translation unit of libfenics_hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fenics_hip.h"

namespace fh {

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define FH_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) {                                                       \
            char b_[256];                                                             \
            std::snprintf(b_, sizeof b_, "HIP error %d (%s) in %s", (int)e_,          \
                          hipGetErrorString(e_), #call);                              \
            throw ::fh::Error(FH_ERR_HIP, b_);                                        \
        }                                                                             \
    } while (0)

#define FH_CHECK(cond, msg)                                   \
    do {                                                      \
        if (!(cond)) throw ::fh::Error(FH_ERR_BAD_INPUT, msg); \
    } while (0)

#define FH_LAUNCH_CHECK(name) FH_HIP(hipGetLastError())

// One process-wide lock held for the whole of every exported call: calls
// share each device's stream (solves capture HIP graphs on it) and the pinned
// CG state mirror, so calls from several threads run one after the other.
// Recursive, so an exported call may call another, or a caller's callback
// may call back into the library on the same thread.
std::recursive_mutex& api_lock();

// Wrap the body of every exported function: never let an exception cross the
// C ABI; record the message for fh_last_error() (per thread).
template <class F>
int guard(F&& f) {
    std::lock_guard<std::recursive_mutex> lk(api_lock());
    try {
        f();
        return FH_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host out of memory");
        return FH_ERR_BAD_INPUT;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return FH_ERR_BAD_INPUT;
    }
}

// ---------------------------------------------------------------- devices
struct DeviceCtx {
    int id = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t prop{};
    int n_cu = 0;
};
// Lazily created per-device context (stream at default priority), cached for
// the life of the process.  Makes `dev` current.
DeviceCtx& device(int dev);
int resolve_device(int dev);  // -1 -> best device (largest free memory)

// ---------------------------------------------------------------- memory
// Per-device caching allocator: freed blocks are kept by exact byte size and
// handed back to later requests of the same size, so repeated solves of one
// problem allocate nothing after the first.
void* dmalloc(int dev, size_t bytes);
void dfree(int dev, void* p, size_t bytes);
void release_cache(int dev);

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    int dev = 0;
    DBuf() = default;
    DBuf(int d, size_t count) { alloc(d, count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n), dev(o.dev) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) {
            reset();
            p = o.p; n = o.n; dev = o.dev;
            o.p = nullptr; o.n = 0;
        }
        return *this;
    }
    ~DBuf() { reset(); }
    void alloc(int d, size_t count) {
        reset();
        dev = d;
        n = count;
        if (count) p = (T*)dmalloc(d, count * sizeof(T));
    }
    void reset() {
        if (p) dfree(dev, p, n * sizeof(T));
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
    T* get() const { return p; }
};

template <class T>
void h2d(T* dst, const T* src, size_t n, hipStream_t s) {
    if (n) FH_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
}
template <class T>
void d2h(T* dst, const T* src, size_t n, hipStream_t s) {
    if (n) FH_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s));
}
template <class T>
void d2d(T* dst, const T* src, size_t n, hipStream_t s) {
    if (n) FH_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, s));
}

inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// Fill kernels instead of hipMemsetAsync so that every piece of the solve can
// be captured into a HIP graph.
void fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);
void fill_f64(double* p, size_t n, double v, hipStream_t s);
void fill_f32(float* p, size_t n, float v, hipStream_t s);
void fill_u8(uint8_t* p, size_t n, uint8_t v, hipStream_t s);
// Exclusive prefix sum of n uint32 counts into out[0..n] (out[n] = total);
// returns the total (synchronises the stream).
uint64_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, int dev, hipStream_t s);

// Event-timed region on a stream.
struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s;
    explicit Timer(hipStream_t st) : s(st) {
        FH_HIP(hipEventCreate(&a));
        FH_HIP(hipEventCreate(&b));
        FH_HIP(hipEventRecord(a, s));
    }
    double stop_ms() {
        FH_HIP(hipEventRecord(b, s));
        FH_HIP(hipEventSynchronize(b));
        float ms = 0;
        FH_HIP(hipEventElapsedTime(&ms, a, b));
        return ms;
    }
    ~Timer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

}  // namespace fh

// ---------------------------------------------------------------- device helpers
#ifdef __HIPCC__
namespace fhd {

// Deterministic wavefront sum (fixed butterfly over 64 lanes).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Deterministic block sum for blockDim.x == 256 (4 waves); result valid in
// every thread.  `sh` must hold 4 doubles.
__device__ __forceinline__ double block_sum256(double v, double* sh) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    const double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
    return r;
}

// Deterministic block sum for NT threads (NT / 64 waves, a power of two);
// result valid in every thread.  `sh` must hold NT / 64 doubles.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sh) {
    constexpr int NW = NT / 64;
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double t[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) t[i] = sh[i];
#pragma unroll
    for (int n = NW; n > 1; n >>= 1)  // adjacent pairs: ((t0 + t1) + (t2 + t3)) + ...
#pragma unroll
        for (int i = 0; i < n / 2; ++i) t[i] = t[2 * i] + t[2 * i + 1];
    __syncthreads();
    return t[0];
}

// Sum of m partials written by a previous kernel, in a fixed order, by an
// NT-thread block; identical in every block that calls it.
template <int NT>
__device__ __forceinline__ double reduce_partials(const double* __restrict__ part, int m,
                                                  double* sh) {
    double s = 0.0;
    int i = threadIdx.x;
    for (; i + 3 * NT < m; i += 4 * NT) {  // four independent loads in flight
        const double a = part[i], b = part[i + NT], c = part[i + 2 * NT], d = part[i + 3 * NT];
        s += (a + b) + (c + d);
    }
    for (; i < m; i += NT) s += part[i];
    return block_sum<NT>(s, sh);
}

// Sum of m partials written by a previous kernel, in a fixed order, by a
// 256-thread block; identical in every block that calls it.
__device__ __forceinline__ double reduce_partials256(const double* __restrict__ part, int m,
                                                     double* sh) {
    double s = 0.0;
    int i = threadIdx.x;
    for (; i + 768 < m; i += 1024) {  // four independent loads in flight
        const double a = part[i], b = part[i + 256], c = part[i + 512], d = part[i + 768];
        s += (a + b) + (c + d);
    }
    for (; i < m; i += 256) s += part[i];
    return block_sum256(s, sh);
}

// XCD-aware block order: the hardware deals consecutive workgroups round-robin
// over the 8 XCDs; remap so that each XCD's L2 sees one contiguous range of
// logical blocks (better reuse of the gathered vector entries).
__device__ __forceinline__ unsigned xcd_block(unsigned bid, unsigned nblk) {
    const unsigned nx = 8;
    const unsigned per = nblk / nx, rem = nblk % nx;
    const unsigned xcd = bid % nx, slot = bid / nx;
    // xcd < rem owns (per + 1) blocks, the others own `per`
    const unsigned base = xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
    return base + slot;
}

}  // namespace fhd
#endif
