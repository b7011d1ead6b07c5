// fh_device.hip — device discovery (replaces the dlopen'd CUDA driver table of
// fenics-rs gpu.rs:36-242), per-device contexts, the caching allocator, fill
// kernels and a deterministic prefix sum.
#include <algorithm>
#include <map>
#include <mutex>
#include <string>

#include "fh_internal.hpp"

namespace fh {

static thread_local std::string g_last_error;
void set_last_error(const std::string& m) { g_last_error = m; }
std::recursive_mutex& api_lock() {
    static std::recursive_mutex m;
    return m;
}

static std::mutex g_dev_mu;
static std::map<int, std::unique_ptr<DeviceCtx>> g_devs;

int resolve_device(int dev) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        throw Error(FH_ERR_NO_DEVICE, "no HIP device available");
    if (dev >= 0) {
        if (dev >= count) throw Error(FH_ERR_BAD_INPUT, "device index out of range");
        return dev;
    }
    // gpu.rs:237-242: the device with the largest free memory
    int best = 0;
    size_t best_free = 0;
    for (int i = 0; i < count; ++i) {
        FH_HIP(hipSetDevice(i));
        size_t fr = 0, tot = 0;
        FH_HIP(hipMemGetInfo(&fr, &tot));
        if (fr > best_free) {
            best_free = fr;
            best = i;
        }
    }
    return best;
}

DeviceCtx& device(int dev) {
    std::lock_guard<std::mutex> lk(g_dev_mu);
    auto it = g_devs.find(dev);
    if (it == g_devs.end()) {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
            throw Error(FH_ERR_NO_DEVICE, "no HIP device available");
        if (dev < 0 || dev >= count) throw Error(FH_ERR_BAD_INPUT, "device index out of range");
        auto c = std::make_unique<DeviceCtx>();
        c->id = dev;
        FH_HIP(hipSetDevice(dev));
        FH_HIP(hipGetDeviceProperties(&c->prop, dev));
        c->n_cu = c->prop.multiProcessorCount;
        FH_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        it = g_devs.emplace(dev, std::move(c)).first;
    }
    FH_HIP(hipSetDevice(dev));
    return *it->second;
}

// ---------------------------------------------------------------- allocator
static std::mutex g_mem_mu;
static std::map<int, std::multimap<size_t, void*>> g_free;

void* dmalloc(int dev, size_t bytes) {
    if (!bytes) return nullptr;
    bytes = (bytes + 255) & ~(size_t)255;
    {
        std::lock_guard<std::mutex> lk(g_mem_mu);
        auto& fl = g_free[dev];
        auto it = fl.find(bytes);
        if (it != fl.end()) {
            void* p = it->second;
            fl.erase(it);
            return p;
        }
    }
    FH_HIP(hipSetDevice(dev));
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        release_cache(dev);
        FH_HIP(hipSetDevice(dev));
        FH_HIP(hipMalloc(&p, bytes));
    }
    return p;
}

void dfree(int dev, void* p, size_t bytes) {
    if (!p) return;
    bytes = (bytes + 255) & ~(size_t)255;
    std::lock_guard<std::mutex> lk(g_mem_mu);
    g_free[dev].emplace(bytes, p);
}

void release_cache(int dev) {
    std::multimap<size_t, void*> fl;
    {
        std::lock_guard<std::mutex> lk(g_mem_mu);
        fl.swap(g_free[dev]);
    }
    if (fl.empty()) return;
    // frees must not race with kernels still using recycled blocks
    (void)hipSetDevice(dev);
    (void)hipDeviceSynchronize();
    for (auto& kv : fl) (void)hipFree(kv.second);
}

// ---------------------------------------------------------------- fills
template <class T>
__global__ void k_fill(T* __restrict__ p, size_t n, T v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}
template <class T>
static void fill(T* p, size_t n, T v, hipStream_t s) {
    if (!n) return;
    const unsigned g = std::min<size_t>(cdiv(n, 256), 65536);
    hipLaunchKernelGGL(k_fill<T>, dim3(g), dim3(256), 0, s, p, n, v);
    FH_LAUNCH_CHECK("fill");
}
void fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) { fill(p, n, v, s); }
void fill_f64(double* p, size_t n, double v, hipStream_t s) { fill(p, n, v, s); }
void fill_f32(float* p, size_t n, float v, hipStream_t s) { fill(p, n, v, s); }
void fill_u8(uint8_t* p, size_t n, uint8_t v, hipStream_t s) { fill(p, n, v, s); }

// ---------------------------------------------------------------- scan
// Three-phase exclusive scan: per-1024 block scan, a one-block scan of the
// block totals (serial over 1024-wide chunks), then the block offsets added.
// Integer arithmetic, so the result is exact and deterministic.
constexpr int SCAN_B = 1024;

__device__ __forceinline__ uint64_t block_incl_scan(uint64_t v, uint64_t* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = 1; o < SCAN_B; o <<= 1) {
        const uint64_t a = t >= o ? sh[t - o] : 0;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    return sh[t];
}

__global__ __launch_bounds__(SCAN_B) void k_scan_blocks(const uint32_t* __restrict__ in,
                                                         uint64_t* __restrict__ tmp,
                                                         uint64_t* __restrict__ bsum, size_t n) {
    __shared__ uint64_t sh[SCAN_B];
    const size_t i = (size_t)blockIdx.x * SCAN_B + threadIdx.x;
    const uint64_t v = i < n ? in[i] : 0;
    const uint64_t inc = block_incl_scan(v, sh);
    if (i < n) tmp[i] = inc - v;
    if (threadIdx.x == SCAN_B - 1) bsum[blockIdx.x] = inc;
}

__global__ __launch_bounds__(SCAN_B) void k_scan_sums(uint64_t* __restrict__ bsum, size_t nb,
                                                       uint64_t* __restrict__ total) {
    __shared__ uint64_t sh[SCAN_B];
    uint64_t carry = 0;
    for (size_t base = 0; base < nb; base += SCAN_B) {
        const size_t i = base + threadIdx.x;
        const uint64_t v = i < nb ? bsum[i] : 0;
        const uint64_t inc = block_incl_scan(v, sh);
        if (i < nb) bsum[i] = carry + inc - v;
        carry += sh[SCAN_B - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void k_scan_add(const uint64_t* __restrict__ tmp, const uint64_t* __restrict__ bsum,
                           uint32_t* __restrict__ out, size_t n, const uint64_t* total) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint32_t)(tmp[i] + bsum[i / SCAN_B]);
    if (i == n) out[n] = (uint32_t)*total;
}

uint64_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, int dev, hipStream_t s) {
    const size_t nb = std::max<size_t>(1, cdiv(n, SCAN_B));
    DBuf<uint64_t> tmp(dev, std::max<size_t>(n, 1)), bsum(dev, nb), total(dev, 1);
    if (n) hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(SCAN_B), 0, s, in, tmp.p, bsum.p, n);
    else fill(bsum.p, 1, (uint64_t)0, s);
    FH_LAUNCH_CHECK("scan_blocks");
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_B), 0, s, bsum.p, n ? nb : 1, total.p);
    FH_LAUNCH_CHECK("scan_sums");
    hipLaunchKernelGGL(k_scan_add, dim3(cdiv(n + 1, 256)), dim3(256), 0, s, tmp.p, bsum.p, out, n,
                       total.p);
    FH_LAUNCH_CHECK("scan_add");
    uint64_t t = 0;
    FH_HIP(hipMemcpyAsync(&t, total.p, 8, hipMemcpyDeviceToHost, s));
    FH_HIP(hipStreamSynchronize(s));
    if (t > 0xffffffffull) throw Error(FH_ERR_UNSUPPORTED, "index range exceeds 32 bits");
    return t;
}

// ---------------------------------------------------------------- ceilings
// Bandwidth ceilings of the device: a read stream (16-byte non-temporal
// loads, the form of the CG matrix stream; one partial per block is written so
// the loads stay live), a copy, and a write stream.  Grid-stride loops over
// 16-byte elements, four independent loads in flight per thread.
typedef double dbl2_t __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(256) void k_stream_read(const dbl2_t* __restrict__ a, uint64_t n,
                                                     double* __restrict__ part) {
    __shared__ double sh[4];
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    uint64_t i = blockIdx.x * 256ull + threadIdx.x;
    dbl2_t s0 = {0.0, 0.0}, s1 = s0, s2 = s0, s3 = s0;
    for (; i + 3 * stride < n; i += 4 * stride) {
        s0 += __builtin_nontemporal_load(a + i);
        s1 += __builtin_nontemporal_load(a + i + stride);
        s2 += __builtin_nontemporal_load(a + i + 2 * stride);
        s3 += __builtin_nontemporal_load(a + i + 3 * stride);
    }
    for (; i < n; i += stride) s0 += __builtin_nontemporal_load(a + i);
    const dbl2_t t = (s0 + s1) + (s2 + s3);
    const double v = fhd::block_sum256(t.x + t.y, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ __launch_bounds__(256) void k_stream_copy(const dbl2_t* __restrict__ a,
                                                     dbl2_t* __restrict__ b, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    uint64_t i = blockIdx.x * 256ull + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const dbl2_t x0 = __builtin_nontemporal_load(a + i);
        const dbl2_t x1 = __builtin_nontemporal_load(a + i + stride);
        const dbl2_t x2 = __builtin_nontemporal_load(a + i + 2 * stride);
        const dbl2_t x3 = __builtin_nontemporal_load(a + i + 3 * stride);
        b[i] = x0;
        b[i + stride] = x1;
        b[i + 2 * stride] = x2;
        b[i + 3 * stride] = x3;
    }
    for (; i < n; i += stride) b[i] = __builtin_nontemporal_load(a + i);
}

__global__ __launch_bounds__(256) void k_stream_write(dbl2_t* __restrict__ b, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    const dbl2_t v = {1.0, 2.0};
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += stride) b[i] = v;
}

}  // namespace fh

using namespace fh;

extern "C" {

const char* fh_last_error(void) { return g_last_error.c_str(); }

int fh_device_count(int* count) {
    return guard([&] {
        FH_CHECK(count, "null output");
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
        *count = c;
    });
}

int fh_get_device_info(int index, fh_device_info* out) {
    return guard([&] {
        FH_CHECK(out, "null output");
        DeviceCtx& d = device(index);
        std::memset(out, 0, sizeof *out);
        out->index = index;
        // some ROCm builds leave the marketing name empty: fall back to the ISA name
        std::snprintf(out->name, sizeof out->name, "%s",
                      d.prop.name[0] ? d.prop.name : d.prop.gcnArchName);
        std::snprintf(out->arch, sizeof out->arch, "%s", d.prop.gcnArchName);
        size_t fr = 0, tot = 0;
        FH_HIP(hipMemGetInfo(&fr, &tot));
        out->total_memory_mib = tot >> 20;
        out->free_memory_mib = fr >> 20;
        out->compute_units = d.n_cu;
        out->clock_mhz = d.prop.clockRate / 1000;
    });
}

int fh_best_device(int* index) {
    return guard([&] {
        FH_CHECK(index, "null output");
        *index = resolve_device(-1);
    });
}

int fh_stream_bandwidth(int dev, uint64_t bytes, int kind, int reps, double* gbps,
                        double* best_ms) {
    return guard([&] {
        FH_CHECK(gbps && best_ms, "null output");
        FH_CHECK(kind >= 0 && kind <= 2, "kind must be 0 (read), 1 (copy) or 2 (write)");
        FH_CHECK(bytes >= 4096 && reps >= 1, "bytes must be at least 4096, reps at least 1");
        DeviceCtx& d = device(dev);
        hipStream_t s = d.stream;
        const uint64_t n = bytes / 16;
        DBuf<dbl2_t> a(dev, n), b;
        if (kind == 1) b.alloc(dev, n);
        const unsigned grid = (unsigned)std::min<uint64_t>(cdiv(n, 256), 8u * 256u * (unsigned)d.n_cu / 4u);
        DBuf<double> part(dev, grid);
        fill_f64(reinterpret_cast<double*>(a.p), 2 * n, 1.0, s);
        hipEvent_t e0, e1;
        FH_HIP(hipEventCreate(&e0));
        FH_HIP(hipEventCreate(&e1));
        float best = 1e30f;
        for (int r = 0; r <= reps; ++r) {  // the first launch is a warm-up
            FH_HIP(hipEventRecord(e0, s));
            if (kind == 0) hipLaunchKernelGGL(k_stream_read, dim3(grid), dim3(256), 0, s, a.p, n, part.p);
            else if (kind == 1) hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, s, a.p, b.p, n);
            else hipLaunchKernelGGL(k_stream_write, dim3(grid), dim3(256), 0, s, a.p, n);
            FH_LAUNCH_CHECK("stream");
            FH_HIP(hipEventRecord(e1, s));
            FH_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            FH_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (r > 0) best = std::min(best, ms);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        const double moved = (double)n * 16.0 * (kind == 1 ? 2.0 : 1.0);
        *best_ms = best;
        *gbps = moved / (best * 1e-3) / 1e9;
    });
}

int fh_synchronize(int dev) {
    return guard([&] {
        DeviceCtx& d = device(dev);
        FH_HIP(hipStreamSynchronize(d.stream));
    });
}

}  // extern "C"
