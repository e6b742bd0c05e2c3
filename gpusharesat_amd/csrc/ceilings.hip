// Device ceilings for the sweep's access patterns: a streaming 16-byte copy and
// random 128-byte line gathers (one line per 16-lane group, as the 16-word
// sweep reads a table row), timed with HIP events on a private stream.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "errors.h"
#include "gss.h"
#include "gss_ceilings.h"

namespace gss {
namespace {

int set_err(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return record_error(code, buf);
}

// Four 16-byte loads in flight per thread, then the four stores; streaming
// (non-temporal) on both sides.  Elements past the last full round of four
// are copied one by one.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_ceiling_copy(const u32x4 *__restrict__ src,
                                                      u32x4 *__restrict__ dst, uint64_t n) {
    const uint64_t nthreads = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t tid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t full = n / (4 * nthreads) * (4 * nthreads);
    for (uint64_t base = tid; base < full; base += 4 * nthreads) {
        u32x4 v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = __builtin_nontemporal_load(src + base + u * nthreads);
#pragma unroll
        for (int u = 0; u < 4; u++) __builtin_nontemporal_store(v[u], dst + base + u * nthreads);
    }
    for (uint64_t i = full + tid; i < n; i += nthreads) dst[i] = src[i];
}

// Uniform line index in [0, nlines) for read r: murmur3's 32-bit finalizer
// of (r, seed), scaled by a multiply-high.
__device__ __forceinline__ uint32_t random_line(uint64_t r, uint32_t seed, uint32_t nlines) {
    uint32_t h = (uint32_t)r * 0x9E3779B1u ^ (uint32_t)(r >> 32) ^ seed;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return (uint32_t)(((uint64_t)h * nlines) >> 32);
}

// Each group of G lanes reads U independent blocks of 8 G bytes per
// iteration (lane q reads word q of each; G = 16 is a 128-byte line, as the
// 16-word sweep reads a table row); reads past n_reads fall on block 0.  The
// xor of everything read is written per thread so the loads are kept.
template <int G, int U>
__global__ __launch_bounds__(256) void k_ceiling_gather(const uint64_t *__restrict__ table,
                                                        uint32_t nblocks, uint64_t nreads,
                                                        uint32_t seed, uint64_t *__restrict__ out) {
    const uint64_t tid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint32_t q = threadIdx.x % G;
    const uint64_t group = tid / G;
    const uint64_t ngroups = ((uint64_t)gridDim.x * blockDim.x) / G;
    uint64_t acc = 0;
    for (uint64_t base = group * U; base < nreads; base += ngroups * U) {
        uint64_t v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t r = base + u;
            const uint32_t b = r < nreads ? random_line(r, seed, nblocks) : 0u;
            v[u] = table[(uint64_t)b * G + q];
        }
#pragma unroll
        for (int u = 0; u < U; u++) acc ^= v[u];
    }
    out[tid] = acc;
}

// Device buffers, a pinned host buffer and the private stream and events of
// one measurement.
struct Scope {
    void *a = nullptr, *b = nullptr, *host = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Scope() {
        if (s) (void)hipStreamSynchronize(s);
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
        if (host) (void)hipHostFree(host);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
    }
};

#define CEIL_HIP(call)                                                                  \
    do {                                                                                \
        const hipError_t he_ = (call);                                                  \
        if (he_ != hipSuccess)                                                          \
            return set_err(GSS_E_HIP, "%s: %s", #call, hipGetErrorString(he_));         \
    } while (0)

int open_scope(int device, Scope &sc) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
        return set_err(GSS_E_INVAL, "no device %d", device);
    CEIL_HIP(hipSetDevice(device));
    CEIL_HIP(hipStreamCreateWithFlags(&sc.s, hipStreamNonBlocking));
    CEIL_HIP(hipEventCreate(&sc.e0));
    CEIL_HIP(hipEventCreate(&sc.e1));
    return GSS_OK;
}

int persistent_blocks(int device) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        cus <= 0)
        cus = 256;
    return cus * 8;  // 8 blocks of 4 waves: 8 waves per SIMD
}

template <typename Launch>
int time_reps(Scope &sc, int reps, Launch &&launch, double *ms) {
    CEIL_HIP(launch());  // untimed: first touch of the buffers, code load
    CEIL_HIP(hipEventRecord(sc.e0, sc.s));
    for (int r = 0; r < reps; r++) CEIL_HIP(launch());
    CEIL_HIP(hipEventRecord(sc.e1, sc.s));
    CEIL_HIP(hipEventSynchronize(sc.e1));
    float t = 0.f;
    CEIL_HIP(hipEventElapsedTime(&t, sc.e0, sc.e1));
    *ms = (double)t / reps;
    return GSS_OK;
}

template <int G>
hipError_t launch_gather(int unroll, int blocks, hipStream_t s, const uint64_t *t,
                                uint32_t nblocks, uint64_t nr, uint32_t seed, uint64_t *o) {
    switch (unroll) {
        case 1: hipLaunchKernelGGL((k_ceiling_gather<G, 1>), dim3(blocks), dim3(256), 0, s, t, nblocks, nr, seed, o); break;
        case 2: hipLaunchKernelGGL((k_ceiling_gather<G, 2>), dim3(blocks), dim3(256), 0, s, t, nblocks, nr, seed, o); break;
        case 4: hipLaunchKernelGGL((k_ceiling_gather<G, 4>), dim3(blocks), dim3(256), 0, s, t, nblocks, nr, seed, o); break;
        default: hipLaunchKernelGGL((k_ceiling_gather<G, 8>), dim3(blocks), dim3(256), 0, s, t, nblocks, nr, seed, o); break;
    }
    return hipGetLastError();
}

}  // namespace
}  // namespace gss

using namespace gss;

extern "C" {

int gss_device_query(int device, gss_device_info *out) {
    if (!out) return set_err(GSS_E_INVAL, "null output");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
        return set_err(GSS_E_INVAL, "no device %d", device);
    hipDeviceProp_t p;
    CEIL_HIP(hipGetDeviceProperties(&p, device));
    std::memset(out, 0, sizeof *out);
    std::snprintf(out->name, sizeof out->name, "%s", p.name);
    std::snprintf(out->arch, sizeof out->arch, "%s", p.gcnArchName);
    out->cus = p.multiProcessorCount;
    out->clock_khz = p.clockRate;
    out->l2_bytes = p.l2CacheSize;
    out->lds_per_cu = (int64_t)p.maxSharedMemoryPerMultiProcessor;
    out->hbm_bytes = (int64_t)p.totalGlobalMem;
    out->mem_clock_khz = p.memoryClockRate;
    out->mem_bus_bits = p.memoryBusWidth;
    return GSS_OK;
}

int gss_ceiling_copy(int device, int64_t bytes, int reps, double *gbps, double *ms) {
    if (bytes <= 0 || bytes % 16 || reps < 1 || !gbps || !ms)
        return set_err(GSS_E_INVAL, "copy: bytes must be a positive multiple of 16, reps >= 1");
    Scope sc;
    int rc = open_scope(device, sc);
    if (rc) return rc;
    if (hipMalloc(&sc.a, (size_t)bytes) != hipSuccess ||
        hipMalloc(&sc.b, (size_t)bytes) != hipSuccess)
        return set_err(GSS_E_NOMEM, "copy: cannot allocate 2 x %lld bytes", (long long)bytes);
    CEIL_HIP(hipMemsetAsync(sc.a, 0x5A, (size_t)bytes, sc.s));
    const uint64_t n = (uint64_t)bytes / 16;
    const int blocks = persistent_blocks(device);
    auto launch = [&]() {
        hipLaunchKernelGGL(k_ceiling_copy, dim3(blocks), dim3(256), 0, sc.s,
                           (const u32x4 *)sc.a, (u32x4 *)sc.b, n);
        return hipGetLastError();
    };
    if ((rc = time_reps(sc, reps, launch, ms))) return rc;
    *gbps = 2.0 * (double)bytes / (*ms * 1e-3) / 1e9;
    return GSS_OK;
}

int gss_ceiling_gather(int device, int64_t table_bytes, int64_t n_reads, int unroll, int reps,
                       double *gbps, double *ms) {
    return gss_ceiling_gather_block(device, table_bytes, 128, n_reads, unroll, reps, gbps, ms);
}

int gss_ceiling_gather_block(int device, int64_t table_bytes, int block_bytes, int64_t n_reads,
                             int unroll, int reps, double *gbps, double *ms) {
    if (block_bytes != 64 && block_bytes != 128 && block_bytes != 256)
        return set_err(GSS_E_INVAL, "gather: blocks of 64, 128 or 256 bytes");
    if (table_bytes <= 0 || table_bytes % block_bytes || table_bytes / block_bytes > 0xFFFFFFFFll ||
        n_reads <= 0 || reps < 1 || !gbps || !ms)
        return set_err(GSS_E_INVAL, "gather: table must be a positive multiple of the block "
                                    "(at most 2^32 blocks), n_reads > 0, reps >= 1");
    if (unroll != 1 && unroll != 2 && unroll != 4 && unroll != 8)
        return set_err(GSS_E_INVAL, "gather: unroll must be 1, 2, 4 or 8");
    Scope sc;
    int rc = open_scope(device, sc);
    if (rc) return rc;
    const int blocks = persistent_blocks(device);
    const size_t out_bytes = (size_t)blocks * 256 * sizeof(uint64_t);
    if (hipMalloc(&sc.a, (size_t)table_bytes) != hipSuccess ||
        hipMalloc(&sc.b, out_bytes) != hipSuccess)
        return set_err(GSS_E_NOMEM, "gather: cannot allocate a %lld-byte table",
                       (long long)table_bytes);
    CEIL_HIP(hipMemsetAsync(sc.a, 0xA5, (size_t)table_bytes, sc.s));
    const uint32_t nblocks = (uint32_t)(table_bytes / block_bytes);
    const uint64_t nr = (uint64_t)n_reads;
    const uint64_t *t = (const uint64_t *)sc.a;
    uint64_t *o = (uint64_t *)sc.b;
    uint32_t seed = 0x1234567u;
    auto launch = [&]() {
        seed = seed * 1664525u + 1013904223u;  // fresh positions every run
        return block_bytes == 64    ? launch_gather<8>(unroll, blocks, sc.s, t, nblocks, nr, seed, o)
               : block_bytes == 128 ? launch_gather<16>(unroll, blocks, sc.s, t, nblocks, nr, seed, o)
                                    : launch_gather<32>(unroll, blocks, sc.s, t, nblocks, nr, seed, o);
    };
    if ((rc = time_reps(sc, reps, launch, ms))) return rc;
    *gbps = (double)block_bytes * (double)n_reads / (*ms * 1e-3) / 1e9;
    return GSS_OK;
}

int gss_ceiling_pcie(int device, int64_t bytes, int reps, double *d2h_gbps, double *h2d_gbps) {
    if (bytes <= 0 || reps < 1 || !d2h_gbps || !h2d_gbps)
        return set_err(GSS_E_INVAL, "pcie: bytes > 0 and reps >= 1");
    Scope sc;
    int rc = open_scope(device, sc);
    if (rc) return rc;
    if (hipMalloc(&sc.a, (size_t)bytes) != hipSuccess ||
        hipHostMalloc(&sc.host, (size_t)bytes, hipHostMallocDefault) != hipSuccess)
        return set_err(GSS_E_NOMEM, "pcie: cannot allocate %lld bytes", (long long)bytes);
    CEIL_HIP(hipMemsetAsync(sc.a, 0x3C, (size_t)bytes, sc.s));
    double ms = 0;
    auto d2h = [&]() { return hipMemcpyAsync(sc.host, sc.a, (size_t)bytes, hipMemcpyDeviceToHost, sc.s); };
    if ((rc = time_reps(sc, reps, d2h, &ms))) return rc;
    *d2h_gbps = (double)bytes / (ms * 1e-3) / 1e9;
    auto h2d = [&]() { return hipMemcpyAsync(sc.a, sc.host, (size_t)bytes, hipMemcpyHostToDevice, sc.s); };
    if ((rc = time_reps(sc, reps, h2d, &ms))) return rc;
    *h2d_gbps = (double)bytes / (ms * 1e-3) / 1e9;
    return GSS_OK;
}

}  // extern "C"
