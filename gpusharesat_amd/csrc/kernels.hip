// CDNA4 (gfx950, wave64) kernels of the clause-vs-assignment check.
//
// The usefulness rule of the reference (gpuShareLib/GpuRunner.cu:34-57,
// ReportComputer) reports slot s of a clause when every literal is false
// ("falsified") or when exactly one literal occurrence is undefined and every
// other one is false ("unit").  The sweep evaluates it in an equivalent two-pass
// form that reads one table row per literal:
//
//   pass 1  with f_l = slots where literal l is false, count the non-false
//           occurrences per slot, saturating at 2:
//               c2 |= c1 & ~f_l;  c1 |= ~f_l
//           falsified = ~c1;  one_nonfalse = c1 & ~c2.  A slot with c2 set can
//           never report, so the clause stops once every slot of every word of
//           its group has c2 set.
//   pass 2  (only when some slot has exactly one non-false occurrence) that
//           occurrence is undefined iff no literal of the clause is true:
//               unit = one_nonfalse & ~OR_l f_{l^1}
//           (literal l is true exactly where its negation l^1 is false).
//
// Duplicate literals are separate occurrences, as in the reference
// (GpuRunner.cu:68-89 walks the literal list), so "x, x" with x undefined is
// not unit.  tests/test_oracle.py checks the forms agree by brute force.
//
// Work mapping.  A group of G = W*S lanes owns one clause: lane (s, w) handles
// table word w and literal sub-lane s (S literals are read per step and
// combined with a butterfly over lanes).  With W = 16 (1024 slots) each
// literal step is one coalesced 128-byte row read by 16 lanes.  A wave holds
// 64/G groups; a task is a run of clauses of one length bucket; waves are
// persistent and take tasks round-robin.
//
// Hits.  A clause/word with a report becomes one 24-byte record.  Each wave
// compacts its records with a ballot + mbcnt prefix and appends them to a
// private chunk of kHitChunk records claimed with one atomic; a chunk is
// abandoned (fill recorded) when the next pass does not fit.  Capacity is
// sized on the host so that nothing is dropped (the reference drops on
// overflow, gpuShareLib/Reporter.cuh:46-48); if it ever is exceeded the claim
// counter still counts and the host re-runs the sweep.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace gss {

namespace {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t prefix_count(unsigned long long bal) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
}

template <int G>
__device__ __forceinline__ bool group_any(bool pred, int g) {
    unsigned long long b = __ballot(pred);
    if constexpr (G == 64) {
        return b != 0;
    } else {
        return ((b >> (g * G)) & ((1ull << G) - 1ull)) != 0;
    }
}

// Evaluates one clause for this lane's word.  All G lanes of the group call it
// together (group-uniform control flow).
template <int W, int S>
__device__ __forceinline__ void check_clause(const uint32_t *__restrict__ cl, int len,
                                             const uint64_t *__restrict__ table, int w, int sub,
                                             int g, uint64_t valid, uint64_t &fm, uint64_t &um) {
    constexpr int G = W * S;
    // Invalid slots start "dead" (two non-false occurrences).
    uint64_t c1 = ~valid, c2 = ~valid;
    for (int i0 = 0; i0 < len; i0 += S) {
        const int i = i0 + sub;
        uint64_t m1 = 0;
        if (i < len) m1 = ~table[(size_t)cl[i] * W + w];
        uint64_t m2 = 0;
#pragma unroll
        for (int d = W; d < G; d <<= 1) {
            const uint64_t p1 = __shfl_xor(m1, d);
            const uint64_t p2 = __shfl_xor(m2, d);
            m2 = m2 | p2 | (m1 & p1);
            m1 |= p1;
        }
        c2 |= m2 | (c1 & m1);
        c1 |= m1;
        if (!group_any<G>(~c2 != 0, g)) break;
    }
    fm = ~c1;
    const uint64_t cand = c1 & ~c2;
    um = 0;
    if (group_any<G>(cand != 0, g)) {
        uint64_t any_true = 0;
        for (int i0 = 0; i0 < len; i0 += S) {
            const int i = i0 + sub;
            uint64_t t = 0;
            if (i < len) t = table[(size_t)(cl[i] ^ 1u) * W + w];
#pragma unroll
            for (int d = W; d < G; d <<= 1) t |= __shfl_xor(t, d);
            any_true |= t;
            if (!group_any<G>((cand & ~any_true) != 0, g)) break;
        }
        um = cand & ~any_true;
    }
}

__device__ __forceinline__ int find_bucket(const BucketDesc *__restrict__ b, int nb, int task) {
    int lo = 0, hi = nb;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if ((int)b[mid].task_begin <= task) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <int W, int S>
__global__ __launch_bounds__(256) void k_sweep(SweepArgs a) {
    constexpr int G = W * S;
    constexpr int CPW = 64 / G;  // clauses per wave pass
    const int lane = (int)lane_id();
    const int g = lane / G;
    const int sub = (lane % G) / W;
    const int w = lane % W;
    const uint64_t valid = a.valid[w];
    const int gw = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int nw = (int)((gridDim.x * blockDim.x) >> 6);

    int chunk = -1;
    int fill = 0;
    for (int task = gw; task < a.ntasks; task += nw) {
        const int bi = find_bucket(a.buckets, a.nbuckets, task);
        const BucketDesc bd = a.buckets[bi];
        const int len = (int)bd.len;
        const int64_t row0 = (int64_t)(task - (int)bd.task_begin) * a.task_clauses;
        const int64_t rem = (int64_t)bd.count - row0;
        const int nrows = rem < a.task_clauses ? (int)rem : a.task_clauses;
        for (int it = 0; it < nrows; it += CPW) {
            const bool active = it + g < nrows;
            const int64_t row = row0 + it + g;
            uint64_t fm = 0, um = 0;
            if (active) {
                check_clause<W, S>(bd.lits + (size_t)row * len, len, a.table, w, sub, g, valid, fm,
                                   um);
            }
            const bool hit = active && sub == 0 && (fm | um) != 0;
            const unsigned long long bal = __ballot(hit);
            if (bal) {
                const int n = __popcll(bal);
                if (chunk < 0 || fill + n > kHitChunk) {
                    if (chunk >= 0 && lane == 0 && (uint32_t)chunk < a.chunk_cap)
                        a.fills[chunk] = (uint32_t)fill;
                    uint32_t c = 0;
                    if (lane == 0) c = atomicAdd(a.alloc, 1u);
                    chunk = (int)__shfl(c, 0);
                    fill = 0;
                }
                if (hit && (uint32_t)chunk < a.chunk_cap) {
                    gss_hit h;
                    h.row = (uint32_t)row;
                    h.len = (uint16_t)len;
                    h.word = (uint16_t)w;
                    h.false_mask = fm;
                    h.unit_mask = um;
                    a.hits[(size_t)chunk * kHitChunk + fill + prefix_count(bal)] = h;
                }
                fill += n;
            }
        }
    }
    if (chunk >= 0 && lane == 0 && (uint32_t)chunk < a.chunk_cap) a.fills[chunk] = (uint32_t)fill;
}

__global__ void k_reset_u32(uint32_t *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_fill_u64(uint64_t *p, int64_t n, uint64_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_scatter(const ScatterSeg *__restrict__ segs, int nseg,
                          const uint32_t *__restrict__ src) {
    for (int s = blockIdx.x; s < nseg; s += gridDim.x) {
        const ScatterSeg sg = segs[s];
        for (uint64_t i = threadIdx.x; i < sg.count; i += blockDim.x) sg.dst[i] = src[sg.src + i];
    }
}

// Row 2v of the table holds the slots where v is false (literal 2v false),
// row 2v+1 the slots where v is true (literal 2v+1 false).
__global__ void k_assig_apply(const gss_var_update *__restrict__ ups, int64_t n,
                              uint64_t *__restrict__ table, int words) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const gss_var_update u = ups[i];
    const size_t base = (size_t)u.var * 2 * words + u.word;
    table[base] = u.is_def & ~u.is_true;
    table[base + words] = u.is_def & u.is_true;
}

__global__ void k_assig_roll(const gss_var_update *__restrict__ ups, int64_t n,
                             uint64_t *__restrict__ table, int words,
                             const int8_t *__restrict__ last_pos) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const gss_var_update u = ups[i];
    const int p = last_pos[u.word];
    if (p < 0) return;
    const size_t base = (size_t)u.var * 2 * words + u.word;
    const uint64_t f = table[base], t = table[base + words];
    table[base] = ((f >> p) & 1ull) ? ~0ull : 0ull;
    table[base + words] = ((t >> p) & 1ull) ? ~0ull : 0ull;
}

template <int W, int S>
hipError_t launch_sweep_t(const SweepArgs &a, int blocks, hipStream_t s) {
    hipLaunchKernelGGL((k_sweep<W, S>), dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

template <int W, int S>
int occupancy_t() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sweep<W, S>, 256, 0) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

}  // namespace

int sweep_clauses_per_pass(int words) { return words <= 16 ? 4 : 64 / words; }

hipError_t launch_sweep(const SweepArgs &a, int words, int blocks, hipStream_t s) {
    switch (words) {
        case 1: return launch_sweep_t<1, 16>(a, blocks, s);
        case 2: return launch_sweep_t<2, 8>(a, blocks, s);
        case 4: return launch_sweep_t<4, 4>(a, blocks, s);
        case 8: return launch_sweep_t<8, 2>(a, blocks, s);
        case 16: return launch_sweep_t<16, 1>(a, blocks, s);
        case 32: return launch_sweep_t<32, 1>(a, blocks, s);
        case 64: return launch_sweep_t<64, 1>(a, blocks, s);
        default: return hipErrorInvalidValue;
    }
}

int sweep_occupancy(int words) {
    switch (words) {
        case 1: return occupancy_t<1, 16>();
        case 2: return occupancy_t<2, 8>();
        case 4: return occupancy_t<4, 4>();
        case 8: return occupancy_t<8, 2>();
        case 16: return occupancy_t<16, 1>();
        case 32: return occupancy_t<32, 1>();
        case 64: return occupancy_t<64, 1>();
        default: return 1;
    }
}

hipError_t launch_reset_u32(uint32_t *p, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reset_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t *p, int64_t n, uint64_t v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_scatter(const ScatterSeg *segs, int nseg, const uint32_t *src, hipStream_t s) {
    if (nseg <= 0) return hipSuccess;
    const int blocks = nseg < 4096 ? nseg : 4096;
    hipLaunchKernelGGL(k_scatter, dim3(blocks), dim3(256), 0, s, segs, nseg, src);
    return hipGetLastError();
}

hipError_t launch_assig_apply(const gss_var_update *ups, int64_t n, uint64_t *table, int words,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_assig_apply, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ups, n,
                       table, words);
    return hipGetLastError();
}

hipError_t launch_assig_roll(const gss_var_update *ups, int64_t n, uint64_t *table, int words,
                             const int8_t *last_pos, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_assig_roll, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ups, n,
                       table, words, last_pos);
    return hipGetLastError();
}

}  // namespace gss
