// Clause residency on the device: per-clause hit statistics folded from the
// sweep's own records, the reference's activity policy, and a learned policy
// (a 12 -> 64 -> 64 -> 1 MLP trained online by the kernels below).
//
// Per-clause state lives in structure-of-arrays indexed by the clause's tag
// (its row tag in the engine's tagged DB).  A round is: the caller appends
// clauses (gss_res_add), sweeps, then gss_res_account maps every hit record
// (len, row) -> tag through the bucket tag arrays and updates the state; when
// the DB is over budget gss_res_reduce scores every clause, finds the
// approximate median from a 20000-bucket histogram (as the reference's
// approxNthAct, gpuShareLib/Clauses.cu:484-525) and compacts the DB on the
// device.
//
// The MLP kernels: training runs one sample per wave (lane i = hidden unit i
// of both 64-unit layers, the parameters staged in LDS) and writes the
// activations and the back-propagated deltas; the weight gradient is a sum of
// per-sample outer products, computed per (parameter, chunk of 32 samples)
// and reduced in a fixed order, so gradients are deterministic.  Scoring runs
// one clause per thread.  ReLU'(0) = 0.  Loss: weighted logistic, mean over
// the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <vector>

#include "engine_internal.h"
#include "errors.h"
#include "gss.h"
#include "gss_residency.h"
#include "kernels.h"

namespace gss {
namespace {

constexpr int F = GSS_RES_FEATURES;
constexpr int H = GSS_RES_HIDDEN;
constexpr int P = GSS_RES_PARAMS;
constexpr int OW1 = 0, OB1 = H * F, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H,
              OB3 = OW3 + H;
static_assert(OB3 + 1 == P, "parameter layout");
constexpr int kBuckets = 20000;
constexpr float kRescale = 1e19f;  // RESCALE_CONST, gpuShareLib/Clauses.cuh:39
constexpr int kGradChunk = 32;     // samples per partial gradient

int set_err(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return record_error(code, buf);
}

#define RES_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t _e = (call);                                                         \
        if (_e != hipSuccess)                                                           \
            return set_err(GSS_E_HIP, "%s failed: %s", #call, hipGetErrorString(_e));   \
    } while (0)

template <typename T>
struct Buf {
    T *p = nullptr;
    size_t cap = 0;
    int reserve(size_t n) {
        if (n <= cap) return GSS_OK;
        if (p) {
            (void)hipDeviceSynchronize();
            (void)hipFree(p);
        }
        p = nullptr;
        cap = 0;
        if (hipMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
            return set_err(GSS_E_NOMEM, "device allocation of %zu bytes failed", n * sizeof(T));
        cap = n;
        return GSS_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Per-clause state, indexed by tag.
struct TagState {
    uint32_t *len;
    int32_t *birth;
    int32_t *last_hit;     // -1 = never
    uint32_t *hit_rounds;  // rounds with at least one record
    uint32_t *cnt;         // records of the round being accounted
    uint32_t *cnt_prev;    // records of the last accounted round
    uint64_t *solv;        // words with a record this round
    uint64_t *solv_cum;    // words with a record, ever
    float *ema_f, *ema_s;  // EMA of "reported this round" (0.5, 0.05)
    float *tot;            // records, ever
    float *act;            // reference activity
    float *score;          // last reduction's score
    uint32_t *gen;         // allocation generation of the tag
    uint8_t *alive;
};

// ---- features ----------------------------------------------------------------

__device__ __forceinline__ void features(const TagState &s, uint32_t t, int32_t round, float incr,
                                         float f[F]) {
    const float len = (float)s.len[t];
    const float age = (float)max(0, round - s.birth[t]);
    const int32_t lh = s.last_hit[t];
    f[0] = log2f(len) * (1.0f / 9.0f);
    f[1] = len <= 2.0f ? 1.0f : 0.0f;
    f[2] = log1pf(age) * 0.125f;
    f[3] = lh == round ? 1.0f : 0.0f;
    f[4] = log1pf((float)s.cnt_prev[t]) * (1.0f / 6.0f);
    f[5] = s.ema_f[t];
    f[6] = s.ema_s[t];
    f[7] = log1pf(s.tot[t]) * (1.0f / 12.0f);
    f[8] = lh < 0 ? 1.0f : fminf(1.0f, log1pf((float)(round - lh)) * 0.125f);
    f[9] = (float)s.hit_rounds[t] / (age + 1.0f);
    f[10] = (float)__popcll(s.solv_cum[t]) * (1.0f / 64.0f);
    const float rel = (logf(fmaxf(s.act[t], FLT_MIN)) - logf(incr)) * (1.0f / 64.0f);
    f[11] = fmaxf(-1.0f, fminf(1.0f, rel));
}

// ---- state kernels -----------------------------------------------------------

__global__ void k_res_init(TagState s, const uint32_t *__restrict__ tags,
                           const int32_t *__restrict__ lens, const float *__restrict__ act0,
                           int64_t n, int32_t round) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t t = tags[i];
    s.len[t] = (uint32_t)lens[i];
    s.birth[t] = round;
    s.last_hit[t] = -1;
    s.hit_rounds[t] = 0;
    s.cnt[t] = 0;
    s.cnt_prev[t] = 0;
    s.solv[t] = 0;
    s.solv_cum[t] = 0;
    s.ema_f[t] = 0.f;
    s.ema_s[t] = 0.f;
    s.tot[t] = 0.f;
    s.act[t] = act0[i];
    s.score[t] = 0.f;
    s.gen[t] += 1;
    s.alive[t] = 1;
}

// Hit records -> tag state, one 256-thread block per quarter of a claimed
// chunk (one record per thread: a sweep claims few chunks, so one block per
// chunk would leave the kernel waiting on four dependent lookup chains in a
// row).  A clause's records (one per assignment word that hit it) are adjacent in
// its chunk, so each wave merges runs of equal (len, row): the run's first
// lane ORs the run's word bits with a segmented suffix scan and issues one
// atomicAdd of the run length and one atomicOr for the whole run.
constexpr uint32_t kRecBlock = 256;  // k_res_records block size
constexpr uint32_t kRecParts = kHitChunk / kRecBlock;
static_assert(kHitChunk % kRecBlock == 0, "a chunk splits into whole blocks");

__global__ void k_res_records(const gss_hit *__restrict__ hits, const uint32_t *__restrict__ fills,
                              uint32_t chunks, const uint32_t *const *__restrict__ tagptr,
                              int ntagptr, TagState s, unsigned long long *pairs) {
    __shared__ unsigned long long blk;
    if (threadIdx.x == 0) blk = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    unsigned long long mine = 0;
    const uint64_t parts = (uint64_t)chunks * kRecParts;
    for (uint64_t g = blockIdx.x; g < parts; g += gridDim.x) {
        const uint32_t c = (uint32_t)(g / kRecParts);
        const uint32_t f = fills[c];  // uniform in the block
        const uint32_t base = (uint32_t)(g % kRecParts) * kRecBlock;
        if (base >= f) continue;
        const uint32_t i = base + threadIdx.x;
        const bool on = i < f;
        uint32_t row = 0, len = 0xffffffffu;
        uint64_t bit = 0;
        const uint32_t *tp = nullptr;
        if (on) {
            const gss_hit h = hits[(size_t)c * kHitChunk + i];
            row = h.row;
            len = h.len;
            bit = 1ull << (h.word & 63);
            tp = (int)len < ntagptr ? tagptr[len] : nullptr;
            if (tp) mine += (unsigned long long)__popcll(h.false_mask | h.unit_mask);
        }
        const uint32_t prow = (uint32_t)__shfl_up((int)row, 1);
        const uint32_t plen = (uint32_t)__shfl_up((int)len, 1);
        const bool head = on && (lane == 0 || prow != row || plen != len);
        // a run ends at the next head or at the first lane past the fill
        const uint64_t bound = __ballot(head || !on);
        const uint64_t above = lane == 63 ? 0ull : bound & (~0ull << (lane + 1));
        const int end = above ? __ffsll((unsigned long long)above) - 1 : 64;
        uint64_t acc = bit;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t lo = (uint32_t)__shfl_down((int)(uint32_t)acc, d);
            const uint32_t hi = (uint32_t)__shfl_down((int)(uint32_t)(acc >> 32), d);
            if (lane + d < end) acc |= ((uint64_t)hi << 32) | lo;
        }
        if (head && tp) {
            const uint32_t t = tp[row];
            atomicAdd(&s.cnt[t], (uint32_t)(end - lane));
            atomicOr((unsigned long long *)&s.solv[t], (unsigned long long)acc);
        }
    }
    // Device-scope atomics on one address are serialised (~10 ns each):
    // blocks with nothing to add issue none.
    if (mine) atomicAdd(&blk, mine);
    __syncthreads();
    if (threadIdx.x == 0 && blk) atomicAdd(pairs, blk);
}

// All of a tag's fields are read before any is written: the state arrays
// are separate allocations the compiler cannot prove apart, so interleaved
// reads and writes would each wait for the store before them.
__global__ void k_res_update(TagState s, int64_t cap, int32_t round, float incr,
                             unsigned int *max_act_bits) {
    unsigned int mx = 0;  // largest activity seen (positive floats order as uints)
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < cap;
         t += (int64_t)gridDim.x * blockDim.x) {
        if (!s.alive[t]) continue;
        const uint32_t c = s.cnt[t];
        const float ef = s.ema_f[t], es = s.ema_s[t], tot = s.tot[t];
        const int32_t lh = s.last_hit[t];
        const uint32_t hr = s.hit_rounds[t];
        const uint64_t sc = s.solv_cum[t], sv = s.solv[t];
        float a = s.act[t];
        const bool hit = c > 0;
        // one bump per record, added one at a time as the reference does
        for (uint32_t k = 0; k < c; k++) a += incr;
        s.ema_f[t] = 0.5f * ef + (hit ? 0.5f : 0.0f);
        s.ema_s[t] = 0.95f * es + (hit ? 0.05f : 0.0f);
        s.tot[t] = tot + (float)c;
        s.last_hit[t] = hit ? round : lh;
        s.hit_rounds[t] = hr + (hit ? 1u : 0u);
        s.solv_cum[t] = sc | sv;
        s.cnt_prev[t] = c;
        s.act[t] = a;
        s.cnt[t] = 0;
        s.solv[t] = 0;
        mx = max(mx, __float_as_uint(a));
    }
    // Device-scope atomics on one address are serialised (~10 ns each), so
    // the block's maximum is formed in LDS and sent only when it exceeds
    // the value already there.  A stale read can only be lower than the
    // true value, so skipping when it is not lower is exact.
    __shared__ unsigned int bmx;
    if (threadIdx.x == 0) bmx = 0;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) mx = max(mx, (unsigned int)__shfl_xor((int)mx, d));
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && mx) atomicMax(&bmx, mx);
    __syncthreads();
    if (threadIdx.x == 0 && bmx > __hip_atomic_load(max_act_bits, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(max_act_bits, bmx);
}

__global__ void k_res_rescale(TagState s, int64_t cap) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < cap;
         t += (int64_t)gridDim.x * blockDim.x)
        if (s.alive[t]) s.act[t] /= kRescale;
}

__global__ void k_res_features(TagState s, const uint32_t *__restrict__ tags, int64_t n,
                               int32_t round, float incr, float *__restrict__ x,
                               uint32_t *__restrict__ gen_out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t t = tags[i];
    float f[F];
    features(s, t, round, incr, f);
#pragma unroll
    for (int k = 0; k < F; k++) x[i * F + k] = f[k];
    if (gen_out) gen_out[i] = s.gen[t];
}

__global__ void k_res_gather_act(TagState s, const uint32_t *__restrict__ tags, int64_t n,
                                 float *__restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = s.act[tags[i]];
}

// Labels of a snapshot taken at round r0: reported after r0; samples whose
// clause left (or whose tag was reused) get weight 0.
__global__ void k_res_labels(TagState s, const uint32_t *__restrict__ tags,
                             const uint32_t *__restrict__ gens, int64_t n, int32_t r0,
                             float *__restrict__ y, float *__restrict__ valid,
                             unsigned long long *counts) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t t = tags[i];
    const bool ok = s.alive[t] && s.gen[t] == gens[i];
    const bool pos = ok && s.last_hit[t] > r0;
    y[i] = pos ? 1.f : 0.f;
    valid[i] = ok ? 1.f : 0.f;
    if (ok) atomicAdd(&counts[0], 1ull);
    if (pos) atomicAdd(&counts[1], 1ull);
}

// ---- MLP ---------------------------------------------------------------------

__device__ __forceinline__ void stage_params(const float *__restrict__ params, float *lds) {
    for (int i = threadIdx.x; i < P; i += blockDim.x) lds[i] = params[i];
    __syncthreads();
}

// Register arrays are only indexed by fully unrolled loop counters, so they
// stay in VGPRs; the loop over the second layer's units stays rolled.
__device__ __forceinline__ float forward(const float *w, const float x[F], float a1[H]) {
#pragma unroll
    for (int i = 0; i < H; i++) {
        float z = w[OB1 + i];
#pragma unroll
        for (int k = 0; k < F; k++) z = fmaf(w[OW1 + i * F + k], x[k], z);
        a1[i] = z > 0.f ? z : 0.f;
    }
    float z3 = w[OB3];
#pragma unroll 1
    for (int j = 0; j < H; j++) {
        float z = w[OB2 + j];
#pragma unroll
        for (int i = 0; i < H; i++) z = fmaf(w[OW2 + j * H + i], a1[i], z);
        if (z > 0.f) z3 = fmaf(w[OW3 + j], z, z3);
    }
    return z3;
}

__device__ __forceinline__ float sigmoidf(float z) { return 1.0f / (1.0f + expf(-z)); }

// Weighted logistic loss of logit z and label y: softplus(z) - y z.
__device__ __forceinline__ float logistic_loss(float z, float y) {
    return fmaxf(z, 0.f) - z * y + log1pf(expf(-fabsf(z)));
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
    return v;
}

// Forward + backward, one sample per wave: lane i is hidden unit i of both
// layers (H = 64 = the wave).  W2 is staged in LDS in both orientations, so
// the forward (row j by lane j) and the backward (column i by lane i) read
// consecutive words.  A wave hands its layer's outputs to all its lanes
// through its own LDS row.  Writes a1, a2 (activations), d1, d2 (deltas at
// the hidden pre-activations), d3 (at the output logit), already multiplied
// by the sample weight.
__global__ __launch_bounds__(256) void k_mlp_fwd_bwd(
    const float *__restrict__ params, const float *__restrict__ x, const float *__restrict__ y,
    const float *__restrict__ wt, const uint32_t *__restrict__ idx, int64_t n, float w_pos,
    float w_neg, float *__restrict__ a1o, float *__restrict__ a2o, float *__restrict__ d1o,
    float *__restrict__ d2o, float *__restrict__ d3o, float *loss_sum) {
    static_assert(H == 64, "one hidden unit per lane of a wave");
    __shared__ float w1[H * F], b1[H], w2[H * H], w2t[H * H], b2[H], w3[H];
    __shared__ float row[4][H];
    __shared__ float b3s, blk;
    const int tid = threadIdx.x;
    for (int i = tid; i < H * F; i += blockDim.x) w1[i] = params[OW1 + i];
    for (int i = tid; i < H * H; i += blockDim.x) {
        const float v = params[OW2 + i];
        w2[i] = v;
        w2t[(i % H) * H + i / H] = v;
    }
    if (tid < H) {
        b1[tid] = params[OB1 + tid];
        b2[tid] = params[OB2 + tid];
        w3[tid] = params[OW3 + tid];
    }
    if (tid == 0) {
        b3s = params[OB3];
        blk = 0.f;
    }
    __syncthreads();
    const int lane = tid & 63, wv = tid >> 6;
    float my_loss = 0.f;
    for (int64_t b = (int64_t)blockIdx.x * 4 + wv; b < n; b += (int64_t)gridDim.x * 4) {
        const int64_t src = idx ? (int64_t)idx[b] : b;
        float z1 = b1[lane];
#pragma unroll
        for (int k = 0; k < F; k++) z1 = fmaf(w1[lane * F + k], x[src * F + k], z1);
        const float a1 = z1 > 0.f ? z1 : 0.f;
        a1o[b * H + lane] = a1;
        row[wv][lane] = a1;
        __builtin_amdgcn_wave_barrier();
        float z2 = b2[lane];
#pragma unroll 16
        for (int i = 0; i < H; i++) z2 = fmaf(w2t[i * H + lane], row[wv][i], z2);
        const float a2 = z2 > 0.f ? z2 : 0.f;
        a2o[b * H + lane] = a2;
        const float z3 = b3s + wave_sum(w3[lane] * a2);
        const float yv = y[src];
        const float sw = (wt ? wt[src] : 1.f) * (yv > 0.5f ? w_pos : w_neg);
        const float d3 = sw * (sigmoidf(z3) - yv);
        if (lane == 0) {
            d3o[b] = d3;
            my_loss += sw * logistic_loss(z3, yv);
        }
        // d2_j = d3 w3_j relu'(z2_j); d1_i = relu'(z1_i) sum_j W2[j][i] d2_j
        const float d2 = a2 > 0.f ? d3 * w3[lane] : 0.f;
        d2o[b * H + lane] = d2;
        __builtin_amdgcn_wave_barrier();  // every lane has read a1 from the row
        row[wv][lane] = d2;
        __builtin_amdgcn_wave_barrier();
        float acc = 0.f;
#pragma unroll 16
        for (int j = 0; j < H; j++) acc = fmaf(w2[j * H + lane], row[wv][j], acc);
        d1o[b * H + lane] = a1 > 0.f ? acc : 0.f;
        __builtin_amdgcn_wave_barrier();  // the next sample rewrites the row
    }
    atomicAdd(&blk, my_loss);
    __syncthreads();
    if (tid == 0) atomicAdd(loss_sum, blk);
}

// Partial gradient of every parameter over one chunk of samples.
__global__ void k_mlp_grad_partial(const float *__restrict__ x, const uint32_t *__restrict__ idx,
                                   const float *__restrict__ a1, const float *__restrict__ a2,
                                   const float *__restrict__ d1, const float *__restrict__ d2,
                                   const float *__restrict__ d3, int64_t n,
                                   float *__restrict__ partial) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t b0 = (int64_t)blockIdx.y * kGradChunk;
    const int64_t b1 = min<int64_t>(n, b0 + kGradChunk);
    if (p >= P) return;
    float g = 0.f;
    if (p < OB1) {
        const int i = p / F, k = p % F;
        for (int64_t b = b0; b < b1; b++) {
            const int64_t src = idx ? (int64_t)idx[b] : b;
            g = fmaf(d1[b * H + i], x[src * F + k], g);
        }
    } else if (p < OW2) {
        const int i = p - OB1;
        for (int64_t b = b0; b < b1; b++) g += d1[b * H + i];
    } else if (p < OB2) {
        const int q = p - OW2, j = q / H, i = q % H;
        for (int64_t b = b0; b < b1; b++) g = fmaf(d2[b * H + j], a1[b * H + i], g);
    } else if (p < OW3) {
        const int j = p - OB2;
        for (int64_t b = b0; b < b1; b++) g += d2[b * H + j];
    } else if (p < OB3) {
        const int j = p - OW3;
        for (int64_t b = b0; b < b1; b++) g = fmaf(d3[b], a2[b * H + j], g);
    } else {
        for (int64_t b = b0; b < b1; b++) g += d3[b];
    }
    partial[(size_t)blockIdx.y * P + p] = g;
}

// grad = scale * sum of the partials (fixed order); with lr != 0 also a
// momentum step: m = mu m + grad, w -= lr m.
__global__ void k_mlp_reduce_step(const float *__restrict__ partial, int nchunks, float scale,
                                  float *__restrict__ grad, float *__restrict__ params,
                                  float *__restrict__ mom, float lr, float mu) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    // loads issued eight at a time, summed in chunk order (deterministic)
    float g = 0.f;
    int c = 0;
    for (; c + 8 <= nchunks; c += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = partial[(size_t)(c + u) * P + p];
#pragma unroll
        for (int u = 0; u < 8; u++) g += v[u];
    }
    for (; c < nchunks; c++) g += partial[(size_t)c * P + p];
    g *= scale;
    grad[p] = g;
    if (lr != 0.f) {
        const float m = mu * mom[p] + g;
        mom[p] = m;
        params[p] -= lr * m;
    }
}

__global__ __launch_bounds__(256) void k_mlp_score(const float *__restrict__ params,
                                                   const float *__restrict__ x, int64_t n,
                                                   float *__restrict__ out) {
    __shared__ float w[P];
    stage_params(params, w);
    const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (b >= n) return;
    float xv[F], a1[H];
#pragma unroll
    for (int k = 0; k < F; k++) xv[k] = x[b * F + k];
    out[b] = sigmoidf(forward(w, xv, a1));
}

// Scores every live clause and adds it to a kBuckets histogram: the MLP's
// probability, uniform buckets over [0, 1] ...
__global__ __launch_bounds__(256) void k_res_score_learned(TagState s, int64_t cap, int32_t round,
                                                           float incr,
                                                           const float *__restrict__ params,
                                                           unsigned int *hist) {
    // one clause per thread: in a grid-stride loop the compiler hoists the
    // loop-invariant weight reads out of it into (spilled) registers
    __shared__ float w[P];
    stage_params(params, w);
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= cap || !s.alive[t]) return;
    float f[F], a1[H];
    features(s, (uint32_t)t, round, incr, f);
    const float sc = sigmoidf(forward(w, f, a1));
    int bucket = (int)floorf(sc * kBuckets);
    bucket = bucket < 0 ? 0 : bucket >= kBuckets ? kBuckets - 1 : bucket;
    s.score[t] = sc;
    atomicAdd(&hist[bucket], 1u);
}

// ... or the activity, buckets log-spaced over the float range as the
// reference's approxNthAct (Clauses.cu:484-525).
__global__ void k_res_score_activity(TagState s, int64_t cap, unsigned int *hist) {
    const float lowest = logf(FLT_MIN), largest = logf(FLT_MAX);
    const float step = (largest - lowest) / kBuckets;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < cap;
         t += (int64_t)gridDim.x * blockDim.x) {
        if (!s.alive[t]) continue;
        const float sc = s.act[t];
        int bucket = (int)floor(((double)logf(sc) - (double)lowest) / (double)step);
        bucket = bucket < 0 ? 0 : bucket >= kBuckets ? kBuckets - 1 : bucket;
        s.score[t] = sc;
        atomicAdd(&hist[bucket], 1u);
    }
}

__global__ void k_res_keep(TagState s, int64_t cap, float thr, uint8_t *__restrict__ keep,
                           uint8_t *__restrict__ evicted) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < cap;
         t += (int64_t)gridDim.x * blockDim.x) {
        const bool was = s.alive[t] != 0;
        const bool k = was && (s.len[t] <= 2 || s.score[t] >= thr);
        keep[t] = k ? 1 : 0;
        evicted[t] = (was && !k) ? 1 : 0;
        if (!k) s.alive[t] = 0;
    }
}

unsigned grid(int64_t n, unsigned cap = 65536) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : b > cap ? cap : b);
}

// splitmix64 -> uniform in [-1, 1)
double uniform_pm1(uint64_t &state) {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

struct Snapshot {
    Buf<float> x;
    Buf<uint32_t> tags, gens;
    int64_t n = 0;
};

}  // namespace
}  // namespace gss

using namespace gss;

struct gss_res {
    gss_engine *e = nullptr;
    int64_t cap = 0;
    float decay = 0.99999f;
    float incr = 1.0f;  // activity increment (the reference's clause act incr)
    std::vector<uint8_t> alive;  // host mirror
    // per-tag arrays
    Buf<uint32_t> len, hit_rounds, cnt, cnt_prev, gen;
    Buf<int32_t> birth, last_hit;
    Buf<uint64_t> solv, solv_cum;
    Buf<float> ema_f, ema_s, tot, act, score;
    Buf<uint8_t> alive_d, keep, evicted;
    // MLP
    Buf<float> params, mom, grad, partial;
    Buf<float> a1, a2, d1, d2, d3, loss;
    Buf<float> xs, ys, ws;
    Buf<uint32_t> idx;
    // scratch
    Buf<uint32_t> tags_in, hist, maxbits;
    Buf<int32_t> lens_in;
    Buf<float> act_in, fout;
    Buf<unsigned long long> counters;
    Buf<const uint32_t *> tagptr;
    std::map<int32_t, Snapshot> snaps;

    TagState state() {
        return TagState{len.p,     birth.p, last_hit.p, hit_rounds.p, cnt.p,   cnt_prev.p,
                        solv.p,    solv_cum.p, ema_f.p, ema_s.p,      tot.p,   act.p,
                        score.p,   gen.p,   alive_d.p};
    }
};

namespace {

int check(gss_res *r) {
    if (!r) return set_err(GSS_E_INVAL, "null scorer");
    return engine_activate(r->e);
}

// Runs forward/backward + gradient (+ optional SGD step) on n samples
// already in r->xs / ys / ws, in the order of idx (null = identity).
int run_batch(gss_res *r, hipStream_t st, const float *x, const float *y, const float *wt,
              const uint32_t *idx, int64_t n, float w_pos, float w_neg, float lr, float *loss) {
    int rc;
    const int nchunks = (int)((n + kGradChunk - 1) / kGradChunk);
    if ((rc = r->a1.reserve((size_t)n * H)) || (rc = r->a2.reserve((size_t)n * H)) ||
        (rc = r->d1.reserve((size_t)n * H)) || (rc = r->d2.reserve((size_t)n * H)) ||
        (rc = r->d3.reserve((size_t)n)) || (rc = r->loss.reserve(1)) ||
        (rc = r->partial.reserve((size_t)nchunks * P)))
        return rc;
    RES_HIP(hipMemsetAsync(r->loss.p, 0, sizeof(float), st));
    const unsigned fb_blocks = (unsigned)std::min<int64_t>((n + 3) / 4, 4096);
    hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(std::max(fb_blocks, 1u)), dim3(256), 0, st, r->params.p, x,
                       y, wt, idx, n, w_pos, w_neg, r->a1.p, r->a2.p, r->d1.p, r->d2.p, r->d3.p,
                       r->loss.p);
    RES_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_grad_partial, dim3((P + 255) / 256, nchunks), dim3(256), 0, st, x,
                       idx, r->a1.p, r->a2.p, r->d1.p, r->d2.p, r->d3.p, n, r->partial.p);
    RES_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_reduce_step, dim3((P + 255) / 256), dim3(256), 0, st, r->partial.p,
                       nchunks, 1.0f / (float)n, r->grad.p, r->params.p, r->mom.p, lr, 0.9f);
    RES_HIP(hipGetLastError());
    if (loss) {
        float l = 0.f;
        RES_HIP(hipMemcpyAsync(&l, r->loss.p, sizeof(float), hipMemcpyDeviceToHost, st));
        RES_HIP(hipStreamSynchronize(st));
        *loss = l / (float)n;
    }
    return GSS_OK;
}

}  // namespace

extern "C" {

int gss_res_create(gss_engine *e, int64_t max_tags, float decay, uint64_t seed, gss_res **out) {
    if (!out || !e || max_tags <= 0 || max_tags > 0xFFFFFFFFll || !(decay > 0.f && decay <= 1.f))
        return set_err(GSS_E_INVAL, "bad scorer arguments");
    *out = nullptr;
    std::vector<const uint32_t *> probe;
    int rc = db_tag_arrays(e, probe);
    if (rc) return rc;
    gss_res *r = new (std::nothrow) gss_res();
    if (!r) return set_err(GSS_E_NOMEM, "scorer allocation failed");
    r->e = e;
    r->cap = max_tags;
    r->decay = decay;
    r->alive.assign((size_t)max_tags, 0);
    const size_t n = (size_t)max_tags;
    if ((rc = r->len.reserve(n)) || (rc = r->hit_rounds.reserve(n)) || (rc = r->cnt.reserve(n)) ||
        (rc = r->cnt_prev.reserve(n)) || (rc = r->gen.reserve(n)) || (rc = r->birth.reserve(n)) ||
        (rc = r->last_hit.reserve(n)) || (rc = r->solv.reserve(n)) ||
        (rc = r->solv_cum.reserve(n)) || (rc = r->ema_f.reserve(n)) ||
        (rc = r->ema_s.reserve(n)) || (rc = r->tot.reserve(n)) || (rc = r->act.reserve(n)) ||
        (rc = r->score.reserve(n)) || (rc = r->alive_d.reserve(n)) || (rc = r->keep.reserve(n)) ||
        (rc = r->evicted.reserve(n)) || (rc = r->params.reserve(P)) || (rc = r->mom.reserve(P)) ||
        (rc = r->grad.reserve(P)) || (rc = r->hist.reserve(kBuckets)) ||
        (rc = r->maxbits.reserve(1)) || (rc = r->counters.reserve(2))) {
        gss_res_destroy(r);
        return rc;
    }
    hipStream_t st = engine_stream(e);
    // fill kernels would do as well; these buffers are never in a graph
    if (hipMemsetAsync(r->alive_d.p, 0, n, st) != hipSuccess ||
        hipMemsetAsync(r->gen.p, 0, n * 4, st) != hipSuccess ||
        hipMemsetAsync(r->cnt.p, 0, n * 4, st) != hipSuccess ||
        hipMemsetAsync(r->solv.p, 0, n * 8, st) != hipSuccess ||
        hipMemsetAsync(r->mom.p, 0, P * 4, st) != hipSuccess) {
        gss_res_destroy(r);
        return set_err(GSS_E_HIP, "scorer initialisation failed");
    }
    // He-uniform weights, zero biases
    std::vector<float> w(P, 0.f);
    uint64_t s = seed;
    const double l1 = std::sqrt(6.0 / F), l2 = std::sqrt(6.0 / H);
    for (int i = 0; i < H * F; i++) w[OW1 + i] = (float)(l1 * uniform_pm1(s));
    for (int i = 0; i < H * H; i++) w[OW2 + i] = (float)(l2 * uniform_pm1(s));
    for (int i = 0; i < H; i++) w[OW3 + i] = (float)(0.1 * l2 * uniform_pm1(s));
    if (hipMemcpyAsync(r->params.p, w.data(), P * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        gss_res_destroy(r);
        return set_err(GSS_E_HIP, "scorer initialisation failed");
    }
    *out = r;
    return GSS_OK;
}

int gss_res_destroy(gss_res *r) {
    if (!r) return GSS_OK;
    if (r->e) (void)hipStreamSynchronize(engine_stream(r->e));
    for (auto *b : {&r->len, &r->hit_rounds, &r->cnt, &r->cnt_prev, &r->gen, &r->tags_in,
                    &r->hist, &r->maxbits, &r->idx})
        b->release();
    r->birth.release();
    r->last_hit.release();
    r->lens_in.release();
    r->solv.release();
    r->solv_cum.release();
    for (auto *b : {&r->ema_f, &r->ema_s, &r->tot, &r->act, &r->score, &r->params, &r->mom,
                    &r->grad, &r->partial, &r->a1, &r->a2, &r->d1, &r->d2, &r->d3, &r->loss,
                    &r->xs, &r->ys, &r->ws, &r->act_in, &r->fout})
        b->release();
    r->alive_d.release();
    r->keep.release();
    r->evicted.release();
    r->counters.release();
    r->tagptr.release();
    for (auto &kv : r->snaps) {
        kv.second.x.release();
        kv.second.tags.release();
        kv.second.gens.release();
    }
    delete r;
    return GSS_OK;
}

int gss_res_add(gss_res *r, int64_t n, const uint32_t *lits, const int32_t *lens,
                const uint32_t *tags, int32_t round) {
    int rc = check(r);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!lits || !lens || !tags))) return set_err(GSS_E_INVAL, "bad batch");
    if (n == 0) return GSS_OK;
    for (int64_t i = 0; i < n; i++) {
        if (tags[i] >= (uint64_t)r->cap)
            return set_err(GSS_E_INVAL, "tag %u outside [0, %lld)", tags[i], (long long)r->cap);
        if (r->alive[tags[i]]) return set_err(GSS_E_INVAL, "tag %u is in use", tags[i]);
    }
    // a tag twice in one batch
    {
        std::vector<uint32_t> sorted(tags, tags + n);
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            return set_err(GSS_E_INVAL, "a tag appears twice in the batch");
    }
    // every allocation first: a batch the device cannot take changes nothing
    if ((rc = r->tags_in.reserve(n)) || (rc = r->lens_in.reserve(n)) ||
        (rc = r->act_in.reserve(n)))
        return rc;
    if ((rc = gss_db_append_tagged(r->e, n, lits, lens, tags, nullptr))) return rc;
    // initial activities: the increment after each clause's own decay step
    std::vector<float> act0((size_t)n);
    float inc = r->incr;
    bool rescale = false;
    for (int64_t i = 0; i < n; i++) {
        inc /= r->decay;
        act0[i] = inc;
        if (inc > kRescale) rescale = true;
    }
    hipStream_t st = engine_stream(r->e);
    RES_HIP(hipMemcpyAsync(r->tags_in.p, tags, n * 4, hipMemcpyHostToDevice, st));
    RES_HIP(hipMemcpyAsync(r->lens_in.p, lens, n * 4, hipMemcpyHostToDevice, st));
    RES_HIP(hipMemcpyAsync(r->act_in.p, act0.data(), n * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_res_init, dim3(grid(n, 1u << 30)), dim3(256), 0, st, r->state(),
                       r->tags_in.p, r->lens_in.p, r->act_in.p, n, round);
    RES_HIP(hipGetLastError());
    r->incr = inc;
    for (int64_t i = 0; i < n; i++) r->alive[tags[i]] = 1;
    if (rescale) {
        hipLaunchKernelGGL(k_res_rescale, dim3(grid(r->cap)), dim3(256), 0, st, r->state(),
                           r->cap);
        RES_HIP(hipGetLastError());
        r->incr /= kRescale;
    }
    RES_HIP(hipStreamSynchronize(st));
    return GSS_OK;
}

int gss_res_account(gss_res *r, int32_t round, int64_t *hit_pairs) {
    int rc = check(r);
    if (rc) return rc;
    const gss_hit *hits = nullptr;
    const uint32_t *fills = nullptr;
    uint32_t chunks = 0;
    if ((rc = sweep_records(r->e, &hits, &fills, &chunks))) return rc;
    std::vector<const uint32_t *> tp;
    if ((rc = db_tag_arrays(r->e, tp))) return rc;
    hipStream_t st = engine_stream(r->e);
    if ((rc = r->tagptr.reserve(std::max<size_t>(tp.size(), 1)))) return rc;
    RES_HIP(hipMemcpyAsync(r->tagptr.p, tp.data(), tp.size() * sizeof(void *),
                           hipMemcpyHostToDevice, st));
    RES_HIP(hipMemsetAsync(r->counters.p, 0, 2 * sizeof(unsigned long long), st));
    RES_HIP(hipMemsetAsync(r->maxbits.p, 0, 4, st));
    if (chunks) {
        const uint64_t parts = (uint64_t)chunks * kRecParts;
        hipLaunchKernelGGL(k_res_records, dim3((unsigned)std::min<uint64_t>(parts, 1u << 20)),
                           dim3(kRecBlock), 0,
                           st, hits, fills, chunks, r->tagptr.p, (int)tp.size(), r->state(),
                           r->counters.p);
        RES_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_res_update, dim3(grid(r->cap)), dim3(256), 0, st, r->state(), r->cap,
                       round, r->incr, r->maxbits.p);
    RES_HIP(hipGetLastError());
    unsigned long long pairs = 0;
    unsigned int mb = 0;
    RES_HIP(hipMemcpyAsync(&pairs, r->counters.p, 8, hipMemcpyDeviceToHost, st));
    RES_HIP(hipMemcpyAsync(&mb, r->maxbits.p, 4, hipMemcpyDeviceToHost, st));
    RES_HIP(hipStreamSynchronize(st));
    float mx;
    std::memcpy(&mx, &mb, 4);
    if (mx > kRescale) {
        hipLaunchKernelGGL(k_res_rescale, dim3(grid(r->cap)), dim3(256), 0, st, r->state(),
                           r->cap);
        RES_HIP(hipGetLastError());
        r->incr /= kRescale;
        RES_HIP(hipStreamSynchronize(st));
    }
    if (hit_pairs) *hit_pairs = (int64_t)pairs;
    return GSS_OK;
}

int gss_res_snapshot(gss_res *r, int32_t round, int64_t n, const uint32_t *tags) {
    int rc = check(r);
    if (rc) return rc;
    if (n <= 0 || !tags) return set_err(GSS_E_INVAL, "bad snapshot");
    for (int64_t i = 0; i < n; i++)
        if (tags[i] >= (uint64_t)r->cap || !r->alive[tags[i]])
            return set_err(GSS_E_INVAL, "snapshot tag %u is not alive", tags[i]);
    Snapshot &sn = r->snaps[round];
    if ((rc = sn.x.reserve((size_t)n * F)) || (rc = sn.tags.reserve(n)) ||
        (rc = sn.gens.reserve(n)))
        return rc;
    sn.n = n;
    hipStream_t st = engine_stream(r->e);
    RES_HIP(hipMemcpyAsync(sn.tags.p, tags, n * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_res_features, dim3(grid(n, 1u << 30)), dim3(256), 0, st, r->state(),
                       sn.tags.p, n, round, r->incr, sn.x.p, sn.gens.p);
    RES_HIP(hipGetLastError());
    RES_HIP(hipStreamSynchronize(st));
    // keep only recent snapshots
    while (r->snaps.size() > 16) {
        auto it = r->snaps.begin();
        it->second.x.release();
        it->second.tags.release();
        it->second.gens.release();
        r->snaps.erase(it);
    }
    return GSS_OK;
}

int gss_res_train(gss_res *r, int32_t round, int horizon, int epochs, int batch, float lr,
                  float *loss, int64_t *samples) {
    int rc = check(r);
    if (rc) return rc;
    if (horizon < 1 || epochs < 1 || batch < 1) return set_err(GSS_E_INVAL, "bad training args");
    if (samples) *samples = 0;
    if (loss) *loss = 0.f;
    auto it = r->snaps.find(round - horizon);
    if (it == r->snaps.end()) return GSS_OK;
    Snapshot &sn = it->second;
    const int64_t n = sn.n;
    hipStream_t st = engine_stream(r->e);
    if ((rc = r->ys.reserve(n)) || (rc = r->ws.reserve(n)) || (rc = r->idx.reserve(n)))
        return rc;
    RES_HIP(hipMemsetAsync(r->counters.p, 0, 2 * sizeof(unsigned long long), st));
    hipLaunchKernelGGL(k_res_labels, dim3(grid(n, 1u << 30)), dim3(256), 0, st, r->state(),
                       sn.tags.p, sn.gens.p, n, round - horizon, r->ys.p, r->ws.p, r->counters.p);
    RES_HIP(hipGetLastError());
    unsigned long long cnt[2] = {0, 0};
    RES_HIP(hipMemcpyAsync(cnt, r->counters.p, 16, hipMemcpyDeviceToHost, st));
    RES_HIP(hipStreamSynchronize(st));
    const double valid = (double)cnt[0], pos = (double)cnt[1];
    if (valid < 1 || pos < 1 || pos >= valid) return GSS_OK;  // nothing to learn from
    // balanced classes: each class carries half of the weight
    const float w_pos = (float)(0.5 * valid / pos), w_neg = (float)(0.5 * valid / (valid - pos));
    std::vector<uint32_t> order((size_t)n);
    uint64_t s = 0x5EEDull ^ (uint64_t)round;
    float last = 0.f;
    for (int ep = 0; ep < epochs; ep++) {
        for (int64_t i = 0; i < n; i++) order[i] = (uint32_t)i;
        for (int64_t i = n - 1; i > 0; i--) {  // Fisher-Yates
            const int64_t j = (int64_t)((uniform_pm1(s) * 0.5 + 0.5) * (double)(i + 1)) % (i + 1);
            std::swap(order[i], order[j]);
        }
        RES_HIP(hipMemcpyAsync(r->idx.p, order.data(), n * 4, hipMemcpyHostToDevice, st));
        double acc = 0.0;
        int nb = 0;
        for (int64_t b0 = 0; b0 < n; b0 += batch) {
            const int64_t m = std::min<int64_t>(batch, n - b0);
            float l = 0.f;
            if ((rc = run_batch(r, st, sn.x.p, r->ys.p, r->ws.p, r->idx.p + b0, m, w_pos, w_neg,
                                lr, ep == epochs - 1 ? &l : nullptr)))
                return rc;
            acc += l;
            nb++;
        }
        if (ep == epochs - 1) last = (float)(acc / std::max(nb, 1));
    }
    RES_HIP(hipStreamSynchronize(st));
    if (loss) *loss = last;
    if (samples) *samples = (int64_t)valid;
    return GSS_OK;
}

int gss_res_reduce(gss_res *r, int policy, int32_t round, uint8_t *evicted, int64_t *kept) {
    int rc = check(r);
    if (rc) return rc;
    if (policy != GSS_RES_POLICY_ACTIVITY && policy != GSS_RES_POLICY_LEARNED)
        return set_err(GSS_E_INVAL, "unknown policy %d", policy);
    hipStream_t st = engine_stream(r->e);
    RES_HIP(hipMemsetAsync(r->hist.p, 0, kBuckets * 4, st));
    if (policy == GSS_RES_POLICY_LEARNED)
        hipLaunchKernelGGL(k_res_score_learned, dim3(grid(r->cap, 1u << 30)), dim3(256), 0, st,
                           r->state(), r->cap, round, r->incr, r->params.p, r->hist.p);
    else
        hipLaunchKernelGGL(k_res_score_activity, dim3(grid(r->cap)), dim3(256), 0, st,
                           r->state(), r->cap, r->hist.p);
    RES_HIP(hipGetLastError());
    std::vector<uint32_t> hist(kBuckets);
    RES_HIP(hipMemcpyAsync(hist.data(), r->hist.p, kBuckets * 4, hipMemcpyDeviceToHost, st));
    RES_HIP(hipStreamSynchronize(st));
    int64_t total = 0;
    for (uint32_t h : hist) total += h;
    // approximate median: the upper edge of the bucket holding the n-th
    // smallest score, n = half of the live clauses
    const int64_t nth = total / 2;
    float thr = 0.f;
    if (nth > 0) {
        int64_t seen = 0;
        int b = 0;
        for (; b < kBuckets; b++) {
            seen += hist[b];
            if (seen >= nth) break;
        }
        if (policy == GSS_RES_POLICY_LEARNED) {
            thr = (float)(b + 1) / (float)kBuckets;
        } else {
            const float lowest = std::log(FLT_MIN), largest = std::log(FLT_MAX);
            const float step = (largest - lowest) / kBuckets;
            thr = std::exp(lowest + (float)(b + 1) * step);
        }
    }
    hipLaunchKernelGGL(k_res_keep, dim3(grid(r->cap)), dim3(256), 0, st, r->state(), r->cap, thr,
                       r->keep.p, r->evicted.p);
    RES_HIP(hipGetLastError());
    if ((rc = db_compact_device(r->e, r->keep.p, r->cap))) return rc;
    std::vector<uint8_t> ev((size_t)r->cap);
    RES_HIP(hipMemcpyAsync(ev.data(), r->evicted.p, (size_t)r->cap, hipMemcpyDeviceToHost, st));
    RES_HIP(hipStreamSynchronize(st));
    for (int64_t t = 0; t < r->cap; t++)
        if (ev[t]) r->alive[t] = 0;
    if (evicted) std::memcpy(evicted, ev.data(), (size_t)r->cap);
    if (kept) {
        int64_t c = 0, l = 0;
        if ((rc = gss_db_totals(r->e, &c, &l))) return rc;
        *kept = c;
    }
    return GSS_OK;
}

int gss_res_features(gss_res *r, int32_t round, int64_t n, const uint32_t *tags, float *out) {
    int rc = check(r);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!tags || !out))) return set_err(GSS_E_INVAL, "bad arguments");
    if (!n) return GSS_OK;
    for (int64_t i = 0; i < n; i++)
        if (tags[i] >= (uint64_t)r->cap) return set_err(GSS_E_INVAL, "tag out of range");
    hipStream_t st = engine_stream(r->e);
    if ((rc = r->tags_in.reserve(n)) || (rc = r->fout.reserve((size_t)n * F))) return rc;
    RES_HIP(hipMemcpyAsync(r->tags_in.p, tags, n * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_res_features, dim3(grid(n, 1u << 30)), dim3(256), 0, st, r->state(),
                       r->tags_in.p, n, round, r->incr, r->fout.p, (uint32_t *)nullptr);
    RES_HIP(hipGetLastError());
    RES_HIP(hipMemcpyAsync(out, r->fout.p, (size_t)n * F * 4, hipMemcpyDeviceToHost, st));
    RES_HIP(hipStreamSynchronize(st));
    return GSS_OK;
}

int gss_res_activity(gss_res *r, int64_t n, const uint32_t *tags, float *out) {
    int rc = check(r);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!tags || !out))) return set_err(GSS_E_INVAL, "bad arguments");
    if (!n) return GSS_OK;
    for (int64_t i = 0; i < n; i++)
        if (tags[i] >= (uint64_t)r->cap) return set_err(GSS_E_INVAL, "tag out of range");
    hipStream_t st = engine_stream(r->e);
    if ((rc = r->tags_in.reserve(n)) || (rc = r->fout.reserve((size_t)n))) return rc;
    RES_HIP(hipMemcpyAsync(r->tags_in.p, tags, n * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_res_gather_act, dim3(grid(n, 1u << 30)), dim3(256), 0, st, r->state(),
                       r->tags_in.p, n, r->fout.p);
    RES_HIP(hipGetLastError());
    RES_HIP(hipMemcpyAsync(out, r->fout.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    RES_HIP(hipStreamSynchronize(st));
    return GSS_OK;
}

int gss_res_get_params(gss_res *r, float *out) {
    int rc = check(r);
    if (rc) return rc;
    if (!out) return set_err(GSS_E_INVAL, "null output");
    hipStream_t st = engine_stream(r->e);
    RES_HIP(hipMemcpyAsync(out, r->params.p, P * 4, hipMemcpyDeviceToHost, st));
    RES_HIP(hipStreamSynchronize(st));
    return GSS_OK;
}

int gss_res_set_params(gss_res *r, const float *in) {
    int rc = check(r);
    if (rc) return rc;
    if (!in) return set_err(GSS_E_INVAL, "null input");
    hipStream_t st = engine_stream(r->e);
    RES_HIP(hipMemcpyAsync(r->params.p, in, P * 4, hipMemcpyHostToDevice, st));
    RES_HIP(hipMemsetAsync(r->mom.p, 0, P * 4, st));
    RES_HIP(hipStreamSynchronize(st));
    return GSS_OK;
}

int gss_res_grad(gss_res *r, int64_t n, const float *x, const float *y, const float *wt,
                 float *loss, float *grad) {
    int rc = check(r);
    if (rc) return rc;
    if (n <= 0 || !x || !y || !grad) return set_err(GSS_E_INVAL, "bad arguments");
    hipStream_t st = engine_stream(r->e);
    if ((rc = r->xs.reserve((size_t)n * F)) || (rc = r->ys.reserve(n)) || (rc = r->ws.reserve(n)))
        return rc;
    RES_HIP(hipMemcpyAsync(r->xs.p, x, (size_t)n * F * 4, hipMemcpyHostToDevice, st));
    RES_HIP(hipMemcpyAsync(r->ys.p, y, (size_t)n * 4, hipMemcpyHostToDevice, st));
    if (wt) RES_HIP(hipMemcpyAsync(r->ws.p, wt, (size_t)n * 4, hipMemcpyHostToDevice, st));
    float l = 0.f;
    if ((rc = run_batch(r, st, r->xs.p, r->ys.p, wt ? r->ws.p : nullptr, nullptr, n, 1.f, 1.f,
                        0.f, &l)))
        return rc;
    RES_HIP(hipMemcpyAsync(grad, r->grad.p, P * 4, hipMemcpyDeviceToHost, st));
    RES_HIP(hipStreamSynchronize(st));
    if (loss) *loss = l;
    return GSS_OK;
}

int gss_res_score(gss_res *r, int64_t n, const float *x, float *out) {
    int rc = check(r);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!x || !out))) return set_err(GSS_E_INVAL, "bad arguments");
    if (!n) return GSS_OK;
    hipStream_t st = engine_stream(r->e);
    if ((rc = r->xs.reserve((size_t)n * F)) || (rc = r->fout.reserve(n))) return rc;
    RES_HIP(hipMemcpyAsync(r->xs.p, x, (size_t)n * F * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_mlp_score, dim3(grid(n, 1u << 30)), dim3(256), 0, st, r->params.p,
                       r->xs.p, n, r->fout.p);
    RES_HIP(hipGetLastError());
    RES_HIP(hipMemcpyAsync(out, r->fout.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    RES_HIP(hipStreamSynchronize(st));
    return GSS_OK;
}

}  // extern "C"
