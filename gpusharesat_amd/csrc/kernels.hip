// CDNA4 (gfx950, wave64) kernels of the clause-vs-assignment check.
//
// The usefulness rule of the reference (gpuShareLib/GpuRunner.cu:34-57,
// ReportComputer) reports slot s of a clause when every literal is false
// ("falsified") or when exactly one literal occurrence is undefined and every
// other one is false ("unit").  The sweep evaluates it in an equivalent form
// over table rows, where row l holds nf_l = the slots in which literal l is
// NOT false (true or undefined); literal l is true exactly where its negation
// l^1 is false, i.e. outside nf_{l^1}:
//
//   count   c1/c2 = "at least one / at least two non-false occurrences",
//           saturating:  c2 |= c1 & nf_l;  c1 |= nf_l
//   true    nt = AND over the clause of nf_{l^1} (no literal is true)
//   report  falsified = ~c1;  unit = c1 & ~c2 & nt
//
// A slot with c2 set, or outside nt, can never report, so a clause stops once
// every valid slot of every word of its group is dead.  Two schedules give the
// same answer:
//   two-pass  pass 1 reads only rows nf_l (1 bit per slot per literal) and
//             stops on c2; pass 2 reads rows nf_{l^1} only if some slot has
//             exactly one non-false occurrence, and stops once those slots are
//             resolved;
//   fused     one pass reads both rows of each literal (adjacent in memory)
//             and stops on c2 | ~nt.
// A clause of at most `fuse_len` literals uses the fused schedule, longer
// ones the two-pass one.  Duplicate literals are separate occurrences, as in
// the reference (GpuRunner.cu:68-89 walks the literal list), so "x, x" with x
// undefined is not unit.  tests/test_oracle.py and the GPU parity tests check
// the forms agree.
//
// Clause literals are stored on the device as row offsets: literal l as
// l << 7, the byte offset of its row in a table of 16 words (scaled by the
// actual width at compile time); negation is "xor 128".  Padding lanes past
// the end of a clause hold the offset of a sentinel row pair (nf = 0, and
// all ones for its negation) that contributes nothing, so they need no mask.
//
// Work mapping.  A group of 16 lanes owns one clause.  Up to 16 words, lane
// (s, w) of the group handles table word w and literal sub-lane s (G = W*S =
// 16); at 32 and 64 words each lane handles K = W/16 consecutive words, read
// with 16-byte loads, so every width checks 4 clauses per wave pass and keeps
// 4*U row reads per wave in flight.  The group loads G literals of the
// clause at once (lane q holds literal q of the chunk; the first chunk of the
// next clause is prefetched while the current one is checked) and hands them
// out with DPP row broadcasts (G = 16: a group is one 16-lane DPP row) or
// shuffles; each step a lane issues U row reads (U independent loads in
// flight), the S sub-lanes' counts are merged with a butterfly, then the group
// votes on early exit.  With W = 16 (1024 slots) a row read by the 16 lanes of
// a group is one coalesced 128-byte line.  A wave holds 64/G groups; a task is
// a run of clauses of one length bucket (its parameters are wave-uniform and
// live in scalar registers); waves are persistent and take tasks round-robin.
//
// Hits.  A clause/word with a report becomes one 24-byte record (gss_hit:
// clause, word and both masks), or, for delivery to the solvers, which never
// read the masks, one 8-byte record (gss_hit_id: the first word of the
// gss_hit).  Each wave compacts its records with a ballot + mbcnt prefix and
// appends them to a private chunk of kHitChunk records claimed with one
// atomic (non-temporal stores: records are never re-read by the sweep); a
// chunk is abandoned (fill recorded) when the next pass does not fit.  Capacity is
// sized on the host so that nothing is dropped (the reference drops on
// overflow, gpuShareLib/Reporter.cuh:46-48); if it ever is exceeded the claim
// counter still counts and the host re-runs the sweep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <hipcub/hipcub.hpp>
#include <utility>

#include "kernels.h"

namespace gss {

// k_sweep writes a record as three 8-byte words: {row, len, word}, false_mask,
// unit_mask.
static_assert(sizeof(gss_hit) == 24 && offsetof(gss_hit, len) == 4 &&
                  offsetof(gss_hit, word) == 6 && offsetof(gss_hit, false_mask) == 8 &&
                  offsetof(gss_hit, unit_mask) == 16,
              "gss_hit layout");
static_assert(sizeof(gss_hit_id) == 8 && offsetof(gss_hit_id, len) == 4 &&
                  offsetof(gss_hit_id, word) == 6,
              "gss_hit_id is the first word of a gss_hit");
static_assert(sizeof(gss_hit_clause) == 16 && offsetof(gss_hit_clause, len) == 4 &&
                  offsetof(gss_hit_clause, words) == 8,
              "gss_hit_clause: {row, len, 0} then the word mask");

namespace {

__device__ __forceinline__ uint32_t prefix_count(unsigned long long bal) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
}

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

template <int G>
__device__ __forceinline__ bool group_any(bool pred, uint64_t gmask) {
    const unsigned long long b = __ballot(pred);
    if constexpr (G == 64) {
        return b != 0;
    } else {
        return (b & gmask) != 0;
    }
}

// Byte offset inside the table of row word w, from a stored row offset
// (literal << 7, i.e. the offset for a 16-word table).
template <int W>
__device__ __forceinline__ uint32_t scale_offset(uint32_t stored) {
    if constexpr (W == 16) return stored;
    else if constexpr (W > 16) return stored << (W == 32 ? 1 : 2);
    else return stored >> (W == 8 ? 1 : W == 4 ? 2 : W == 2 ? 3 : 4);
}

// Address of row word w of the literal at chunk position J, for every lane of
// the group.  With G = 16 a group is one DPP row, so the literal is a row
// broadcast (no LDS), fused with the add of the lane's word offset.
template <int W, int G, int J>
__device__ __forceinline__ uint32_t chunk_addr(uint32_t litreg, int gbase, uint32_t w8) {
    if constexpr (G == 16 && W == 16) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)litreg, 0x150 + J, 0xF, 0xF, false) | w8;
    } else if constexpr (G == 16) {
        return scale_offset<W>(
                   (uint32_t)__builtin_amdgcn_mov_dpp((int)litreg, 0x150 + J, 0xF, 0xF, false)) |
               w8;
    } else {
        return scale_offset<W>((uint32_t)__shfl((int)litreg, gbase + J)) | w8;
    }
}

typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

// K consecutive table words of one row (K = 1, 2 or 4), read with 8- or
// 16-byte loads.
template <int K>
struct Words {
    uint64_t v[K];
};

template <int K>
__device__ __forceinline__ Words<K> load_words(const uint64_t *__restrict__ table, uint32_t off) {
    Words<K> r;
    const char *p = (const char *)table + off;
    if constexpr (K == 1) {
        r.v[0] = *(const uint64_t *)p;
    } else {
#pragma unroll
        for (int i = 0; i < K; i += 2) {
            const u64x2 x = *(const u64x2 *)(p + 8 * i);
            r.v[i] = x.x;
            r.v[i + 1] = x.y;
        }
    }
    return r;
}

// Clause state of one lane for its K words: saturating count of non-false
// occurrences per slot (c1: >= 1, c2: >= 2) and the slots where no literal is
// true (nt).
template <int K>
struct ClauseState {
    uint64_t c1[K], c2[K], nt[K];
};

// Offset of the negated literal's row: stored offsets are literal << 7, so the
// negation flips bit 7, scaled like the offset.
template <int W>
__device__ __forceinline__ uint32_t neg_bit() {
    return scale_offset<W>(128u);
}

// One step: U literals per lane.  The chunk holds the group's literals
// [base, base + G), sentinel past the clause end; this step covers chunk
// positions [ST, ST + S*U).  The U row reads are issued back to back, folded
// into the state, and for S > 1 the S sub-lanes are merged with a butterfly.
template <int W, int S, int U, int K, int ST, bool ROWS_F, bool ROWS_T>
__device__ __forceinline__ void step_rows(uint32_t litreg, int gbase, int sub, uint32_t w8,
                                          const uint64_t *__restrict__ table,
                                          ClauseState<K> &cs, bool live) {
    constexpr int LW = W / K;  // lanes across one row
    constexpr int G = LW * S;
    Words<K> f[U], t[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        uint32_t off;
        if constexpr (S == 1 && U <= 4) {
            off = u == 0 ? chunk_addr<W, G, ST + 0>(litreg, gbase, w8)
                : u == 1 ? chunk_addr<W, G, (ST + 1) % G>(litreg, gbase, w8)
                : u == 2 ? chunk_addr<W, G, (ST + 2) % G>(litreg, gbase, w8)
                         : chunk_addr<W, G, (ST + 3) % G>(litreg, gbase, w8);
        } else if constexpr (S == 1) {
            off = scale_offset<W>((uint32_t)__shfl((int)litreg, gbase + ST + u)) | w8;
        } else {
            off = scale_offset<W>((uint32_t)__shfl((int)litreg, gbase + ST + sub * U + u)) | w8;
        }
        // A lane whose slots are all settled skips its reads (wide rows:
        // lines of settled lanes are not fetched; at one word per lane a
        // partial line is fetched whole, so 16-word rows do not skip);
        // all-ones rows leave its state's report masks unchanged.
        if (live) {
            if (ROWS_F) f[u] = load_words<K>(table, off);
            if (ROWS_T) t[u] = load_words<K>(table, off ^ neg_bit<W>());
        } else {
#pragma unroll
            for (int k = 0; k < K; k++) {
                if (ROWS_F) f[u].v[k] = ~0ull;
                if (ROWS_T) t[u].v[k] = ~0ull;
            }
        }
    }
    if constexpr (S == 1) {
#pragma unroll
        for (int u = 0; u < U; u++) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                if (ROWS_F) {
                    cs.c2[k] |= cs.c1[k] & f[u].v[k];
                    cs.c1[k] |= f[u].v[k];
                }
                if (ROWS_T) cs.nt[k] &= t[u].v[k];
            }
        }
    } else {
        static_assert(K == 1, "sub-lanes split narrow rows only");
        uint64_t m1 = 0, m2 = 0, tt = ~0ull;
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (ROWS_F) {
                m2 |= m1 & f[u].v[0];
                m1 |= f[u].v[0];
            }
            if (ROWS_T) tt &= t[u].v[0];
        }
#pragma unroll
        for (int d = W; d < G; d <<= 1) {
            if (ROWS_F) {
                const uint64_t p1 = __shfl_xor(m1, d);
                const uint64_t p2 = __shfl_xor(m2, d);
                m2 = m2 | p2 | (m1 & p1);
                m1 |= p1;
            }
            if (ROWS_T) tt &= __shfl_xor(tt, d);
        }
        if (ROWS_F) {
            cs.c2[0] |= m2 | (cs.c1[0] & m1);
            cs.c1[0] |= m1;
        }
        if (ROWS_T) cs.nt[0] &= tt;
    }
}

// Runs the steps of one preloaded chunk while they start before the clause
// end and `more(cs)` holds.  Returns false if it stopped early.
template <int W, int S, int U, int K, bool ROWS_F, bool ROWS_T, typename More, typename Live>
__device__ __forceinline__ bool chunk_steps(uint32_t litreg, int gbase, int sub, uint32_t w8,
                                            int base, int len,
                                            const uint64_t *__restrict__ table,
                                            ClauseState<K> &cs, More &&more, Live &&live) {
    constexpr int G = (W / K) * S;
    constexpr int LPS = S * U;
    bool go = true;
    [&]<int... I>(std::integer_sequence<int, I...>) {
        ((go = go && (base + I * LPS < len) &&
               [&] {
                   step_rows<W, S, U, K, I * LPS, ROWS_F, ROWS_T>(
                       litreg, gbase, sub, w8, table, cs, K > 1 ? live(cs) : true);
                   return more(cs);
               }()),
         ...);
    }(std::make_integer_sequence<int, G / LPS>{});
    return go;
}

// Evaluates one clause for this lane's K words.  All G lanes of the group call
// it together (group-uniform control flow).  first_chunk holds literal q of
// the clause in lane q of the group (the sentinel past the end), already
// loaded by the caller.
template <int W, int S, int U, int K, bool FUSE>
__device__ __forceinline__ void check_clause(const uint32_t *__restrict__ cl, int len,
                                             uint32_t first_chunk, uint32_t sentinel,
                                             const uint64_t *__restrict__ table, int q,
                                             uint64_t gmask, int gbase, uint32_t w8, int sub,
                                             const uint64_t (&valid)[K], int fuse_len,
                                             uint64_t (&fm)[K], uint64_t (&um)[K]) {
    constexpr int G = (W / K) * S;
    // invalid slots start dead
    ClauseState<K> cs;
#pragma unroll
    for (int k = 0; k < K; k++) {
        cs.c1[k] = ~valid[k];
        cs.c2[k] = ~valid[k];
        cs.nt[k] = ~0ull;
    }
    // FUSE = false compiles the two-pass schedule only (its registers are
    // then not sized for the fused path's two rows per literal)
    const bool fused = FUSE && len <= fuse_len;
    auto lane_alive = [&](const ClauseState<K> &c) {
        uint64_t x = 0;
#pragma unroll
        for (int k = 0; k < K; k++) x |= c.nt[k] & ~c.c2[k];
        return x != 0;
    };
    auto alive = [&](const ClauseState<K> &c) { return group_any<G>(lane_alive(c), gmask); };
    uint32_t litreg = first_chunk;
    for (int base = 0; base < len; base += G) {
        if (base) litreg = base + q < len ? cl[base + q] : sentinel;
        const bool go =
            fused ? chunk_steps<W, S, U, K, true, true>(litreg, gbase, sub, w8, base, len, table,
                                                       cs, alive, lane_alive)
                  : chunk_steps<W, S, U, K, true, false>(litreg, gbase, sub, w8, base, len, table,
                                                        cs, alive, lane_alive);
        if (!go) break;  // dead, or the clause ended inside this chunk
    }
    uint64_t cand[K], any = 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        fm[k] = ~cs.c1[k];
        cand[k] = cs.c1[k] & ~cs.c2[k] & cs.nt[k];
        any |= cand[k];
    }
    if (!fused && group_any<G>(any != 0, gmask)) {
        auto lane_open = [&](const ClauseState<K> &c) {
            uint64_t x = 0;
#pragma unroll
            for (int k = 0; k < K; k++) x |= cand[k] & c.nt[k];
            return x != 0;
        };
        auto open = [&](const ClauseState<K> &c) { return group_any<G>(lane_open(c), gmask); };
        litreg = first_chunk;
        for (int base = 0; base < len; base += G) {
            if (base) litreg = base + q < len ? cl[base + q] : sentinel;
            if (!chunk_steps<W, S, U, K, false, true>(litreg, gbase, sub, w8, base, len, table,
                                                      cs, open, lane_open))
                break;
        }
    }
#pragma unroll
    for (int k = 0; k < K; k++) um[k] = cand[k] & cs.nt[k];
}

// Serving steps, one clause per lane: the request groups in which fewer
// than two literal occurrences of the clause (its row of len stored
// offsets) are non-false in every valid slot (SweepArgs::agg), i.e. the only
// groups where it can report; 0 once none is left.  `vote` literals and then
// their entries are read together per vote (1, 2, 4 or 8), so a wave keeps
// up to 64 * vote loads in flight.
__device__ __forceinline__ uint32_t lane_agg(const uint32_t *__restrict__ row, int len,
                                             const uint8_t *__restrict__ agg, int vote,
                                             uint32_t gdead) {
    constexpr int kMaxVote = 8;
    uint32_t a1 = 0, a2 = gdead;  // groups with >= 1 / >= 2 such occurrences (empty: dead)
    for (int j = 0; j < len; j += vote) {
        uint32_t l[kMaxVote], x[kMaxVote];
#pragma unroll
        for (int u = 0; u < kMaxVote; u++) l[u] = (u < vote && j + u < len) ? row[j + u] : 0u;
#pragma unroll
        for (int u = 0; u < kMaxVote; u++) x[u] = (u < vote && j + u < len) ? agg[l[u] >> 7] : 0u;
#pragma unroll
        for (int u = 0; u < kMaxVote; u++) {
            a2 |= a1 & x[u];
            a1 |= x[u];
        }
        if ((a2 & 0xFFu) == 0xFFu) return 0;
    }
    return ~a2 & 0xFFu;
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

// MODE: 0 = every clause checked, records per (clause, word) (full or id
// form, a.rec_words); 1 = clauses screened against a.agg first (serving and
// screened steps), same record forms; 2 = every clause checked, one record
// per clause with its words (a.rec_words = 2); 3 = screened, one record per
// clause (screened steps).
template <int W, int S, int U, bool FUSE, int MODE>
__global__ __launch_bounds__(256) void k_sweep(SweepArgs a) {
    constexpr bool AGG = MODE == 1 || MODE == 3;
    constexpr bool CLAUSE_RECORDS = MODE == 2 || MODE == 3;
    constexpr int K = W > 16 ? W / 16 : 1;  // table words per lane
    constexpr int LW = W / K;               // lanes across one row
    constexpr int G = LW * S;
    constexpr int CPW = 64 / G;  // clauses per wave pass
    const int lane = (int)__lane_id();
    const int g = lane / G;
    const int q = lane % G;
    const int sub = q / LW;
    const int lw = lane % LW;
    const int gbase = g * G;
    const uint64_t gmask = G == 64 ? ~0ull : (((1ull << G) - 1ull) << gbase);
    const uint32_t w8 = (uint32_t)(lw * K) * 8u;
    uint64_t valid[K];
#pragma unroll
    for (int k = 0; k < K; k++) valid[k] = a.valid[lw * K + k];
    uint32_t gdead = 0xFFu;  // serving steps: request groups without a word
    if constexpr (AGG)
        for (int b = 0; b < 8; b++)
            if (a.agg_words[b]) gdead &= ~(1u << b);
    // wave id and count are wave-uniform: keep the task walk scalar
    const int gw = uniform((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    const int nw = (int)((gridDim.x * blockDim.x) >> 6);

    int nbuckets = a.nbuckets, ntasks = a.ntasks;
    if (a.dyn) {
        nbuckets = uniform(a.dyn[0]);
        ntasks = uniform(a.dyn[1]);
    }
    int chunk = -1;
    int fill = 0;
    // Appends this pass's records (all lanes call it; `active` marks the
    // groups that checked a clause).
    auto claim = [&](unsigned long long bal) {
        const int n = __popcll(bal);
        if (chunk < 0 || fill + n > kHitChunk) {
            if (chunk >= 0 && lane == 0 && (uint32_t)chunk < a.chunk_cap)
                a.fills[chunk] = (uint32_t)fill;
            uint32_t c = 0;
            if (lane == 0) c = atomicAdd(a.alloc, 1u);
            chunk = uniform((int)__shfl(c, 0));
            fill = 0;
        }
    };
    auto emit = [&](bool active, int64_t row, int len, const uint64_t (&fm)[K],
                    const uint64_t (&um)[K]) {
        if constexpr (CLAUSE_RECORDS) {
            // one record per clause: the words with a report of the group
            uint64_t wm = 0;
            if constexpr (K == 1) {
                // word lw is lane gbase + lw (sub-lane 0): one ballot
                const unsigned long long bw =
                    __ballot(active && sub == 0 && (fm[0] | um[0]) != 0);
                wm = (bw >> gbase) & (LW == 64 ? ~0ull : ((1ull << LW) - 1ull));
            } else {
                // each lane holds K consecutive words: OR over the group's lanes
#pragma unroll
                for (int k = 0; k < K; k++)
                    if (active && (fm[k] | um[k]) != 0) wm |= 1ull << (lw * K + k);
#pragma unroll
                for (int d = 1; d < G; d <<= 1) wm |= __shfl_xor(wm, d);
            }
            const bool hit = active && q == 0 && wm != 0;
            const unsigned long long bal = __ballot(hit);
            if (bal) {
                claim(bal);
                if (hit && (uint32_t)chunk < a.chunk_cap) {
                    const size_t r = (size_t)chunk * kHitChunk + fill + prefix_count(bal);
                    uint64_t *dst = (uint64_t *)a.hits + r * 2;
                    __builtin_nontemporal_store((uint64_t)(uint32_t)row | ((uint64_t)len << 32),
                                                dst);
                    __builtin_nontemporal_store(wm, dst + 1);
                }
                fill += __popcll(bal);
            }
        } else {
#pragma unroll
            for (int k = 0; k < K; k++) {
                const bool hit = active && sub == 0 && (fm[k] | um[k]) != 0;
                const unsigned long long bal = __ballot(hit);
                if (bal) {
                    const int n = __popcll(bal);
                    claim(bal);
                    if (hit && (uint32_t)chunk < a.chunk_cap) {
                        // Records are streamed out and never read by the
                        // kernel: non-temporal stores keep them from
                        // displacing table rows in L2 and the memory-side
                        // cache.
                        const uint64_t id = (uint64_t)(uint32_t)row | ((uint64_t)len << 32) |
                                            ((uint64_t)(lw * K + k) << 48);
                        const size_t r = (size_t)chunk * kHitChunk + fill + prefix_count(bal);
                        uint64_t *dst = (uint64_t *)a.hits + r * (size_t)a.rec_words;
                        __builtin_nontemporal_store(id, dst);
                        if (a.rec_words == 3) {
                            __builtin_nontemporal_store(fm[k], dst + 1);
                            __builtin_nontemporal_store(um[k], dst + 2);
                        }
                    }
                    fill += n;
                }
            }
        }
    };
    for (int task = gw; task < ntasks; task += nw) {
        const int bi = uniform(find_bucket(a.buckets, nbuckets, task));
        const BucketDesc bd = a.buckets[bi];
        const int len = uniform((int)bd.len);
        const int64_t row0 = (int64_t)(task - uniform((int)bd.task_begin)) * a.task_clauses;
        const int64_t rem = (int64_t)uniform((int)bd.count) - row0;
        const int nrows = rem < a.task_clauses ? (int)rem : a.task_clauses;
        const uint32_t *rows_base = bd.lits + (size_t)row0 * len;
        if constexpr (AGG) {
            // Serving steps: each lane screens one clause against the
            // aggregate entries; the groups then check the survivors exactly,
            // CPW at a time, on the words of their surviving request groups.
            for (int it = 0; it < nrows; it += 64) {
                const int c = it + lane;
                const uint32_t grp =
                    c < nrows ? lane_agg(rows_base + (size_t)c * len, len, a.agg, a.agg_vote, gdead)
                              : 0u;
                unsigned long long left = __ballot(grp != 0);
                while (left) {
                    unsigned long long m = left;
                    for (int k = 0; k < g && m; k++) m &= m - 1;  // group g: the g-th survivor
                    const int src = m ? (int)__builtin_ctzll(m) : 0;
                    const bool active = m != 0;
                    for (int k = 0; k < CPW && left; k++) left &= left - 1;
                    const uint32_t gr = (uint32_t)__shfl((int)grp, src);  // the survivor's lane
                    const int64_t row = row0 + it + src;
                    uint64_t fm[K], um[K];
#pragma unroll
                    for (int k = 0; k < K; k++) fm[k] = um[k] = 0;
                    if (active) {
                        uint64_t alive = 0;
                        for (int b = 0; b < 8; b++)
                            if ((gr >> b) & 1u) alive |= a.agg_words[b];
                        uint64_t vk[K];
#pragma unroll
                        for (int k = 0; k < K; k++)
                            vk[k] = ((alive >> (lw * K + k)) & 1ull) ? valid[k] : 0ull;
                        const uint32_t *cl = bd.lits + (size_t)row * len;
                        const uint32_t chunk0 = q < len ? cl[q] : a.sentinel;
                        check_clause<W, S, U, K, FUSE>(cl, len, chunk0, a.sentinel, a.table, q,
                                                       gmask, gbase, w8, sub, vk, a.fuse_len, fm,
                                                       um);
                    }
                    emit(active, row, len, fm, um);
                }
            }
        } else {
            // literal q of the group's first clause chunk, prefetched one pass ahead
            uint32_t next_chunk =
                (g < nrows && q < len) ? rows_base[(size_t)g * len + q] : a.sentinel;
            for (int it = 0; it < nrows; it += CPW) {
                const bool active = it + g < nrows;
                const int64_t row = row0 + it + g;
                const uint32_t chunk0 = next_chunk;
                const int nxt = it + CPW + g;
                next_chunk =
                    (nxt < nrows && q < len) ? rows_base[(size_t)nxt * len + q] : a.sentinel;
                uint64_t fm[K], um[K];
#pragma unroll
                for (int k = 0; k < K; k++) fm[k] = um[k] = 0;
                if (active)
                    check_clause<W, S, U, K, FUSE>(bd.lits + (size_t)row * len, len, chunk0,
                                                   a.sentinel, a.table, q, gmask, gbase, w8, sub,
                                                   valid, a.fuse_len, fm, um);
                emit(active, row, len, fm, um);
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

// Copies segments of uint32; with shift > 0 each element is shifted left
// (literal -> stored row offset on clause upload).
__global__ void k_scatter(const ScatterSeg *__restrict__ segs, int nseg,
                          const uint32_t *__restrict__ src, int shift) {
    for (int s = blockIdx.x; s < nseg; s += gridDim.x) {
        const ScatterSeg sg = segs[s];
        for (uint64_t i = threadIdx.x; i < sg.count; i += blockDim.x)
            sg.dst[i] = src[sg.src + i] << shift;
    }
}

// Row 2v holds the slots where literal 2v (v) is not false, i.e. v is not
// false; row 2v+1 the slots where v is not true.
__global__ void k_assig_apply(const gss_var_update *__restrict__ ups, int64_t n,
                              uint64_t *__restrict__ table, int words) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const gss_var_update u = ups[i];
    const size_t base = (size_t)u.var * 2 * words + u.word;
    table[base] = ~(u.is_def & ~u.is_true);
    table[base + words] = ~(u.is_def & u.is_true);
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

// ---- batched steps: fixed launch shapes, counts from device memory --------

// Copies a step's staged inputs from pinned host memory to the device.  The
// region sizes are read once per block from the pinned header.
__global__ void k_stage_in(StageArgs a) {
    if (*a.halt) {
        // halted: a header with no work (and `halted` set) instead of the step
        if (blockIdx.x == 0) {
            uint32_t *d = a.d[kRegHeader];
            constexpr int kFlag = (int)(offsetof(StepHeader, halted) / 4);
            for (int i = (int)threadIdx.x; i < (int)(sizeof(StepHeader) / 4); i += (int)blockDim.x)
                d[i] = i == kFlag ? 1u : 0u;
        }
        return;
    }
    __shared__ int64_t n[kStageRegions];
    if (threadIdx.x < kStageRegions)
        n[threadIdx.x] = ((const StepHeader *)a.h[kRegHeader])->region_words[threadIdx.x];
    __syncthreads();
    const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
#pragma unroll
    for (int r = 0; r < kStageRegions; r++) {
        const uint32_t *src = a.h[r];
        uint32_t *dst = a.d[r];
        for (int64_t i = tid; i < n[r]; i += stride) dst[i] = src[i];
    }
}

__global__ void k_base_apply_dyn(const gss_base_update *__restrict__ u, const int64_t *n_p,
                                 uint64_t *bases, int64_t vw) {
    const int64_t n = *n_p;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const gss_base_update x = u[i];
        const uint64_t bit = 1ull << (x.var & 63);
        unsigned long long *def = (unsigned long long *)bases + (size_t)(2 * x.base) * vw + x.var / 64;
        unsigned long long *tru = def + vw;
        if (x.is_def) atomicOr(def, bit);
        else atomicAnd(def, ~bit);
        if (x.is_def && x.is_true) atomicOr(tru, bit);
        else atomicAnd(tru, ~bit);
    }
}

// Element i covers variable i / nloads and load i % nloads, so the lanes of
// a wave write neighbouring words of the same rows.
__global__ void k_words_load_dyn(const gss_word_load *__restrict__ loads, const int64_t *n_p,
                                 const uint64_t *__restrict__ bases, int64_t vw,
                                 uint64_t *__restrict__ table, int words, int64_t vars) {
    const int64_t nl = *n_p;
    const int64_t total = nl * vars;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = i / nl;
        const gss_word_load ld = loads[i - v * nl];
        const uint64_t *def = bases + (size_t)(2 * ld.base) * vw;
        const uint64_t bit = 1ull << (v & 63);
        const bool d = (def[v >> 6] & bit) != 0;
        const bool t = (def[vw + (v >> 6)] & bit) != 0;
        const size_t row = (size_t)v * 2 * words + ld.word;
        table[row] = (d && !t) ? 0ull : ~0ull;     // literal v is false where v is false
        table[row + words] = (d && t) ? 0ull : ~0ull;  // literal ~v is false where v is true
    }
}

// Serving steps: load k (0 <= k < n) sets word k from loads[k].base (the
// admission hands out words 0..n-1 in order), and the words a serving step
// does not load are not valid in it.  So each row is written in whole
// 128-byte lines: words [0, L), L = n rounded up to 16 (at most `words`),
// with words n..L-1 set to all ones.  Written word by word (k_words_load_dyn),
// 1 to 12 words of a 64-word table cost 0.19-0.29 ms at 2M variables: each
// row line was written in part.  Whole lines, 16-byte stores: a serving
// step's device time fell from 0.435 to 0.327 ms at 2M variables, ~1.3
// requests per step.  One workgroup per 64 variables: the
// bases' 64-variable words go through LDS, then each lane writes 16 bytes
// (2 words) of a variable's F row and of its T row, neighbouring lanes
// neighbouring words.  `words` is even (16-byte stores stay aligned).
__global__ __launch_bounds__(256) void k_words_load_lines(const gss_word_load *__restrict__ loads,
                                                          const int64_t *n_p,
                                                          const uint64_t *__restrict__ bases,
                                                          int64_t vw, uint64_t *__restrict__ table,
                                                          int words, int64_t vars) {
    const int nl = (int)*n_p;
    if (nl <= 0) return;
    __shared__ uint64_t s_bits[2 * GSS_MAX_WORDS];  // [2k] def, [2k + 1] true of load k
    const int64_t blk = blockIdx.x;                   // variables [64 blk, 64 blk + 64)
    const int tid = (int)threadIdx.x;
    if (tid < 2 * nl) {
        const gss_word_load ld = loads[tid >> 1];
        s_bits[tid] = bases[(size_t)(2 * ld.base + (tid & 1)) * vw + blk];
    }
    __syncthreads();
    int L = (nl + 15) & ~15;
    if (L > words) L = words;
    const int pairs = L >> 1;  // 16-byte stores per row
    for (int it = tid; it < 64 * pairs; it += 256) {
        const int vi = it / pairs, p = it - vi * pairs;
        const int64_t v = blk * 64 + vi;
        if (v >= vars) break;  // it grows with vi
        const uint64_t bit = 1ull << vi;
        uint64_t f[2], t[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int w = 2 * p + h;
            f[h] = ~0ull;
            t[h] = ~0ull;
            if (w < nl) {
                const bool d = (s_bits[2 * w] & bit) != 0;
                const bool tr = (s_bits[2 * w + 1] & bit) != 0;
                if (d && !tr) f[h] = 0ull;  // literal v is false where v is false
                if (d && tr) t[h] = 0ull;   // literal ~v is false where v is true
            }
        }
        uint64_t *row = table + (size_t)v * 2 * words + 2 * p;
        *(ulonglong2 *)row = make_ulonglong2(f[0], f[1]);
        *(ulonglong2 *)(row + words) = make_ulonglong2(t[0], t[1]);
    }
}

// Slots in mask take the patch's value; the others keep theirs.
__global__ void k_assig_patch_dyn(const gss_var_patch *__restrict__ p, const int64_t *n_p,
                                  uint64_t *__restrict__ table, int words) {
    const int64_t n = *n_p;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const gss_var_patch x = p[i];
        const size_t base = (size_t)x.var * 2 * words + x.word;
        const uint64_t m = x.mask;
        table[base] = (table[base] & ~m) | (~(x.is_def & ~x.is_true) & m);
        table[base + words] = (table[base + words] & ~m) | (~(x.is_def & x.is_true) & m);
    }
}

// Eight lanes per staged clause: lane j copies literals j, j + 8, ...
__global__ void k_scatter_clauses(const ClauseSeg *__restrict__ cls, const int32_t *n_p,
                                  const uint32_t *__restrict__ src, int shift) {
    const int64_t n = *n_p;
    const int lane = threadIdx.x & 7;
    const int64_t ng = ((int64_t)gridDim.x * blockDim.x) >> 3;
    for (int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 3; i < n; i += ng) {
        const ClauseSeg c = cls[i];
        for (uint32_t j = lane; j < c.len; j += 8) c.dst[j] = src[(uint64_t)c.src + j] << shift;
    }
}

__global__ void k_assig_apply_dyn(const gss_var_update *__restrict__ ups, const int64_t *n_p,
                                  uint64_t *__restrict__ table, int words) {
    const int64_t n = *n_p;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const gss_var_update u = ups[i];
        const size_t base = (size_t)u.var * 2 * words + u.word;
        table[base] = ~(u.is_def & ~u.is_true);
        table[base + words] = ~(u.is_def & u.is_true);
    }
}

__global__ void k_assig_roll_dyn(const gss_var_update *__restrict__ ups, const int64_t *n_p,
                                 uint64_t *__restrict__ table, int words,
                                 const int8_t *__restrict__ last_pos, const uint32_t *alloc,
                                 uint32_t cap, uint32_t *halt) {
    if (*alloc > cap) {
        // records were dropped: keep the table as the sweep read it, and halt
        if (blockIdx.x == 0 && threadIdx.x == 0) *halt = 1u;
        return;
    }
    const int64_t n = *n_p;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const gss_var_update u = ups[i];
        const int p = last_pos[u.word];
        if (p < 0) continue;
        const size_t base = (size_t)u.var * 2 * words + u.word;
        const uint64_t f = table[base], t = table[base + words];
        table[base] = ((f >> p) & 1ull) ? ~0ull : 0ull;
        table[base + words] = ((t >> p) & 1ull) ? ~0ull : 0ull;
    }
}

// Writes the claim count and the fill of every claimed chunk to pinned host
// memory: h_out[0] = claimed chunks, h_out[1 + c] = fill of chunk c.
__global__ void k_stage_out(const uint32_t *alloc, const uint32_t *__restrict__ fills,
                            uint32_t cap, uint32_t *h_out) {
    const uint32_t used = *alloc;
    const uint32_t n = used < cap ? used : cap;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid == 0) h_out[0] = used;
    for (uint32_t i = tid; i < n; i += gridDim.x * blockDim.x) h_out[1 + i] = fills[i];
}

// Exclusive prefix of the chunk fills: offs[c] = records before chunk c for
// the chunks [0, n) a sweep claimed (n = min(*alloc, cap), read on the device
// so a graph replay needs no host count), offs[n] = all records.  One block:
// each thread sums a contiguous run of chunks, the 1024 sums are scanned in
// LDS, then each thread writes its run's prefixes.
__global__ __launch_bounds__(1024) void k_chunk_offsets(const uint32_t *alloc,
                                                        const uint32_t *__restrict__ fills,
                                                        uint32_t cap, uint64_t *__restrict__ offs) {
    __shared__ uint64_t part[1024];
    const uint32_t used = *alloc;
    const uint32_t n = used < cap ? used : cap;
    const uint32_t t = threadIdx.x;
    const uint32_t per = (n + 1023u) / 1024u;
    const uint32_t b = (uint64_t)t * per < n ? t * per : n;
    const uint32_t e = b + per < n ? b + per : n;
    uint64_t sum = 0;
    for (uint32_t i = b; i < e; i++) sum += fills[i];
    part[t] = sum;
    __syncthreads();
    for (uint32_t d = 1; d < 1024u; d <<= 1) {
        const uint64_t v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint64_t acc = t ? part[t - 1] : 0;
    for (uint32_t i = b; i < e; i++) {
        offs[i] = acc;
        acc += fills[i];
    }
    if (t == 1023u) offs[n] = part[1023];
}

// Copies the records of every claimed chunk to its place in a dense array
// (three 8-byte words per record), or nothing when they would not all fit in
// dense_cap records (the host then compacts them itself).
__global__ __launch_bounds__(256) void k_gather_chunks(const uint32_t *alloc,
                                                       const uint32_t *__restrict__ fills,
                                                       uint32_t cap,
                                                       const uint64_t *__restrict__ offs,
                                                       const uint64_t *__restrict__ hits,
                                                       uint64_t *__restrict__ dense,
                                                       uint64_t dense_cap, uint32_t rec_words) {
    const uint32_t used = *alloc;
    const uint32_t n = used < cap ? used : cap;
    if (offs[n] > dense_cap) return;
    for (uint32_t c = blockIdx.x; c < n; c += gridDim.x) {
        const uint32_t words = fills[c] * rec_words;
        const uint64_t *src = hits + (uint64_t)c * kHitChunk * rec_words;
        uint64_t *dst = dense + offs[c] * rec_words;
        for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
    }
}

// Claim counter and record total of a sweep to pinned host memory:
// out[0] = chunks claimed, out[1] / out[2] = low / high word of the records
// in the chunks that fit (one block).
__global__ __launch_bounds__(1024) void k_count_records(const uint32_t *alloc,
                                                        const uint32_t *__restrict__ fills,
                                                        uint32_t cap, uint32_t *out) {
    __shared__ uint64_t part[1024];
    const uint32_t used = *alloc;
    const uint32_t n = used < cap ? used : cap;
    uint64_t sum = 0;
    for (uint32_t i = threadIdx.x; i < n; i += 1024u) sum += fills[i];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t d = 512; d > 0; d >>= 1) {
        if (threadIdx.x < d) part[threadIdx.x] += part[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = used;
        out[1] = (uint32_t)part[0];
        out[2] = (uint32_t)(part[0] >> 32);
    }
}

// ---- records grouped by assignment word (serving steps) ----------------------

// Records per word over the claimed chunks (word = the top 16 bits of a
// record's first word); counts[64] must start at zero.
__global__ __launch_bounds__(256) void k_word_count(const uint32_t *alloc,
                                                    const uint32_t *__restrict__ fills,
                                                    uint32_t cap, const uint64_t *__restrict__ hits,
                                                    uint32_t rec_words, uint32_t *counts) {
    __shared__ uint32_t h[GSS_MAX_WORDS];
    if (threadIdx.x < GSS_MAX_WORDS) h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t used = *alloc;
    const uint32_t n = used < cap ? used : cap;
    for (uint32_t c = blockIdx.x; c < n; c += gridDim.x) {
        const uint32_t f = fills[c];
        const uint64_t *src = hits + (uint64_t)c * kHitChunk * rec_words;
        for (uint32_t i = threadIdx.x; i < f; i += blockDim.x)
            atomicAdd(&h[(uint32_t)(src[(uint64_t)i * rec_words] >> 48) & (GSS_MAX_WORDS - 1)], 1u);
    }
    __syncthreads();
    if (threadIdx.x < GSS_MAX_WORDS && h[threadIdx.x]) atomicAdd(&counts[threadIdx.x], h[threadIdx.x]);
}

// Moves the records into `dense` grouped by word, ascending (within a word in
// no set order): a block takes one chunk at a time, counts its records per
// word in LDS, reserves each word's share with one atomic on that word's
// cursor, then places them.  group[w] = first record of word w, group[64] =
// all records, group[65] = 1 when they fitted in dense_cap (else nothing is
// moved and the host groups the records itself).  cursor[64] starts at zero.
__global__ __launch_bounds__(256) void k_word_scatter(const uint32_t *alloc,
                                                      const uint32_t *__restrict__ fills,
                                                      uint32_t cap,
                                                      const uint64_t *__restrict__ hits,
                                                      uint32_t rec_words,
                                                      const uint32_t *__restrict__ counts,
                                                      uint32_t *cursor, uint64_t *__restrict__ dense,
                                                      uint64_t dense_cap, uint64_t *group) {
    __shared__ uint64_t base[GSS_MAX_WORDS + 1];
    __shared__ uint32_t lc[GSS_MAX_WORDS], lbase[GSS_MAX_WORDS];
    if (threadIdx.x == 0) {
        uint64_t acc = 0;
        for (int w = 0; w < GSS_MAX_WORDS; w++) {
            base[w] = acc;
            acc += counts[w];
        }
        base[GSS_MAX_WORDS] = acc;
    }
    __syncthreads();
    const bool fits = base[GSS_MAX_WORDS] <= dense_cap;
    if (blockIdx.x == 0 && threadIdx.x <= GSS_MAX_WORDS + 1)
        group[threadIdx.x] = threadIdx.x <= GSS_MAX_WORDS ? base[threadIdx.x] : (fits ? 1ull : 0ull);
    if (!fits) return;  // the same for every thread of every block
    const uint32_t used = *alloc;
    const uint32_t n = used < cap ? used : cap;
    for (uint32_t c = blockIdx.x; c < n; c += gridDim.x) {
        if (threadIdx.x < GSS_MAX_WORDS) lc[threadIdx.x] = 0;
        __syncthreads();
        const uint32_t f = fills[c];
        const uint64_t *src = hits + (uint64_t)c * kHitChunk * rec_words;
        for (uint32_t i = threadIdx.x; i < f; i += blockDim.x)
            atomicAdd(&lc[(uint32_t)(src[(uint64_t)i * rec_words] >> 48) & (GSS_MAX_WORDS - 1)], 1u);
        __syncthreads();
        if (threadIdx.x < GSS_MAX_WORDS) {
            const uint32_t k = lc[threadIdx.x];
            lbase[threadIdx.x] = k ? atomicAdd(&cursor[threadIdx.x], k) : 0u;
            lc[threadIdx.x] = 0;
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < f; i += blockDim.x) {
            const uint64_t *r = src + (uint64_t)i * rec_words;
            const uint32_t w = (uint32_t)(r[0] >> 48) & (GSS_MAX_WORDS - 1);
            const uint64_t pos = base[w] + lbase[w] + atomicAdd(&lc[w], 1u);
            for (uint32_t q = 0; q < rec_words; q++) dense[pos * rec_words + q] = r[q];
        }
        __syncthreads();
    }
}

// ---- serving: requests admitted on the device ---------------------------------

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t x, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    return x;
}

__global__ __launch_bounds__(64) void k_serve_admit(ServeRing r, StepHeader *dh,
                                                    gss_word_load *loads, ServeReport *report,
                                                    int words) {
    if (dh->halted) return;  // the ring waits for the step's replay
    const int lane = (int)threadIdx.x;
    __shared__ uint64_t s_head;
    if (lane == 0) {
        // published by the host with a release store after the entries
        s_head = __hip_atomic_load(r.head, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");  // later reads see the published entries
    const uint64_t tail = *r.tail;
    const uint64_t head = s_head;
    const uint64_t pend = head > tail ? head - tail : 0;
    const int avail = (int)(pend < (uint64_t)kServeMaxRequests ? pend : kServeMaxRequests);
    uint32_t solver = 0, k = 0, np = 0;
    uint64_t poff = 0;
    if (lane < avail) {
        const volatile ServeReq *q = r.reqs + (tail + lane) % r.req_cap;
        solver = q->solver;
        k = q->k;
        np = q->n_patches;
        poff = q->patch_off;
    }
    const uint64_t nw = lane < avail ? (k + 63u) / 64u : 0;
    const uint64_t end = wave_incl_scan(nw, lane);
    // end grows with the lane (every request takes at least one word), so
    // the requests that fit are a prefix
    const bool fits = lane < avail && end <= (uint64_t)words;
    const unsigned long long bal = __ballot(fits);
    const int count = __popcll(bal);
    const uint64_t pend_p = wave_incl_scan(fits ? (uint64_t)np : 0, lane);
    const uint64_t used = __shfl(end, count > 0 ? count - 1 : 0);
    const uint64_t total_words = count > 0 ? used : 0;
    const uint64_t total_patches = __shfl(pend_p, count > 0 ? count - 1 : 0);
    if (fits) {
        const uint32_t w0 = (uint32_t)(end - nw);
        ServeAdmitted a;
        a.patch_off = poff;
        a.patch_pfx = pend_p - np;
        a.n_patches = np;
        a.w0 = w0;
        r.adm[lane] = a;
        for (uint32_t j = 0; j < (uint32_t)nw; j++) {
            gss_word_load ld;
            ld.word = (uint16_t)(w0 + j);
            ld.base = (uint16_t)solver;
            loads[w0 + j] = ld;
            const uint32_t last = k - 64u * j;  // slots of this word
            dh->valid[w0 + j] = last >= 64u ? ~0ull : ((1ull << last) - 1ull);
        }
    }
    // words no request took are not checked
    if ((uint64_t)lane < (uint64_t)GSS_MAX_WORDS && (uint64_t)lane >= total_words)
        dh->valid[lane] = 0ull;
    if (lane == 0) {
        *r.adm_count = (uint32_t)count;
        dh->nloads = (int64_t)total_words;
        dh->npatches = count > 0 ? (int64_t)total_patches : 0;
        if (count == 0) dh->ntasks = 0;  // nothing to check
        *r.tail = tail + (uint64_t)count;
        report->first = tail;
        report->count = (uint32_t)count;
        report->words = (uint32_t)total_words;
    }
}

// The admitted request (index into r.adm) of the step's i-th patch.
__device__ __forceinline__ int admitted_of(const ServeRing &r, int cnt, int64_t i) {
    int lo = 0, hi = cnt;  // last admitted request whose patches start at or before i
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if ((int64_t)r.adm[mid].patch_pfx <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void k_serve_stage_patches(ServeRing r, const StepHeader *__restrict__ dh) {
    constexpr int kWords = (int)(sizeof(gss_var_patch) / 4);
    const int64_t n = dh->npatches * kWords;
    const int cnt = (int)*r.adm_count;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = k / kWords;
        const int f = (int)(k - i * kWords);
        const ServeAdmitted a = r.adm[admitted_of(r, cnt, i)];
        const uint64_t j = (uint64_t)(i - (int64_t)a.patch_pfx);
        const uint32_t *src = (const uint32_t *)(r.patches + (a.patch_off + j) % r.patch_cap);
        ((uint32_t *)(r.dpatches + i))[f] = src[f];
    }
}

// Batched-step aggregate: bit b of a literal row's entry is set if the row is
// all ones over every slot of word group b (valid or not, so that the entry
// follows the table alone and a step's updates refresh it).  All ones on
// every slot implies all ones on the valid ones, which is what the screening
// needs.  agg_words[b] = the group's words with a valid slot in this step.
template <int W>
struct TableGroups {
    static constexpr int NG = W < 8 ? W : 8;  // word groups
    static constexpr int GW = W / NG;         // words per group
};

template <int W>
__device__ __forceinline__ void table_agg_words(const StepHeader *__restrict__ dh,
                                                uint64_t *__restrict__ agg_words) {
    constexpr int NG = TableGroups<W>::NG, GW = TableGroups<W>::GW;
    if (blockIdx.x == 0 && (int)threadIdx.x < 8) {
        uint64_t m = 0;
        if ((int)threadIdx.x < NG)
            for (int w = (int)threadIdx.x * GW; w < ((int)threadIdx.x + 1) * GW; w++)
                if (dh->valid[w]) m |= 1ull << w;
        agg_words[threadIdx.x] = m;
    }
}

template <int W>
__device__ __forceinline__ bool group_all_ones(const uint64_t *__restrict__ row, int b) {
    constexpr int GW = TableGroups<W>::GW;
    uint64_t x = ~0ull;
#pragma unroll
    for (int w = 0; w < GW; w++) x &= row[b * GW + w];
    return x == ~0ull;
}

template <int W>
__global__ void k_table_agg(const uint64_t *__restrict__ table, const StepHeader *__restrict__ dh,
                            int64_t rows, uint8_t *__restrict__ agg,
                            uint64_t *__restrict__ agg_words) {
    constexpr int NG = TableGroups<W>::NG;
    if (dh->halted) return;  // the aggregate stays the halted step's
    table_agg_words<W>(dh, agg_words);
    for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < rows;
         l += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t *row = table + (size_t)l * W;
        uint32_t bits = 0xFFu;
#pragma unroll
        for (int b = 0; b < NG; b++)
            if (!group_all_ones<W>(row, b)) bits &= ~(1u << b);
        agg[l] = (uint8_t)bits;
    }
}

// Refreshes the aggregate bits that the step's updates (and patches) may have
// changed: for each (var, word), the bit of the word's group in the entries
// of both literals of var, from the table as it is now.  Entries are bytes,
// four to a 32-bit word: bits change with 32-bit atomics.  Two updates of one
// (var, group) write the same bit.
template <int W>
__global__ void k_table_agg_update(const uint64_t *__restrict__ table,
                                   const StepHeader *__restrict__ dh,
                                   const gss_var_update *__restrict__ ups,
                                   const gss_var_patch *__restrict__ patches, int with_patches,
                                   uint32_t *__restrict__ agg32,
                                   uint64_t *__restrict__ agg_words) {
    constexpr int GW = TableGroups<W>::GW;
    if (dh->halted) return;
    if (agg_words) table_agg_words<W>(dh, agg_words);
    const int64_t nu = dh->nups, n = nu + (with_patches ? dh->npatches : 0);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t var, word;
        if (i < nu) {
            var = ups[i].var;
            word = ups[i].word;
        } else {
            var = patches[i - nu].var;
            word = patches[i - nu].word;
        }
        const int b = (int)(word / GW);
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const uint64_t l = 2ull * var + s;
            const uint32_t bit = 1u << (b + 8 * (int)(l & 3));
            if (group_all_ones<W>(table + l * W, b)) atomicOr(agg32 + (l >> 2), bit);
            else atomicAnd(agg32 + (l >> 2), ~bit);
        }
    }
}

// Aggregate entries of a serving step: bit b of agg[2v] (agg[2v + 1]) is set
// if literal v (~v) is not false in the bases of all of group b's requests
// (r % 8 = b), or if the group is empty.  A request's words are loaded whole
// from its solver's base, so until its patches (k_serve_patch clears what
// they make false) the base value holds in all its slots.
__global__ void k_serve_agg(ServeRing r, const StepHeader *__restrict__ dh,
                            const gss_word_load *__restrict__ loads,
                            const uint64_t *__restrict__ bases, int64_t vw, int64_t vars,
                            uint8_t *__restrict__ agg, uint64_t *__restrict__ agg_words) {
    if (dh->halted) return;  // the aggregate stays the halted step's
    __shared__ uint32_t s_base[kServeMaxRequests];
    const int cnt = (int)*r.adm_count;
    if ((int)threadIdx.x < cnt) s_base[threadIdx.x] = loads[r.adm[threadIdx.x].w0].base;
    if (blockIdx.x == 0 && threadIdx.x < 8) {
        const int64_t nl = dh->nloads;
        uint64_t m = 0;
        for (int q = (int)threadIdx.x; q < cnt; q += 8) {
            const uint32_t w0 = r.adm[q].w0;
            const uint32_t w1 = q + 1 < cnt ? r.adm[q + 1].w0 : (uint32_t)nl;
            for (uint32_t w = w0; w < w1 && w < 64u; w++) m |= 1ull << w;
        }
        agg_words[threadIdx.x] = m;
    }
    __syncthreads();
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v <= vars;
         v += (int64_t)gridDim.x * blockDim.x) {
        uint32_t pos = 0xFFu, neg = 0xFFu;
        if (v < vars) {
            const uint64_t bit = 1ull << (v & 63);
            const int64_t i = v >> 6;
            for (int q = 0; q < cnt; q++) {
                const uint64_t *def = bases + (size_t)(2 * s_base[q]) * vw;
                const bool d = (def[i] & bit) != 0;
                const bool t = (def[vw + i] & bit) != 0;
                const uint32_t gb = 1u << (q & 7);
                if (d && !t) pos &= ~gb;  // v is false in this base
                if (d && t) neg &= ~gb;   // ~v is false in this base
            }
        } else {  // sentinel pair: the padding entry contributes nothing
            pos = 0u;
        }
        ((uint16_t *)agg)[v] = (uint16_t)(pos | (neg << 8));
    }
}

__global__ void k_serve_patch(ServeRing r, const StepHeader *__restrict__ dh,
                              uint64_t *__restrict__ table, int words,
                              uint32_t *__restrict__ agg32) {
    const int64_t n = dh->npatches;
    const int cnt = (int)*r.adm_count;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int q = admitted_of(r, cnt, i);
        const uint32_t w0 = r.adm[q].w0;
        const gss_var_patch pp = r.dpatches[i];
        const uint32_t var = pp.var;
        const uint32_t word = pp.word + w0;
        const uint64_t m = pp.mask, d = pp.is_def, t = pp.is_true;
        const size_t base = (size_t)var * 2 * words + word;
        table[base] = (table[base] & ~m) | (~(d & ~t) & m);
        table[base + words] = (table[base + words] & ~m) | (~(d & t) & m);
        if (agg32 && word < 64u) {
            const uint64_t vm = m & dh->valid[word];
            const uint32_t gb = 1u << (q & 7);
            // entries 2 var and 2 var + 1 are bytes 2 (var & 1) and +1 of
            // 4-byte word var / 2
            const uint32_t sh = (var & 1u) * 16u;
            if (d & ~t & vm) atomicAnd(&agg32[var >> 1], ~(gb << sh));
            if (d & t & vm) atomicAnd(&agg32[var >> 1], ~(gb << (sh + 8u)));
        }
    }
}

// ---- tagged DB compaction --------------------------------------------------

__device__ __forceinline__ int find_ref(const CompactRef *__restrict__ r, int nb, int64_t g) {
    int lo = 0, hi = nb;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (r[mid].start <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void k_compact_flags(const CompactRef *__restrict__ refs, int nb, int64_t n,
                                const uint8_t *__restrict__ keep, int64_t ntags,
                                uint32_t *__restrict__ flags) {
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int b = find_ref(refs, nb, g);
        const uint32_t tag = refs[b].tags[g - refs[b].start];
        flags[g] = tag < (uint64_t)ntags ? (keep[tag] ? 1u : 0u) : 1u;
    }
}

__global__ void k_compact_bounds(const CompactRef *__restrict__ refs, int nb, int64_t n,
                                 const uint32_t *__restrict__ flags,
                                 const uint32_t *__restrict__ pos, uint32_t *__restrict__ kept) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int64_t st = refs[b].start;
    const int64_t en = b + 1 < nb ? refs[b + 1].start : n;
    const uint32_t pe = en < n ? pos[en] : pos[n - 1] + flags[n - 1];
    kept[b] = pe - pos[st];
}

// One thread per source row: kept rows move to their rank within the bucket.
__global__ void k_compact_move(const CompactRef *__restrict__ refs, int nb, int64_t n,
                               const uint32_t *__restrict__ flags,
                               const uint32_t *__restrict__ pos) {
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
         g += (int64_t)gridDim.x * blockDim.x) {
        if (!flags[g]) continue;
        const int b = find_ref(refs, nb, g);
        const CompactRef r = refs[b];
        const int64_t row = g - r.start;
        const int64_t dst = (int64_t)(pos[g] - pos[r.start]);
        const uint32_t *src = r.lits + row * r.len;
        uint32_t *d = r.dst_lits + dst * r.len;
        for (uint32_t i = 0; i < r.len; i++) d[i] = src[i];
        r.dst_tags[dst] = r.tags[row];
    }
}

template <int W, int S, int U>
hipError_t launch_sweep_t(const SweepArgs &a, int blocks, hipStream_t s) {
    const dim3 g(blocks), b(256);
    const bool fused = a.fuse_len > 0;
    const int mode = (a.agg ? 1 : 0) + (a.rec_words == 2 ? 2 : 0);
#define GSS_LAUNCH(F_, M_)                                                     \
    if (fused == F_ && mode == M_) {                                           \
        hipLaunchKernelGGL((k_sweep<W, S, U, F_, M_>), g, b, 0, s, a);         \
        return hipGetLastError();                                              \
    }
    GSS_LAUNCH(true, 0)
    GSS_LAUNCH(false, 0)
    GSS_LAUNCH(true, 1)
    GSS_LAUNCH(false, 1)
    GSS_LAUNCH(true, 2)
    GSS_LAUNCH(false, 2)
    GSS_LAUNCH(true, 3)
    GSS_LAUNCH(false, 3)
#undef GSS_LAUNCH
    return hipErrorInvalidValue;
}

template <int W, int S, int U>
int occupancy_t(bool fused, int mode) {
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
#define GSS_OCC(F_, M_)                                                                     \
    if (fused == F_ && mode == M_)                                                          \
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sweep<W, S, U, F_, M_>, 256, 0);
    GSS_OCC(true, 0)
    GSS_OCC(false, 0)
    GSS_OCC(true, 1)
    GSS_OCC(false, 1)
    GSS_OCC(true, 2)
    GSS_OCC(false, 2)
    GSS_OCC(true, 3)
    GSS_OCC(false, 3)
#undef GSS_OCC
    if (e != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

// Kernel variants: (W, S, U), each with and without the fused schedule
// (fuse_len <= 0 selects the one without), in the four modes of k_sweep.  `unroll` selects U for W >= 16
// (0 = default).
#define GSS_SWEEP_VARIANTS(X)                  \
    X(1, 16, 1)                                \
    X(2, 8, 2)                                 \
    X(4, 4, 4)                                 \
    X(8, 2, 4)                                 \
    X(16, 1, 4)                                \
    X(16, 1, 1)                                \
    X(16, 1, 2)                                \
    X(16, 1, 8)                                \
    X(32, 1, 2)                                \
    X(32, 1, 4)                                \
    X(64, 1, 1)                                \
    X(64, 1, 2)                                \
    X(64, 1, 4)

struct Variant {
    int w, s, u;
};

Variant pick(int words, int unroll) {
    switch (words) {
        case 1: return {1, 16, 1};
        case 2: return {2, 8, 2};
        case 4: return {4, 4, 4};
        case 8: return {8, 2, 4};
        case 16:
            if (unroll == 1 || unroll == 2 || unroll == 8) return {16, 1, unroll};
            return {16, 1, 4};
        case 32: return {32, 1, unroll == 2 ? 2 : 4};
        case 64: return {64, 1, unroll == 1 || unroll == 4 ? unroll : 2};
        default: return {0, 0, 0};
    }
}

}  // namespace

// Every variant gives a clause a group of 16 lanes (W*S = 16 for W <= 16, 16
// lanes of W/16 words each above), so a wave pass checks 4 clauses.
int sweep_clauses_per_pass(int) { return 4; }

hipError_t launch_sweep(const SweepArgs &a, int words, int unroll, int blocks, hipStream_t s) {
    const Variant v = pick(words, unroll);
#define X(W_, S_, U_) \
    if (v.w == W_ && v.s == S_ && v.u == U_) return launch_sweep_t<W_, S_, U_>(a, blocks, s);
    GSS_SWEEP_VARIANTS(X)
#undef X
    return hipErrorInvalidValue;
}

int sweep_occupancy(int words, int unroll, bool fused, int mode) {
    const Variant v = pick(words, unroll);
#define X(W_, S_, U_) \
    if (v.w == W_ && v.s == S_ && v.u == U_) return occupancy_t<W_, S_, U_>(fused, mode);
    GSS_SWEEP_VARIANTS(X)
#undef X
    return 1;
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

hipError_t launch_scatter(const ScatterSeg *segs, int nseg, const uint32_t *src, int shift,
                          hipStream_t s) {
    if (nseg <= 0) return hipSuccess;
    const int blocks = nseg < 4096 ? nseg : 4096;
    hipLaunchKernelGGL(k_scatter, dim3(blocks), dim3(256), 0, s, segs, nseg, src, shift);
    return hipGetLastError();
}

hipError_t launch_assig_apply(const gss_var_update *ups, int64_t n, uint64_t *table, int words,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_assig_apply, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ups, n,
                       table, words);
    return hipGetLastError();
}

hipError_t launch_stage_in(const StageArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_stage_in, dim3(512), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_base_apply_dyn(const gss_base_update *u, const int64_t *n, uint64_t *bases,
                                 int64_t vw, hipStream_t s) {
    hipLaunchKernelGGL(k_base_apply_dyn, dim3(256), dim3(256), 0, s, u, n, bases, vw);
    return hipGetLastError();
}

hipError_t launch_words_load_dyn(const gss_word_load *loads, const int64_t *n,
                                 const uint64_t *bases, int64_t vw, uint64_t *table, int words,
                                 int64_t vars, hipStream_t s) {
    hipLaunchKernelGGL(k_words_load_dyn, dim3(4096), dim3(256), 0, s, loads, n, bases, vw, table,
                       words, vars);
    return hipGetLastError();
}

hipError_t launch_words_load_lines(const gss_word_load *loads, const int64_t *n,
                                   const uint64_t *bases, int64_t vw, uint64_t *table, int words,
                                   int64_t vars, hipStream_t s) {
    if (words % 2 || vw != (vars + 63) / 64) return hipErrorInvalidValue;
    if (vars <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_words_load_lines, dim3((unsigned)vw), dim3(256), 0, s, loads, n, bases,
                       vw, table, words, vars);
    return hipGetLastError();
}

hipError_t launch_assig_patch_dyn(const gss_var_patch *p, const int64_t *n, uint64_t *table,
                                  int words, hipStream_t s) {
    hipLaunchKernelGGL(k_assig_patch_dyn, dim3(1024), dim3(256), 0, s, p, n, table, words);
    return hipGetLastError();
}

hipError_t launch_scatter_clauses(const ClauseSeg *cls, const int32_t *n, const uint32_t *src,
                                  int shift, hipStream_t s) {
    hipLaunchKernelGGL(k_scatter_clauses, dim3(1024), dim3(256), 0, s, cls, n, src, shift);
    return hipGetLastError();
}

hipError_t launch_assig_apply_dyn(const gss_var_update *ups, const int64_t *n, uint64_t *table,
                                  int words, hipStream_t s) {
    hipLaunchKernelGGL(k_assig_apply_dyn, dim3(1024), dim3(256), 0, s, ups, n, table, words);
    return hipGetLastError();
}

hipError_t launch_assig_roll_dyn(const gss_var_update *ups, const int64_t *n, uint64_t *table,
                                 int words, const int8_t *last_pos, const uint32_t *alloc,
                                 uint32_t cap, uint32_t *halt, hipStream_t s) {
    hipLaunchKernelGGL(k_assig_roll_dyn, dim3(1024), dim3(256), 0, s, ups, n, table, words,
                       last_pos, alloc, cap, halt);
    return hipGetLastError();
}

hipError_t launch_stage_out(const uint32_t *alloc, const uint32_t *fills, uint32_t cap,
                            uint32_t *h_out, hipStream_t s) {
    hipLaunchKernelGGL(k_stage_out, dim3(256), dim3(256), 0, s, alloc, fills, cap, h_out);
    return hipGetLastError();
}

hipError_t launch_count_records(const uint32_t *alloc, const uint32_t *fills, uint32_t cap,
                                uint32_t *h_out, hipStream_t s) {
    hipLaunchKernelGGL(k_count_records, dim3(1), dim3(1024), 0, s, alloc, fills, cap, h_out);
    return hipGetLastError();
}

hipError_t launch_compact_records(const uint32_t *alloc, const uint32_t *fills, uint32_t cap,
                                  uint64_t *offs, const void *hits, void *dense,
                                  uint64_t dense_cap, int rec_words, hipStream_t s) {
    if (rec_words < 1 || rec_words > 3) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_chunk_offsets, dim3(1), dim3(1024), 0, s, alloc, fills, cap, offs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_gather_chunks, dim3(2048), dim3(256), 0, s, alloc, fills, cap,
                       (const uint64_t *)offs, (const uint64_t *)hits, (uint64_t *)dense,
                       dense_cap, (uint32_t)rec_words);
    return hipGetLastError();
}

hipError_t launch_group_by_word(const uint32_t *alloc, const uint32_t *fills, uint32_t cap,
                                const void *hits, int rec_words, uint32_t *counts,
                                void *dense, uint64_t dense_cap, uint64_t *group, hipStream_t s) {
    if (rec_words != 1 && rec_words != 3) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_reset_u32, dim3(1), dim3(256), 0, s, counts, 2 * GSS_MAX_WORDS);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_word_count, dim3(1024), dim3(256), 0, s, alloc, fills, cap,
                       (const uint64_t *)hits, (uint32_t)rec_words, counts);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_word_scatter, dim3(1024), dim3(256), 0, s, alloc, fills, cap,
                       (const uint64_t *)hits, (uint32_t)rec_words, counts,
                       counts + GSS_MAX_WORDS, (uint64_t *)dense, dense_cap, group);
    return hipGetLastError();
}

hipError_t launch_serve_admit(const ServeRing &r, StepHeader *dh, gss_word_load *loads,
                              ServeReport *report, int words, hipStream_t s) {
    hipLaunchKernelGGL(k_serve_admit, dim3(1), dim3(64), 0, s, r, dh, loads, report, words);
    return hipGetLastError();
}

hipError_t launch_table_agg(const uint64_t *table, const StepHeader *dh, int64_t vars, int words,
                            uint8_t *agg, uint64_t *agg_words, hipStream_t s) {
    const int64_t rows = 2 * vars + 2;
    const unsigned blocks = (unsigned)std::min<int64_t>((rows + 255) / 256, 8192);
#define GSS_TABLE_AGG(W_)                                                                   \
    if (words == W_) {                                                                      \
        hipLaunchKernelGGL(k_table_agg<W_>, dim3(blocks), dim3(256), 0, s, table, dh, rows, \
                           agg, agg_words);                                                 \
        return hipGetLastError();                                                           \
    }
    GSS_TABLE_AGG(1)
    GSS_TABLE_AGG(2)
    GSS_TABLE_AGG(4)
    GSS_TABLE_AGG(8)
    GSS_TABLE_AGG(16)
    GSS_TABLE_AGG(32)
    GSS_TABLE_AGG(64)
#undef GSS_TABLE_AGG
    return hipErrorInvalidValue;
}

hipError_t launch_table_agg_update(const uint64_t *table, const StepHeader *dh,
                                   const gss_var_update *ups, const gss_var_patch *patches,
                                   bool with_patches, int words, uint8_t *agg,
                                   uint64_t *agg_words, hipStream_t s) {
    uint32_t *agg32 = (uint32_t *)agg;
    const int wp = with_patches ? 1 : 0;
#define GSS_TABLE_AGG_UP(W_)                                                                 \
    if (words == W_) {                                                                       \
        hipLaunchKernelGGL(k_table_agg_update<W_>, dim3(1024), dim3(256), 0, s, table, dh,   \
                           ups, patches, wp, agg32, agg_words);                              \
        return hipGetLastError();                                                            \
    }
    GSS_TABLE_AGG_UP(1)
    GSS_TABLE_AGG_UP(2)
    GSS_TABLE_AGG_UP(4)
    GSS_TABLE_AGG_UP(8)
    GSS_TABLE_AGG_UP(16)
    GSS_TABLE_AGG_UP(32)
    GSS_TABLE_AGG_UP(64)
#undef GSS_TABLE_AGG_UP
    return hipErrorInvalidValue;
}

hipError_t launch_serve_stage_patches(const ServeRing &r, const StepHeader *dh, hipStream_t s) {
    hipLaunchKernelGGL(k_serve_stage_patches, dim3(1024), dim3(256), 0, s, r, dh);
    return hipGetLastError();
}

hipError_t launch_serve_agg(const ServeRing &r, const StepHeader *dh, const gss_word_load *loads,
                            const uint64_t *bases, int64_t vw, int64_t vars, uint8_t *agg,
                            uint64_t *agg_words, hipStream_t s) {
    hipLaunchKernelGGL(k_serve_agg, dim3(2048), dim3(256), 0, s, r, dh, loads, bases, vw, vars,
                       agg, agg_words);
    return hipGetLastError();
}

hipError_t launch_serve_patch(const ServeRing &r, const StepHeader *dh, uint64_t *table,
                              int words, uint8_t *agg, hipStream_t s) {
    hipLaunchKernelGGL(k_serve_patch, dim3(1024), dim3(256), 0, s, r, dh, table, words,
                       (uint32_t *)agg);
    return hipGetLastError();
}

static unsigned grid_for(int64_t n, unsigned cap = 65536) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : b > cap ? cap : b);
}

hipError_t launch_compact_flags(const CompactRef *refs, int nb, int64_t n, const uint8_t *keep,
                                int64_t ntags, uint32_t *flags, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_compact_flags, dim3(grid_for(n)), dim3(256), 0, s, refs, nb, n, keep,
                       ntags, flags);
    return hipGetLastError();
}

hipError_t scan_exclusive_u32(const uint32_t *in, uint32_t *out, int64_t n, void *temp,
                              size_t &temp_bytes, hipStream_t s) {
    return hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, in, out, (int)n, s);
}

hipError_t launch_compact_bounds(const CompactRef *refs, int nb, int64_t n, const uint32_t *flags,
                                 const uint32_t *pos, uint32_t *kept, hipStream_t s) {
    if (nb <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_compact_bounds, dim3((nb + 255) / 256), dim3(256), 0, s, refs, nb, n,
                       flags, pos, kept);
    return hipGetLastError();
}

hipError_t launch_compact_move(const CompactRef *refs, int nb, int64_t n, const uint32_t *flags,
                               const uint32_t *pos, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_compact_move, dim3(grid_for(n)), dim3(256), 0, s, refs, nb, n, flags,
                       pos);
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
