// Device-side data structures and kernel launchers shared by the engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gss.h"

namespace gss {

// Hit records are claimed by each wave in chunks of this many records.
constexpr int kHitChunk = 1024;

// One non-empty length bucket, as seen by the sweep kernel.
struct BucketDesc {
    const uint32_t *lits;  // row-major: literal i of row r at lits[r * len + i], stored as
                           // row offsets (literal << 7)
    uint32_t len;
    uint32_t count;        // rows
    uint32_t task_begin;   // first task index of this bucket
    uint32_t pad;
};

struct SweepArgs {
    const BucketDesc *buckets;
    int nbuckets;
    int ntasks;
    int task_clauses;        // clauses per task (multiple of clauses per wave pass)
    const uint64_t *table;   // [2 * vars + 2][words]: row l = slots where literal l is not
                             // false; rows 2*vars (zeros), 2*vars+1 (ones) are sentinels
    const uint64_t *valid;   // [words]
    gss_hit *hits;           // chunk_cap * kHitChunk records
    uint32_t *fills;         // [chunk_cap] records used in each claimed chunk
    uint32_t *alloc;         // number of chunks claimed (may exceed chunk_cap)
    uint32_t chunk_cap;
    int fuse_len;            // clauses of at most this many literals use the fused schedule
    uint32_t sentinel;       // stored row offset of the sentinel literal
};

struct ScatterSeg {
    uint32_t *dst;
    uint64_t src;    // offset in the staging buffer (uint32 units)
    uint64_t count;  // uint32 elements
};

// words is the table row width (power of two, 1..64); unroll selects the
// number of rows in flight per lane for 16 words (0 = default).
hipError_t launch_sweep(const SweepArgs &a, int words, int unroll, int blocks, hipStream_t s);
int sweep_clauses_per_pass(int words);
// Max resident 256-thread blocks per CU for the sweep kernel variant.
int sweep_occupancy(int words, int unroll);

hipError_t launch_reset_u32(uint32_t *p, int n, hipStream_t s);
hipError_t launch_scatter(const ScatterSeg *segs, int nseg, const uint32_t *src, int shift,
                          hipStream_t s);
hipError_t launch_assig_apply(const gss_var_update *ups, int64_t n, uint64_t *table, int words,
                              hipStream_t s);
hipError_t launch_assig_roll(const gss_var_update *ups, int64_t n, uint64_t *table, int words,
                             const int8_t *last_pos, hipStream_t s);
hipError_t launch_fill_u64(uint64_t *p, int64_t n, uint64_t v, hipStream_t s);

}  // namespace gss
