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
    const int32_t *dyn;      // if set, {nbuckets, ntasks} are read here (graph replays)
};

// Per-step inputs of a batched step (gss_step_*), written by the host into a
// pinned staging slot and copied to the device by k_stage_in.
struct StepHeader {
    int32_t nbuckets;
    int32_t ntasks;
    int32_t nsegs;
    int32_t pad0;
    int64_t nlits;
    int64_t nups;
    uint64_t valid[GSS_MAX_WORDS];
    int8_t last_pos[GSS_MAX_WORDS];
};

struct StageArgs {
    const StepHeader *h_hdr;  // pinned host
    const BucketDesc *h_desc;
    const void *h_segs;
    const uint32_t *h_lits;
    const gss_var_update *h_ups;
    StepHeader *d_hdr;        // device
    BucketDesc *d_desc;
    void *d_segs;
    uint32_t *d_lits;
    gss_var_update *d_ups;
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

// Fixed-shape launches of a batched step (counts read from the device header,
// so a captured graph can be replayed with new data).
hipError_t launch_stage_in(const StageArgs &a, hipStream_t s);
hipError_t launch_scatter_dyn(const ScatterSeg *segs, const int32_t *nseg, const uint32_t *src,
                              int shift, hipStream_t s);
hipError_t launch_assig_apply_dyn(const gss_var_update *ups, const int64_t *n, uint64_t *table,
                                  int words, hipStream_t s);
hipError_t launch_assig_roll_dyn(const gss_var_update *ups, const int64_t *n, uint64_t *table,
                                 int words, const int8_t *last_pos, hipStream_t s);
hipError_t launch_stage_out(const uint32_t *alloc, const uint32_t *fills, uint32_t cap,
                            uint32_t *h_out, hipStream_t s);

}  // namespace gss
