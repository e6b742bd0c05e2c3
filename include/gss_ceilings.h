/*
 * gss_ceilings.h -- what the device reaches for the sweep's memory access
 * patterns, measured on the device itself.
 *
 * No counterpart in the reference: these entry points exist to put the
 * sweep's measured traffic beside the device's own ceilings (a plain copy,
 * and uniformly random 128-byte line gathers from a table of a chosen size,
 * the access pattern of the 16-word sweep) instead of the data-sheet peak
 * alone.  They allocate their own buffers, run on a private stream, and are
 * used by `bench.py --workload ceilings`.  Return GSS_* codes (gss.h).
 */
#ifndef GSS_CEILINGS_H
#define GSS_CEILINGS_H

#include <stdint.h>

#include "gss.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gss_device_info {
    char name[64];
    char arch[64];      /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    int32_t cus;        /* compute units */
    int32_t clock_khz;  /* peak engine clock */
    int64_t l2_bytes;   /* L2 of one XCD as reported by the runtime */
    int64_t lds_per_cu; /* LDS bytes per compute unit */
    int64_t hbm_bytes;  /* device memory */
    int32_t mem_clock_khz;
    int32_t mem_bus_bits;
} gss_device_info;

int gss_device_query(int device, gss_device_info *out);

/* Copies `bytes` (a positive multiple of 16) from one device buffer to
 * another with a grid-stride kernel of 16-byte loads and stores, `reps`
 * times after one untimed copy.  *gbps = 2 * bytes / average kernel time
 * (bytes read + written), *ms = average kernel time. */
int gss_ceiling_copy(int device, int64_t bytes, int reps, double *gbps, double *ms);

/* Reads n_reads 128-byte lines at uniformly random line positions of a table
 * of table_bytes (a positive multiple of 128, at most 2^32 lines): each group
 * of 16 lanes reads one line as sixteen 8-byte words, `unroll` (1, 2, 4 or 8)
 * independent lines in flight per lane, as the 16-word sweep reads table
 * rows.  `reps` timed runs after one untimed run.  *gbps = 128 * n_reads /
 * average kernel time, *ms = average kernel time. */
int gss_ceiling_gather(int device, int64_t table_bytes, int64_t n_reads, int unroll, int reps,
                       double *gbps, double *ms);
/* The same with blocks of block_bytes (64, 128 or 256; table_bytes a positive
 * multiple of it, at most 2^32 blocks): block_bytes / 8 lanes read one block
 * at a uniformly random block position.  *gbps = block_bytes * n_reads /
 * average kernel time.  Whether a read of part of a 128-byte line costs the
 * device a whole line shows in the rate of 64-byte reads. */
int gss_ceiling_gather_block(int device, int64_t table_bytes, int block_bytes, int64_t n_reads,
                             int unroll, int reps, double *gbps, double *ms);

/* Copies `bytes` (positive) between a device buffer and pinned host memory
 * (hipHostMalloc) with hipMemcpyAsync, `reps` times each way after one
 * untimed copy: *d2h_gbps and *h2d_gbps are bytes / average copy time.  The
 * floor under any transfer of hit records to the host. */
int gss_ceiling_pcie(int device, int64_t bytes, int reps, double *d2h_gbps, double *h2d_gbps);

#ifdef __cplusplus
}
#endif

#endif /* GSS_CEILINGS_H */
