/*
 * gss.h -- C ABI of the MI355X clause-vs-assignment check engine.
 *
 * This is the thin, torch-free device-module layer under the C++
 * GpuShare::GpuClauseSharer implementation (include/gpuShareLib/GpuClauseSharer.h).
 * Every entry point takes plain pointers and sizes, returns an int status
 * (0 = ok, negative = error, see GSS_E_*), and never throws.  The message of the
 * last error on the calling thread is available from gss_last_error().
 *
 * Reference interfaces each group replaces (GpuShareSat tree, file:line):
 *   engine / stream     gpuShareLib/GpuUtils.cuh:33-67 (StreamPointer/EventPointer),
 *                       gpuShareLib/GpuClauseSharerImpl.cu:40-65 (device set-up)
 *   clause DB           gpuShareLib/Clauses.cu:128-145,200-217,318-356
 *                       (PerSizeKeeper / DClauses::update / initClauses),
 *                       Clauses.cu:249-282 (full bucket re-upload after reduceDb)
 *   assignment table    gpuShareLib/Assigs.cu:100-116 (dUpdateAssigs),
 *                       Assigs.cu:127-141 (dSetAllAssigsToLast),
 *                       Assigs.cu:326-372 (fillAssigsAsync)
 *   sweep               gpuShareLib/GpuRunner.cu:34-177 (ReportComputer,
 *                       dFindClauses), GpuRunner.cu:323-345 (launchGpu)
 *   hits                gpuShareLib/Reporter.cuh:41-58,77-127 (DReporter /
 *                       Reporter::getCopiedToHost), GpuRunner.cu:347-383
 *
 * Data model
 *   literal    uint32 2*var + sign (sign 1 = negative), as the reference's Lit.
 *   clause DB  clauses live in length buckets; a clause is addressed by
 *              (len, row) where row is its position in bucket `len`.
 *   assignment table
 *              `words` uint64 words per literal; bit b of word w is assignment
 *              slot 64*w + b.  Internally the row of literal l holds the mask of
 *              slots where l is NOT false, and the DB stores each literal as
 *              its row offset; gss_assig_* and gss_db_read take and give the
 *              reference's per-variable isDef/isTrue view and plain literals.
 *              At most 2^24 - 1 variables (32-bit row offsets).
 *   hit        one record per (clause, word) whose false or unit mask is
 *              nonzero; nothing is ever dropped.
 */
#ifndef GSS_H
#define GSS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSS_OK 0
#define GSS_E_INVAL (-1)    /* invalid argument */
#define GSS_E_NOMEM (-2)    /* device or host allocation failed / over budget */
#define GSS_E_HIP (-3)      /* HIP runtime error */
#define GSS_E_STATE (-4)    /* call not valid in the current state */

#define GSS_MAX_WORDS 64    /* max uint64 words per literal row (4096 slots) */
#define GSS_MAX_CLAUSE_LEN 65535
/* variables 0 .. GSS_MAX_VARS - 1 (fewer for tables wider than 16 words,
 * whose size is capped at 4 GiB) */
#define GSS_MAX_VARS ((1u << 24) - 1)

typedef struct gss_engine gss_engine;

typedef struct gss_hit {
    uint32_t row;        /* row within the length bucket */
    uint16_t len;        /* clause length = bucket */
    uint16_t word;       /* assignment word */
    uint64_t false_mask; /* slots where every literal is false */
    uint64_t unit_mask;  /* slots where exactly one literal is undefined, rest false */
} gss_hit;

/* The first word of a gss_hit: which clause and which assignment word (the
 * solver), without the masks.  A batched step can record hits in this form
 * (GSS_RECORDS_IDS) for delivery: the reference's host side never reads a
 * report's slot mask either (Reported.cu:190-204 folds it into a field that is
 * never read), so a third of the bytes carries everything a solver receives. */
typedef struct gss_hit_id {
    uint32_t row;
    uint16_t len;
    uint16_t word;
} gss_hit_id;

/* One record per clause for delivery (GSS_RECORDS_CLAUSES): the clause and
 * the set of assignment words in which it has a report.  In the drop-in
 * interface word s holds solver s's slots, so `words` is the set of solvers
 * that receive the clause; a clause reported in many words costs 16 bytes
 * instead of 8 per word. */
typedef struct gss_hit_clause {
    uint32_t row;
    uint16_t len;
    uint16_t pad;   /* 0 */
    uint64_t words; /* bit w: a report in word w */
} gss_hit_clause;

typedef struct gss_var_update {
    uint32_t var;
    uint32_t word;
    uint64_t is_def;     /* slots where the variable is assigned */
    uint64_t is_true;    /* slots where it is true (ignored where !is_def) */
} gss_var_update;

/* Slots in mask take (is_def, is_true); the word's other slots are kept. */
typedef struct gss_var_patch {
    uint32_t var;
    uint32_t word;
    uint64_t mask;
    uint64_t is_def;
    uint64_t is_true;
} gss_var_patch;

/* Variable var of solver base `base` becomes (is_def, is_true). */
typedef struct gss_base_update {
    uint32_t var;
    uint16_t base;
    uint8_t is_def;
    uint8_t is_true;
} gss_base_update;

/* Every slot of table word `word` takes base `base`'s value, for every
 * variable. */
typedef struct gss_word_load {
    uint16_t word;
    uint16_t base;
} gss_word_load;

/* Last error message of the calling thread ("" if none). */
const char *gss_last_error(void);

/* ---- engine ---------------------------------------------------------------
 * Replaces the reference's device set-up + single stream
 * (GpuClauseSharerImpl.cu:40-65, GpuClauseSharerImpl.cuh:34). */
int gss_engine_create(int device, gss_engine **out);
int gss_engine_destroy(gss_engine *e);
/* Number of compute units of the engine's device. */
int gss_engine_cu_count(gss_engine *e, int *out);
/* Blocks until all work queued on the engine's stream is done. */
int gss_engine_sync(gss_engine *e);
/* Host buffers that are only copy endpoints (the plain path's staging, the
 * steps' records on the host) are allocated pageable instead of page-locked
 * when one allocation reaches `bytes`, or when page-locking fails; bytes < 0
 * (the default): no limit.  Replaces maxPageLockedMemory
 * (GpuClauseSharer.h:47, gpuShareLib/CorrespArr.cu:212-235).  Staging that the
 * step's kernels read in place stays page-locked. */
int gss_engine_set_page_locked_limit(gss_engine *e, int64_t bytes);
/* Free and total device memory (replaces cudaMemGetInfo in
 * GpuClauseSharerImpl::getGpuMemInfo, GpuClauseSharerImpl.cu:205-207). */
int gss_engine_mem_info(gss_engine *e, size_t *free_bytes, size_t *total_bytes);

/* ---- clause DB -------------------------------------------------------------
 * Append n clauses (lens[i] literals each, concatenated in lits).  Rows are
 * assigned in order within each length bucket and returned in out_rows (may be
 * NULL).  Replaces HostClauses::getUpdatesForDevice + initClauses
 * (Clauses.cu:318-346, GpuRunner.cu:64-66,314-316): one H2D copy of the packed
 * batch and one scatter kernel. */
int gss_db_append(gss_engine *e, int64_t n, const uint32_t *lits,
                  const int32_t *lens, uint32_t *out_rows);
/* Replace the whole content of bucket `len` by n rows (n*len literals);
 * replaces the full re-upload after compaction (Clauses.cu:249-282). */
int gss_db_set_bucket(gss_engine *e, int len, int64_t n, const uint32_t *lits);
/* Number of rows in bucket `len`. */
int gss_db_bucket_count(gss_engine *e, int len, int64_t *out);
/* Totals over all buckets. */
int gss_db_totals(gss_engine *e, int64_t *clauses, int64_t *literals);
/* Copy rows [row0, row0+n) of bucket len to host (n*len literals). */
int gss_db_read(gss_engine *e, int len, int64_t row0, int64_t n, uint32_t *out);
/* Remove every clause. */
int gss_db_clear(gss_engine *e);
/* Tagged DB: every row carries a caller uint32 tag (a stable clause id) that
 * moves with it.  Enable on an empty DB; then append with
 * gss_db_append_tagged (gss_db_set_bucket and clause-carrying steps are
 * refused).  Tags stand in for the reference's GpuClauseId / ClMetadata
 * association (Clauses.cu:318-356, Clauses.cuh:120-160). */
int gss_db_enable_tags(gss_engine *e);
int gss_db_append_tagged(gss_engine *e, int64_t n, const uint32_t *lits,
                         const int32_t *lens, const uint32_t *tags, uint32_t *out_rows);
/* Copy the tags of rows [row0, row0+n) of bucket len to host. */
int gss_db_read_tags(gss_engine *e, int len, int64_t row0, int64_t n, uint32_t *out);
/* Keep the rows whose tag t has keep[t] != 0 (host array of ntags flags; tags
 * >= ntags are kept), in order, compacting every bucket on the device in one
 * scan.  Replaces the host-side compaction and full re-upload of
 * HostClauses::tryRemoveClausesAsync (Clauses.cu:249-282). */
int gss_db_compact(gss_engine *e, const uint8_t *keep, int64_t ntags);

/* ---- assignment table ------------------------------------------------------
 * (Re)shape the table to var_count variables x words words; every slot of
 * every variable becomes undefined.  words must be in [1, GSS_MAX_WORDS]. */
int gss_assig_reset(gss_engine *e, int64_t var_count, int words);
/* Grow to var_count variables keeping existing rows (new ones undefined). */
int gss_assig_grow_vars(gss_engine *e, int64_t var_count);
/* Upload whole planes, var-major: is_def[var*words + w], is_true[...]. */
int gss_assig_upload(gss_engine *e, const uint64_t *is_def, const uint64_t *is_true);
/* Download whole planes (same layout).  Synchronous. */
int gss_assig_download(gss_engine *e, uint64_t *is_def, uint64_t *is_true);
/* Apply n per-(var, word) updates (replaces dUpdateAssigs, Assigs.cu:100-116).
 * At most one update per (var, word) per call.  Asynchronous on the engine's
 * stream (the updates are copied before the call returns). */
int gss_assig_apply(gss_engine *e, int64_t n, const gss_var_update *ups);
/* Roll forward the (var, word) pairs of the last gss_assig_apply, as
 * gss_assig_roll_forward does, without uploading them again; asynchronous.
 * A run enqueues apply -> sweep -> this, so the roll-forward follows the
 * check as dSetAllAssigsToLast follows dFindClauses (GpuRunner.cu:339-344). */
int gss_assig_roll_forward_applied(gss_engine *e, const int8_t *last_pos);
/* For each (var, word) of ups, set every slot of that word to the value held
 * by slot last_pos[word] (replaces dSetAllAssigsToLast, Assigs.cu:127-141).
 * Words with last_pos < 0 are left alone. */
int gss_assig_roll_forward(gss_engine *e, int64_t n, const gss_var_update *ups,
                           const int8_t *last_pos);

/* ---- sweep -----------------------------------------------------------------
 * Check every clause against every slot set in valid[w] (w < words) and
 * record the hits on the device.  Replaces launchGpu/dFindClauses
 * (GpuRunner.cu:134-177,323-345).  Asynchronous on the engine's stream. */
int gss_sweep_async(gss_engine *e, const uint64_t *valid);
/* Wait for the last sweep; returns its number of hit records.  If the hit
 * buffer was too small the sweep is re-run with a larger one (no drops). */
int gss_sweep_wait(gss_engine *e, int64_t *n_hits);
/* gss_sweep_async + gss_sweep_wait. */
int gss_sweep(gss_engine *e, const uint64_t *valid, int64_t *n_hits);
/* Copy the hit records of the last (waited) sweep to host: up to cap records,
 * *n = total.  Record order is unspecified. */
int gss_hits_copy(gss_engine *e, gss_hit *out, int64_t cap, int64_t *n);
/* No-allocation form of gss_hits_copy (no counterpart in the reference): for
 * callers that cannot allocate device memory -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  Writes the records of
 * the last waited sweep densely into d_out, caller DEVICE memory of cap
 * records (record order unspecified, as gss_hits_copy), and *n = their
 * number; with d_out NULL and cap 0 only *n is set.  Calls no hipMalloc,
 * hipFree or other allocating routine: its only working memory is a static
 * per-device array of 262,145 uint64 (2 MiB, held by every device the library
 * runs on until the process ends) for the chunk offsets, so a sweep that
 * claimed more than 262,144 hit chunks of 1024 records is refused with
 * GSS_E_INVAL, as is cap below the record count.  Synchronous.  Every call on
 * one device shares that array: calls that overlap on one device (from two
 * threads or two engines) race. */
int gss_hits_gather(gss_engine *e, gss_hit *d_out, int64_t cap, int64_t *n);
/* Duration in ms of the sweep kernel of the last waited sweep, measured with
 * HIP events recorded on the engine's stream around the kernel launch. */
int gss_sweep_kernel_ms(gss_engine *e, float *ms);
/* Sum of sweep-kernel durations (ms) and count since the last reset. */
int gss_sweep_kernel_stats(gss_engine *e, double *total_ms, int64_t *count, int reset);

/* Tuning knobs (0 = default).  Exposed for the autotuner and the tests;
 * results never depend on them. */
int gss_set_tuning(gss_engine *e, int key, int value);
#define GSS_TUNE_WAVES_PER_CU 1   /* persistent waves per CU (0 = occupancy) */
#define GSS_TUNE_TASK_CLAUSES 2   /* clauses per task */
#define GSS_TUNE_MAX_HIT_CHUNKS 3 /* cap on the hit buffer a sweep starts with
                                     (0 = sized to the worst case, for steps
                                     by the records seen); a sweep that needs
                                     more re-runs with a larger one.  Steps are
                                     held to it on every submit */
#define GSS_TUNE_UNROLL 4         /* table rows in flight per lane: 1, 2, 4 or 8
                                     at 16 words (default 4), 2 or 4 at 32
                                     (default 4), 1, 2 or 4 at 64 (default 2);
                                     fixed below 16 words */
#define GSS_TUNE_FUSE_LEN 5       /* clauses of at most this many literals read
                                     both rows of a literal in one pass
                                     (0 = default 4; -1 = none, which runs the
                                     kernel built without that pass and its
                                     registers) */
#define GSS_TUNE_SCHEDULE 6       /* wave-to-bucket schedule: tasks of long
                                     clauses first (0 = default, 1), or short
                                     clauses first (2) */

/* Autotuner: times the sweep on the current clause DB and assignment table
 * under a grid of (rows in flight, fuse length, clauses per task, waves per
 * CU, schedule) settings, `reps` sweeps each, and keeps the fastest.  The
 * chosen values are written to best[0..4] in that order (may be NULL) and
 * the best average kernel time to *best_ms.  Replaces the reference's fixed
 * launch guidelines (GpuClauseSharerImpl.cu:50-57). */
int gss_autotune(gss_engine *e, const uint64_t *valid, int reps, int *best, float *best_ms);

/* ---- batched steps ---------------------------------------------------------
 * One step = append n clauses (lits/lens as gss_db_append; their rows to
 * out_rows if non-null) -> apply n_ups updates -> sweep every clause against
 * the slots in valid -> roll forward (last_pos as gss_assig_roll_forward, may
 * be null = none) -> stage the hit counts to pinned host memory.  The device
 * side of a step is one captured HIP graph per ping-pong slot, replayed with
 * new data (the counts are read from device memory), so two steps can be in
 * flight and the host fills one slot while the other runs.  Replaces the
 * per-run sequence of GpuRunner::wholeRun / startGpuRunAsync / launchGpu /
 * gatherGpuRunResults (GpuRunner.cu:211-345,360) with one graph launch.
 * Clause literals must have table rows (grow the table before the step). */
int gss_step_enqueue(gss_engine *e, int64_t n, const uint32_t *lits, const int32_t *lens,
                     uint32_t *out_rows, int64_t n_ups, const gss_var_update *ups,
                     const uint64_t *valid, const int8_t *last_pos);
/* The general step.  In order, on the device: append the clauses, apply the
 * base updates, load the listed words from their bases, apply the full
 * updates, apply the patches, sweep, roll forward, stage the hit counts.
 * Serving uses the base planes: a request's words are loaded from its
 * solver's base and its assignments are patches over it, so no host-side
 * table state is needed per request.  At most one update per (var, word)
 * among ups, one patch per (var, word), one base update per (base, var) and
 * one load per word in a step. */
typedef struct gss_step_desc {
    int64_t n_clauses;
    const uint32_t *lits;
    const int32_t *lens;
    uint32_t *out_rows;                /* may be NULL */
    int64_t n_ups;
    const gss_var_update *ups;
    int64_t n_patches;
    const gss_var_patch *patches;
    int64_t n_base_ups;
    const gss_base_update *base_ups;
    int64_t n_loads;
    const gss_word_load *loads;
    const uint64_t *valid;             /* [words] */
    const int8_t *last_pos;            /* [words] or NULL */
    int32_t records;                   /* GSS_RECORDS_FULL (0): gss_hit records, read with
                                          gss_step_hits_map / gss_step_hits_copy;
                                          GSS_RECORDS_IDS (1): gss_hit_id records, read
                                          with gss_step_ids_map;
                                          GSS_RECORDS_CLAUSES (2): gss_hit_clause records,
                                          read with gss_step_clauses_map */
    int32_t flags;                     /* 0, or GSS_STEP_SCREEN */
} gss_step_desc;
#define GSS_RECORDS_FULL 0
#define GSS_RECORDS_IDS 1
#define GSS_RECORDS_CLAUSES 2
/* Screens every clause, one per lane, against an aggregate of the table
 * (per literal, a bit per group of words: non-false in every slot of the
 * group) and runs the exact check only on the clauses and word groups the
 * screening leaves.  The records are the same with or without it, in any
 * record form.  The aggregate is read from the whole table when the table
 * changed otherwise since the last screened step (and on that step's word
 * loads); else only the entries of the step's updates and patches are
 * refreshed, after they are applied and again after the roll forward.  It
 * pays where the slots of a word group share most of their values
 * (consecutive assignments of one solver). */
#define GSS_STEP_SCREEN 1
int gss_step_submit(gss_engine *e, const gss_step_desc *d);
/* (Re)creates nbases solver base planes over the current variable count,
 * every variable undefined (2 bits per variable per base).  A table that
 * grows past a multiple of 64 variables needs a new reset before base
 * updates or loads.  The reference keeps one MultiLBool array per solver
 * (Assigs.cuh:39-58); here a base is the solver's last assignment, packed. */
int gss_bases_reset(gss_engine *e, int nbases);
/* Wait for the oldest step in flight: its number of hit records and the time
 * (ms) between events recorded around its graph launch.  The step's records
 * start travelling to pinned host memory at once, on a stream of the slot's
 * own, while the caller goes on (typically enqueueing the next step); the
 * slot's next step does not overwrite them on the device before that copy is
 * done.
 * A step's hit buffer holds twice the most records a step of the engine has
 * made (at least 4M records, 16M on a serving engine), never more than
 * clauses x valid words.  A step whose records overflow it keeps the table
 * as its sweep read it and halts the engine: the step in flight behind it
 * then stages nothing.  Its wait -- or a call that reads or changes the DB
 * or the table while it is in flight -- grows the buffer, re-runs the sweep
 * and the roll forward, and replays the later step, so no record is lost
 * (GSS_E_NOMEM if the buffer cannot grow: the step's records are lost, the
 * table still rolls forward).  The reference drops the reports past a full
 * buffer (Reporter.cuh:46-48) and doubles it for later runs
 * (GpuRunner.cu:363-365). */
int gss_step_wait(gss_engine *e, int64_t *n_hits, float *step_ms);
/* Counters of the engine's batched steps, readable from any thread (also
 * one other than the thread stepping the engine, e.g. a server's). */
typedef struct gss_step_counters {
    int64_t reruns;          /* steps re-run: their records overflowed the hit buffer */
    int64_t hit_chunks;      /* chunks of 1024 records the step hit buffers hold */
    int64_t host_grows;      /* pinned host record buffers grown by gss_step_wait */
    double host_grow_ms;     /* time those allocations held the waiting thread, in all */
    double host_grow_ms_max; /* the longest of them */
} gss_step_counters;
int gss_step_stats(gss_engine *e, gss_step_counters *out);
/* Copy the hit records of the last waited step to host (up to cap; *n =
 * total), waiting for their transfer if it is still running.  Valid until
 * the next gss_step_wait of that step's slot (two waits later), also when
 * the slot has been enqueued again meanwhile. */
int gss_step_hits_copy(gss_engine *e, gss_hit *out, int64_t cap, int64_t *n);
/* Zero-copy form: *out points at the records of the last waited step in the
 * engine's pinned host memory (*n records), valid until the next
 * gss_step_wait of that step's slot.  Same source as gss_step_hits_copy;
 * replaces Reporter::getCopiedToHost (Reporter.cuh:103). */
int gss_step_hits_map(gss_engine *e, const gss_hit **out, int64_t *n);
/* The same for a step submitted with GSS_RECORDS_IDS: *out points at its
 * gss_hit_id records in pinned host memory, valid as long as
 * gss_step_hits_map's.  Each is the first word of the gss_hit the full form
 * records for the same (clause, word). */
int gss_step_ids_map(gss_engine *e, const gss_hit_id **out, int64_t *n);
/* The same for a step submitted with GSS_RECORDS_CLAUSES: one gss_hit_clause
 * per clause with a report in some word (the words of its gss_hit_id records,
 * as a mask), in no particular order. */
int gss_step_clauses_map(gss_engine *e, const gss_hit_clause **out, int64_t *n);
/* Makes room for n more clauses (lens[i] literals each) in their length
 * buckets without adding any: a step that appends them then needs no bucket
 * growth.  GSS_E_NOMEM when the device (or the DB byte cap) cannot hold them,
 * with nothing else changed.  Stands in for the capacity check the reference
 * makes before a run's clause copy (GpuRunner.cu:243-260). */
int gss_db_reserve(gss_engine *e, int64_t n, const int32_t *lens);

#ifdef __cplusplus
}
#endif

#endif /* GSS_H */
