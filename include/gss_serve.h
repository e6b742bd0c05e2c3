/*
 * gss_serve.h -- C ABI of the serving loop: continuous batching of
 * assignment-set requests from many solvers over one device-resident clause
 * DB, with clause growth admitted under an HBM budget.
 *
 * The reference checks the assignments its solvers sent since the last run,
 * once per gpuRun() (GpuClauseSharerImpl.cu:83-94, GpuRunner.cu:211-261);
 * this loop extends that run loop to requests that arrive at any time.  A
 * request is solver s's base assignment (its packed base plane on the
 * device, kept current by gss_server_update_base) plus k assignments given
 * as patches over it, in relative words 0 .. ceil(k / 64) - 1.  Clients
 * publish requests into a ring of fine-grained pinned host memory; a serving
 * thread of the library keeps up to max_inflight batched steps on the device,
 * and each step admits, on the device when it starts, the published requests
 * whose words fit the table (in arrival order), loads their words from their
 * solvers' bases, applies their patches straight from the ring, and checks
 * every clause against them (the usefulness rule of GpuRunner.cu:34-57).  So
 * requests that arrive while other steps are on the GPU join the next step
 * without waiting for the host.  The thread then hands each request its own
 * records (clause id, word, masks) and its submit-to-report latency.
 *
 * A request's records depend only on its solver's base at its step, its
 * patches and the DB at its step, never on the requests around it.
 *
 * All entry points return GSS_* codes (gss.h); messages via gss_last_error().
 * Any thread may call them; one server drives one engine of its own.
 */
#ifndef GSS_SERVE_H
#define GSS_SERVE_H

#include <stdint.h>

#include "gss.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gss_server gss_server;

typedef struct gss_server_config {
    int64_t var_count;       /* variables of the table */
    int32_t solvers;         /* solver bases (at most 65535) */
    int32_t words;           /* table words, a power of two <= 64: 64 * words slots per step */
    double hbm_budget;       /* bytes of clause literals the DB may hold */
    int32_t max_inflight;    /* steps on the device at once: 1 or 2 */
    int32_t records;         /* GSS_RECORDS_FULL (masks kept) or GSS_RECORDS_IDS */
    int32_t request_ring;    /* requests published and not yet completed, at most */
    int32_t pad;
    int64_t patch_ring;      /* patches published and not yet completed, at most */
} gss_server_config;

/* One hit of a request: clause id (the order of gss_server_add_clauses),
 * word of the request (slot 64 * word + b is its assignment 64 * word + b)
 * and, for GSS_RECORDS_FULL, the slots where the clause is falsified / unit
 * (zero with GSS_RECORDS_IDS). */
typedef struct gss_serve_hit {
    int64_t clause;
    uint32_t word;
    uint32_t pad;
    uint64_t false_mask;
    uint64_t unit_mask;
} gss_serve_hit;

typedef struct gss_serve_result {
    int64_t request;          /* id from gss_server_submit */
    int64_t step;             /* the serving step that checked it (0, 1, ...) */
    int64_t clauses;          /* clauses in the DB at that step: ids [0, clauses) */
    int64_t base_updates;     /* base updates applied before that step (their order of
                                 gss_server_update_base) */
    int32_t solver;
    int32_t k;
    int32_t first_word;       /* table word the request's slots started at in its step */
    int32_t pad;
    double latency_ms;        /* submit -> its records grouped for it on the host */
    int64_t n_records;
    const gss_serve_hit *records; /* NULL until gss_server_records has been called */
} gss_serve_result;

typedef struct gss_server_stats {
    int64_t steps, requests_done, requests_pending, records;
    int64_t clauses, clauses_refused, literals;
    double checks;            /* literal x assignment checks of the steps so far */
    double device_ms;         /* sum of the steps' device times */
    int64_t host_grouped;     /* steps whose records outgrew the device grouping buffer and
                                 were grouped by the serving thread */
    int64_t graph_captures;   /* step graphs captured (a new one whenever buffers grow) */
    double submit_ms_total;   /* serving thread: host time staging and launching steps */
    double submit_ms_max;     /* the longest of them */
    double wait_ms_max;       /* the longest wait for a step to end */
} gss_server_stats;

/* Defaults: 64 words, 2 steps in flight, full records, 4096 requests and
 * 8M patches in the rings, a budget of 120 GB. */
int gss_server_default_config(gss_server_config *c);
/* Creates the server on `device` (its own engine: table of var_count x words,
 * `solvers` bases, every variable undefined) and starts its thread. */
int gss_server_create(int device, const gss_server_config *c, gss_server **out);
/* Stops the thread once the steps in flight are done; requests not yet
 * admitted are dropped. */
int gss_server_destroy(gss_server *s);
/* The server's engine, for tuning (gss_set_tuning) before the first request.
 * Its table is the serving steps' scratch: a step writes the words it loads
 * and sets the rest of their 16-word groups to all ones. */
int gss_server_engine(gss_server *s, gss_engine **out);
/* Base updates (a later one for the same (base, var) wins): the next step the
 * serving thread launches applies them, before it loads its requests' words.
 * A request published after a step was launched may still be admitted by
 * that step, and then sees the bases as of its launch
 * (gss_serve_result.base_updates says how many updates it saw). */
int gss_server_update_base(gss_server *s, int64_t n, const gss_base_update *ups);
/* Clauses (lens[i] literals each, 2 * var + sign): admitted whole if the DB
 * then fits the budget, appended by the next step and checked by it; their
 * ids are *first_id .. *first_id + n - 1, or *first_id = -1 if refused. */
int gss_server_add_clauses(gss_server *s, int64_t n, const uint32_t *lits, const int32_t *lens,
                           int64_t *first_id);
/* A request of k assignments (1 <= k <= 64 * words) of `solver`: slot j is the
 * base with the patches' slots j applied (patch words relative to the request,
 * each word's slots past k ignored).  The patches are copied.  GSS_E_STATE if
 * a ring is full (retry once requests complete). */
int gss_server_submit(gss_server *s, int solver, int k, int64_t n_patches,
                      const gss_var_patch *patches, int64_t *out_id);
/* Up to max ids of completed requests, oldest first; waits up to timeout_us
 * for one if none is ready. */
int gss_server_poll(gss_server *s, int64_t max, int64_t *ids, int64_t *n, int64_t timeout_us);
int gss_server_result(gss_server *s, int64_t id, gss_serve_result *out);
/* The request's records with clause ids (looked up on this call, once), valid
 * until gss_server_release. */
int gss_server_records(gss_server *s, int64_t id, const gss_serve_hit **out, int64_t *n);
int gss_server_release(gss_server *s, int64_t id);
int gss_server_stats_get(gss_server *s, gss_server_stats *out);

#ifdef __cplusplus
}
#endif

#endif /* GSS_SERVE_H */
