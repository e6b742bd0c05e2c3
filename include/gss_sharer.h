/*
 * gss_sharer.h -- C ABI over GpuShare::GpuClauseSharer
 * (include/gpuShareLib/GpuClauseSharer.h), for callers that bind C rather than
 * the C++ vtable (ctypes, cgo, JNI, N-API).
 *
 * One entry point per virtual of the reference interface
 * (gpuShareLib/GpuClauseSharer.h:74-165 of GpuShareSat), in the same order,
 * with the same argument meaning.  Results come back through out-parameters;
 * the return value is a status (0 ok, GSS_E_* from gss.h), and the message of
 * a failure is available from gss_last_error().  Solver ids are checked
 * (GSS_E_INVAL) where the C++ interface, like the reference, does not check.
 */
#ifndef GSS_SHARER_H
#define GSS_SHARER_H

#include <stddef.h>
#include <stdint.h>

#include "gss.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gss_sharer gss_sharer;

/* Mirrors GpuClauseSharerOptions (GpuClauseSharer.h:25-59); negative values
 * select the defaults. */
typedef struct gss_sharer_options {
    int gpu_block_count_guideline;
    int gpu_threads_per_block_guideline;
    int min_gpu_latency_micros;
    int verbosity;
    double clause_activity_decay;
    int quick_prof;
    int init_report_count_per_category;
    int max_page_locked_memory;
} gss_sharer_options;

typedef void (*gss_log_fn)(const char *message, void *user);

/* Fills opts with the reference defaults (GpuClauseSharer.h:50-59). */
int gss_sharer_default_options(gss_sharer_options *opts);
/* makeGpuClauseSharerPtr (GpuClauseSharer.h:165).  log may be NULL. */
int gss_sharer_create(const gss_sharer_options *opts, gss_log_fn log, void *log_user,
                      gss_sharer **out);
int gss_sharer_destroy(gss_sharer *s);

/* GpuClauseSharer.h:83-122 */
int gss_sharer_gpu_run(gss_sharer *s);
int gss_sharer_reduce_db(gss_sharer *s);
int gss_sharer_get_added_clause_count(gss_sharer *s, int64_t *out);
int gss_sharer_get_added_clause_count_at_last_reduce_db(gss_sharer *s, int64_t *out);
int gss_sharer_has_run_out_of_gpu_memory_once(gss_sharer *s, int *out);
int gss_sharer_get_gpu_mem_info(gss_sharer *s, size_t *free_bytes, size_t *total_bytes);
int gss_sharer_get_global_stat_count(gss_sharer *s, int *out);
int gss_sharer_get_global_stat(gss_sharer *s, int stat, int64_t *out);
/* writeClausesInCnf(FILE*): takes a path here; "-" writes to stdout. */
int gss_sharer_write_clauses_in_cnf(gss_sharer *s, const char *path);
int gss_sharer_set_var_count(gss_sharer *s, int count);
/* addClause: *out_id = clause id or -1 if rejected. */
int gss_sharer_add_clause(gss_sharer *s, int solver_id, const int *lits, int count,
                          int64_t *out_id);
int gss_sharer_set_cpu_solver_count(gss_sharer *s, int count);
int gss_sharer_get_one_solver_stat_name(gss_sharer *s, int stat, const char **out);
int gss_sharer_get_global_stat_name(gss_sharer *s, int stat, const char **out);
int gss_sharer_get_one_solver_stat_count(gss_sharer *s, int *out);

/* GpuClauseSharer.h:129-159 */
int gss_sharer_try_set_solver_values(gss_sharer *s, int solver_id, const int *lits, int count,
                                     int *ok);
int gss_sharer_unset_solver_values(gss_sharer *s, int solver_id, const int *lits, int count);
/* *out_id = assignment id, or -1 if no slot is free. */
int gss_sharer_try_send_assignment(gss_sharer *s, int solver_id, int64_t *out_id);
/* *ok = 1 and lits/count/id set if a clause was popped; *lits stays valid
 * until the next pop for this solver. */
int gss_sharer_pop_reported_clause(gss_sharer *s, int solver_id, const int **lits, int *count,
                                   int64_t *gpu_clause_id, int *ok);
int gss_sharer_get_last_assig_all_reported(gss_sharer *s, int solver_id, int64_t *out);
/* assig must hold one byte per variable of the sharer's count (the last
 * gss_sharer_set_var_count): 0 true, 1 false, 2 undef; exactly that many
 * bytes are written. */
int gss_sharer_get_current_assignment(gss_sharer *s, int solver_id, uint8_t *assig);
int gss_sharer_get_one_solver_stat(gss_sharer *s, int solver_id, int stat, int64_t *out);

/* Not in the reference interface: assignment slots per solver (64 natively,
 * 32 with GPUSHARESAT_COMPAT=1) and the largest accepted clause length. */
int gss_sharer_limits(gss_sharer *s, int *slots_per_solver, int *max_clause_size);
/* Not in the reference interface: test hook (the reference's tests reach the
 * same state through HostClauses, glucose-syrup/test/ClauseActivityLbdTest.cu).
 * Activity and lbd of the resident clause with this id; *found = 0 if it is
 * not resident (deleted, or still staged before the next gpu run). */
int gss_sharer_debug_clause(gss_sharer *s, int64_t gpu_clause_id, float *activity, int *lbd,
                            int *found);

#ifdef __cplusplus
}
#endif

#endif /* GSS_SHARER_H */
