// extern "C" entry points of include/gss_sharer.h: thin wrappers that catch
// every exception and turn it into a status + gss_last_error() message.
#include <cstdio>
#include <exception>
#include <new>
#include <string>

#include "errors.h"
#include "gss_sharer.h"
#include "sharer_impl.h"

using GpuShare::EngineClauseSharer;

struct gss_sharer {
    EngineClauseSharer *impl = nullptr;
    gss_log_fn log = nullptr;
    void *user = nullptr;
};

namespace {

template <typename F>
int guard(F &&f) {
    try {
        f();
        return GSS_OK;
    } catch (const std::bad_alloc &) {
        return gss::record_error(GSS_E_NOMEM, "host allocation failed");
    } catch (const std::invalid_argument &e) {
        return gss::record_error(GSS_E_INVAL, e.what());
    } catch (const std::exception &e) {
        return gss::record_error(GSS_E_STATE, e.what());
    } catch (...) {
        return gss::record_error(GSS_E_STATE, "unknown error");
    }
}

int bad(const char *what) { return gss::record_error(GSS_E_INVAL, what); }

bool solver_ok(gss_sharer *s, int solver) {
    return s && s->impl && solver >= 0 && solver < s->impl->solver_count();
}

#define NEED(cond, msg) \
    do {                \
        if (!(cond)) return bad(msg); \
    } while (0)

}  // namespace

extern "C" {

int gss_sharer_default_options(gss_sharer_options *o) {
    NEED(o, "null options");
    GpuShare::GpuClauseSharerOptions d;
    o->gpu_block_count_guideline = d.gpuBlockCountGuideline;
    o->gpu_threads_per_block_guideline = d.gpuThreadsPerBlockGuideline;
    o->min_gpu_latency_micros = d.minGpuLatencyMicros;
    o->verbosity = d.verbosity;
    o->clause_activity_decay = d.clauseActivityDecay;
    o->quick_prof = d.quickProf ? 1 : 0;
    o->init_report_count_per_category = d.initReportCountPerCategory;
    o->max_page_locked_memory = d.maxPageLockedMemory;
    return GSS_OK;
}

int gss_sharer_create(const gss_sharer_options *o, gss_log_fn log, void *user, gss_sharer **out) {
    NEED(out, "null out");
    *out = nullptr;
    GpuShare::GpuClauseSharerOptions opts;
    if (o) {
        opts.gpuBlockCountGuideline = o->gpu_block_count_guideline;
        opts.gpuThreadsPerBlockGuideline = o->gpu_threads_per_block_guideline;
        opts.minGpuLatencyMicros = o->min_gpu_latency_micros;
        opts.verbosity = o->verbosity;
        opts.clauseActivityDecay = o->clause_activity_decay;
        opts.quickProf = o->quick_prof != 0;
        opts.initReportCountPerCategory = o->init_report_count_per_category;
        opts.maxPageLockedMemory = o->max_page_locked_memory;
    }
    gss_sharer *s = new (std::nothrow) gss_sharer();
    if (!s) return gss::record_error(GSS_E_NOMEM, "host allocation failed");
    s->log = log;
    s->user = user;
    int rc = guard([&] {
        s->impl = new EngineClauseSharer(opts, [s](const std::string &m) {
            if (s->log) s->log(m.c_str(), s->user);
        });
    });
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return GSS_OK;
}

int gss_sharer_destroy(gss_sharer *s) {
    if (!s) return GSS_OK;
    int rc = guard([&] { delete s->impl; });
    delete s;
    return rc;
}

int gss_sharer_gpu_run(gss_sharer *s) {
    NEED(s, "null sharer");
    return guard([&] { s->impl->gpuRun(); });
}

int gss_sharer_reduce_db(gss_sharer *s) {
    NEED(s, "null sharer");
    return guard([&] { s->impl->reduceDb(); });
}

int gss_sharer_get_added_clause_count(gss_sharer *s, int64_t *out) {
    NEED(s && out, "null argument");
    return guard([&] { *out = s->impl->getAddedClauseCount(); });
}

int gss_sharer_get_added_clause_count_at_last_reduce_db(gss_sharer *s, int64_t *out) {
    NEED(s && out, "null argument");
    return guard([&] { *out = s->impl->getAddedClauseCountAtLastReduceDb(); });
}

int gss_sharer_has_run_out_of_gpu_memory_once(gss_sharer *s, int *out) {
    NEED(s && out, "null argument");
    return guard([&] { *out = s->impl->hasRunOutOfGpuMemoryOnce() ? 1 : 0; });
}

int gss_sharer_get_gpu_mem_info(gss_sharer *s, size_t *free_bytes, size_t *total_bytes) {
    NEED(s && free_bytes && total_bytes, "null argument");
    return guard([&] { s->impl->getGpuMemInfo(*free_bytes, *total_bytes); });
}

int gss_sharer_get_global_stat_count(gss_sharer *s, int *out) {
    NEED(s && out, "null argument");
    return guard([&] { *out = s->impl->getGlobalStatCount(); });
}

int gss_sharer_get_global_stat(gss_sharer *s, int stat, int64_t *out) {
    NEED(s && out, "null argument");
    NEED(stat >= 0 && stat < s->impl->getGlobalStatCount(), "stat out of range");
    return guard([&] { *out = s->impl->getGlobalStat((GpuShare::GlobalStats)stat); });
}

int gss_sharer_write_clauses_in_cnf(gss_sharer *s, const char *path) {
    NEED(s && path, "null argument");
    const bool out = path[0] == '-' && path[1] == 0;
    FILE *f = out ? stdout : std::fopen(path, "w");
    if (!f) return gss::record_error(GSS_E_INVAL, (std::string("cannot open ") + path).c_str());
    int rc = guard([&] { s->impl->writeClausesInCnf(f); });
    if (out) std::fflush(f);
    else std::fclose(f);
    return rc;
}

int gss_sharer_set_var_count(gss_sharer *s, int count) {
    NEED(s && count >= 0, "bad argument");
    return guard([&] { s->impl->setVarCount(count); });
}

int gss_sharer_add_clause(gss_sharer *s, int solver_id, const int *lits, int count,
                          int64_t *out_id) {
    NEED(s && out_id, "null argument");
    NEED(solver_id == -1 || solver_ok(s, solver_id), "solver id out of range");
    NEED(count >= 0 && (count == 0 || lits), "bad clause");
    return guard([&] { *out_id = s->impl->addClause(solver_id, const_cast<int *>(lits), count); });
}

int gss_sharer_set_cpu_solver_count(gss_sharer *s, int count) {
    NEED(s, "null sharer");
    return guard([&] { s->impl->setCpuSolverCount(count); });
}

int gss_sharer_get_one_solver_stat_name(gss_sharer *s, int stat, const char **out) {
    NEED(s && out, "null argument");
    NEED(stat >= 0 && stat < s->impl->getOneSolverStatCount(), "stat out of range");
    return guard([&] { *out = s->impl->getOneSolverStatName((GpuShare::OneSolverStats)stat); });
}

int gss_sharer_get_global_stat_name(gss_sharer *s, int stat, const char **out) {
    NEED(s && out, "null argument");
    NEED(stat >= 0 && stat < s->impl->getGlobalStatCount(), "stat out of range");
    return guard([&] { *out = s->impl->getGlobalStatName((GpuShare::GlobalStats)stat); });
}

int gss_sharer_get_one_solver_stat_count(gss_sharer *s, int *out) {
    NEED(s && out, "null argument");
    return guard([&] { *out = s->impl->getOneSolverStatCount(); });
}

int gss_sharer_try_set_solver_values(gss_sharer *s, int solver_id, const int *lits, int count,
                                     int *ok) {
    NEED(solver_ok(s, solver_id) && ok && count >= 0 && (count == 0 || lits), "bad argument");
    for (int i = 0; i < count; i++)
        NEED(lits[i] >= 0 && (unsigned)(lits[i] >> 1) < GSS_MAX_VARS, "literal out of range");
    return guard([&] {
        *ok = s->impl->trySetSolverValues(solver_id, const_cast<int *>(lits), count) ? 1 : 0;
    });
}

int gss_sharer_unset_solver_values(gss_sharer *s, int solver_id, const int *lits, int count) {
    NEED(solver_ok(s, solver_id) && count >= 0 && (count == 0 || lits), "bad argument");
    for (int i = 0; i < count; i++)
        NEED(lits[i] >= 0 && (unsigned)(lits[i] >> 1) < GSS_MAX_VARS, "literal out of range");
    return guard([&] { s->impl->unsetSolverValues(solver_id, const_cast<int *>(lits), count); });
}

int gss_sharer_try_send_assignment(gss_sharer *s, int solver_id, int64_t *out_id) {
    NEED(solver_ok(s, solver_id) && out_id, "bad argument");
    return guard([&] { *out_id = s->impl->trySendAssignment(solver_id); });
}

int gss_sharer_pop_reported_clause(gss_sharer *s, int solver_id, const int **lits, int *count,
                                   int64_t *gpu_clause_id, int *ok) {
    NEED(solver_ok(s, solver_id) && lits && count && gpu_clause_id && ok, "bad argument");
    return guard([&] {
        int *l = nullptr;
        int c = 0;
        long id = 0;
        const bool got = s->impl->popReportedClause(solver_id, l, c, id);
        *ok = got ? 1 : 0;
        *lits = got ? l : nullptr;
        *count = got ? c : 0;
        *gpu_clause_id = got ? id : -1;
    });
}

int gss_sharer_get_last_assig_all_reported(gss_sharer *s, int solver_id, int64_t *out) {
    NEED(solver_ok(s, solver_id) && out, "bad argument");
    return guard([&] { *out = s->impl->getLastAssigAllReported(solver_id); });
}

int gss_sharer_get_current_assignment(gss_sharer *s, int solver_id, uint8_t *assig) {
    NEED(solver_ok(s, solver_id) && assig, "bad argument");
    return guard([&] { s->impl->getCurrentAssignment(solver_id, assig); });
}

int gss_sharer_get_one_solver_stat(gss_sharer *s, int solver_id, int stat, int64_t *out) {
    NEED(solver_ok(s, solver_id) && out, "bad argument");
    NEED(stat >= 0 && stat < s->impl->getOneSolverStatCount(), "stat out of range");
    return guard([&] {
        *out = s->impl->getOneSolverStat(solver_id, (GpuShare::OneSolverStats)stat);
    });
}

int gss_sharer_debug_clause(gss_sharer *s, int64_t gpu_clause_id, float *activity, int *lbd,
                            int *found) {
    NEED(s && activity && lbd && found, "null argument");
    return guard([&] {
        *found = s->impl->debug_clause((long)gpu_clause_id, *activity, *lbd) ? 1 : 0;
    });
}

int gss_sharer_limits(gss_sharer *s, int *slots_per_solver, int *max_clause_size) {
    NEED(s, "null sharer");
    if (slots_per_solver) *slots_per_solver = s->impl->slots_per_solver();
    if (max_clause_size) *max_clause_size = s->impl->max_clause_size();
    return GSS_OK;
}

}  // extern "C"
