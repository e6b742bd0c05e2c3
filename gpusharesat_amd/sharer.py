"""Python mirror of GpuShare::GpuClauseSharer (include/gpuShareLib/GpuClauseSharer.h),
through the C ABI of include/gss_sharer.h.

Method names, argument meaning and return values follow the reference
interface (gpuShareLib/GpuClauseSharer.h:74-165 of GpuShareSat), so tests read
like the reference's own: literals are Minisat-encoded ints (2 * var + sign,
sign 1 = negative), ``trySendAssignment`` returns -1 when no slot is free,
``popReportedClause`` returns ``(lits, gpuClauseId)`` or ``None``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib

GLOBAL_STATS = ["gpuClauses", "gpuClauseLengthSum", "gpuClausesAdded", "gpuRuns",
                "clauseTestsOnGroups", "clauseTestsOnAssigs", "totalAssigClauseTested",
                "gpuReduceDbs", "gpuReports", "timeSpentTestingClauses", "timeSpentFillingAssigs",
                "timeSpentFillingReported", "timeSpentReduceGpuDb"]
ONE_SOLVER_STATS = ["varUpdatesSentToGpu", "assigsSentToGpu", "failuresToFindAssig",
                    "reportedClauses", "reportedClausesUnit", "reportedClausesBinary"]


def mkLit(var: int, sign: bool = False) -> int:  # noqa: N802 - reference name
    return 2 * var + int(sign)


class Options(ctypes.Structure):
    """GpuClauseSharerOptions (negative = default)."""
    _fields_ = [("gpu_block_count_guideline", ctypes.c_int),
                ("gpu_threads_per_block_guideline", ctypes.c_int),
                ("min_gpu_latency_micros", ctypes.c_int),
                ("verbosity", ctypes.c_int),
                ("clause_activity_decay", ctypes.c_double),
                ("quick_prof", ctypes.c_int),
                ("init_report_count_per_category", ctypes.c_int),
                ("max_page_locked_memory", ctypes.c_int)]


LOG_FN = ctypes.CFUNCTYPE(None, ctypes.c_char_p, ctypes.c_void_p)


def default_options(**overrides) -> Options:
    o = Options()
    _lib.call("gss_sharer_default_options", ctypes.byref(o))
    for k, v in overrides.items():
        setattr(o, k, v)
    return o


def _ints(lits) -> tuple:
    a = np.ascontiguousarray(np.asarray(lits, dtype=np.int32).ravel())
    return a, a.ctypes.data, a.size


class GpuClauseSharer:
    def __init__(self, options: Options | None = None, log=None):
        self._h = ctypes.c_void_p()
        self.logs: list[str] = []

        def _log(msg, _user):
            text = msg.decode(errors="replace")
            self.logs.append(text)
            if log is not None:
                log(text)

        self._log_cb = LOG_FN(_log)
        opts = options if options is not None else default_options()
        _lib.call("gss_sharer_create", ctypes.byref(opts), self._log_cb, None,
                  ctypes.byref(self._h))
        slots, maxlen = ctypes.c_int(), ctypes.c_int()
        _lib.call("gss_sharer_limits", self._h, ctypes.byref(slots), ctypes.byref(maxlen))
        self.slots_per_solver = slots.value
        self.max_clause_size = maxlen.value
        self._var_count = 0

    # lifetime --------------------------------------------------------------
    def close(self) -> None:
        if self._h:
            _lib.call("gss_sharer_destroy", self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # GpuClauseSharer.h:83-122 -----------------------------------------------
    def gpuRun(self) -> None:  # noqa: N802
        _lib.call("gss_sharer_gpu_run", self._h)

    def reduceDb(self) -> None:  # noqa: N802
        _lib.call("gss_sharer_reduce_db", self._h)

    def getAddedClauseCount(self) -> int:  # noqa: N802
        v = ctypes.c_int64()
        _lib.call("gss_sharer_get_added_clause_count", self._h, ctypes.byref(v))
        return v.value

    def getAddedClauseCountAtLastReduceDb(self) -> int:  # noqa: N802
        v = ctypes.c_int64()
        _lib.call("gss_sharer_get_added_clause_count_at_last_reduce_db", self._h, ctypes.byref(v))
        return v.value

    def hasRunOutOfGpuMemoryOnce(self) -> bool:  # noqa: N802
        v = ctypes.c_int()
        _lib.call("gss_sharer_has_run_out_of_gpu_memory_once", self._h, ctypes.byref(v))
        return bool(v.value)

    def getGpuMemInfo(self) -> tuple[int, int]:  # noqa: N802
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        _lib.call("gss_sharer_get_gpu_mem_info", self._h, ctypes.byref(f), ctypes.byref(t))
        return f.value, t.value

    def getGlobalStatCount(self) -> int:  # noqa: N802
        v = ctypes.c_int()
        _lib.call("gss_sharer_get_global_stat_count", self._h, ctypes.byref(v))
        return v.value

    def getGlobalStat(self, stat) -> int:  # noqa: N802
        idx = GLOBAL_STATS.index(stat) if isinstance(stat, str) else int(stat)
        v = ctypes.c_int64()
        _lib.call("gss_sharer_get_global_stat", self._h, idx, ctypes.byref(v))
        return v.value

    def writeClausesInCnf(self, path: str) -> None:  # noqa: N802
        _lib.call("gss_sharer_write_clauses_in_cnf", self._h, path.encode())

    def setVarCount(self, count: int) -> None:  # noqa: N802
        self._var_count = count
        _lib.call("gss_sharer_set_var_count", self._h, count)

    def addClause(self, solverId: int, lits) -> int:  # noqa: N802,N803
        keep, p, n = _ints(lits)
        v = ctypes.c_int64()
        _lib.call("gss_sharer_add_clause", self._h, solverId, p, n, ctypes.byref(v))
        return v.value

    def setCpuSolverCount(self, count: int) -> None:  # noqa: N802
        _lib.call("gss_sharer_set_cpu_solver_count", self._h, count)

    def getOneSolverStatName(self, stat: int) -> str:  # noqa: N802
        v = ctypes.c_char_p()
        _lib.call("gss_sharer_get_one_solver_stat_name", self._h, stat, ctypes.byref(v))
        return v.value.decode()

    def getGlobalStatName(self, stat: int) -> str:  # noqa: N802
        v = ctypes.c_char_p()
        _lib.call("gss_sharer_get_global_stat_name", self._h, stat, ctypes.byref(v))
        return v.value.decode()

    def getOneSolverStatCount(self) -> int:  # noqa: N802
        v = ctypes.c_int()
        _lib.call("gss_sharer_get_one_solver_stat_count", self._h, ctypes.byref(v))
        return v.value

    # GpuClauseSharer.h:129-159 ----------------------------------------------
    def trySetSolverValues(self, solverId: int, lits) -> bool:  # noqa: N802,N803
        keep, p, n = _ints(lits)
        ok = ctypes.c_int()
        _lib.call("gss_sharer_try_set_solver_values", self._h, solverId, p, n, ctypes.byref(ok))
        return bool(ok.value)

    def unsetSolverValues(self, solverId: int, lits) -> None:  # noqa: N802,N803
        keep, p, n = _ints(lits)
        _lib.call("gss_sharer_unset_solver_values", self._h, solverId, p, n)

    def trySendAssignment(self, solverId: int) -> int:  # noqa: N802,N803
        v = ctypes.c_int64()
        _lib.call("gss_sharer_try_send_assignment", self._h, solverId, ctypes.byref(v))
        return v.value

    def popReportedClause(self, solverId: int):  # noqa: N802,N803
        lits = ctypes.POINTER(ctypes.c_int)()
        count, ok = ctypes.c_int(), ctypes.c_int()
        cid = ctypes.c_int64()
        _lib.call("gss_sharer_pop_reported_clause", self._h, solverId, ctypes.byref(lits),
                  ctypes.byref(count), ctypes.byref(cid), ctypes.byref(ok))
        if not ok.value:
            return None
        return [lits[i] for i in range(count.value)], cid.value

    def popAll(self, solverId: int) -> list:  # noqa: N802,N803
        out = []
        while (r := self.popReportedClause(solverId)) is not None:
            out.append(r)
        return out

    def getLastAssigAllReported(self, solverId: int) -> int:  # noqa: N802,N803
        v = ctypes.c_int64()
        _lib.call("gss_sharer_get_last_assig_all_reported", self._h, solverId, ctypes.byref(v))
        return v.value

    def getCurrentAssignment(self, solverId: int, var_count: int | None = None):  # noqa: N802,N803
        n = var_count if var_count is not None else self._var_count
        out = np.full(max(n, 1), 255, dtype=np.uint8)
        _lib.call("gss_sharer_get_current_assignment", self._h, solverId, out.ctypes.data)
        return out[:n]

    def getOneSolverStat(self, solverId: int, stat) -> int:  # noqa: N802,N803
        idx = ONE_SOLVER_STATS.index(stat) if isinstance(stat, str) else int(stat)
        v = ctypes.c_int64()
        _lib.call("gss_sharer_get_one_solver_stat", self._h, solverId, idx, ctypes.byref(v))
        return v.value

    # not in the reference interface ------------------------------------------
    def debug_clause(self, gpu_clause_id: int):
        a, l, f = ctypes.c_float(), ctypes.c_int(), ctypes.c_int()
        _lib.call("gss_sharer_debug_clause", self._h, gpu_clause_id, ctypes.byref(a),
                  ctypes.byref(l), ctypes.byref(f))
        return (a.value, l.value) if f.value else None

    def execute(self) -> None:
        """Two runs, so that everything sent so far has been reported
        (glucose-syrup/testUtils/TestHelper.cu:55-59)."""
        self.gpuRun()
        self.gpuRun()


@dataclass
class Snapshot:
    """A solver's var -> {0 true, 1 false, 2 undef} state, for test models."""
    values: np.ndarray
