"""ctypes binding of libgpusharesat.so (the C ABI in include/gss.h and
include/gss_sharer.h).

The library is built in-tree by gpusharesat_amd.build.  Importing this module
without it raises ImportError: there is no CPU fallback for the check path.
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libgpusharesat.so")

c_int = ctypes.c_int
c_i64 = ctypes.c_int64
c_u64 = ctypes.c_uint64
c_p = ctypes.c_void_p
c_float = ctypes.c_float
c_double = ctypes.c_double
c_char_p = ctypes.c_char_p


class GssError(RuntimeError):
    def __init__(self, func: str, code: int, msg: str):
        super().__init__(f"{func} failed ({code}): {msg}")
        self.code = code


# Every symbol exported by the headers under include/ (gss.h, gss_sharer.h,
# gss_residency.h, gss_ceilings.h, gss_serve.h), with its argument types (all return int status unless listed in _RESTYPES).
SIGNATURES: dict[str, list] = {
    "gss_last_error": [],
    "gss_engine_create": [c_int, c_p],
    "gss_engine_destroy": [c_p],
    "gss_engine_cu_count": [c_p, c_p],
    "gss_engine_sync": [c_p],
    "gss_db_append": [c_p, c_i64, c_p, c_p, c_p],
    "gss_db_set_bucket": [c_p, c_int, c_i64, c_p],
    "gss_db_bucket_count": [c_p, c_int, c_p],
    "gss_db_totals": [c_p, c_p, c_p],
    "gss_db_read": [c_p, c_int, c_i64, c_i64, c_p],
    "gss_db_clear": [c_p],
    "gss_db_enable_tags": [c_p],
    "gss_db_append_tagged": [c_p, c_i64, c_p, c_p, c_p, c_p],
    "gss_db_read_tags": [c_p, c_int, c_i64, c_i64, c_p],
    "gss_db_compact": [c_p, c_p, c_i64],
    "gss_assig_reset": [c_p, c_i64, c_int],
    "gss_assig_grow_vars": [c_p, c_i64],
    "gss_assig_upload": [c_p, c_p, c_p],
    "gss_assig_download": [c_p, c_p, c_p],
    "gss_assig_apply": [c_p, c_i64, c_p],
    "gss_assig_roll_forward": [c_p, c_i64, c_p, c_p],
    "gss_assig_roll_forward_applied": [c_p, c_p],
    "gss_engine_mem_info": [c_p, c_p, c_p],
    "gss_engine_set_page_locked_limit": [c_p, c_i64],
    "gss_sweep_async": [c_p, c_p],
    "gss_sweep_wait": [c_p, c_p],
    "gss_sweep": [c_p, c_p, c_p],
    "gss_hits_copy": [c_p, c_p, c_i64, c_p],
    "gss_hits_gather": [c_p, c_p, c_i64, c_p],
    "gss_sweep_kernel_ms": [c_p, c_p],
    "gss_sweep_kernel_stats": [c_p, c_p, c_p, c_int],
    "gss_set_tuning": [c_p, c_int, c_int],
    "gss_autotune": [c_p, c_p, c_int, c_p, c_p],
    "gss_step_enqueue": [c_p, c_i64, c_p, c_p, c_p, c_i64, c_p, c_p, c_p],
    "gss_step_wait": [c_p, c_p, c_p],
    "gss_step_hits_copy": [c_p, c_p, c_i64, c_p],
    "gss_step_hits_map": [c_p, c_p, c_p],
    "gss_step_ids_map": [c_p, c_p, c_p],
    "gss_step_clauses_map": [c_p, c_p, c_p],
    "gss_step_stats": [c_p, c_p],
    "gss_db_reserve": [c_p, c_i64, c_p],
    "gss_server_default_config": [c_p],
    "gss_server_create": [c_int, c_p, c_p],
    "gss_server_destroy": [c_p],
    "gss_server_engine": [c_p, c_p],
    "gss_server_update_base": [c_p, c_i64, c_p],
    "gss_server_add_clauses": [c_p, c_i64, c_p, c_p, c_p],
    "gss_server_submit": [c_p, c_int, c_int, c_i64, c_p, c_p],
    "gss_server_poll": [c_p, c_i64, c_p, c_p, c_i64],
    "gss_server_result": [c_p, c_i64, c_p],
    "gss_server_records": [c_p, c_i64, c_p, c_p],
    "gss_server_release": [c_p, c_i64],
    "gss_server_stats_get": [c_p, c_p],
    "gss_step_submit": [c_p, c_p],
    "gss_bases_reset": [c_p, c_int],
    # include/gss_residency.h
    "gss_res_create": [c_p, c_i64, ctypes.c_float, ctypes.c_uint64, c_p],
    "gss_res_destroy": [c_p],
    "gss_res_add": [c_p, c_i64, c_p, c_p, c_p, ctypes.c_int32],
    "gss_res_account": [c_p, ctypes.c_int32, c_p],
    "gss_res_snapshot": [c_p, ctypes.c_int32, c_i64, c_p],
    "gss_res_train": [c_p, ctypes.c_int32, c_int, c_int, c_int, ctypes.c_float, c_p, c_p],
    "gss_res_reduce": [c_p, c_int, ctypes.c_int32, c_p, c_p],
    "gss_res_features": [c_p, ctypes.c_int32, c_i64, c_p, c_p],
    "gss_res_activity": [c_p, c_i64, c_p, c_p],
    "gss_res_get_params": [c_p, c_p],
    "gss_res_set_params": [c_p, c_p],
    "gss_res_grad": [c_p, c_i64, c_p, c_p, c_p, c_p, c_p],
    "gss_res_score": [c_p, c_i64, c_p, c_p],
    # include/gss_ceilings.h
    "gss_device_query": [c_int, c_p],
    "gss_ceiling_copy": [c_int, c_i64, c_int, c_p, c_p],
    "gss_ceiling_gather": [c_int, c_i64, c_i64, c_int, c_int, c_p, c_p],
    "gss_ceiling_gather_block": [c_int, c_i64, c_int, c_i64, c_int, c_int, c_p, c_p],
    "gss_ceiling_pcie": [c_int, c_i64, c_int, c_p, c_p],
    # include/gss_sharer.h
    "gss_sharer_default_options": [c_p],
    "gss_sharer_create": [c_p, c_p, c_p, c_p],
    "gss_sharer_destroy": [c_p],
    "gss_sharer_gpu_run": [c_p],
    "gss_sharer_reduce_db": [c_p],
    "gss_sharer_get_added_clause_count": [c_p, c_p],
    "gss_sharer_get_added_clause_count_at_last_reduce_db": [c_p, c_p],
    "gss_sharer_has_run_out_of_gpu_memory_once": [c_p, c_p],
    "gss_sharer_get_gpu_mem_info": [c_p, c_p, c_p],
    "gss_sharer_get_global_stat_count": [c_p, c_p],
    "gss_sharer_get_global_stat": [c_p, c_int, c_p],
    "gss_sharer_write_clauses_in_cnf": [c_p, c_char_p],
    "gss_sharer_set_var_count": [c_p, c_int],
    "gss_sharer_add_clause": [c_p, c_int, c_p, c_int, c_p],
    "gss_sharer_set_cpu_solver_count": [c_p, c_int],
    "gss_sharer_get_one_solver_stat_name": [c_p, c_int, c_p],
    "gss_sharer_get_global_stat_name": [c_p, c_int, c_p],
    "gss_sharer_get_one_solver_stat_count": [c_p, c_p],
    "gss_sharer_try_set_solver_values": [c_p, c_int, c_p, c_int, c_p],
    "gss_sharer_unset_solver_values": [c_p, c_int, c_p, c_int],
    "gss_sharer_try_send_assignment": [c_p, c_int, c_p],
    "gss_sharer_pop_reported_clause": [c_p, c_int, c_p, c_p, c_p, c_p],
    "gss_sharer_get_last_assig_all_reported": [c_p, c_int, c_p],
    "gss_sharer_get_current_assignment": [c_p, c_int, c_p],
    "gss_sharer_get_one_solver_stat": [c_p, c_int, c_int, c_p],
    "gss_sharer_limits": [c_p, c_p, c_p],
    "gss_sharer_debug_clause": [c_p, c_i64, c_p, c_p, c_p],
}
_RESTYPES = {"gss_last_error": c_char_p}

_lib = None


def load() -> ctypes.CDLL:
    """Loads the library once; raises ImportError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} not found: build it with `python -m gpusharesat_amd.build` "
            "(there is no CPU fallback for the check path)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        try:
            f = getattr(lib, name)
        except AttributeError as exc:
            raise ImportError(f"{LIB_PATH} does not export {name}") from exc
        f.argtypes = args
        f.restype = _RESTYPES.get(name, c_int)
    _lib = lib
    return lib


def check(name: str, rc: int) -> None:
    if rc != 0:
        msg = load().gss_last_error().decode(errors="replace")
        raise GssError(name, rc, msg)


def call(name: str, *args) -> None:
    check(name, getattr(load(), name)(*args))
