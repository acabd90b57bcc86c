"""Loads libghive_mi355x.so (built in-tree by ghive_amd.build) and declares the C ABI.

There is no fallback: if the library is missing or cannot be loaded every operator raises.
"""
import ctypes as C
import os

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libghive_mi355x.so")

_lib = None

P = C.c_void_p
u32, u64, i32, i64 = C.c_uint32, C.c_uint64, C.c_int32, C.c_int64
PU64 = C.POINTER(C.c_uint64)

# name -> (restype, argtypes); mirrors include/ghive/ghive.h
SIGNATURES = {
    "ghive_abi_version": (C.c_int, []),
    "ghive_init": (C.c_int, [C.c_int, u64, C.POINTER(P)]),
    "ghive_shutdown": (None, [P]),
    "ghive_last_error": (C.c_char_p, [P]),
    "ghive_set_stream": (C.c_int, [P, P]),
    "ghive_get_stream": (P, [P]),
    "ghive_use_own_stream": (C.c_int, [P]),
    "ghive_batch_create": (C.c_int, [P, P, P, u32, P, C.c_int, C.POINTER(P)]),
    "ghive_batch_run": (C.c_int, [P, P, P]),
    "ghive_batch_results": (C.c_int, [P, P, P, P]),
    "ghive_batch_info": (C.c_int, [P, C.POINTER(u32), PU64, PU64, C.POINTER(i32), C.POINTER(u32),
                                   C.POINTER(C.c_float)]),
    "ghive_batch_free": (None, [P, P]),
    "ghive_sync": (C.c_int, [P]),
    "ghive_device_query": (C.c_int, [P, C.POINTER(abi.DeviceInfo)]),
    "ghive_malloc": (C.c_int, [P, u64, C.POINTER(P)]),
    "ghive_free": (C.c_int, [P, P]),
    "ghive_memcpy": (C.c_int, [P, P, P, u64, C.c_int]),
    "ghive_mem_stats": (C.c_int, [P, PU64, PU64, PU64, PU64]),
    "ghive_gen_uniform_i32": (C.c_int, [P, u64, u32, u64, u64, i32, i32, P]),
    "ghive_gen_uniform_i64": (C.c_int, [P, u64, u32, u64, u64, i64, i64, P]),
    "ghive_gen_sequence_i32": (C.c_int, [P, u64, u64, i32, P]),
    "ghive_gen_sequence_i64": (C.c_int, [P, u64, u64, i64, P]),
    "ghive_gen_uniform_f32": (C.c_int, [P, u64, u32, u64, u64, P]),
    "ghive_gen_normal_f64": (C.c_int, [P, u64, u32, u64, u64, P]),
    "ghive_parse_predicate": (C.c_int, [C.c_char_p, C.POINTER(C.c_char_p), u32, C.POINTER(abi.Pred),
                                        u32, C.POINTER(u32), C.POINTER(u32), C.c_char_p, C.c_size_t]),
    "ghive_predicate_to_string": (C.c_size_t, [C.POINTER(abi.Pred), u32, u32, C.c_char_p, C.c_size_t]),
    "ghive_filter": (C.c_int, [P, C.POINTER(abi.Table), C.POINTER(abi.Pred), u32, u32, P, PU64]),
    "ghive_filter_async": (C.c_int, [P, C.POINTER(abi.Table), C.POINTER(abi.Pred), u32, u32, P, P]),
    "ghive_filter_compact": (C.c_int, [P, C.POINTER(abi.Table), C.POINTER(abi.Pred), u32, u32,
                                       C.POINTER(abi.Column), PU64]),
    "ghive_gather": (C.c_int, [P, C.POINTER(abi.Table), P, u64, C.POINTER(abi.Column)]),
    "ghive_hash_build": (C.c_int, [P, C.POINTER(abi.Column), P, u64, C.POINTER(abi.Column), C.c_float,
                                   C.POINTER(P)]),
    "ghive_hash_build_async": (C.c_int, [P, C.POINTER(abi.Column), P, P, u64, C.POINTER(abi.Column),
                                         C.c_float, C.POINTER(P)]),
    "ghive_hash_build_filtered": (C.c_int, [P, C.POINTER(abi.FilteredBuild), u32, C.POINTER(P)]),
    "ghive_hash_table_info": (C.c_int, [P, PU64, PU64, PU64, C.POINTER(i32)]),
    "ghive_hash_table_free": (None, [P, P]),
    "ghive_hash_probe": (C.c_int, [P, P, C.POINTER(abi.Column), P, u64, P, P, u64, PU64]),
    "ghive_join": (C.c_int, [P, C.POINTER(abi.Table), u32, P, C.POINTER(abi.JoinCond), u32, P, u64,
                             PU64]),
    "ghive_agg_output_type": (C.c_int, [i32, i32]),
    "ghive_groupby": (C.c_int, [P, C.POINTER(abi.Table), P, u64, P, u32, C.POINTER(abi.Agg), u32,
                                C.POINTER(abi.Column), u64, PU64]),
    "ghive_sort": (C.c_int, [P, C.POINTER(abi.Table), P, C.c_char_p, u32, P]),
    "ghive_fragment_execute": (C.c_int, [P, C.POINTER(abi.Fragment), C.POINTER(abi.Column), u64, PU64]),
    "ghive_fragment_execute_noalloc": (C.c_int, [P, C.POINTER(abi.Fragment), C.POINTER(abi.Column), u64, PU64]),
    "ghive_fragment_execute_late": (C.c_int, [P, C.POINTER(abi.Fragment), C.POINTER(abi.Column), u64,
                                              PU64]),
    "ghive_set_option": (C.c_int, [P, C.c_char_p, i64]),
    "ghive_get_option": (C.c_int, [P, C.c_char_p, C.POINTER(i64)]),
    "ghive_last_stage_times": (C.c_int, [P, C.POINTER(abi.StageTimes)]),
    "ghive_last_plan_info": (C.c_int, [P, C.POINTER(abi.PlanExecInfo)]),
    "ghive_gpu_process": (C.c_int, [P, P, C.POINTER(abi.GTables), C.c_char_p, C.c_char_p, C.c_uint64,
                                    C.POINTER(P)]),
    "ghive_gpu_result_info": (C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                        C.POINTER(C.POINTER(C.c_int32))]),
    "ghive_gpu_result_long": (C.POINTER(C.c_int64), [P, C.c_uint32]),
    "ghive_gpu_result_double": (C.POINTER(C.c_double), [P, C.c_uint32]),
    "ghive_gpu_result_int": (C.POINTER(C.c_int32), [P, C.c_uint32]),
    "ghive_gpu_result_string": (C.c_int, [P, C.c_uint32, C.POINTER(C.c_char_p), C.POINTER(C.c_uint64),
                                          C.POINTER(C.POINTER(C.c_int32))]),
    "ghive_gpu_result_free": (None, [P]),
    "ghive_last_process_times": (C.c_int, [P, C.POINTER(abi.ProcessTimes)]),
    "ghive_last_sort_info": (C.c_int, [P, C.POINTER(abi.SortInfo)]),
}


class GhiveError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"ghive status {abi.STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def lib():
    """The loaded library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (python -m ghive_amd.build); "
                               "the HIP library is required, there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib

SIGNATURES_EXTRA = {
    "ghive_bench_stream": (C.c_int, [P, P, P, u64, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "ghive_gen_rank_keys": (C.c_int, [P, u64, u64, u64, P]),
    "ghive_gen_zipf_keys": (C.c_int, [P, u64, u32, u64, u64, u64, C.c_double, C.c_double, P]),
    "ghive_join_result": (C.c_int, [P, C.POINTER(abi.Table), u32, C.POINTER(abi.JoinCond), u32, P, P, P, u64,
                                    C.POINTER(abi.Column), u32, C.POINTER(u32)]),
    "ghive_plan_parse": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, u64, C.POINTER(P)]),
    "ghive_plan_describe": (C.c_int, [P, C.c_char_p, u64, PU64]),
    "ghive_plan_free": (None, [P]),
    "ghive_plan_execute": (C.c_int, [P, P, C.c_char_p, C.POINTER(abi.PlanInput), u32, C.c_char_p, u64,
                                     C.POINTER(P)]),
    "ghive_result_info": (C.c_int, [P, PU64, C.POINTER(u32)]),
    "ghive_result_column": (C.c_int, [P, u32, C.POINTER(abi.Column), C.POINTER(C.c_char_p)]),
    "ghive_result_free": (None, [P, P]),
    "ghive_gtables_stage": (C.c_int, [P, C.POINTER(abi.GTables), C.POINTER(P)]),
    "ghive_staged_inputs": (C.c_int, [P, C.POINTER(C.POINTER(abi.PlanInput)), C.POINTER(u32)]),
    "ghive_staged_free": (None, [P, P]),
    "ghive_gpuresult_sequence": (C.c_int, [C.POINTER(i32), u32, C.POINTER(i32), C.POINTER(u32)]),
    "ghive_select_arith": (C.c_int, [P, C.POINTER(abi.Column), C.POINTER(abi.Column), C.c_char, u64,
                                     C.POINTER(abi.Column)]),
    "ghive_select_cast": (C.c_int, [P, C.POINTER(abi.Column), i32, u64, C.POINTER(abi.Column)]),
    "ghive_select_substr": (C.c_int, [P, C.POINTER(abi.Column), i32, i32, u64, C.POINTER(abi.Column)]),
    "ghive_server_create": (C.c_int, [C.c_int, C.POINTER(abi.ServerConfig), C.POINTER(P)]),
    "ghive_server_submit": (C.c_int, [P, C.POINTER(abi.Request), PU64]),
    "ghive_server_wait": (C.c_int, [P, u64, C.POINTER(abi.Response)]),
    "ghive_server_get_stats": (C.c_int, [P, C.POINTER(abi.ServerStats)]),
    "ghive_server_slot_context": (P, [P, u32]),
    "ghive_server_destroy": (None, [P]),
}
SIGNATURES.update(SIGNATURES_EXTRA)
