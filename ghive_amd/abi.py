"""ctypes mirror of include/ghive/ghive.h (structs and constants only; loads nothing).

Each structure has the field order and widths of its C declaration; tests/test_abi.py checks the
sizes and offsets against the header through the compiled library's own sizeof/offsetof report.
"""
import ctypes as C

ABI_VERSION = 1

# status codes
OK, ERR_INVALID, ERR_UNSUPPORTED, ERR_OOM, ERR_CAPACITY, ERR_HIP, ERR_ADMISSION, ERR_NO_DEVICE = range(8)
STATUS_NAMES = {
    0: "OK", 1: "INVALID", 2: "UNSUPPORTED", 3: "OOM", 4: "CAPACITY", 5: "HIP", 6: "ADMISSION",
    7: "NO_DEVICE",
}

# ColumnType (gpu_impl/inc/DataFlow/Column.hpp:5-11) + FLOAT extension
LONG, DOUBLE, INT, STRING, FLOAT = 0, 1, 2, 3, 4

# PredicateMode (gpu_impl/inc/Operator/FilterPredicate.hpp:17-58)
(FILTER_EQ, FILTER_LE, FILTER_GE, FILTER_LT, FILTER_GT, FILTER_RANGE, FILTER_EQ_RANGE,
 FILTER_NOT_EQ, FILTER_NOT_RANGE, FILTER_NOT_EQ_RANGE, FILTER_NOT_NULL, FILTER_AND,
 FILTER_OR) = range(13)

# BaseJoinType (gpu_impl/inc/Operator/BaseJoinPredicate.hpp:11-25)
(INNER_JOIN, LEFT_OUTER_JOIN, RIGHT_OUTER_JOIN, FULL_OUTER_JOIN, LEFT_SEMI_JOIN,
 CROSS_JOIN) = range(6)

# AggregationType (gpu_impl/inc/Operator/AggregationDesc.hpp:6-27)
AGG_SUM, AGG_MAX, AGG_MIN, AGG_AVG, AGG_CNT, AGG_RANK, AGG_UNKNOWN = range(7)
COUNT_STAR = 0xFFFFFFFF

SRC_FACT, SRC_JOIN = 0, 1
EXPR_COL, EXPR_MUL, EXPR_ONE = 0, 1, 2

# NULL sentinels (gpu_impl/inc/DataFlow/Table.hpp:29-52)
INT_NULL = 2**31 - 1
LONG_NULL = 2**63 - 1


class Column(C.Structure):
    _fields_ = [("type", C.c_int32), ("reserved", C.c_uint32), ("data", C.c_void_p),
                ("aux", C.c_void_p), ("row_num", C.c_uint64), ("char_size", C.c_uint64)]


class Table(C.Structure):
    _fields_ = [("columns", C.POINTER(Column)), ("ncols", C.c_uint32), ("key_num", C.c_uint32),
                ("row_num", C.c_uint64)]


class FilteredBuild(C.Structure):
    _fields_ = [("table", C.c_void_p), ("nodes", C.c_void_p), ("nnodes", C.c_uint32), ("root", C.c_uint32),
                ("key_col", C.c_uint32), ("payload_col", C.c_int32), ("load_factor", C.c_float),
                ("reserved", C.c_uint32)]


MAX_FILTERED_BUILDS = 8

PRED_STR_MAX = 64


class Pred(C.Structure):
    _fields_ = [("mode", C.c_int32), ("pred_col_num", C.c_int32), ("filter_col", C.c_uint32 * 2),
                ("param_num", C.c_int32), ("data_type", C.c_int32),
                ("long_params", C.c_int64 * 2), ("int_params", C.c_int32 * 2),
                ("double_params", C.c_double * 2), ("children", C.c_int32 * 2),
                ("string_sizes", C.c_int32 * 2), ("string_params", (C.c_char * PRED_STR_MAX) * 2)]


class JoinCond(C.Structure):
    _fields_ = [("type", C.c_int32), ("left", C.c_uint32), ("right", C.c_uint32)]


class PlanInput(C.Structure):
    _fields_ = [("name", C.c_char_p), ("table", C.POINTER(Table)), ("column_names", C.POINTER(C.c_char_p))]


class GTables(C.Structure):
    _fields_ = [("ntables", C.c_uint32), ("types", C.POINTER(C.POINTER(C.c_int32))),
                ("ncols", C.POINTER(C.c_uint32)), ("row_cnt", C.POINTER(C.c_int32)),
                ("key_cnt", C.POINTER(C.c_int32)), ("vertex_name", C.POINTER(C.c_char_p)),
                ("long_cols", C.POINTER(C.c_void_p)), ("nlong", C.c_uint32),
                ("double_cols", C.POINTER(C.c_void_p)), ("ndouble", C.c_uint32),
                ("int_cols", C.POINTER(C.c_void_p)), ("nint", C.c_uint32),
                ("string_cols", C.POINTER(C.c_void_p)), ("string_idx_cols", C.POINTER(C.c_void_p)),
                ("nstring", C.c_uint32)]


class Agg(C.Structure):
    _fields_ = [("type", C.c_int32), ("index", C.c_uint32)]


class FragmentJoin(C.Structure):
    _fields_ = [("fact_key_col", C.c_uint32), ("table", C.c_void_p)]


class FragmentKey(C.Structure):
    _fields_ = [("source", C.c_int32), ("index", C.c_uint32)]


class FragmentAgg(C.Structure):
    _fields_ = [("type", C.c_int32), ("expr", C.c_int32), ("a", C.c_uint32), ("b", C.c_uint32)]


class Fragment(C.Structure):
    _fields_ = [("fact", Table), ("filter", C.POINTER(Pred)), ("filter_nnodes", C.c_uint32),
                ("filter_root", C.c_uint32), ("joins", C.POINTER(FragmentJoin)),
                ("njoins", C.c_uint32), ("keys", C.POINTER(FragmentKey)), ("nkeys", C.c_uint32),
                ("aggs", C.POINTER(FragmentAgg)), ("naggs", C.c_uint32)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("compute_units", C.c_int32),
                ("wavefront_size", C.c_int32), ("lds_per_block", C.c_uint64),
                ("hbm_bytes", C.c_uint64), ("l2_bytes", C.c_uint64), ("clock_khz", C.c_int32),
                ("mem_clock_khz", C.c_int32), ("mem_bus_width", C.c_int32)]


class StageTimes(C.Structure):
    _fields_ = [("total_ms", C.c_float), ("main_kernel_ms", C.c_float),
                ("finalize_ms", C.c_float), ("rows_in", C.c_uint64), ("rows_joined", C.c_uint64)]


class PlanExecInfo(C.Structure):
    _fields_ = [("fused", C.c_int32), ("joins", C.c_uint32), ("narrowed_columns", C.c_uint32),
                ("has_filter", C.c_int32), ("fused_kernel_ms", C.c_float), ("reserved", C.c_uint32),
                ("not_fused", C.c_char * 160)]


class ProcessTimes(C.Structure):
    _fields_ = [("stage_ms", C.c_double), ("execute_ms", C.c_double), ("result_ms", C.c_double),
                ("total_ms", C.c_double), ("h2d_bytes", C.c_uint64), ("narrowed_columns", C.c_uint32),
                ("fused", C.c_int32)]


class SortInfo(C.Structure):
    _fields_ = [("rows", C.c_uint64), ("key_words", C.c_uint32), ("passes", C.c_uint32),
                ("embedded_ids", C.c_uint32), ("reserved", C.c_uint32), ("scatter_bytes", C.c_uint64)]


# Oracle-side fragment description (oracle/ghive_oracle.h), kept here so both sides share layouts.
class OrcDim(C.Structure):
    _fields_ = [("table", Table), ("filter", C.POINTER(Pred)), ("filter_nnodes", C.c_uint32),
                ("filter_root", C.c_uint32), ("key_col", C.c_uint32), ("payload_col", C.c_int32),
                ("fact_key_col", C.c_uint32)]


class OrcFragment(C.Structure):
    _fields_ = [("fact", Table), ("filter", C.POINTER(Pred)), ("filter_nnodes", C.c_uint32),
                ("filter_root", C.c_uint32), ("dims", C.POINTER(OrcDim)), ("ndims", C.c_uint32),
                ("keys", C.POINTER(FragmentKey)), ("nkeys", C.c_uint32),
                ("aggs", C.POINTER(FragmentAgg)), ("naggs", C.c_uint32)]


def agg_output_type(agg_type, input_type):
    """Hive output types (same rule as ghive_agg_output_type)."""
    if agg_type == AGG_CNT:
        return LONG
    if agg_type == AGG_AVG:
        return DOUBLE
    if agg_type == AGG_SUM:
        return LONG if input_type in (INT, LONG) else DOUBLE
    return input_type


def make_pred(mode, col=0, params=(), data_type=None, col2=None, children=(-1, -1)):
    """Build a Pred node. params are Python ints (long literal) or floats (double literal)."""
    p = Pred()
    p.mode = mode
    p.children[0], p.children[1] = children
    if mode in (FILTER_AND, FILTER_OR):
        p.pred_col_num, p.param_num, p.data_type = 0, 0, 4
        return p
    p.filter_col[0] = col
    if col2 is not None:
        p.pred_col_num, p.filter_col[1], p.param_num, p.data_type = 2, col2, 0, 4
        return p
    p.pred_col_num = 1
    p.param_num = len(params)
    if mode == FILTER_NOT_NULL:
        p.data_type = 4
        return p
    if data_type is None:
        if any(isinstance(v, (str, bytes)) for v in params):
            data_type = 3
        else:
            data_type = 1 if any(isinstance(v, float) for v in params) else 0
    p.data_type = data_type
    for i, v in enumerate(params):
        if data_type == 3:
            b = v.encode() if isinstance(v, str) else bytes(v)
            if len(b) > PRED_STR_MAX:
                raise ValueError(f"string literal longer than {PRED_STR_MAX} bytes")
            p.string_sizes[i] = len(b)
            C.memmove(C.addressof(p.string_params[i]), b, len(b))
        elif data_type == 1:
            p.double_params[i] = float(v)
        elif data_type == 2:
            p.int_params[i] = int(v)
        else:
            p.long_params[i] = int(v)
    return p


def pred_array(nodes):
    arr = (Pred * max(len(nodes), 1))()
    for i, n in enumerate(nodes):
        arr[i] = n
    return arr


class BatchItem(C.Structure):
    _fields_ = [("key", C.c_void_p), ("a", C.c_void_p), ("b", C.c_void_p), ("rows", C.c_uint64),
                ("filter_lo", C.c_int32), ("filter_hi", C.c_int32)]


class BatchSpec(C.Structure):
    _fields_ = [("dim", C.c_void_p), ("filter_shift", C.c_uint32), ("filter_bits", C.c_uint32),
                ("group_shift", C.c_uint32), ("group_bits", C.c_uint32)]


class BatchTuning(C.Structure):
    _fields_ = [("tile_rows", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("probe_mode", C.c_int32),
                ("reserved", C.c_int32)]


# ---- serving form (ghive_server_*) ----------------------------------------------------------
REQ_FRAGMENT, REQ_GROUPBY, REQ_SORT = 1, 2, 3


class ServerConfig(C.Structure):
    _fields_ = [("slots", C.c_uint32), ("reserved", C.c_uint32), ("hbm_budget_bytes", C.c_uint64),
                ("slot_mem_limit_bytes", C.c_uint64)]


class Request(C.Structure):
    _fields_ = [("kind", C.c_int32), ("host_io", C.c_int32), ("fragment", C.POINTER(Fragment)),
                ("table", C.POINTER(Table)), ("keys", C.POINTER(C.c_uint32)), ("orders", C.c_char_p),
                ("nkeys", C.c_uint32), ("naggs", C.c_uint32), ("aggs", C.POINTER(Agg)),
                ("out_cols", C.POINTER(Column)), ("capacity", C.c_uint64), ("out_perm", C.c_void_p),
                ("footprint_bytes", C.c_uint64)]


class Response(C.Structure):
    _fields_ = [("status", C.c_int32), ("slot", C.c_uint32), ("groups", C.c_uint64),
                ("submit_us", C.c_double), ("start_us", C.c_double), ("end_us", C.c_double),
                ("error", C.c_char * 256)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_uint64), ("rejected", C.c_uint64), ("queued", C.c_uint64),
                ("inflight", C.c_uint32), ("peak_inflight", C.c_uint32), ("inflight_bytes", C.c_uint64),
                ("peak_inflight_bytes", C.c_uint64)]
