"""ctypes image of include/turbogp_gfx950.h and loader of the in-tree library.

The library is required: there is no Python or CPU fallback for any operator.
Loading fails loudly when libturbogp_gfx950.so has not been built.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libturbogp_gfx950.so")

# status codes
OK = 0
ERR_INVALID = -1
ERR_HIP = -2
ERR_OOM = -3
ERR_BUDGET = -4
ERR_UNSUPPORTED = -5
ERR_NO_DEVICE = -6
ERR_CAPACITY = -7
ERR_CANCELLED = -8
ERR_INJECTED = -9

# dtypes
I8, U8, I16, U16, I32, U32, I64, U64 = range(8)
DTYPE_WIDTH = {I8: 1, U8: 1, I16: 2, U16: 2, I32: 4, U32: 4, I64: 8, U64: 8}
# comparisons (bitmap.rs filter_{eq,ne,lt,le,gt,ge})
EQ, NE, LT, LE, GT, GE = range(6)
BIT_AND, BIT_OR, BIT_ANDNOT = range(3)
AGG_SUM, AGG_COUNT_NONZERO = 0, 1
AGG_MIN, AGG_MAX, AGG_MIN_U64, AGG_MAX_U64 = 2, 3, 4, 5
DISTINCT_AUTO, DISTINCT_EXACT, DISTINCT_HLL = 0, 1, 2
AGG_MINMAX = {"min": AGG_MIN, "max": AGG_MAX, "min_u64": AGG_MIN_U64, "max_u64": AGG_MAX_U64}
JOIN_GROUP_BUILD_PAYLOAD, JOIN_GROUP_PROBE_COL, JOIN_GROUP_NONE = 0, 1, 2
TUNE_WG_PER_CU, TUNE_GRID_MAX, TUNE_SPECIALIZE, TUNE_SPEC_KERNEL, TUNE_CHUNK = 0, 1, 2, 3, 4
TUNE_SLICES_PER_WAVE, TUNE_JOIN_ATOMICS, TUNE_JOIN_PACK_BITS, TUNE_DISTINCT_SHRINK = 5, 6, 7, 8
TUNE_GROUP_TAG_BITS, TUNE_JOIN_HOT, TUNE_JIT, TUNE_AGG_PART = 9, 10, 11, 12
(EX_COL, EX_INT, EX_FLOAT, EX_DATE, EX_NULL, EX_ADD, EX_SUB, EX_MUL, EX_DIV, EX_MOD, EX_EQ, EX_NE,
 EX_LT, EX_LE, EX_GT, EX_GE, EX_AND, EX_OR, EX_NOT, EX_NEG, EX_BETWEEN, EX_NOT_BETWEEN, EX_IN,
 EX_NOT_IN, EX_IF) = range(25)
MAX_EXPR_OPS, MAX_EXPR_DEPTH = 64, 8
SPEC_SHAPE_RUNTIME = 1000
JOIN_INNER, JOIN_LEFT = 0, 1
JOIN_NO_MATCH = 0xFFFFFFFF
MAX_JOIN_KEYS = 4
GATHER_NO_ROW = 0xFFFFFFFFFFFFFFFF
BATCH_PERSISTENT, BATCH_PER_QUERY = 0, 1

MAX_COLS, MAX_PREDS, MAX_AGGS, MAX_FACTORS, MAX_GROUP_PARTS = 8, 8, 6, 3, 2


class Column(C.Structure):
    _fields_ = [("data", C.c_void_p), ("len", C.c_uint64), ("dtype", C.c_int32),
                ("_pad", C.c_int32)]


class ExprOp(C.Structure):
    _fields_ = [("code", C.c_int32), ("arg", C.c_int32), ("imm", C.c_int64)]


class Pred(C.Structure):
    _fields_ = [("col", C.c_int32), ("op", C.c_int32), ("is_signed", C.c_int32),
                ("_pad", C.c_int32), ("lit", C.c_int64)]


class GroupPart(C.Structure):
    _fields_ = [("col", C.c_int32), ("divisor", C.c_int32), ("base", C.c_int64),
                ("card", C.c_int64)]


class Agg(C.Structure):
    _fields_ = [("kind", C.c_int32), ("nfactors", C.c_int32), ("fcol", C.c_int32 * 3),
                ("_pad", C.c_int32), ("coef", C.c_int64), ("fa", C.c_int64 * 3),
                ("fb", C.c_int64 * 3)]


class AggQuery(C.Structure):
    _fields_ = [("row_begin", C.c_uint64), ("row_end", C.c_uint64),
                ("cols", Column * MAX_COLS),
                ("ncols", C.c_int32), ("npreds", C.c_int32), ("nparts", C.c_int32),
                ("naggs", C.c_int32),
                ("preds", Pred * MAX_PREDS), ("parts", GroupPart * MAX_GROUP_PARTS),
                ("aggs", Agg * MAX_AGGS), ("d_sel_bitmap", C.c_void_p)]


MAX_GROUP_KEYS = 4


class HashGroupQuery(C.Structure):
    _fields_ = [("agg", AggQuery), ("nkeys", C.c_int32), ("key_col", C.c_int32 * MAX_GROUP_KEYS),
                ("_pad", C.c_int32 * 3), ("d_key_nulls", C.c_void_p * MAX_GROUP_KEYS)]


class HashGroupResult(C.Structure):
    _fields_ = [("d_keys", C.c_void_p), ("d_key_nulls", C.c_void_p), ("d_count", C.c_void_p),
                ("d_sums", C.c_void_p), ("capacity", C.c_uint64)]


class GroupResult(C.Structure):
    _fields_ = [("d_count", C.c_void_p), ("d_sums", C.c_void_p),
                ("d_n_out_of_range", C.c_void_p)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("cu_count", C.c_int32),
                ("wave_size", C.c_int32), ("hbm_bytes", C.c_int64),
                ("lds_bytes_per_cu", C.c_int64), ("l2_bytes", C.c_int64),
                ("clock_khz", C.c_int32), ("mem_clock_khz", C.c_int32)]


class JoinAggQuery(C.Structure):
    _fields_ = [("table", C.c_void_p), ("probe_keys", Column), ("value", Column),
                ("group_col", Column), ("group_by", C.c_int32), ("build_rows", C.c_uint32),
                ("key_base", C.c_int64), ("ngroups", C.c_uint64),
                ("row_begin", C.c_uint64), ("row_end", C.c_uint64)]


class KernelParams(C.Structure):
    """repr(C) image of KernelParams (kernel/mod.rs:96-119)."""
    _fields_ = [("target_u64", C.c_uint64), ("low_u64", C.c_uint64), ("high_u64", C.c_uint64),
                ("max_distance", C.c_uint32), ("_pad0", C.c_uint32),
                ("cell_count", C.c_uint64), ("target2_u64", C.c_uint64),
                ("target3_u64", C.c_uint64), ("pred1_op", C.c_int32), ("pred2_op", C.c_int32),
                ("pred3_op", C.c_int32), ("predicate_count", C.c_uint8),
                ("_pad1", C.c_uint8 * 3)]


class KernelResult(C.Structure):
    """repr(C) image of KernelResult (kernel/mod.rs:141-148)."""
    _fields_ = [("count", C.c_uint64), ("sum", C.c_double), ("mask", C.c_uint64)]


P = C.c_void_p
I32_ = C.c_int32
U64_ = C.c_uint64
I64_ = C.c_int64

class BatchTuning(C.Structure):
    _fields_ = [("wg_per_cu", C.c_uint32), ("chunk_slices", C.c_uint32), ("order", C.c_int32),
                ("cached_loads", C.c_int32), ("use_graph", C.c_int32),
                ("items_per_grab", C.c_uint32)]


SERVE_AGG, SERVE_JOIN_AGG = 0, 1


class ServeJoin(C.Structure):
    _fields_ = [("build_keys", Column), ("build_payload", Column), ("probe", JoinAggQuery)]


class ServeRequest(C.Structure):
    _fields_ = [("kind", C.c_int32), ("_pad", C.c_int32), ("agg", AggQuery), ("join", ServeJoin)]


class ServeConfig(C.Structure):
    _fields_ = [("max_batch_queries", C.c_uint32), ("batch_slots", C.c_uint32),
                ("join_slots", C.c_uint32), ("batch_wait_us", C.c_uint32),
                ("heavy_rows", C.c_uint64),
                ("hash_budget_bytes", C.c_uint64), ("wg_per_cu", C.c_uint32),
                ("chunk_slices", C.c_uint32), ("order", C.c_int32), ("cached_loads", C.c_int32),
                ("per_query_launch", C.c_int32), ("share_build_tables", C.c_int32)]


class ServeTiming(C.Structure):
    _fields_ = [("arrival_ns", C.c_int64), ("start_ns", C.c_int64), ("done_ns", C.c_int64),
                ("status", C.c_int32), ("deferred", C.c_int32), ("result_words", C.c_uint64)]


class ServeStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "submitted", "completed", "failed", "rejected", "deferred", "admitted_joins",
        "batches", "batched_requests", "max_batch", "heavy_in_flight",
        "submit_ns", "launch_ns", "harvest_ns", "table_builds")]


# name -> (restype, argtypes)
SIGNATURES = {
    "tgp_serve_now_ns": (I64_, []),
    "tgp_serve_create": (I32_, [P, C.POINTER(ServeConfig), C.POINTER(C.c_void_p)]),
    "tgp_serve_destroy": (I32_, [P]),
    "tgp_serve_submit": (I32_, [P, C.POINTER(ServeRequest), I64_, C.POINTER(C.c_uint64)]),
    "tgp_serve_wait": (I32_, [P, C.c_uint64, I64_, C.POINTER(C.c_int32)]),
    "tgp_serve_result": (I32_, [P, C.c_uint64, C.POINTER(C.c_uint64), C.c_uint64,
                                C.POINTER(ServeTiming)]),
    "tgp_serve_release": (I32_, [P, C.c_uint64]),
    "tgp_serve_replay": (I32_, [P, C.POINTER(ServeRequest), C.POINTER(C.c_int64), C.c_uint32,
                                C.POINTER(C.c_uint64), C.POINTER(C.c_int64)]),
    "tgp_serve_get_stats": (I32_, [P, C.POINTER(ServeStats)]),
    "tgp_last_error": (C.c_char_p, []),
    "tgp_abi_version": (I32_, []),
    "tgp_ctx_create": (I32_, [I32_, U64_, C.POINTER(P)]),
    "tgp_ctx_destroy": (I32_, [P]),
    "tgp_device_info_get": (I32_, [P, C.POINTER(DeviceInfo)]),
    "tgp_stream_create": (I32_, [P, C.POINTER(P)]),
    "tgp_stream_destroy": (I32_, [P, P]),
    "tgp_stream_sync": (I32_, [P, P]),
    "tgp_hbm_alloc": (I32_, [P, U64_, C.POINTER(P)]),
    "tgp_hbm_free": (I32_, [P, P]),
    "tgp_hbm_usage": (I32_, [P, C.POINTER(U64_), C.POINTER(U64_), C.POINTER(U64_)]),
    "tgp_hbm_set_budget": (I32_, [P, U64_]),
    "tgp_memcpy_h2d": (I32_, [P, P, P, U64_, P]),
    "tgp_memcpy_d2h": (I32_, [P, P, P, U64_, P]),
    "tgp_memcpy_d2d": (I32_, [P, P, P, U64_, P]),
    "tgp_fill_u64": (I32_, [P, P, U64_, U64_, P]),
    "tgp_column_minmax_sync": (I32_, [P, C.POINTER(Column), C.POINTER(I64_), C.POINTER(I64_)]),
    "tgp_bitmap_words": (U64_, [U64_]),
    "tgp_scan_filter": (I32_, [P, C.POINTER(Column), I32_, C.POINTER(Pred), I32_, U64_, U64_,
                               P, P, P]),
    "tgp_bitmap_combine": (I32_, [P, I32_, P, P, P, U64_, P]),
    "tgp_bitmap_not": (I32_, [P, P, P, U64_, P]),
    "tgp_bitmap_count_range": (I32_, [P, P, U64_, U64_, U64_, P, P]),
    "tgp_bitmap_to_indices_sync": (I32_, [P, P, U64_, U64_, I32_, P, U64_, C.POINTER(U64_), P]),
    "tgp_bitmap_to_indices_noalloc_sync": (I32_, [P, P, U64_, U64_, I32_, P, U64_,
                                                  C.POINTER(U64_), P]),
    "tgp_group_agg": (I32_, [P, C.POINTER(AggQuery), C.POINTER(GroupResult), P]),
    "tgp_group_agg_ngroups": (I32_, [C.POINTER(AggQuery), C.POINTER(U64_)]),
    "tgp_group_agg_explain": (I32_, [C.POINTER(AggQuery), C.POINTER(I32_), C.POINTER(I32_)]),
    "tgp_jit_info": (I32_, [C.POINTER(U64_), C.POINTER(U64_), C.POINTER(U64_),
                            C.POINTER(C.c_double)]),
    "tgp_scan_filter_f64": (I32_, [P, C.POINTER(Column), I32_, I32_, C.c_double, U64_, U64_, P,
                                   P, P]),
    "tgp_scan_expr": (I32_, [P, C.POINTER(Column), I32_, C.POINTER(ExprOp), I32_, U64_, U64_, P,
                             P]),
    "tgp_join_residual_sync": (I32_, [P, C.POINTER(Column), I32_, C.POINTER(Column), I32_,
                                      C.POINTER(ExprOp), I32_, I32_, U64_, U64_, P, P, U64_, P,
                                      P, C.POINTER(U64_), P]),
    "tgp_hash_group_agg_sync": (I32_, [P, C.POINTER(HashGroupQuery), C.POINTER(HashGroupResult),
                                       C.POINTER(U64_), P]),
    "tgp_synth_uniform": (I32_, [P, C.POINTER(Column), U64_, U64_, I64_, I64_, U64_, U64_, P]),
    "tgp_synth_lookup_cdf": (I32_, [P, C.POINTER(Column), U64_, U64_, P, U64_, P, U64_, U64_,
                                    P]),
    "tgp_debug_inject_failure": (I32_, [P, I32_]),
    "tgp_kernel_table_new": (I32_, [P, C.POINTER(P)]),
    "tgp_kernel_table_destroy": (I32_, [P]),
    "tgp_kernel_table_select": (I32_, [P, I32_, I32_, C.POINTER(P)]),
    "tgp_kernel_table_len": (I32_, [P, C.POINTER(I32_)]),
    "tgp_kernel_table_entry": (I32_, [P, I32_, C.POINTER(I32_), C.POINTER(I32_),
                                      C.POINTER(I32_), C.POINTER(C.c_char_p)]),
    "tgp_kernel_info": (I32_, [P, C.POINTER(I32_), C.POINTER(I32_), C.POINTER(I32_),
                               C.POINTER(C.c_char_p)]),
    "tgp_kernel_execute": (I32_, [P, P, P, C.POINTER(KernelParams), C.POINTER(KernelResult)]),
    "tgp_batch_create": (I32_, [P, C.c_uint32, C.POINTER(P)]),
    "tgp_batch_destroy": (I32_, [P]),
    "tgp_batch_set_queries": (I32_, [P, C.POINTER(AggQuery), C.c_uint32, P]),
    "tgp_batch_set_tuning": (I32_, [P, C.POINTER(BatchTuning)]),
    "tgp_batch_get_tuning": (I32_, [P, C.POINTER(BatchTuning)]),
    "tgp_batch_autotune": (I32_, [P, P, C.POINTER(BatchTuning), C.POINTER(C.c_float)]),
    "tgp_batch_run": (I32_, [P, I32_, P]),
    "tgp_batch_run_timed": (I32_, [P, I32_, P, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "tgp_batch_info": (I32_, [P, C.POINTER(U64_), C.POINTER(U64_), C.POINTER(C.c_uint32),
                              C.POINTER(C.c_uint32)]),
    "tgp_batch_result_layout": (I32_, [P, C.c_uint32, C.POINTER(U64_), C.POINTER(U64_),
                                       C.POINTER(U64_), C.POINTER(U64_)]),
    "tgp_batch_results_sync": (I32_, [P, C.POINTER(U64_), U64_, P]),
    "tgp_hash_build": (I32_, [P, C.POINTER(Column), C.POINTER(Column), C.POINTER(P), P]),
    "tgp_hash_table_destroy": (I32_, [P, P]),
    "tgp_hash_table_footprint": (I32_, [P, C.POINTER(U64_)]),
    "tgp_sum_f64_sync": (I32_, [P, P, U64_, P, I32_, C.POINTER(C.c_double), P]),
    "tgp_count_distinct": (I32_, [P, C.POINTER(Column), U64_, U64_, P, I32_,
                                  C.POINTER(U64_), P]),
    "tgp_hll_sketch": (I32_, [P, C.POINTER(Column), U64_, U64_, P, P, P]),
    "tgp_hll_estimate": (U64_, [P]),
    "tgp_select_rows": (I32_, [P, U64_, U64_, P, C.POINTER(Column), I32_, U64_, P, U64_,
                               C.POINTER(U64_), P]),
    "tgp_gather": (I32_, [P, C.POINTER(Column), P, U64_, I32_, P, P]),
    "tgp_int128_to_numeric": (I32_, [U64_, U64_, I32_, C.c_char_p, U64_, C.POINTER(U64_)]),
    "tgp_int128_avg_to_numeric": (I32_, [U64_, U64_, U64_, I32_, I32_, C.c_char_p, U64_,
                                         C.POINTER(U64_)]),
    "tgp_column_from_cells": (I32_, [P, P, U64_, I32_, P, C.POINTER(U64_), P]),
    "tgp_null_bitmap_from_flags": (I32_, [P, P, U64_, P, P]),
    "tgp_hash_build_in": (I32_, [P, C.POINTER(Column), C.POINTER(Column), P, U64_,
                                 C.POINTER(C.c_void_p), P]),
    "tgp_hash_table_bytes_for": (U64_, [U64_]),
    "tgp_hash_table_info": (I32_, [P, C.POINTER(U64_), C.POINTER(U64_), C.POINTER(I32_)]),
    "tgp_hash_probe_count": (I32_, [P, P, C.POINTER(Column), U64_, U64_, P, P]),
    "tgp_hash_probe_mark": (I32_, [P, P, C.POINTER(Column), U64_, U64_, P, P]),
    "tgp_hash_probe_pairs_sync": (I32_, [P, P, C.POINTER(Column), U64_, U64_, P, P, U64_,
                                         C.POINTER(U64_), P]),
    "tgp_hash_probe_agg": (I32_, [P, C.POINTER(JoinAggQuery), C.POINTER(GroupResult), P]),
    "tgp_hash_table_add_bloom": (I32_, [P, P, C.c_uint32, P]),
    "tgp_hash_probe_agg_batch": (I32_, [P, C.POINTER(JoinAggQuery), C.c_uint32, C.POINTER(GroupResult),
                                        P]),
    "tgp_scan_compare_columns": (I32_, [P, C.POINTER(Column), I32_, C.POINTER(Column), I32_,
                                        U64_, U64_, P, P]),
    "tgp_hash_build_multi": (I32_, [P, C.POINTER(Column), I32_, C.POINTER(Column),
                                    C.POINTER(C.c_void_p), P]),
    "tgp_hash_join_pairs_sync": (I32_, [P, P, C.POINTER(Column), I32_, I32_, U64_, U64_, P, P,
                                        U64_, C.POINTER(U64_), P]),
    "tgp_hash_join_mark": (I32_, [P, P, C.POINTER(Column), I32_, U64_, U64_, P, P]),
    "tgp_event_create": (I32_, [P, C.POINTER(P)]),
    "tgp_tune_set": (I32_, [P, I32_, I64_]),
    "tgp_tune_get": (I32_, [P, I32_, C.POINTER(I64_)]),
    "tgp_event_destroy": (I32_, [P, P]),
    "tgp_event_record": (I32_, [P, P, P]),
    "tgp_event_elapsed_ms": (I32_, [P, P, P, C.POINTER(C.c_float)]),
}


class TgpError(RuntimeError):
    def __init__(self, code, where, msg):
        super().__init__(f"{where} failed with status {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Loads the operator library (once).  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
            "g.build()'` (make -C turbogp_amd/csrc).  There is no CPU fallback.")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name, None)
        if f is None:
            continue  # declared in a later ABI; checked by the export test
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(code, where):
    if code != OK:
        msg = lib().tgp_last_error()
        raise TgpError(code, where, msg.decode() if msg else "")
    return code
