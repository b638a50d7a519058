"""ctypes mirror of include/hjoin.h (plan/table types and the engine's entry points)."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhjoin.so")

PAGE_SIZE = 8192
INT32, INT64, FP64, VARCHAR = 0, 1, 2, 3
NODE_SCAN, NODE_JOIN = 0, 1

HJ_OK = 0
HJ_E_INVALID = -1
HJ_E_TYPE = -2
HJ_E_OOM = -3
HJ_E_HIP = -4
HJ_E_RANGE = -5
HJ_E_BUSY = -6
HJ_E_NODEV = -7

NOALLOC_SCRATCH_BYTES = 64 << 20  # include/hjoin.h HJ_NOALLOC_SCRATCH_BYTES
NOALLOC_MAX_PROBE_KEYS = 3_300_000  # include/hjoin.h HJ_NOALLOC_MAX_PROBE_KEYS


class hj_column(C.Structure):
    _fields_ = [("type", C.c_int32), ("num_pages", C.c_uint32),
                ("pages", C.POINTER(C.c_void_p))]


class hj_table(C.Structure):
    _fields_ = [("num_rows", C.c_uint64), ("num_columns", C.c_uint32),
                ("reserved", C.c_uint32), ("columns", C.POINTER(hj_column))]


class hj_attr(C.Structure):
    _fields_ = [("index", C.c_uint64), ("type", C.c_int32), ("reserved", C.c_int32)]


class hj_filter(C.Structure):
    _fields_ = [("column", C.c_uint32), ("lo", C.c_int32), ("hi", C.c_int32),
                ("reserved", C.c_int32)]


class hj_pred(C.Structure):
    _fields_ = [("op", C.c_int32), ("column", C.c_uint32), ("ival", C.c_int64),
                ("sval", C.c_char_p), ("slen", C.c_uint64)]


PRED_OPS = {"eq": 0, "neq": 1, "lt": 2, "gt": 3, "leq": 4, "geq": 5, "like": 6, "not_like": 7,
            "is_null": 8, "is_not_null": 9, "and": 16, "or": 17, "not": 18}


class hj_node(C.Structure):
    _fields_ = [("kind", C.c_int32), ("build_left", C.c_int32),
                ("base_table_id", C.c_uint64), ("left", C.c_uint64), ("right", C.c_uint64),
                ("left_attr", C.c_uint64), ("right_attr", C.c_uint64),
                ("num_output_attrs", C.c_uint64), ("output_attrs", C.POINTER(hj_attr)),
                ("num_filters", C.c_uint64), ("filters", C.POINTER(hj_filter)),
                ("num_preds", C.c_uint64), ("preds", C.POINTER(hj_pred))]


class hj_plan(C.Structure):
    _fields_ = [("num_nodes", C.c_uint64), ("nodes", C.POINTER(hj_node)),
                ("num_inputs", C.c_uint64), ("inputs", C.POINTER(hj_table)),
                ("root", C.c_uint64)]


class hj_out_column(C.Structure):
    _fields_ = [("type", C.c_int32), ("num_pages", C.c_uint32), ("pages", C.c_void_p)]


class hj_result(C.Structure):
    _fields_ = [("num_rows", C.c_uint64), ("num_columns", C.c_uint32),
                ("reserved", C.c_uint32), ("columns", C.POINTER(hj_out_column))]


class hj_batch_params(C.Structure):
    _fields_ = [("probe_item_groups", C.c_uint32), ("insert_item_rows", C.c_uint32),
                ("radix_bits", C.c_uint32), ("load_shift", C.c_uint32),
                ("large_build_rows", C.c_uint32), ("large_load_shift", C.c_uint32),
                ("unroll", C.c_uint32), ("coop", C.c_uint32), ("use_graph", C.c_uint32),
                ("plain_probe_rows", C.c_uint32), ("plain_build_rows", C.c_uint32),
                ("reserved", C.c_uint32 * 1)]


P = C.POINTER
_SIGS = {
    "hj_context_create": (C.c_int, [C.c_uint64, P(C.c_void_p)]),
    "hj_context_destroy": (C.c_int, [C.c_void_p]),
    "hj_last_error": (C.c_char_p, [C.c_void_p]),
    "hj_global_error": (C.c_char_p, []),
    "hj_hash_algo": (C.c_char_p, []),
    "hj_execute": (C.c_int, [C.c_void_p, P(hj_plan), P(hj_result)]),
    "hj_result_free": (None, [P(hj_result)]),
    "hj_table_upload": (C.c_int, [C.c_void_p, P(hj_table), P(C.c_void_p)]),
    "hj_table_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hj_table_num_rows": (C.c_uint64, [C.c_void_p]),
    "hj_execute_device": (C.c_int, [C.c_void_p, P(hj_plan), P(C.c_void_p), P(C.c_void_p)]),
    "hj_dresult_num_rows": (C.c_uint64, [C.c_void_p]),
    "hj_dresult_to_host": (C.c_int, [C.c_void_p, C.c_void_p, P(hj_result)]),
    "hj_dresult_column": (C.c_int, [C.c_void_p, C.c_uint32, P(C.c_int32), P(C.c_void_p),
                                    P(C.c_void_p), P(C.c_void_p)]),
    "hj_dresult_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hj_context_stream": (C.c_void_p, [C.c_void_p]),
    "hj_synchronize": (C.c_int, [C.c_void_p]),
    "hj_encode_int32": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, P(hj_out_column)]),
    "hj_encode_varchar": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                    P(hj_out_column)]),
    "hj_out_column_free": (None, [P(hj_out_column)]),
    "hj_column_num_rows": (C.c_uint64, [P(hj_column)]),
    "hj_column_num_chars": (C.c_uint64, [P(hj_column)]),
    "hj_decode_int32": (C.c_int, [P(hj_column), C.c_void_p, C.c_void_p]),
    "hj_decode_varchar": (C.c_int, [P(hj_column), C.c_void_p, C.c_void_p, C.c_void_p]),
    "hj_directory_build": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, P(C.c_void_p)]),
    "hj_directory_probe": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                     C.c_void_p, C.c_void_p]),
    "hj_directory_probe_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                             C.c_void_p, C.c_void_p]),
    "hj_directory_probe_pairs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                           P(C.c_void_p), P(C.c_uint64)]),
    "hj_directory_info": (C.c_int, [C.c_void_p, P(C.c_uint64), P(C.c_uint64), P(C.c_uint64)]),
    "hj_directory_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hj_free": (None, [C.c_void_p]),
    "hj_cache_open": (C.c_int, [C.c_char_p, P(C.c_void_p)]),
    "hj_cache_table": (P(hj_table), [C.c_void_p]),
    "hj_cache_close": (None, [C.c_void_p]),
    "hj_cache_write": (C.c_int, [P(hj_table), C.c_char_p]),
    "hj_table_upload_cache": (C.c_int, [C.c_void_p, C.c_char_p, P(C.c_void_p)]),
    "hj_server_create": (C.c_int, [C.c_uint32, C.c_uint64, P(C.c_void_p)]),
    "hj_server_destroy": (C.c_int, [C.c_void_p]),
    "hj_server_submit": (C.c_int, [C.c_void_p, P(hj_plan), P(C.c_void_p), P(C.c_uint64)]),
    "hj_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, P(hj_result), P(C.c_uint64),
                                 P(C.c_uint64)]),
    "hj_server_slots": (C.c_uint32, [C.c_void_p]),
    "hj_server_last_error": (C.c_char_p, [C.c_void_p]),
    "hj_batch_create": (C.c_int, [C.c_void_p, C.c_uint32, P(P(hj_plan)), P(P(C.c_void_p)),
                                  P(hj_batch_params), P(C.c_void_p)]),
    "hj_batch_run": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hj_batch_size": (C.c_uint32, [C.c_void_p]),
    "hj_batch_info": (C.c_int, [C.c_void_p, C.c_uint32, P(C.c_int32), P(C.c_uint64)]),
    "hj_batch_result": (C.c_int, [C.c_void_p, C.c_uint32, P(C.c_void_p)]),
    "hj_batch_result_to_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, P(hj_result)]),
    "hj_batch_device_bytes": (C.c_uint64, [C.c_void_p]),
    "hj_batch_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hj_set_option": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64]),
    "hj_dir_cache_stats": (C.c_int, [C.c_void_p, P(C.c_uint64), P(C.c_uint64), P(C.c_uint64),
                                     P(C.c_uint64)]),
    "hj_server_set_option": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64]),
    "hj_server_cache_stats": (C.c_int, [C.c_void_p, P(C.c_uint64), P(C.c_uint64), P(C.c_uint64),
                                        P(C.c_uint64)]),
    "hj_profile_enable": (C.c_int, [C.c_void_p, C.c_char_p]),
    "hj_profile_report": (C.c_int, [C.c_void_p, C.c_char_p, C.c_uint64]),
}

EXPORTED = tuple(_SIGS)

_lib = None


class EngineMissing(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Load the engine library. Raises EngineMissing when it has not been built: there is no
    CPU fallback for the engine."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise EngineMissing(
            f"{path} is missing: build it with `python hash-join-optimizer_amd/build.py`")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib
