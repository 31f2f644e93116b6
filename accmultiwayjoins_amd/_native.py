"""ctypes declarations of libamj's C ABI (include/amj.h) and libamj_gen (include/amj_gen.h).

The product path is the HIP library; loading fails loudly when it has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libamj.so")
GEN_PATH = os.path.join(_HERE, "libamj_gen.so")

AMJ_MAX_ATTRS = 8
AMJ_MAX_BUILD_TABLES = 7

AMJ_OK = 0
STATUS = {
    0: "AMJ_OK", -1: "AMJ_ERR_ARG", -2: "AMJ_ERR_HIP", -3: "AMJ_ERR_NOMEM",
    -4: "AMJ_ERR_NODEV", -5: "AMJ_ERR_STATE",
}
PATHS = {1: "plain", 2: "fib", 3: "ws", 4: "dro", 5: "lftj-dfs", 6: "lftj-bfs", 7: "ws-dro", 8: "batch",
         9: "ooc", 10: "lftj-ooc", 11: "tpt", 12: "open-addr", 13: "pw"}

QUERY = {"NORMAL": 0, "TRI": 1, "FOUR": 2}
AMJ_REL_HOST = 2
ALGO = {"BFS": 0, "DFS": 1}


class AmjRel(C.Structure):
    _fields_ = [
        ("num_cols", C.c_uint32),
        ("rows", C.c_uint64),
        ("cols", C.POINTER(C.c_void_p)),
        ("attr_ids", C.POINTER(C.c_uint32)),
        ("on_device", C.c_int),
    ]


class AmjOpts(C.Structure):
    _fields_ = [
        ("ooc", C.c_int),
        ("ws", C.c_int),
        ("dro", C.c_int),
        ("fib", C.c_int),
        ("materialize", C.c_int),
        ("checksum", C.c_int),
        ("hbm_budget_bytes", C.c_uint64),
        ("ooc_slice_rows", C.c_uint64),
        ("out_chunk_rows", C.c_uint64),
        ("dro_capacity_rows", C.c_uint64),
        ("ws_split_threshold", C.c_uint32),
        ("fib_suffix", C.c_int),
        ("probe_order", C.c_int),
        ("engine", C.c_uint32),
        ("ws_stall_ms", C.c_uint32),
        ("ws_claim_units", C.c_uint32),
        ("reserved", C.c_uint32 * 2),
    ]


class AmjResult(C.Structure):
    _fields_ = [
        ("count", C.c_uint64),
        ("checksum", C.c_uint64),
        ("num_cols", C.c_uint32),
        ("cols", C.c_void_p * AMJ_MAX_ATTRS),
        ("rows_written", C.c_uint64),
        ("attr_ids", C.c_uint32 * AMJ_MAX_ATTRS),
        ("build_ms", C.c_double),
        ("probe_ms", C.c_double),
        ("probe_kernel_ms", C.c_double),
        ("probe_kernel_launches", C.c_uint64),
        ("path", C.c_uint32),
        ("reserved", C.c_uint32 * 3),
        ("detail", C.c_uint32),
        ("passes", C.c_uint32),
    ]


# (name, restype, argtypes) of every exported C-ABI symbol.
SIGNATURES = [
    ("amj_version", C.c_char_p, []),
    ("amj_default_opts", None, [C.POINTER(AmjOpts)]),
    ("amj_ctx_create", C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    ("amj_ctx_destroy", None, [C.c_void_p]),
    ("amj_last_error", C.c_char_p, [C.c_void_p]),
    ("amj_sync", C.c_int, [C.c_void_p]),
    ("amj_host_alloc", C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    ("amj_host_free", None, [C.c_void_p]),
    ("amj_memory_info", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("amj_device_info", C.c_int,
     [C.c_void_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_uint64), C.POINTER(C.c_int)]),
    ("amj_relation_create", C.c_int, [C.c_void_p, C.POINTER(AmjRel), C.POINTER(C.c_void_p)]),
    ("amj_relation_destroy", None, [C.c_void_p, C.c_void_p]),
    ("amj_hash_table_build", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.POINTER(C.c_void_p)]),
    ("amj_hash_table_destroy", None, [C.c_void_p, C.c_void_p]),
    ("amj_hash_table_info", C.c_int,
     [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
      C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    ("amj_execute", C.c_int,
     [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(AmjOpts),
      C.POINTER(AmjResult)]),
    ("amj_device_alloc", C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    ("amj_device_free", None, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("amj_memcpy", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("amj_gen_device_uniform", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int64, C.c_int64, C.c_uint64, C.c_uint64]),
    ("amj_gen_device_zipf", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_double, C.c_uint64, C.c_uint64]),
    ("amj_gen_device_rmat", C.c_int,
     [C.c_void_p, C.c_int, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint64,
      C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64)]),
    ("amj_execute_noalloc", C.c_int,
     [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(AmjOpts),
      C.POINTER(AmjResult)]),
    ("amj_chain_join", C.c_int,
     [C.c_void_p, C.POINTER(AmjRel), C.c_uint32, C.POINTER(AmjOpts), C.POINTER(AmjResult)]),
    ("amj_pattern_join", C.c_int,
     [C.c_void_p, C.POINTER(AmjRel), C.c_int, C.POINTER(AmjOpts), C.POINTER(AmjResult)]),
    ("amj_execute_pattern", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(AmjOpts), C.POINTER(AmjResult)]),
    ("amj_result_free", C.c_int, [C.c_void_p, C.POINTER(AmjResult)]),
    ("amj_result_copy_col", C.c_int,
     [C.c_void_p, C.POINTER(AmjResult), C.c_uint32, C.c_uint64, C.c_uint64, C.c_void_p]),
    ("amj_autotune", C.c_int,
     [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(AmjOpts),
      C.POINTER(C.c_double)]),
    ("amj_autotune_forget", None, [C.c_void_p]),
    ("amj_batch_create", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("amj_batch_destroy", None, [C.c_void_p]),
    ("amj_batch_add", C.c_int,
     [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(AmjOpts),
      C.POINTER(C.c_uint32)]),
    ("amj_batch_add_pattern", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(AmjOpts), C.POINTER(C.c_uint32)]),
    ("amj_batch_set_budget", C.c_int, [C.c_void_p, C.c_uint64]),
    ("amj_batch_run", C.c_int, [C.c_void_p]),
    ("amj_batch_capture", C.c_int, [C.c_void_p, C.c_uint64]),
    ("amj_batch_replay", C.c_int, [C.c_void_p]),
    ("amj_batch_result", C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(AmjResult)]),
    ("amj_batch_info", C.c_int,
     [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    ("amj_server_create", C.c_int,
     [C.c_int, C.c_uint64, C.c_uint32, C.c_int, C.POINTER(C.c_void_p)]),
    ("amj_server_destroy", None, [C.c_void_p]),
    ("amj_server_last_error", C.c_char_p, [C.c_void_p]),
    ("amj_server_add_relation", C.c_int, [C.c_void_p, C.POINTER(AmjRel), C.POINTER(C.c_uint32)]),
    ("amj_server_submit", C.c_int,
     [C.c_void_p, C.POINTER(C.c_uint32), C.c_uint32, C.c_int, C.POINTER(C.c_uint64)]),
    ("amj_server_submit_pattern", C.c_int,
     [C.c_void_p, C.c_uint32, C.c_int, C.c_int, C.POINTER(C.c_uint64)]),
    ("amj_server_wait", C.c_int,
     [C.c_void_p, C.c_uint64, C.POINTER(AmjResult), C.POINTER(C.c_double)]),
    ("amj_server_wait_ex", C.c_int,
     [C.c_void_p, C.c_uint64, C.POINTER(AmjResult), C.POINTER(C.c_double),
      C.POINTER(C.c_double)]),
    ("amj_server_set_budget", C.c_int, [C.c_void_p, C.c_uint64]),
    ("amj_server_set_autotune", C.c_int, [C.c_void_p, C.c_uint64]),
    ("amj_server_tuning", C.c_int,
     [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("amj_server_set_triejoin", C.c_int, [C.c_void_p, C.c_uint64]),
    ("amj_server_triejoin_runs", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    ("amj_server_memory", C.c_int,
     [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("amj_server_release", C.c_int, [C.c_void_p, C.POINTER(AmjResult)]),
    ("amj_server_copy_col", C.c_int,
     [C.c_void_p, C.POINTER(AmjResult), C.c_uint32, C.c_uint64, C.c_uint64, C.c_void_p]),
    ("amj_server_stats", C.c_int,
     [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
      C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    ("alftj_trie_build", C.c_int, [C.c_void_p, C.POINTER(AmjRel), C.POINTER(C.c_void_p)]),
    ("alftj_trie_build_resident", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    ("alftj_trie_destroy", None, [C.c_void_p, C.c_void_p]),
    ("alftj_trie_info", C.c_int,
     [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
      C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
    ("alftj_trie_copy_level", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    ("alftj_run", C.c_int,
     [C.c_void_p, C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(C.c_uint32), C.c_uint32, C.c_int,
      C.POINTER(AmjOpts), C.POINTER(AmjResult)]),
    ("alftj_run_aliased", C.c_int,
     [C.c_void_p, C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(C.c_uint32),
      C.POINTER(C.c_uint32), C.c_uint32, C.c_int, C.POINTER(AmjOpts), C.POINTER(AmjResult)]),
    ("alftj_pattern", C.c_int,
     [C.c_void_p, C.POINTER(AmjRel), C.c_int, C.c_int, C.POINTER(AmjOpts), C.POINTER(AmjResult)]),
    ("alftj_trie_from_csr", C.c_int,
     [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
      C.POINTER(C.c_void_p)]),
    ("alftj_trie_save", C.c_int, [C.c_void_p, C.c_void_p, C.c_char_p]),
    ("alftj_trie_offload", C.c_int, [C.c_void_p, C.c_void_p]),
    ("alftj_trie_load", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    ("amj_db_parse_name", C.c_int,
     [C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
      C.c_uint32]),
    ("amj_db_read", C.c_int, [C.c_char_p, C.c_uint64, C.c_uint32, C.POINTER(C.c_void_p)]),
    ("amj_db_write", C.c_int, [C.c_char_p, C.POINTER(AmjRel), C.c_char_p, C.c_size_t]),
    ("amj_file_error", C.c_char_p, []),
    ("amj_result_md5", C.c_int, [C.c_void_p, C.POINTER(AmjResult), C.c_char_p]),
    ("amj_res_stat_md5", C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.c_uint64, C.c_char_p]),
]

GEN_SIGNATURES = [
    ("amj_gen_uniform", None,
     [C.c_void_p, C.c_uint64, C.c_int64, C.c_int64, C.c_uint64, C.c_uint64]),
    ("amj_gen_zipf", C.c_int,
     [C.c_void_p, C.c_uint64, C.c_uint64, C.c_double, C.c_uint64, C.c_uint64]),
    ("amj_gen_zipf_mapped", C.c_int,
     [C.c_void_p, C.c_uint64, C.c_uint64, C.c_double, C.c_uint64, C.c_uint64, C.c_int]),
    ("amj_gen_unique_column", C.c_int,
     [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64]),
    ("amj_gen_unique_pairs", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
      C.c_uint64, C.c_int]),
    ("amj_gen_zipf_pairs", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
      C.c_uint64, C.c_int, C.c_double, C.c_int]),
    ("amj_gen_seq_pairs", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int64, C.c_int64]),
    ("amj_gen_rmat", None,
     [C.c_void_p, C.c_void_p, C.c_int, C.c_uint64, C.c_double, C.c_double, C.c_double,
      C.c_uint64]),
    ("amj_graph_orient", C.c_uint64,
     [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p]),
]

_lib = None
_gen = None


def _bind(lib, sigs):
    for name, res, args in sigs:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def lib():
    """The HIP library. Raises if it has not been built (there is no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with "
                "`make -C accmultiwayjoins_amd/csrc` (or __graft_entry__.build())")
        _lib = _bind(C.CDLL(LIB_PATH), SIGNATURES)
    return _lib


def gen():
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_PATH):
            raise RuntimeError(f"{GEN_PATH} is missing: run `make -C accmultiwayjoins_amd/csrc`")
        _gen = _bind(C.CDLL(GEN_PATH), GEN_SIGNATURES)
    return _gen
