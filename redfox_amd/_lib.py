"""Loader for librfx_hip.so, the gfx950 relational-algebra operator library.

The shared library is built in-tree (``make -C redfox_amd/csrc``) and loaded
from this package directory.  There is no fallback: if the library is missing
or the HIP runtime cannot be loaded, importing :mod:`redfox_amd` raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_uint32, c_uint64, c_void_p

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librfx_hip.so")

P = c_void_p
U64 = c_uint64
U32 = c_uint32
I = c_int
U64P = POINTER(c_uint64)

# name -> (restype, argtypes).  Mirrors include/rfx_ra.h.
SIGNATURES = {
    "rfx_strerror": (c_char_p, [I]),
    "rfx_last_error": (c_char_p, []),
    "rfx_version": (c_char_p, []),
    "rfx_ctx_create": (I, [POINTER(c_void_p), I, P]),
    "rfx_ctx_destroy": (I, [P]),
    "rfx_ctx_set_stream": (I, [P, P]),
    "rfx_ctx_stream": (P, [P]),
    "rfx_ctx_reserve": (I, [P, U64]),
    "rfx_ctx_workspace_bytes": (U64, [P]),
    "rfx_ctx_tune": (I, [P, c_char_p, I]),
    "rfx_ctx_tuning": (I, [P, c_char_p, POINTER(c_int)]),
    "rfx_ctx_synchronize": (I, [P]),
    "rfx_default_ctx": (P, []),
    "rfx_device_count": (I, [POINTER(c_int)]),
    "rfx_set_device": (I, [I]),
    "rfx_malloc": (I, [POINTER(c_void_p), U64]),
    "rfx_free": (I, [P]),
    "rfx_memcpy_htod": (I, [P, P, U64]),
    "rfx_memcpy_dtoh": (I, [P, P, U64]),
    "rfx_memcpy_dtod": (I, [P, P, U64]),
    "rfx_memset": (I, [P, I, U64]),
    "rfx_device_synchronize": (I, []),
    "rfx_mem_info": (I, [U64P, U64P]),
    "rfx_ctx_profile": (I, [P, I]),
    "rfx_ctx_profile_read": (I, [P, c_char_p, U64P, POINTER(c_double)]),
    "rfx_event_create": (I, [POINTER(c_void_p)]),
    "rfx_event_destroy": (I, [P]),
    "rfx_event_record": (I, [P, P]),
    "rfx_event_elapsed": (I, [POINTER(c_float), P, P]),
    "rfx_gen_u32": (I, [P, P, U64, U64]),
    "rfx_gen_u64": (I, [P, P, U64, U64]),
    "rfx_gen_iota_u32": (I, [P, P, U64]),
    "rfx_gen_zipf_u64": (I, [P, P, U64, U64, c_double, U64]),
    "rfx_gen_lineitem": (I, [P, P, P, P, P, P, P, P, U64, U64]),
    "rfx_stream_create": (I, [POINTER(c_void_p)]),
    "rfx_stream_destroy": (I, [P]),
    "rfx_stream_synchronize": (I, [P]),
    "rfx_graph_begin": (I, [P]),
    "rfx_graph_end": (I, [P, POINTER(c_void_p)]),
    "rfx_graph_launch": (I, [P, P]),
    "rfx_graph_destroy": (I, [P]),
    "rfx_copy": (I, [P, P, P, U64]),
    "rfx_segment_keys": (I, [P, P, P, P, U32, U64, U32]),
    "rfx_gen_fk_batch": (I, [P, P, P, P, P, P, U32, U64, U64, c_double]),
    "rfx_sort_pairs": (I, [P, P, P, U64, I, I, I]),
    "rfx_sort_workspace": (U64, [U64, I, I, I]),
    "rfx_join_find_bounds": (I, [P, P, P, P, P, U64, P, U64, I]),
    "rfx_join_find_bounds_noalloc": (I, [P, P, P, P, P, U64, P, U64, I]),
    "rfx_join_main": (I, [P, P, P, P, U64, P, P, U64]),
    "rfx_join_gather": (I, [P, P, I, P, I, P, I, P, I, P, P, P, U64, U32, U32]),
    "rfx_join_expand_gather": (I, [P, P, I, P, I, P, I, P, I, P, P, U64, P, U64, U32, U32]),
    "rfx_join_expand_project": (I, [P, P, P, U64, P, U64, P, I, P, I, P, I, P, I, P, I, I, P]),
    "rfx_reduce_by_key_f64": (I, [P, P, P, P, P, P, P, U64, I, I]),
    "rfx_reduce_by_key_u32": (I, [P, P, P, P, P, P, P, U64, I]),
    "rfx_reduce_count_by_key_f64": (I, [P, P, P, P, U64, P, P, P, P, U64, I, I]),
    "rfx_count_by_key": (I, [P, P, P, P, P, P, U64, I]),
    "rfx_total_f64": (I, [P, P, P, U64]),
    "rfx_total_u32": (I, [P, P, P, U64]),
    "rfx_max_u64": (I, [P, P, P, U64]),
    "rfx_set_op": (I, [P, I, P, P, P, U64, P, U64, I]),
    "rfx_unique": (I, [P, P, P, P, U64, I]),
    "rfx_select": (I, [P, P, P, P, U64, I, U32, U32, I, U64]),
    "rfx_project": (I, [P, P, I, P, I, U64, I, P, P, P]),
    "rfx_project_sized": (I, [P, P, I, P, I, P, U64, I, P, P, P]),
    "rfx_server_create": (I, [POINTER(c_void_p), I, I, U64]),
    "rfx_server_destroy": (I, [P]),
    "rfx_server_add_schema": (I, [P, I, P, I, POINTER(c_uint32)]),
    "rfx_server_request_bytes": (I, [P, P, U64P]),
    "rfx_server_run": (I, [P, P, U64, I, P]),
    "rfx_split": (I, [P, P, I, P, I, P, I, U64, U32]),
    "rfx_merge": (I, [P, P, I, P, I, P, I, U64, U32]),
    "rfx_generate_ones": (I, [P, P, U64]),
    "rfx_q1": (I, [P, P, P, P, P, P, P, P, P, U64, c_int]),
    "set_b40csortpair": (None, [P, P, U64, U64, U64]),
    "set_b40csortkey": (None, [P, U64, U64]),
    "set_mgpusortpair": (None, [P, P, U64, U64, U64]),
    "set_mgpusortkey": (None, [P, U64, U64]),
    "mgpu_join_find_bounds": (None, [P, P, P, P, U64, P, U64, U64]),
    "mgpu_join_main": (None, [P, P, U64, P, P, U64]),
    "join_get_result_size": (None, [P, P, U64]),
    "mgpu_join_gather": (None, [P, P, P, P, P, P, P, P, U64]),
    "set_reduce": (None, [P, P, P, P, P, P, U64]),
    "set_count": (None, [P, P, P, P, P, P, U64]),
    "set_single_reduce": (None, [P, P, U64, U64]),
    "set_union": (None, [P, P, P, U64, P, U64, U64]),
    "set_difference": (None, [P, P, P, U64, P, U64, U64]),
    "set_intersection": (None, [P, P, P, U64, P, U64, U64]),
    "set_unique": (None, [P, P, U64]),
    "rfx_harmony_call": (I, [c_char_p, POINTER(c_uint64), U32]),
    "rfx_harmony_launch": (I, [c_char_p, I, c_char_p, U64, POINTER(c_uint64), U32]),
    "rfx_mapfilter": (I, [P, P, I, U64, P, P, I, P]),
    "rfx_tbl_parse": (I, [P, P, U64, ctypes.c_char, P, I, P, U64, P, P]),
    "rfx_join_find_bounds_string": (I, [P, P, P, P, P, U64, P, U64]),
    "rfx_sort_strings": (I, [P, P, P, U64]),
    "rfx_sort_strings_rows": (I, [P, P, P, P, U64]),
    "rfx_string_groups": (I, [P, P, U64, P, P, P]),
    "rfx_hir_load": (I, [POINTER(c_void_p), P, U64]),
    "rfx_hir_destroy": (None, [P]),
    "rfx_hir_info": (I, [P, P]),
    "rfx_hir_schedule": (I, [P, U32, c_char_p, U64]),
    "rfx_hir_set_input": (I, [P, U32, P, U64]),
    "rfx_hir_run": (I, [P]),
    "rfx_hir_variable": (I, [P, U32, POINTER(c_void_p), U64P]),
    "rfx_hir_read": (I, [P, U32, P, U64, U64P]),
    "rfx_hir_run_tests": (I, [P, POINTER(c_uint32), POINTER(c_uint32), c_char_p, U64]),
}


class RfxError(RuntimeError):
    """A librfx_hip call returned a negative status."""

    def __init__(self, fn: str, status: int, detail: str):
        super().__init__(f"{fn} failed: status {status}: {detail}")
        self.status = status


def load(path: str = LIB_PATH, strict: bool = True) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build it with `make -C redfox_amd/csrc` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    missing = []
    for name, (res, args) in SIGNATURES.items():
        try:
            fn = getattr(lib, name)
        except AttributeError:
            missing.append(name)
            continue
        fn.restype = res
        fn.argtypes = args
    if missing and strict:
        raise ImportError(f"{path} lacks symbols: {', '.join(missing)}")
    return lib
