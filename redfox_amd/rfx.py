"""Python host interface to librfx_hip.

Mirrors the reference's operator interface for the RA hot path (the
``redfox::`` functions behind Harmony's binary kernels and the PTX device
templates) over device buffers owned by this module.  Every call goes through
the C-ABI in include/rfx_ra.h; a failing call raises :class:`RfxError` where
the reference would abort.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_double, c_float, c_int, c_uint64, c_void_p
from typing import Optional

import numpy as np

from ._lib import RfxError, load

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        # Symbols that are absent raise AttributeError at first use;
        # tests/test_abi.py checks that every declared entry point exists.
        _lib = load(strict=False)
    return _lib


def _check(fn: str, status: int) -> None:
    if status != 0:
        L = lib()
        raise RfxError(fn, status, (L.rfx_last_error() or b"").decode(errors="replace"))


def device_count() -> int:
    n = c_int(0)
    _check("rfx_device_count", lib().rfx_device_count(byref(n)))
    return n.value


# ---------------------------------------------------------------------------
# Device memory
# ---------------------------------------------------------------------------
class DeviceBuffer:
    """A hipMalloc'd byte range, freed on close()/garbage collection."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = c_void_p()
        _check("rfx_malloc", lib().rfx_malloc(byref(p), max(1, self.nbytes)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        buf.upload(arr)
        return buf

    def upload(self, arr: np.ndarray, offset: int = 0) -> None:
        arr = np.ascontiguousarray(arr)
        assert offset + arr.nbytes <= self.nbytes
        if arr.nbytes:
            _check("rfx_memcpy_htod",
                   lib().rfx_memcpy_htod(self.ptr + offset, arr.ctypes.data, arr.nbytes))

    def download(self, dtype, count: Optional[int] = None, offset: int = 0) -> np.ndarray:
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        if out.nbytes:
            _check("rfx_memcpy_dtoh",
                   lib().rfx_memcpy_dtoh(out.ctypes.data, self.ptr + offset, out.nbytes))
        return out

    def zero(self) -> None:
        _check("rfx_memset", lib().rfx_memset(self.ptr, 0, self.nbytes))

    def close(self) -> None:
        if getattr(self, "ptr", None):
            lib().rfx_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def u64_cell(value: int = 0) -> DeviceBuffer:
    """A device u64 holding `value` (a Harmony size variable)."""
    return DeviceBuffer.from_numpy(np.array([value], dtype=np.uint64))


def read_u64(buf: DeviceBuffer, index: int = 0) -> int:
    return int(buf.download(np.uint64, 1, offset=8 * index)[0])


# ---------------------------------------------------------------------------
# Context
# ---------------------------------------------------------------------------
class Context:
    """An rfx_ctx: one device, one stream, one cached workspace."""

    def __init__(self, device: int = 0, stream: Optional[int] = None):
        p = c_void_p()
        _check("rfx_ctx_create", lib().rfx_ctx_create(byref(p), device, stream))
        self.handle = p.value

    def synchronize(self) -> None:
        _check("rfx_ctx_synchronize", lib().rfx_ctx_synchronize(self.handle))

    def reserve(self, nbytes: int) -> None:
        _check("rfx_ctx_reserve", lib().rfx_ctx_reserve(self.handle, nbytes))

    def tune(self, knob: str, value: int) -> None:
        """Select a tile shape (rfx_ctx_tune); results do not depend on it."""
        _check("rfx_ctx_tune", lib().rfx_ctx_tune(self.handle, knob.encode(), int(value)))

    def tuning(self, knob: str) -> int:
        v = ctypes.c_int(0)
        _check("rfx_ctx_tuning", lib().rfx_ctx_tuning(self.handle, knob.encode(), byref(v)))
        return int(v.value)

    @property
    def workspace_bytes(self) -> int:
        return int(lib().rfx_ctx_workspace_bytes(self.handle))

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().rfx_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- data generation ---------------------------------------------------
    def gen_u32(self, buf: DeviceBuffer, n: int, seed: int) -> None:
        _check("rfx_gen_u32", lib().rfx_gen_u32(self.handle, buf.ptr, n, seed))

    def gen_u64(self, buf: DeviceBuffer, n: int, seed: int) -> None:
        _check("rfx_gen_u64", lib().rfx_gen_u64(self.handle, buf.ptr, n, seed))

    def iota_u32(self, buf: DeviceBuffer, n: int) -> None:
        _check("rfx_gen_iota_u32", lib().rfx_gen_iota_u32(self.handle, buf.ptr, n))

    def gen_zipf_u64(self, buf: DeviceBuffer, n: int, seed: int, s: float, n_max: int) -> None:
        _check("rfx_gen_zipf_u64", lib().rfx_gen_zipf_u64(self.handle, buf.ptr, n, seed, s, n_max))

    def gen_lineitem(self, cols, n: int, seed: int) -> None:
        _check("rfx_gen_lineitem", lib().rfx_gen_lineitem(
            self.handle, *[c.ptr for c in cols], n, seed))

    # -- streams, graphs, batching -------------------------------------------
    def copy(self, dst: DeviceBuffer, src: DeviceBuffer, nbytes: int, dst_off: int = 0,
             src_off: int = 0) -> None:
        _check("rfx_copy", lib().rfx_copy(self.handle, dst.ptr + dst_off, src.ptr + src_off, nbytes))

    def segment_keys(self, out, key, seg_offsets, nseg, n, key_bits) -> None:
        _check("rfx_segment_keys", lib().rfx_segment_keys(
            self.handle, out.ptr, key.ptr, seg_offsets.ptr, nseg, n, key_bits))

    def gen_fk_batch(self, fk, measure, row_off, dim_off, perm, nseg, n, seed, s) -> None:
        _check("rfx_gen_fk_batch", lib().rfx_gen_fk_batch(
            self.handle, fk.ptr, measure.ptr, row_off.ptr, dim_off.ptr, perm.ptr, nseg, n, seed, s))

    def graph_begin(self) -> None:
        _check("rfx_graph_begin", lib().rfx_graph_begin(self.handle))

    def graph_end(self) -> "Graph":
        g = c_void_p()
        _check("rfx_graph_end", lib().rfx_graph_end(self.handle, byref(g)))
        return Graph(self, g.value)

    # -- per-launch timing -------------------------------------------------
    def profile(self, enable: bool) -> None:
        _check("rfx_ctx_profile", lib().rfx_ctx_profile(self.handle, int(enable)))

    def profile_read(self, tag: str):
        """(launches, total milliseconds) recorded for `tag` since profile()."""
        n = c_uint64(0)
        ms = c_double(0.0)
        _check("rfx_ctx_profile_read",
               lib().rfx_ctx_profile_read(self.handle, tag.encode(), byref(n), byref(ms)))
        return int(n.value), float(ms.value)

    # -- sort (b40cSort.cu / ModernGPUSort.cu) ------------------------------
    def sort_pairs(self, keys: DeviceBuffer, vals: Optional[DeviceBuffer], n: int,
                   key_bytes: int, key_bits: int, val_bytes: int) -> None:
        _check("rfx_sort_pairs", lib().rfx_sort_pairs(
            self.handle, keys.ptr, vals.ptr if vals is not None else None, n,
            key_bytes, key_bits, val_bytes))

    # -- join (ModernGPUJoin.cu / ModernGPUJoin.h) --------------------------
    def join_find_bounds(self, lower_bound, left_count, total, left, nl, right, nr,
                         key_bytes) -> None:
        _check("rfx_join_find_bounds", lib().rfx_join_find_bounds(
            self.handle, lower_bound.ptr, left_count.ptr, total.ptr, left.ptr, nl,
            right.ptr, nr, key_bytes))

    def join_find_bounds_string(self, lower_bound, left_count, total, left, nl, right,
                                nr) -> None:
        """left/right: device arrays of string addresses (u64), each sorted."""
        _check("rfx_join_find_bounds_string", lib().rfx_join_find_bounds_string(
            self.handle, lower_bound.ptr, left_count.ptr, total.ptr, left.ptr, nl, right.ptr, nr))

    def sort_strings(self, addresses, scratch, n) -> None:
        """Sort n string addresses (u64) in place into compare_string2's
        order; scratch holds n + 1 u64."""
        _check("rfx_sort_strings", lib().rfx_sort_strings(self.handle, addresses.ptr,
                                                          scratch.ptr, n))

    def sort_strings_rows(self, addresses, rows, scratch, n) -> None:
        """Sort n row numbers (u32) stably by the string each row's address
        points to; scratch holds n + 1 u64."""
        _check("rfx_sort_strings_rows", lib().rfx_sort_strings_rows(
            self.handle, addresses.ptr, rows.ptr, scratch.ptr, n))

    def string_groups(self, sorted_addresses, n, group, scratch, n_groups) -> None:
        """Group numbers (u32) of the runs of equal strings in a sorted
        address array; scratch holds n int32, n_groups is a device u64."""
        _check("rfx_string_groups", lib().rfx_string_groups(
            self.handle, sorted_addresses.ptr, n, group.ptr, scratch.ptr, n_groups.ptr))

    def join_main(self, left_idx, right_idx, total, total_cap, lower_bound, left_count,
                  nl) -> None:
        _check("rfx_join_main", lib().rfx_join_main(
            self.handle, left_idx.ptr, right_idx.ptr, total.ptr, total_cap,
            lower_bound.ptr, left_count.ptr, nl))

    def join_gather(self, out, out_bytes, lkey, lkey_bytes, lval, lval_bytes, rval,
                    rval_bytes, left_idx, right_idx, total, total_cap, key_shift,
                    lval_shift) -> None:
        _check("rfx_join_gather", lib().rfx_join_gather(
            self.handle, out.ptr, out_bytes, lkey.ptr, lkey_bytes,
            lval.ptr if lval is not None else None, lval_bytes,
            rval.ptr if rval is not None else None, rval_bytes, left_idx.ptr,
            right_idx.ptr, total.ptr, total_cap, key_shift, lval_shift))

    def join_expand_gather(self, out, out_bytes, lkey, lkey_bytes, lval, lval_bytes,
                           rval, rval_bytes, lower_bound, left_count, nl, total,
                           total_cap, key_shift, lval_shift) -> None:
        _check("rfx_join_expand_gather", lib().rfx_join_expand_gather(
            self.handle, out.ptr, out_bytes, lkey.ptr, lkey_bytes,
            lval.ptr if lval is not None else None, lval_bytes,
            rval.ptr if rval is not None else None, rval_bytes, lower_bound.ptr,
            left_count.ptr, nl, total.ptr, total_cap, key_shift, lval_shift))

    def join_expand_project(self, lb, lc, nl, total, total_cap, lkey, lkey_bytes, lval, lval_bytes,
                            rval, rval_bytes, out0, out0_bytes, out1, out1_bytes, fields) -> None:
        """fields: list of (column, source, in_shift, bits, out_shift); source 0 = left
        key, 1 = left value, 2 = right value."""
        f = np.ascontiguousarray(np.array(fields, dtype=np.uint32).reshape(-1))
        _check("rfx_join_expand_project", lib().rfx_join_expand_project(
            self.handle, lb.ptr, lc.ptr, nl, total.ptr, total_cap,
            lkey.ptr, lkey_bytes, lval.ptr if lval is not None else None, lval_bytes,
            rval.ptr if rval is not None else None, rval_bytes, out0.ptr, out0_bytes,
            out1.ptr if out1 is not None else None, out1_bytes, len(fields), f.ctypes.data))

    # -- reductions (Reduce.cu) --------------------------------------------
    def reduce_by_key_f64(self, out_keys, out_vals, n_out, keys, vals, n, n_cap,
                          key_bytes, op=0) -> None:
        _check("rfx_reduce_by_key_f64", lib().rfx_reduce_by_key_f64(
            self.handle, out_keys.ptr, out_vals.ptr, n_out.ptr, keys.ptr, vals.ptr,
            n.ptr, n_cap, key_bytes, op))

    def reduce_by_key_u32(self, out_keys, out_vals, n_out, keys, vals, n, n_cap,
                          key_bytes) -> None:
        _check("rfx_reduce_by_key_u32", lib().rfx_reduce_by_key_u32(
            self.handle, out_keys.ptr, out_vals.ptr, n_out.ptr, keys.ptr, vals.ptr,
            n.ptr, n_cap, key_bytes))

    def count_by_key(self, out_keys, out_counts, n_out, keys, n, n_cap, key_bytes) -> None:
        _check("rfx_count_by_key", lib().rfx_count_by_key(
            self.handle, out_keys.ptr, out_counts.ptr, n_out.ptr, keys.ptr, n.ptr,
            n_cap, key_bytes))

    def reduce_count_by_key_f64(self, out_keys, out_vals, out_counts, n_out, keys, vals, n,
                                n_cap, key_bytes, op=0, out_cap=None) -> None:
        """out_cap: rows the outputs hold (default n_cap); runs past it are
        counted in n_out but not written."""
        cap = n_cap if out_cap is None else out_cap
        _check("rfx_reduce_count_by_key_f64", lib().rfx_reduce_count_by_key_f64(
            self.handle, out_keys.ptr, out_vals.ptr, out_counts.ptr, cap, n_out.ptr, keys.ptr,
            vals.ptr, n.ptr, n_cap, key_bytes, op))

    def total_f64(self, result, v, n) -> None:
        _check("rfx_total_f64", lib().rfx_total_f64(self.handle, result.ptr, v.ptr, n))

    def total_u32(self, result, v, n) -> None:
        _check("rfx_total_u32", lib().rfx_total_u32(self.handle, result.ptr, v.ptr, n))

    def max_u64(self, result, v, n) -> None:
        _check("rfx_max_u64", lib().rfx_max_u64(self.handle, result.ptr, v.ptr, n))

    # -- set operations (Union.cu / Difference.cu / Unique.cu) --------------
    def set_op(self, op, out, n_out, a, na, b, nb, elem_bytes) -> None:
        _check("rfx_set_op", lib().rfx_set_op(
            self.handle, op, out.ptr, n_out.ptr, a.ptr, na, b.ptr, nb, elem_bytes))

    def unique(self, out, n_out, data, n, elem_bytes) -> None:
        _check("rfx_unique", lib().rfx_unique(
            self.handle, out.ptr, n_out.ptr, data.ptr, n, elem_bytes))

    # -- select / project / split / merge ----------------------------------
    def select(self, out, n_out, inp, n, tuple_bytes, field_shift, field_bits, op,
               constant) -> None:
        _check("rfx_select", lib().rfx_select(
            self.handle, out.ptr, n_out.ptr, inp.ptr, n, tuple_bytes, field_shift,
            field_bits, op, constant & 0xFFFFFFFFFFFFFFFF))

    def project(self, out, out_bytes, inp, in_bytes, n, in_shift, bits, out_shift) -> None:
        a = np.ascontiguousarray(in_shift, dtype=np.uint32)
        b = np.ascontiguousarray(bits, dtype=np.uint32)
        c = np.ascontiguousarray(out_shift, dtype=np.uint32)
        _check("rfx_project", lib().rfx_project(
            self.handle, out.ptr, out_bytes, inp.ptr, in_bytes, n, len(a),
            a.ctypes.data, b.ctypes.data, c.ctypes.data))

    def split(self, key, key_bytes, val, val_bytes, inp, in_bytes, n, key_shift) -> None:
        _check("rfx_split", lib().rfx_split(
            self.handle, key.ptr, key_bytes, val.ptr if val is not None else None,
            val_bytes, inp.ptr, in_bytes, n, key_shift))

    def merge(self, out, out_bytes, key, key_bytes, val, val_bytes, n, key_shift) -> None:
        _check("rfx_merge", lib().rfx_merge(
            self.handle, out.ptr, out_bytes, key.ptr, key_bytes, val.ptr, val_bytes, n,
            key_shift))

    def generate_ones(self, out, n) -> None:
        _check("rfx_generate_ones", lib().rfx_generate_ones(self.handle, out.ptr, n))

    # -- fused Q1 -----------------------------------------------------------
    def q1(self, out, shipdate, flag, status, qty, price, disc, tax, n, cutoff) -> None:
        _check("rfx_q1", lib().rfx_q1(
            self.handle, out.ptr, shipdate.ptr, flag.ptr, status.ptr, qty.ptr, price.ptr,
            disc.ptr, tax.ptr, n, cutoff))


class Stream:
    """A non-blocking HIP stream (for contexts that capture graphs)."""

    def __init__(self):
        p = c_void_p()
        _check("rfx_stream_create", lib().rfx_stream_create(byref(p)))
        self.handle = p.value

    def synchronize(self) -> None:
        _check("rfx_stream_synchronize", lib().rfx_stream_synchronize(self.handle))

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().rfx_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """An instantiated HIP graph captured from a context's stream."""

    def __init__(self, ctx: "Context", handle: int):
        self.ctx = ctx
        self.handle = handle

    def launch(self) -> None:
        _check("rfx_graph_launch", lib().rfx_graph_launch(self.ctx.handle, self.handle))

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().rfx_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Timer:
    """HIP events recorded on a context's stream."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        a, b = c_void_p(), c_void_p()
        _check("rfx_event_create", lib().rfx_event_create(byref(a)))
        _check("rfx_event_create", lib().rfx_event_create(byref(b)))
        self.start, self.stop = a.value, b.value

    def begin(self) -> None:
        _check("rfx_event_record", lib().rfx_event_record(self.start, self.ctx.handle))

    def end(self) -> float:
        _check("rfx_event_record", lib().rfx_event_record(self.stop, self.ctx.handle))
        ms = c_float(0.0)
        _check("rfx_event_elapsed", lib().rfx_event_elapsed(byref(ms), self.start, self.stop))
        return float(ms.value)

    def close(self) -> None:
        if self.start:
            lib().rfx_event_destroy(self.start)
            lib().rfx_event_destroy(self.stop)
            self.start = self.stop = None
