"""Delimited text (TPC-H .tbl rows, the reference's TC.tbl) to typed device
columns through rfx_tbl_parse, and lineitem text to the column relations
the reference's Q1 plan reads (q1.man.cpp).

    cols, n = parse_tbl(ctx, text, ["u32", "skip", "skip", "u8", "f64"], delim="|")
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import mapfilter as M
from . import rfx

TBL = {"skip": 0, "i32": 1, "i64": 2, "u32": 3, "u8": 4, "f64": 5, "date": 6, "char": 7,
       "str": 8}
WIDTH = {"skip": 0, "i32": 4, "i64": 8, "u32": 4, "u8": 1, "f64": 8, "date": 8, "char": 1,
         "str": 8}
NUMPY = {"i32": np.int32, "i64": np.int64, "u32": np.uint32, "u8": np.uint8, "f64": np.float64,
         "date": np.int64, "char": np.uint8, "str": np.uint64}


def _call(ctx, text, nbytes, delim, spec, columns, capacity):
    types = (ctypes.c_int * max(len(spec), 1))(*[TBL[s] for s in spec])
    ptrs = (ctypes.c_void_p * max(len(spec), 1))(*[c.ptr if c is not None else None
                                                   for c in columns])
    rows, short = rfx.u64_cell(), rfx.u64_cell()
    rfx._check("rfx_tbl_parse", rfx.lib().rfx_tbl_parse(
        ctx.handle, text.ptr, nbytes, delim.encode(), types, len(spec), ptrs, capacity,
        rows.ptr, short.ptr))
    return rfx.read_u64(rows), rfx.read_u64(short)


def parse_tbl(ctx, text, spec: Sequence[str], delim: str = "|",
              nbytes: Optional[int] = None) -> Tuple[List[Optional[object]], int, int]:
    """Parse rows of `text` (host bytes, or a DeviceBuffer holding `nbytes`)
    into one device column per non-skipped field.  Returns (columns, rows,
    rows with missing fields).  "str" fields become device addresses of
    NUL-terminated strings in the text buffer, which they modify."""
    for s in spec:
        if s not in TBL:
            raise ValueError(f"unknown field type {s!r}")
    if isinstance(text, (bytes, bytearray, memoryview)):
        raw = np.frombuffer(bytes(text), dtype=np.uint8)
        nbytes = len(raw)
        buf = rfx.DeviceBuffer.from_numpy(raw) if nbytes else rfx.DeviceBuffer(16)
    else:
        buf = text
        if nbytes is None:
            raise ValueError("nbytes is needed with a device buffer")
    # Count the rows first, then parse into columns of that size.
    n, _ = _call(ctx, buf, nbytes, delim, spec, [None] * len(spec), 0)
    cols = [rfx.DeviceBuffer(max(8, n * WIDTH[s])) if s != "skip" else None for s in spec]
    n2, short = _call(ctx, buf, nbytes, delim, spec, cols, n)
    assert n2 == n
    return cols, n, short


# lineitem.tbl (dbgen): the columns the Q1 plan reads.
LINEITEM_SPEC = ["u32", "skip", "skip", "u8", "f64", "f64", "f64", "f64", "char", "char", "date"]


def lineitem_q1_relations(ctx, text, nbytes: Optional[int] = None):
    """(rows, {relation name: [(device column, type), ...]}) for the Q1 plan
    from lineitem.tbl text: keys (LINENUMBER u8, ORDER u32); RETURNFLAG
    A/N/R -> 0/1/2 and LINESTATUS F/O -> 0/1 as entity ids."""
    cols, n, short = parse_tbl(ctx, text, LINEITEM_SPEC, "|", nbytes)
    if short:
        raise ValueError(f"{short} lineitem rows have fewer than {len(LINEITEM_SPEC)} fields")
    order, ln, qty, price, disc, tax, rf, ls, ship = [c for c in cols if c is not None]
    rf_id, ls_id = rfx.DeviceBuffer(max(n, 8)), rfx.DeviceBuffer(max(n, 8))
    if n:
        x = M.MapFilter([(rf, "u8"), (ls, "u8")])
        r, s = x.cols
        x.emit(r.eq(ord("N")) + 2 * r.eq(ord("R")), "u8")
        x.emit(s.eq(ord("O")), "u8")
        x.run(ctx, n, [rf_id, ls_id], rfx.u64_cell())
    key = [(ln, "u8"), (order, "u32")]
    rel = {"L_SHIPDATE": key + [(ship, "i64")], "L_RETURNFLAG": key + [(rf_id, "u8")],
           "L_LINESTATUS": key + [(ls_id, "u8")], "L_QUANTITY": key + [(qty, "f64")],
           "L_EXTENDEDPRICE": key + [(price, "f64")], "L_DISCOUNT": key + [(disc, "f64")],
           "L_TAX": key + [(tax, "f64")]}
    return n, rel
