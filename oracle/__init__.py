"""ctypes binding of the CPU oracle (librfx_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and only as the checker or the reported CPU
baseline.  Nothing in redfox_amd/ imports it.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_int32, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "librfx_oracle.so")

P, U64, U32, I = c_void_p, c_uint64, c_uint32, c_int

_SIG = {
    "rfxo_tuple_extract": (U64, [P, P, I, I]),
    "rfxo_tuple_insert": (None, [P, P, I, I, U64]),
    "rfxo_bits_after": (U32, [P, I, I]),
    "rfxo_tuple_bytes": (U32, [P, I]),
    "rfxo_sort_pairs": (I, [P, P, U64, I, I, I]),
    "rfxo_join_find_bounds": (U64, [P, P, P, U64, P, U64, I]),
    "rfxo_join_main": (None, [P, P, U64, P, P, U64]),
    "rfxo_join_gather": (None, [P, I, P, I, P, I, P, I, P, P, U64, U32, U32]),
    "rfxo_cpujoin": (U64, [P, P, U64, P, U64, P, U64, I]),
    "rfxo_reduce_by_key_f64": (U64, [P, P, P, P, U64, I, I, I]),
    "rfxo_reduce_by_key_u32": (U64, [P, P, P, P, U64, I]),
    "rfxo_total_double": (c_double, [P, U64, I]),
    "rfxo_count_u32": (U32, [P, U64]),
    "rfxo_max_double_bits": (U64, [P, U64]),
    "rfxo_set_union": (U64, [P, P, U64, P, U64, I]),
    "rfxo_set_difference": (U64, [P, P, U64, P, U64, I]),
    "rfxo_set_intersection": (U64, [P, P, U64, P, U64, I]),
    "rfxo_unique": (U64, [P, U64, I]),
    "rfxo_select": (U64, [P, P, U64, I, U32, U32, I, U64]),
    "rfxo_project": (None, [P, I, P, I, U64, I, P, P, P]),
    "rfxo_split": (None, [P, I, P, I, P, I, U64, U32]),
    "rfxo_merge": (None, [P, I, P, I, P, I, U64, U32]),
    "rfxo_q1": (None, [P, P, P, P, P, P, P, P, U64, c_int32, I]),
    "rfxo_date2int": (U64, [I, I, I]),
    "rfxo_int2date": (None, [U64, P, P, P]),
    "rfxo_date_op": (None, [I, P, P, U64, ctypes.c_int64]),
}

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} missing: run `make -C oracle`")
        _lib = ctypes.CDLL(LIB)
        for name, (res, args) in _SIG.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


# dtype for an element width; 16-byte elements are (lo, hi) u64 pairs.
def dtype_of(width: int):
    return {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64, 16: np.uint64}[width]


def count_of(arr: np.ndarray, width: int) -> int:
    return arr.nbytes // width


# ---- wrappers returning new numpy arrays --------------------------------
def sort_pairs(keys, vals, key_bytes, key_bits, val_bytes):
    k = np.array(keys, copy=True)
    v = None if vals is None else np.array(vals, copy=True)
    n = count_of(k, key_bytes)
    rc = lib().rfxo_sort_pairs(_p(k), _p(v), n, key_bytes, key_bits, val_bytes)
    assert rc == 0
    return k, v


def join_find_bounds(left, right, key_bytes):
    nl, nr = count_of(left, key_bytes), count_of(right, key_bytes)
    lb = np.zeros(nl, dtype=np.int32)
    lc = np.zeros(nl, dtype=np.int32)
    total = lib().rfxo_join_find_bounds(_p(lb), _p(lc), _p(left), nl, _p(right), nr, key_bytes)
    return lb, lc, int(total)


def join_main(lb, lc, total):
    li = np.zeros(total, dtype=np.int32)
    ri = np.zeros(total, dtype=np.int32)
    lib().rfxo_join_main(_p(li), _p(ri), total, _p(lb), _p(lc), len(lb))
    return li, ri


def join_gather(out_bytes, lkey, lkey_bytes, lval, lval_bytes, rval, rval_bytes, li, ri,
                key_shift, lval_shift):
    total = len(li)
    out = np.zeros(total * out_bytes // dtype_of(out_bytes)().itemsize, dtype=dtype_of(out_bytes))
    lib().rfxo_join_gather(_p(out), out_bytes, _p(lkey), lkey_bytes, _p(lval), lval_bytes,
                           _p(rval), rval_bytes, _p(li), _p(ri), total, key_shift, lval_shift)
    return out


def cpujoin(a, b, key_bytes):
    na, nb = count_of(a, key_bytes), count_of(b, key_bytes)
    total = lib().rfxo_cpujoin(None, None, 0, _p(a), na, _p(b), nb, key_bytes)
    ai = np.zeros(total, dtype=np.int32)
    bi = np.zeros(total, dtype=np.int32)
    lib().rfxo_cpujoin(_p(ai), _p(bi), total, _p(a), na, _p(b), nb, key_bytes)
    return ai, bi


def reduce_by_key_f64(keys, vals, key_bytes, op=0, compensated=1):
    n = count_of(keys, key_bytes)
    ok = np.zeros_like(keys)
    ov = np.zeros(n, dtype=np.float64)
    g = lib().rfxo_reduce_by_key_f64(_p(ok), _p(ov), _p(keys), _p(vals), n, key_bytes, op,
                                     compensated)
    words = g * key_bytes // keys.itemsize
    return ok[:words], ov[:g]


def reduce_by_key_u32(keys, vals, key_bytes):
    n = count_of(keys, key_bytes)
    ok = np.zeros_like(keys)
    ov = np.zeros(n, dtype=np.uint32)
    g = lib().rfxo_reduce_by_key_u32(_p(ok), _p(ov), _p(keys), _p(vals), n, key_bytes)
    return ok[:g * key_bytes // keys.itemsize], ov[:g]


def total_double(v, compensated=1):
    return float(lib().rfxo_total_double(_p(v), len(v), compensated))


def count_u32(v):
    return int(lib().rfxo_count_u32(_p(v), len(v)))


def max_double_bits(v):
    return int(lib().rfxo_max_double_bits(_p(v), len(v)))


def set_op(op, a, b, eb):
    na, nb = count_of(a, eb), count_of(b, eb)
    out = np.zeros((na + nb) * eb // a.itemsize if a.size or b.size else 0, dtype=a.dtype)
    fn = {0: "rfxo_set_union", 1: "rfxo_set_difference", 2: "rfxo_set_intersection"}[op]
    m = getattr(lib(), fn)(_p(out), _p(a), na, _p(b), nb, eb)
    return out[:m * eb // a.itemsize]


def unique(a, eb):
    d = np.array(a, copy=True)
    m = lib().rfxo_unique(_p(d), count_of(d, eb), eb)
    return d[:m * eb // d.itemsize]


def select(inp, tuple_bytes, field_shift, field_bits, op, constant):
    n = count_of(inp, tuple_bytes)
    out = np.zeros_like(inp)
    m = lib().rfxo_select(_p(out), _p(inp), n, tuple_bytes, field_shift, field_bits, op,
                          constant & 0xFFFFFFFFFFFFFFFF)
    return out[:m * tuple_bytes // inp.itemsize]


def project(out_bytes, inp, in_bytes, in_shift, bits, out_shift):
    n = count_of(inp, in_bytes)
    a = np.ascontiguousarray(in_shift, dtype=np.uint32)
    b = np.ascontiguousarray(bits, dtype=np.uint32)
    c = np.ascontiguousarray(out_shift, dtype=np.uint32)
    out = np.zeros(n * out_bytes // dtype_of(out_bytes)().itemsize, dtype=dtype_of(out_bytes))
    lib().rfxo_project(_p(out), out_bytes, _p(inp), in_bytes, n, len(a), _p(a), _p(b), _p(c))
    return out


def split(inp, in_bytes, key_bytes, val_bytes, key_shift):
    n = count_of(inp, in_bytes)
    k = np.zeros(n * key_bytes // dtype_of(key_bytes)().itemsize, dtype=dtype_of(key_bytes))
    v = (np.zeros(n * val_bytes // dtype_of(val_bytes)().itemsize, dtype=dtype_of(val_bytes))
         if val_bytes else None)
    lib().rfxo_split(_p(k), key_bytes, _p(v), val_bytes, _p(inp), in_bytes, n, key_shift)
    return k, v


def merge(out_bytes, key, key_bytes, val, val_bytes, key_shift):
    n = count_of(key, key_bytes)
    out = np.zeros(n * out_bytes // dtype_of(out_bytes)().itemsize, dtype=dtype_of(out_bytes))
    lib().rfxo_merge(_p(out), out_bytes, _p(key), key_bytes, _p(val), val_bytes, n, key_shift)
    return out


def q1(shipdate, flag, status, qty, price, disc, tax, cutoff, compensated=1):
    out = np.zeros(48, dtype=np.float64)
    lib().rfxo_q1(_p(out), _p(shipdate), _p(flag), _p(status), _p(qty), _p(price), _p(disc),
                  _p(tax), len(shipdate), cutoff, compensated)
    return out.reshape(6, 8)


# ---- dates and MAPFILTER expressions -------------------------------------
def date2int(year, month, day):
    return int(lib().rfxo_date2int(int(year), int(month), int(day)))


def int2date(seconds):
    y, m, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib().rfxo_int2date(int(seconds) & 0xFFFFFFFFFFFFFFFF, ctypes.byref(y), ctypes.byref(m),
                        ctypes.byref(d))
    return y.value, m.value, d.value


def date_op(op, values, k=0):
    """op 0 AddMonth(k), 1 AddYear(k), 2 PartYear over an int64 array."""
    v = np.ascontiguousarray(values, dtype=np.int64)
    out = np.empty_like(v)
    lib().rfxo_date_op(op, _p(v), _p(out), v.size, int(k))
    return out


_COL_KIND = {np.dtype(np.uint64): "u", np.dtype(np.float32): "f", np.dtype(np.float64): "f"}


def _trunc_div(a, b):
    """C division of int64 arrays: truncating; b == 0 gives 0, b == -1 negates
    (two's complement)."""
    a = np.asarray(a, dtype=np.int64)
    b = np.asarray(b, dtype=np.int64)
    with np.errstate(all="ignore"):
        safe = np.where((b == 0) | (b == -1), 1, b)
        # Magnitudes as uint64, so that |INT64_MIN| = 2^63 is exact.
        ma = np.where(a < 0, (~a.view(np.uint64)) + np.uint64(1), a.view(np.uint64))
        mb = np.where(safe < 0, (~safe.view(np.uint64)) + np.uint64(1), safe.view(np.uint64))
        qm = ma // mb
        q = np.where((a < 0) != (safe < 0), (~qm) + np.uint64(1), qm).view(np.int64)
        q = np.where(b == 0, 0, q)
        q = np.where(b == -1, (~a.view(np.uint64) + np.uint64(1)).view(np.int64), q)
    return q.astype(np.int64)


def _c_mod(a, b):
    with np.errstate(all="ignore"):
        q = _trunc_div(a, b)
        r = a - q * np.where(b == 0, 0, b)
        return np.where((b == 0) | (b == -1), 0, r).astype(np.int64)


def _f2i(x):
    x = np.asarray(x, dtype=np.float64)
    with np.errstate(all="ignore"):
        t = np.trunc(np.nan_to_num(x, nan=0.0))
        out = np.where(t >= 9.223372036854775807e18, np.iinfo(np.int64).max,
                       np.where(t <= -9.223372036854775808e18, np.iinfo(np.int64).min, 0))
        mid = (t > -9.223372036854775808e18) & (t < 9.223372036854775807e18)
        out = np.where(mid, np.where(mid, t, 0).astype(np.int64), out)
    return out.astype(np.int64)


def mapfilter_eval(tree, cols):
    """Evaluate an expression tree over columns (numpy arrays) row by row in
    C's semantics on 64-bit values: integers wrap, divide truncating (a zero
    divisor gives 0), doubles round per operation; a double operand makes
    an arithmetic or comparison a double one, an unsigned 64-bit one an
    unsigned comparison.  Returns (values, kind) with kind "i", "u" or "f";
    integer values as int64 bit patterns.  This restates the per-tuple C
    expressions of set_arith and set_select (RelationalAlgebraKernel.cpp:
    1822-1939, 2001-2215, 2301-2593) for columns."""
    tag = tree[0]
    with np.errstate(all="ignore"):
        if tag == "col":
            c = cols[tree[1]]
            kind = _COL_KIND.get(c.dtype, "i")
            if kind == "f":
                return c.astype(np.float64), "f"
            return c.astype(np.int64) if c.dtype != np.uint64 else c.view(np.int64), kind
        if tag == "field":
            _, idx, shift, bits, kind = tree
            c = cols[idx]
            if c.dtype == np.uint64 and c.ndim == 2:  # 16-byte tuples as (lo, hi) rows
                lo, hi = c[:, 0], c[:, 1]
                v = (lo >> np.uint64(shift)) if shift < 64 else (hi >> np.uint64(shift - 64))
                if 0 < shift < 64:
                    v = v | (hi << np.uint64(64 - shift))
            else:
                v = c.astype(np.uint64) >> np.uint64(shift)
            if bits < 64:
                v = v & np.uint64((1 << bits) - 1)
            if kind == "double":
                return v.view(np.float64), "f"
            if kind == "signed" and bits < 64:
                m = np.uint64(1 << (bits - 1))
                v = (v ^ m) - m
            return v.view(np.int64), ("u" if bits == 64 and kind == "unsigned" else "i")
        if tag == "gather":
            # Column tree[1] at the rows tree[2] gives; past its end, 0.
            rows, _ = mapfilter_eval(tree[2], cols)
            c = cols[tree[1]]
            r = np.atleast_1d(np.asarray(rows, dtype=np.int64)).view(np.uint64)
            inside = r < np.uint64(len(c))
            picked = np.zeros((len(r),) + c.shape[1:], dtype=c.dtype)
            picked[inside] = c[r[inside].astype(np.int64)]
            if len(tree) > 3:
                sub = list(cols)
                sub[tree[1]] = picked
                v, k = mapfilter_eval(("field", tree[1]) + tuple(tree[3]), sub)
                return np.where(inside, v, 0), k
            sub = list(cols)
            sub[tree[1]] = picked
            return mapfilter_eval(("col", tree[1]), sub)
        if tag == "schunk":
            # Chunk k of each row's string: an object column of bytes (None:
            # the empty string) read through ("col", i).
            if tree[1][0] != "col" or cols[tree[1][1]].dtype != object:
                raise ValueError("schunk reads a column of byte strings")
            k = tree[2]
            out = []
            for s in cols[tree[1][1]]:
                s = (s or b"").split(b"\0", 1)[0]
                codes = [(b ^ 0x80) + (1 if b >= 0x80 else 0) for b in s] + [0] * (8 * k + 8)
                out.append(int.from_bytes(bytes(codes[8 * k:8 * k + 8]), "big"))
            return np.array(out, dtype=np.uint64).view(np.int64), "u"
        if tag == "like":
            if tree[1][0] != "col" or cols[tree[1][1]].dtype != object:
                raise ValueError("like reads a column of byte strings")
            return np.array([int(like(s or b"", tree[2])) for s in cols[tree[1][1]]],
                            dtype=np.int64), "i"
        if tag == "int":
            return np.array(tree[1] & 0xFFFFFFFFFFFFFFFF, dtype=np.uint64).view(np.int64), "i"
        if tag == "float":
            return np.float64(tree[1]), "f"
        if tag in ("add", "sub", "mul", "div", "mod"):
            (x, kx), (y, ky) = mapfilter_eval(tree[1], cols), mapfilter_eval(tree[2], cols)
            if "f" in (kx, ky):
                x = _to_f(x, kx)
                y = _to_f(y, ky)
                return {"add": x + y, "sub": x - y, "mul": x * y, "div": x / y}[tag], "f"
            k = "u" if "u" in (kx, ky) else "i"
            x = np.asarray(x, dtype=np.int64)
            y = np.asarray(y, dtype=np.int64)
            if tag == "div":
                return _trunc_div(x, y), k
            if tag == "mod":
                return _c_mod(x, y), k
            return {"add": x + y, "sub": x - y, "mul": x * y}[tag], k
        if tag in ("lt", "le", "gt", "ge", "eq", "ne"):
            (x, kx), (y, ky) = mapfilter_eval(tree[1], cols), mapfilter_eval(tree[2], cols)
            if "f" in (kx, ky):
                x, y = _to_f(x, kx), _to_f(y, ky)
            elif "u" in (kx, ky):
                x = np.asarray(x, dtype=np.int64).view(np.uint64)
                y = np.asarray(y, dtype=np.int64).view(np.uint64)
            r = {"lt": x < y, "le": x <= y, "gt": x > y, "ge": x >= y, "eq": x == y,
                 "ne": x != y}[tag]
            return np.asarray(r).astype(np.int64), "i"
        if tag in ("and", "or"):
            (x, _), (y, _) = mapfilter_eval(tree[1], cols), mapfilter_eval(tree[2], cols)
            r = (x != 0) & (y != 0) if tag == "and" else (x != 0) | (y != 0)
            return np.asarray(r).astype(np.int64), "i"
        if tag == "not":
            x, _ = mapfilter_eval(tree[1], cols)
            return np.asarray(x == 0).astype(np.int64), "i"
        if tag == "i2f":
            x, k = mapfilter_eval(tree[1], cols)
            return _to_f(x, k), "f"
        if tag == "f2i":
            x, k = mapfilter_eval(tree[1], cols)
            return (_f2i(x) if k == "f" else x), "i"
        if tag in ("add_month", "add_year", "part_year"):
            x, _ = mapfilter_eval(tree[1], cols)
            n = max(len(c) for c in cols)
            x = np.broadcast_to(np.asarray(x, dtype=np.int64), (n,))
            op = {"add_month": 0, "add_year": 1, "part_year": 2}[tag]
            return date_op(op, x, tree[2] if op < 2 else 0), "i"
    raise ValueError(f"unknown node {tag}")


def _to_f(x, k):
    if k == "f":
        return np.asarray(x, dtype=np.float64)
    if k == "u":
        return np.asarray(x, dtype=np.int64).view(np.uint64).astype(np.float64)
    return np.asarray(x, dtype=np.int64).astype(np.float64)


def mapfilter(filter_tree, outputs, cols, n):
    """Kept-row mask and the output columns (numpy, by output type) of a
    MAPFILTER: the rows whose filter is nonzero, in input order."""
    keep = np.ones(n, dtype=bool)
    if filter_tree is not None:
        f, _ = mapfilter_eval(filter_tree, cols)
        keep = np.broadcast_to(np.asarray(f) != 0, (n,)).copy()
    outs = []
    np_t = {"u8": np.uint8, "u16": np.uint16, "u32": np.uint32, "u64": np.uint64, "i8": np.int8,
            "i16": np.int16, "i32": np.int32, "i64": np.int64, "f32": np.float32, "f64": np.float64}
    for tree, t in outputs:
        v, k = mapfilter_eval(tree, cols)
        v = np.broadcast_to(v, (n,))[keep]
        if t in ("f32", "f64"):
            v = _to_f(v, k).astype(np_t[t])
        else:  # integers keep their low bits
            bits = (_f2i(v) if k == "f" else np.asarray(v, dtype=np.int64)).view(np.uint64)
            v = bits.astype(np.dtype(f"u{np.dtype(np_t[t]).itemsize}")).view(np_t[t])
        outs.append(np.ascontiguousarray(v))
    return keep, outs


def _atoi(field: bytes) -> int:
    """C atoi/atol on one field: leading spaces, a sign, digits up to the
    first other byte (TC_data_process.cu:60-64 calls atoi)."""
    i, n = 0, len(field)
    while i < n and field[i:i + 1] == b" ":
        i += 1
    neg = False
    if i < n and field[i:i + 1] in (b"-", b"+"):
        neg = field[i:i + 1] == b"-"
        i += 1
    v = 0
    while i < n and 48 <= field[i] <= 57:
        v = v * 10 + field[i] - 48
        i += 1
    return -v if neg else v


def _decimal(field: bytes) -> float:
    """digits[.digits] with an optional sign: Python's correctly rounded
    conversion of the same digits."""
    i, n = 0, len(field)
    while i < n and field[i:i + 1] == b" ":
        i += 1
    j = i
    if j < n and field[j:j + 1] in (b"-", b"+"):
        j += 1
    seen_point = False
    while j < n and (48 <= field[j] <= 57 or (field[j:j + 1] == b"." and not seen_point)):
        seen_point |= field[j:j + 1] == b"."
        j += 1
    text = field[i:j].decode()
    if text in ("", "-", "+", ".", "-.", "+."):
        return -0.0 if text.startswith("-") else 0.0
    return float(text + ("0" if text.endswith(".") else ""))


def parse_tbl(text: bytes, spec, delim: bytes = b"|"):
    """Rows of delimited text to numpy columns (one per non-skipped spec
    entry): rows end at a newline (a final row needs none, a carriage
    return before the newline is dropped); missing fields read 0.  Types:
    i32/i64/u32/u8 (atoi, wrapping to the width), f64 (decimal), date
    (YYYY-MM-DD -> seconds, Date.h), char (first byte).  Returns (columns,
    rows with missing fields)."""
    lines = text.split(b"\n")
    if lines and lines[-1] == b"":
        lines.pop()
    starts, at = [], 0  # offset of every row in the text
    for line in lines:
        starts.append(at)
        at += len(line) + 1
    np_t = {"i32": np.int32, "i64": np.int64, "u32": np.uint32, "u8": np.uint8,
            "f64": np.float64, "date": np.int64, "char": np.uint8}
    out = {f: [] for f, s in enumerate(spec) if s != "skip"}
    short = 0
    for row, line in enumerate(lines):
        if line.endswith(b"\r"):
            line = line[:-1]
        fields = line.split(delim)
        field_at = [starts[row]]
        for fld_ in fields[:-1]:
            field_at.append(field_at[-1] + len(fld_) + 1)
        if len(fields) < len(spec):
            short += 1
        for f, s in enumerate(spec):
            if s == "skip":
                continue
            fld = fields[f] if f < len(fields) else None
            if fld is None:
                v = 0
            elif s == "f64":
                v = _decimal(fld)
            elif s == "char":
                v = fld[0] if fld else 0
            elif s == "str":
                # Offset of the field; None where it has no terminator left
                # (it runs to the end of the text) or is missing.
                v = None if fld is None or field_at[f] + len(fld) >= len(text) else field_at[f]
            elif s == "date":
                parts = (fld.split(b"-") + [b"", b"", b""])[:3]
                v = date2int(*[_atoi(p) for p in parts])
            else:
                v = _atoi(fld)
            out[f].append(v)
    cols = []
    for f, s in enumerate(spec):
        if s == "skip":
            continue
        if s == "str":
            cols.append(out[f])
            continue
        if s == "f64":
            cols.append(np.array(out[f], dtype=np.float64))
        else:
            bits = np.array([v & 0xFFFFFFFFFFFFFFFF for v in out[f]], dtype=np.uint64)
            cols.append(bits.astype(np.dtype(np_t[s]).newbyteorder("=")) if s in ("u8", "char")
                        else bits.view(np.int64).astype(np_t[s]) if s != "u32"
                        else bits.astype(np.uint32))
    return cols, short


def string_less(a: bytes, b: bytes) -> bool:
    """compare_string2 (ModernGPUJoin.cu:72-97): bytes as signed char up to
    the first NUL of either; a proper prefix is less."""
    i = 0
    while i < len(a) and i < len(b) and a[i] != 0 and b[i] != 0:
        x = a[i] - 256 if a[i] >= 128 else a[i]
        y = b[i] - 256 if b[i] >= 128 else b[i]
        if x != y:
            return x < y
        i += 1
    end_a = i >= len(a) or a[i] == 0
    end_b = i >= len(b) or b[i] == 0
    return end_a and not end_b


def string_key(a: bytes):
    """A sort key giving string_less's order (signed bytes up to a NUL)."""
    a = a.split(b"\0", 1)[0]
    return [x - 256 if x >= 128 else x for x in a]


def find_bounds_string(left, right):
    """find_bounds_string over sorted string lists: lower bounds, exclusive
    prefix of the equal-run counts, and the total (ModernGPUJoin.cu:
    213-312)."""
    import bisect

    rk = [string_key(r) for r in right]
    lb, cnt = [], []
    for x in left:
        k = string_key(x)
        lo = bisect.bisect_left(rk, k)
        lb.append(lo)
        cnt.append(bisect.bisect_right(rk, k) - lo)
    lc = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int64) if cnt else np.zeros(0, np.int64)
    return np.array(lb, dtype=np.int64), lc, int(sum(cnt))


def like(s: bytes, p: bytes) -> bool:
    """The reference's LIKE (Comparisons.h:12-372), '%' the only wildcard:
    the pattern's head must start the string and its tail end it (compared
    from the ends, not past the head), then each middle word is found
    leftmost from where the previous one ended."""
    s = s.split(b"\0", 1)[0]
    p = p.split(b"\0", 1)[0]
    s0 = p0 = 0
    if not p.startswith(b"%"):
        i = 0
        while i < len(s) and i < len(p) and p[i:i + 1] != b"%":
            if s[i] != p[i]:
                return False
            i += 1
        if i < len(p) and p[i:i + 1] != b"%":
            return False
        s0 = p0 = i
    if p0 >= len(p):
        return s0 >= len(s)
    p1 = len(p)
    if not p.endswith(b"%"):
        i, j = len(s), len(p)
        while i > s0 and j > p0:
            i -= 1
            j -= 1
            if p[j:j + 1] == b"%":
                break
            if s[i] != p[j]:
                return False
        if p[j:j + 1] != b"%":
            return False
        p1 = j + 1
    i, j = s0, p0
    while j < p1:
        while p[j:j + 1] == b"%":
            j += 1
        if j >= p1:
            break
        w0 = j
        while j < p1 and p[j:j + 1] != b"%":
            j += 1
        at = s.find(p[w0:j], i)
        if at < 0:
            return False
        i = at + (j - w0)
    return j == p1
