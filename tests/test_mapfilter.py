"""MAPFILTER without a GPU: the oracle pinned against independent
computations (Python's datetime for Date.h's calendar over the years it
covers, exact Python integers for C division and remainder, per-row Python
float evaluation for the expression semantics), and the expression
compiler's limits and errors."""
import ctypes
import datetime
import math
import random

import numpy as np
import pytest

import oracle
from redfox_amd import mapfilter as M


def test_date2int_is_unix_time_for_1992_to_1999():
    day = datetime.date(1992, 1, 1)
    while day.year < 2000:
        want = (day - datetime.date(1970, 1, 1)).days * 86400
        assert oracle.date2int(day.year, day.month, day.day) == want, day
        day += datetime.timedelta(days=1)


def test_int2date_inverts_date2int_over_1992_to_1998():
    day = datetime.date(1992, 1, 1)
    while day.year < 1999:
        assert oracle.int2date(oracle.date2int(day.year, day.month, day.day)) == \
            (day.year, day.month, day.day)
        day += datetime.timedelta(days=1)


def test_date_ops_against_datetime_and_known_answers():
    days = np.arange(8035, 8035 + 2557, 37, dtype=np.int64)  # 1992-01-01 onward
    sec = days * 86400
    years = oracle.date_op(2, sec)
    plus = oracle.date_op(1, sec, 1)
    for s, y, p in zip(sec, years, plus):
        d = datetime.date(1970, 1, 1) + datetime.timedelta(days=int(s) // 86400)
        assert y == d.year
        if not (d.month == 2 and d.day == 29) and d.year < 1999:
            e = datetime.date(d.year + 1, d.month, d.day)
            assert p == (e - datetime.date(1970, 1, 1)).days * 86400
    # AddMonth within a year matches the calendar (days that exist).
    for k in (1, 3, 5):
        out = oracle.date_op(0, sec, k)
        for s, o in zip(sec, out):
            d = datetime.date(1970, 1, 1) + datetime.timedelta(days=int(s) // 86400)
            if d.month + k <= 12 and d.day <= 28:
                e = datetime.date(d.year, d.month + k, d.day)
                assert o == (e - datetime.date(1970, 1, 1)).days * 86400
    # Date.h's wrap: 1995-12-10 plus 12 months is month 24 -> year 1996,
    # month 0 (24 % 12), which date2int counts as January.
    s = oracle.date2int(1995, 12, 10)
    assert oracle.date_op(0, np.array([s]), 12)[0] == oracle.date2int(1996, 0, 10)
    assert oracle.date2int(1996, 0, 10) == oracle.date2int(1996, 1, 10)
    # 1994-11-03 plus 2 months: month 13 -> 1995, month 1.
    assert oracle.date_op(0, np.array([oracle.date2int(1994, 11, 3)]), 2)[0] == \
        oracle.date2int(1995, 1, 3)


def c_div(a, b):
    if b == 0:
        return 0
    if b == -1:
        return (-a + 2**63) % 2**64 - 2**63
    q = abs(a) // abs(b)
    return q if (a < 0) == (b < 0) else -q


def test_c_division_and_remainder():
    rnd = random.Random(5)
    vals = [0, 1, -1, 2, -2, 7, -7, 2**63 - 1, -2**63] + [rnd.randint(-2**63, 2**63 - 1)
                                                           for _ in range(300)]
    a = np.array([x for x in vals for _ in vals], dtype=np.int64)
    b = np.array([y for _ in vals for y in vals], dtype=np.int64)
    q = oracle._trunc_div(a, b)
    r = oracle._c_mod(a, b)
    for x, y, qq, rr in zip(a.tolist(), b.tolist(), q.tolist(), r.tolist()):
        assert qq == c_div(x, y), (x, y)
        want_r = 0 if y in (0, -1) else x - c_div(x, y) * y
        assert rr == want_r, (x, y)


def test_evaluator_matches_per_row_python():
    """Mixed int/double expressions, comparisons and logic, evaluated per
    row in Python floats and ints with C's conversions."""
    rnd = np.random.default_rng(9)
    n = 500
    i = rnd.integers(-1000, 1000, n).astype(np.int32)
    f = rnd.normal(size=n) * 100
    g = rnd.normal(size=n)
    tree = ("add", ("mul", ("col", 1), ("sub", ("int", 1), ("col", 2))), ("div", ("col", 0), ("int", 7)))
    v, k = oracle.mapfilter_eval(tree, [i, f, g])
    assert k == "f"
    for r in range(n):
        want = f[r] * (1.0 - g[r]) + float(c_div(int(i[r]), 7))
        assert v[r] == want  # one rounding per operation on both sides
    cond = ("and", ("le", ("col", 0), ("int", 10)), ("gt", ("col", 1), ("float", 0.5)))
    c, _ = oracle.mapfilter_eval(cond, [i, f, g])
    assert c.tolist() == [int(i[r] <= 10 and f[r] > 0.5) for r in range(n)]
    keep, outs = oracle.mapfilter(cond, [(("col", 0), "i16"), (("col", 2), "f32")], [i, f, g], n)
    assert np.array_equal(outs[0], i[keep].astype(np.int16))
    assert np.array_equal(outs[1], g[keep].astype(np.float32))


def test_f2i_saturates_and_zeroes_nan():
    x = np.array([1.9, -1.9, np.nan, np.inf, -np.inf, 1e30, -1e30, 2.0**62])
    assert oracle._f2i(x).tolist() == [1, -1, 0, 2**63 - 1, -2**63, 2**63 - 1, -2**63, 2**62]


class _Buf:
    ptr = 0


def test_compiler_program_shape():
    x = M.MapFilter([(_Buf(), "i32"), (_Buf(), "f64")])
    a, p = x.cols
    x.where((a > 3) & ~(p < 0.5))
    x.emit(p * 2 + a, "f64")
    prog = x.program()
    codes = [prog.ops[k].code for k in range(prog.n_ops)]
    assert prog.n_filter < prog.n_ops and codes[-1] == M.OP["OUT"]
    assert M.OP["I2F"] in codes[prog.n_filter:]  # a widened to double
    assert M.OP["ILT"] in codes[:prog.n_filter]  # a > 3 as 3 < a
    assert x.trees()[0] == ("and", ("gt", ("col", 0), ("int", 3)),
                            ("not", ("lt", ("col", 1), ("float", 0.5))))


def test_compiler_errors():
    x = M.MapFilter([(_Buf(), "u64"), (_Buf(), "f64"), (_Buf(), "tuple8")])
    u, f, t = x.cols
    with pytest.raises(ValueError):
        u / 3
    with pytest.raises(ValueError):
        f & 1
    with pytest.raises(ValueError):
        M.add_month(f, 1)
    with pytest.raises(ValueError):
        x.emit(t, "u64").program()
    deep = x.cols[1]
    for k in range(20):
        deep = 1.0 + deep * (k + 2.0) if k % 2 else (deep + 1.0) * 0.5
    y = M.MapFilter([(_Buf(), "f64")])
    e = y.cols[0]
    for _ in range(17):
        e = e * (e + 1.0)
    with pytest.raises(ValueError, match="registers"):
        y.emit(e, "f64").program()
    z = M.MapFilter([(_Buf(), "f64")])
    with pytest.raises(ValueError, match="integer"):
        z.where(z.cols[0] * 2.0).program()


def test_abi_validates_before_touching_the_device():
    """Bad programs are refused before any device work (a placeholder
    context pointer is never dereferenced on these paths)."""
    import redfox_amd
    L = redfox_amd.lib()
    prog = M.Program()
    prog.n_ops = 1
    prog.ops[0].code = 99
    cols = (M.Col * 1)()
    cols[0].data, cols[0].type = 1, M.TYPES["i32"]
    fake_ctx = ctypes.c_void_p(8)
    cell = ctypes.c_void_p(16)
    assert L.rfx_mapfilter(fake_ctx, cols, 1, 10, ctypes.byref(prog), None, 0, cell) == -1
    assert b"code 99" in L.rfx_last_error()
    prog.ops[0].code = M.OP["OUT"]
    prog.n_filter = 1
    assert L.rfx_mapfilter(fake_ctx, cols, 1, 10, ctypes.byref(prog), None, 0, cell) == -1
    prog.n_filter = 0
    prog.ops[0].code, prog.ops[0].d = M.OP["LDI"], 16
    assert L.rfx_mapfilter(fake_ctx, cols, 1, 10, ctypes.byref(prog), None, 0, cell) == -1
    assert b"register" in L.rfx_last_error()
    cols[0].type = 42
    prog.ops[0].d = 0
    assert L.rfx_mapfilter(fake_ctx, cols, 1, 10, ctypes.byref(prog), None, 0, cell) == -1
    assert L.rfx_mapfilter(fake_ctx, cols, 17, 10, ctypes.byref(prog), None, 0, cell) == -1


def test_gather_known_answers():
    """The oracle's gather: rows past either end read 0; a 16-byte tuple
    field straddling the words ((lo >> 60) | (hi << 4), 8 bits)."""
    c = np.array([5, 6, 7], dtype=np.int32)
    idx = np.array([2, 0, 3, -1, 1], dtype=np.int32)
    v, k = oracle.mapfilter_eval(("gather", 1, ("col", 0)), [idx, c])
    assert v.tolist() == [7, 5, 0, 0, 6] and k == "i"
    t = np.array([[1 << 61, 3], [7, 1]], dtype=np.uint64)
    v, _ = oracle.mapfilter_eval(("gather", 1, ("col", 0), (60, 8, "unsigned")),
                                 [np.array([0, 1, 2]), t])
    assert v.tolist() == [50, 16, 0]
    x = M.MapFilter([(_Buf(), "i32"), (_Buf(), "f64", 10)])
    x.emit(x.gather(1, x.cols[0] + 1), "f64")
    p = x.program()
    assert [p.ops[i].code for i in range(p.n_ops)][-2:] == [M.OP["LDX"], M.OP["OUT"]]


def test_struct_layouts_match_the_header():
    """rfx_mf_op 16 B, rfx_mf_col 24 B (data, type, reserved, rows),
    rfx_mf_out 16 B, rfx_mf_program 16 + 16 * 96 B (static_assert in
    mapfilter.hip)."""
    assert ctypes.sizeof(M.Op) == 16 and ctypes.sizeof(M.Col) == 24
    assert ctypes.sizeof(M.Out) == 16 and ctypes.sizeof(M.Program) == 16 + 16 * M.MAX_OPS
    assert M.Col.rows.offset == 16


def test_string_compare_against_compare_string2():
    """str_compare (SCHUNK chunks, unsigned chunk order) evaluated by the
    oracle equals the compare_string2 transliteration for every operator,
    constants of 0 to 17 bytes, bytes >= 0x80 and prefixes."""
    rng = np.random.default_rng(1)
    alpha = [b"A", b"B", b"AB", b"\x80", b"\xff", b"Z", b"MAIL", b"TRUCK", b"AUTOMOBILE"]
    strs = [b"".join(alpha[int(i)] for i in rng.integers(0, len(alpha), int(rng.integers(0, 4))))
            for _ in range(400)] + [None]
    col = np.empty(len(strs), dtype=object)
    col[:] = strs
    lt = oracle.string_less
    for const in [b"", b"MAIL", b"AUTOMOBI", b"AUTOMOBILE", b"\x80AB", b"Z" * 17]:
        for op in ("eq", "ne", "lt", "le", "gt", "ge"):
            x = M.MapFilter([(_Buf(), "u64")])
            v, _ = oracle.mapfilter_eval(M.str_compare(x.cols[0], const, op).tree(), [col])
            want = []
            for s in strs:
                s = s or b""
                want.append(int({"eq": not lt(s, const) and not lt(const, s),
                                 "ne": lt(s, const) or lt(const, s), "lt": lt(s, const),
                                 "le": not lt(const, s), "gt": lt(const, s),
                                 "ge": not lt(s, const)}[op]))
            assert v.tolist() == want, (const, op)


def test_schunk_chunk_index_is_bounded():
    import redfox_amd
    L = redfox_amd.lib()
    prog = M.Program()
    prog.n_ops = 1
    prog.ops[0].code, prog.ops[0].imm = M.OP["SCHUNK"], 5000
    cols = (M.Col * 1)()
    cols[0].data, cols[0].type = 1, M.TYPES["u64"]
    assert L.rfx_mapfilter(ctypes.c_void_p(8), cols, 1, 10, ctypes.byref(prog), None, 0,
                           ctypes.c_void_p(16)) == -1
    assert b"chunk" in L.rfx_last_error()


def test_like_against_regex_where_they_agree():
    """The oracle's LIKE (the reference's '%' algorithm) equals a regular
    expression match for patterns without a tail (no '%', or ending in
    '%').  With a tail the reference differs from SQL in two ways, and so
    does the oracle (known answers): a middle word may be found overlapping
    the tail, and a tail that consumes the whole string fails (its loop
    stops before reaching the '%')."""
    import re
    rng = np.random.default_rng(3)
    alpha = b"ab%"
    for _ in range(3000):
        s = rng.choice(list(b"abc"), int(rng.integers(0, 7))).astype(np.uint8).tobytes()
        p = rng.choice(list(alpha), int(rng.integers(0, 7))).astype(np.uint8).tobytes()
        words = p.split(b"%")
        middle = [w for w in words[1:-1] if w] if len(words) > 2 else []
        if words[-1] and len(words) > 1:
            continue  # a tail: the reference's tail check has its own quirks (below)
        rx = b".*".join(re.escape(w) for w in words)
        assert oracle.like(s, p) == bool(re.fullmatch(rx, s, re.S)), (s, p)
    assert oracle.like(b"abc", b"a%bc%c")          # the middle word overlaps the tail
    assert not re.fullmatch(b"a.*bc.*c", b"abc")
    assert oracle.like(b"caba", b"%a%aba") and not re.fullmatch(b".*a.*aba", b"caba")
    assert not oracle.like(b"ab", b"%ab") and oracle.like(b"xab", b"%ab")
    assert oracle.like(b"abxcd", b"ab%cd") and not oracle.like(b"abcd", b"ab%cd")
    assert oracle.like(b"", b"%") and oracle.like(b"", b"") and not oracle.like(b"x", b"")
    x = M.MapFilter([(_Buf(), "u64"), (_Buf(), "u64", 1)])
    x.where(M.like(x.cols[0], x.gather(1, 0), b"a%"))
    p = x.program()
    assert M.OP["LIKE"] in [p.ops[i].code for i in range(p.n_ops)]
