"""Text loading on the GPU (rfx_tbl_parse) against the oracle's parser on
the same bytes: the reference's TC.tbl fixture, edge rows (no final
newline, CRLF, short and empty rows, signs and spaces), texts across the
8192-byte tile edges, an unaligned text pointer, the capacity limit, and
the Q1 plan run from lineitem.tbl text."""
import ctypes
import os

import numpy as np
import pytest

import oracle
from redfox_amd import datagen, loaders, raplan

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def check(rf, ctx, text, spec, delim="|"):
    cols, n, short = loaders.parse_tbl(ctx, text, spec, delim)
    want, wshort = oracle.parse_tbl(text, spec, delim.encode())
    assert short == wshort
    got = [c.download(loaders.NUMPY[s], n) for c, s in zip(cols, spec) if s != "skip"]
    for g, w, s in zip(got, want, [s for s in spec if s != "skip"]):
        assert len(g) == len(w) == n
        if s == "f64":
            assert np.array_equal(g.view(np.uint64), w.view(np.uint64)), s
        else:
            assert np.array_equal(g, w), s
    return got, n


def test_reference_tc_fixture(rf, ctx):
    text = open(os.path.join(GOLDEN, "TC.tbl"), "rb").read()
    (a, b), n = check(rf, ctx, text, ["i32", "i32"], "\t")
    assert n == 5 and a.tolist() == [1, 2, 3, 4, 5] and b.tolist() == [6, 7, 8, 9, 10]


@pytest.mark.parametrize("text", [
    b"", b"\n", b"42", b"1|a|2.5|\r\n-3|b\n\n7|c|0.125|extra|\n8",
    b"  -12|  +7|-0.50|1996-02-29|R|\n|||||\n9999999999|1|123456789.123|1998-12-31|N|",
    b"1|2|3.5|19a96-0 2-+3x|Z\n1|2|3.5| 1996- 02-29|Z\n",
])
def test_edge_rows(rf, ctx, text):
    check(rf, ctx, text, ["i64", "i32", "f64", "date", "char"])


@pytest.mark.parametrize("n_orders", [1, 60, 61, 62, 200, 4000, 60_000])
def test_lineitem_text_sizes(rf, ctx, n_orders):
    text = datagen.lineitem_tbl(n_orders, n_orders)[0]
    check(rf, ctx, text, loaders.LINEITEM_SPEC + ["skip", "skip", "skip", "char"])


def test_tile_edge_newlines(rf, ctx):
    # Newlines at 8191/8192/8193, at the ends of 16-byte words, and a block
    # whose rows span more text than the field pass stages in LDS.
    rows = []
    for L in (8190, 1, 14, 15, 16, 17, 8191, 31, 50_000, 3):  # 50000: past the LDS stage
        rows.append(b"7" * L)
    text = b"\n".join(rows) + b"\n"
    check(rf, ctx, text, ["u32"])


@pytest.mark.parametrize("delim,text", [
    # Empty fields between spaces: a field is its own bytes only, never the
    # next field's digits after the delimiter.
    (" ", b"12 34 5.5 1996-01-02 x\n 7  8\n-1 +2 .25 1998-12-31\n  \n"),
    (".", b"1.2.3.1996-01-02.x\n7..8\n"),
    ("+", b"1+-2+3.5+1992-01-01+y\n+5\n"),
    ("5", b"1525352.51996-01-025x\n555\n"),
])
def test_delimiters_that_numbers_use(rf, ctx, delim, text):
    check(rf, ctx, text, ["i32", "i32", "f64", "date", "char"], delim)


def test_unaligned_text_and_capacity(rf, ctx):
    text = datagen.lineitem_tbl(3000, 5)[0]
    buf = rf.DeviceBuffer(len(text) + 16)
    buf.upload(np.frombuffer(text, dtype=np.uint8), offset=3)

    class At:
        ptr = buf.ptr + 3
    cols, n, short = loaders.parse_tbl(ctx, At(), ["u32", "skip", "skip", "u8"], "|", len(text))
    want, _ = oracle.parse_tbl(text, ["u32", "skip", "skip", "u8"])
    assert n == len(want[0]) and short == 0
    assert np.array_equal(cols[0].download(np.uint32, n), want[0])
    assert np.array_equal(cols[3].download(np.uint8, n), want[1])
    # Capacity below the row count: every row counted, the first `cap` stored.
    cap = 1000
    out = rf.DeviceBuffer(4 * cap)
    types = (ctypes.c_int * 1)(loaders.TBL["u32"])
    ptrs = (ctypes.c_void_p * 1)(out.ptr)
    rows, sh = rf.u64_cell(), rf.u64_cell()
    assert rf.lib().rfx_tbl_parse(ctx.handle, At.ptr, len(text), b"|", types, 1, ptrs, cap,
                                  rows.ptr, sh.ptr) == 0
    assert rf.read_u64(rows) == n
    assert np.array_equal(out.download(np.uint32, cap), want[0][:cap])


def test_bad_arguments_are_refused(rf, ctx):
    types = (ctypes.c_int * 1)(42)
    ptrs = (ctypes.c_void_p * 1)(None)
    cell = rf.u64_cell()
    buf = rf.DeviceBuffer(16)
    assert rf.lib().rfx_tbl_parse(ctx.handle, buf.ptr, 4, b"|", types, 1, ptrs, 0, cell.ptr,
                                  cell.ptr) == -1
    types[0] = loaders.TBL["u32"]
    assert rf.lib().rfx_tbl_parse(ctx.handle, buf.ptr, 4, b"\n", types, 1, ptrs, 0, cell.ptr,
                                  cell.ptr) == -1


def test_q1_plan_from_lineitem_text(rf, ctx):
    """lineitem.tbl text -> column relations -> the reference's Q1 plan."""
    text, order, ln, qty, price, disc, tax, rfl, lsl, ship = datagen.lineitem_tbl(30_000, 9)
    n, rels = loaders.lineitem_q1_relations(ctx, text)
    assert n == len(order)
    p = raplan.Plan(ctx, raplan.q1_plan())
    p.set("+$logicQ1:_dateDelta", [np.array([90 * 86400])])
    for name, cols in rels.items():
        p.set_columns(name, cols, n)
    p.run()
    keep = ship <= 912470400 - 90 * 86400
    fid = (rfl == ord("N")) * 1 + (rfl == ord("R")) * 2
    g = fid.astype(np.int64) * 2 + (lsl == ord("O"))
    groups = sorted(np.unique(g[keep]).tolist())
    k0, k1, cnt = p.get("+$logicQ1:_count_order")
    assert (k0.astype(np.int64) * 2 + k1).tolist() == groups
    assert cnt.tolist() == [int((keep & (g == k)).sum()) for k in groups]
    _, _, charge = p.get("+$logicQ1:_sum_charge")
    for k, got in zip(groups, charge):
        sel = keep & (g == k)
        want = (price * (1 - disc) * (1 + tax))[sel].sum()
        assert abs(got - want) <= 1e-12 * abs(want)
