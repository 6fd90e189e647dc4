"""MAPFILTER on the GPU against the oracle's row-by-row C-semantics
evaluation: bit-exact integers and doubles (one IEEE rounding per
operation, no fused multiply-add on either side), kept rows in input
order.  Every case runs twice: with the program compiled for its types,
and through the interpreter."""
import numpy as np
import pytest

from redfox_amd import datagen
from redfox_amd import mapfilter as M

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[1, 2], ids=["compiled", "interpreted"])
def mctx(ctx, request):
    ctx.tune("mapfilter.mode", request.param)
    yield ctx
    ctx.tune("mapfilter.mode", 0)

NP = {"u8": np.uint8, "u16": np.uint16, "u32": np.uint32, "u64": np.uint64, "i8": np.int8,
      "i16": np.int16, "i32": np.int32, "i64": np.int64, "f32": np.float32, "f64": np.float64}


def run(rf, ctx, orc, arrays, types, build, n=None):
    """Build a MapFilter over device copies of `arrays`, run it, and check
    every output and the kept count against the oracle."""
    n = len(arrays[0]) if n is None else n
    bufs = [rf.DeviceBuffer.from_numpy(np.ascontiguousarray(a)) if a.size else rf.DeviceBuffer(8)
            for a in arrays]
    x = M.MapFilter(list(zip(bufs, types)))
    build(x, *x.cols)
    outs = [rf.DeviceBuffer(max(n * np.dtype(NP[t]).itemsize, 8)) for _, t in x.outputs]
    cell = rf.u64_cell(12345)
    x.run(ctx, n, outs, cell)
    ctx.synchronize()
    kept = rf.read_u64(cell)
    ft, ot = x.trees()
    keep, want = orc.mapfilter(ft, ot, list(arrays), n)
    assert kept == int(keep.sum())
    got = [o.download(NP[t], kept) for o, (_, t) in zip(outs, x.outputs)]
    for g, w, (_, t) in zip(got, want, x.outputs):
        if t in ("f32", "f64"):
            assert np.array_equal(g.view(f"u{g.itemsize}"), w.view(f"u{w.itemsize}")), t
        else:
            assert np.array_equal(g, w), t
    return kept, got


def test_q1_projection(rf, mctx, orc):
    """Q1's filter and computed columns (q1.man.cpp:959-2180): shipdate
    <= 1998-12-01 - 90 days, disc_price = price * (1 - disc), charge =
    disc_price * (1 + tax)."""
    n = 1_000_003
    ship, flag, status, qty, price, disc, tax = datagen.lineitem(datagen.BASE_SEED + 5, n)
    rows = np.arange(n, dtype=np.uint32)

    def build(x, s, f, st, q, p, d, t, r):
        x.where(s <= 2436)
        x.emit(p * (1 - d), "f64")
        x.emit(p * (1 - d) * (1 + t), "f64")
        x.emit(q, "i32")
        x.emit(f * 2 + st, "u8")
        x.emit(r, "u32")

    kept, got = run(rf, mctx, orc, [ship, flag, status, qty, price, disc, tax, rows],
                    ["i32", "u8", "u8", "i32", "f64", "f64", "f64", "u32"], build)
    assert 0 < kept < n
    assert np.all(np.diff(got[4].astype(np.int64)) > 0)  # input order


@pytest.mark.parametrize("n", [0, 1, 63, 511, 512, 8191, 8192, 8193, 100_000, (1 << 22) + 3])
def test_sizes_and_order(rf, mctx, orc, n):
    v = datagen.splitmix64(7 + n, n)
    rows = np.arange(n, dtype=np.uint64)

    def build(x, a, r):
        x.where((a % 3).eq(1))
        x.emit(r, "u64")
        x.emit(a, "u64")

    run(rf, mctx, orc, [v.view(np.int64), rows], ["i64", "u64"], build, n)


def test_keep_all_and_none(rf, mctx, orc):
    v = np.arange(50_000, dtype=np.int32)
    run(rf, mctx, orc, [v], ["i32"], lambda x, a: x.emit(a * 3, "i64"))
    run(rf, mctx, orc, [v], ["i32"], lambda x, a: (x.where(a < -1), x.emit(a, "i32")))


def test_integer_edges(rf, mctx, orc):
    """Truncating division and C remainder with zero and -1 divisors, the
    most negative value, wrapping products, signed against unsigned
    comparison of 64-bit columns."""
    n = 20_000
    a = datagen.splitmix64(11, n).view(np.int64).copy()
    b = (datagen.splitmix64(12, n) % np.uint64(7)).astype(np.int64) - 3  # -3 .. 3
    lo, hi = np.iinfo(np.int64).min, np.iinfo(np.int64).max
    a[:8] = [lo, hi, -7, 7, lo, 1, lo, hi]
    b[:8] = [-1, 0, 2, -2, 2, lo, lo, -3]
    u = datagen.splitmix64(13, n)

    def build(x, a_, b_, u_):
        x.emit(a_ / b_, "i64")
        x.emit(a_ % b_, "i64")
        x.emit(a_ * b_ + 12345, "i64")
        x.emit((u_ < 0x8000000000000000) + 2 * (a_ < 0) + 4 * (u_ > a_), "u8")
        x.emit(~(a_ > b_) | (b_.eq(0)), "i16")

    run(rf, mctx, orc, [a, b, u], ["i64", "i64", "u64"], build)


def test_float_ops_and_conversions(rf, mctx, orc):
    n = 30_000
    f = (datagen.uniform01(21, n) - 0.5) * 1e6
    g = datagen.uniform01(22, n) - 0.25
    f[:6] = [np.nan, np.inf, -np.inf, 0.0, -0.0, 9.3e18]
    g[:6] = [1.0, 0.0, 2.0, 0.0, np.nan, -1.0]
    i = (datagen.splitmix64(23, n) % np.uint64(1000)).astype(np.int32) - 500
    h = (datagen.uniform01(24, n) * 100).astype(np.float32)

    def build(x, f_, g_, i_, h_):
        x.where(~(f_ < g_ * 1000.5) | (i_ > 100))
        x.emit(f_ / g_ - i_ * 0.3, "f64")
        x.emit(M.to_int(f_), "i64")
        x.emit(f_ * g_, "f32")
        x.emit(h_ + i_, "f64")
        x.emit(M.to_float(i_) / 7, "f64")

    run(rf, mctx, orc, [f, g, i, h], ["f64", "f64", "i32", "f32"], build)


def test_date_operations(rf, mctx, orc):
    """AddMonth (with Date.h's month wrap), AddYear and PartYear on dates
    in seconds since 1970 over 1992-1998, and a date-range filter."""
    n = 50_000
    day = (datagen.splitmix64(31, n) % np.uint64(2557)).astype(np.int64) + 8035
    sec = day * 86400

    def build(x, s):
        x.where((s >= orc.date2int(1993, 1, 1)) & (s < M.add_year(orc.date2int(1993, 1, 1), 4)))
        x.emit(M.add_month(s, 3), "i64")
        x.emit(M.add_month(s, 12), "i64")
        x.emit(M.add_month(s, 25), "i64")
        x.emit(M.add_year(s, 1), "i64")
        x.emit(M.part_year(s), "i32")
        x.emit(M.part_year(M.add_month(s, -1)), "i32")

    run(rf, mctx, orc, [sec], ["i64"], build)


def test_packed_tuple_fields(rf, mctx, orc):
    """Fields of 8- and 16-byte packed tuples: unsigned, signed and double,
    one field straddling the two words of a 16-byte tuple."""
    n = 40_000
    t8 = datagen.splitmix64(41, n)
    t16 = np.stack([datagen.splitmix64(42, n), datagen.splitmix64(43, n)], axis=1).copy()
    dbl = datagen.uniform01(44, n) * 10.0
    t16[:, 1] = dbl.view(np.uint64)

    def build(x, a, b):
        lo = x.field(0, 0, 20, "unsigned")
        sg = x.field(0, 20, 12, "signed")
        hi = x.field(0, 32, 32, "unsigned")
        mid = x.field(1, 40, 40, "unsigned")
        d = x.field(1, 64, 64, "double")
        x.where(sg < 100)
        x.emit(lo + sg, "i64")
        x.emit(hi, "u32")
        x.emit(mid, "u64")
        x.emit(d * 2.5, "f64")

    run(rf, mctx, orc, [t8, t16], ["tuple8", "tuple16"], build)


def test_many_wide_tuple_columns(rf, mctx, orc):
    """Ten 16-byte tuple columns (160 bytes a row), a field of each."""
    n = 30_000
    cols = [np.stack([datagen.splitmix64(50 + 2 * c, n), datagen.splitmix64(51 + 2 * c, n)],
                     axis=1).copy() for c in range(10)]

    def build(x, *cs):
        f = [x.field(c, 60, 8, "unsigned") for c in range(10)]
        x.where(f[0] < 200)
        acc = f[1]
        for c in range(2, 10):
            acc = acc + f[c] * c
        x.emit(acc, "i64")
        x.emit(x.field(9, 0, 64, "unsigned"), "u64")

    run(rf, mctx, orc, cols, ["tuple16"] * 10, build)


def test_bad_programs_are_refused(rf, mctx):
    a = rf.DeviceBuffer.from_numpy(np.arange(10, dtype=np.int32))
    x = M.MapFilter([(a, "i32")])
    x.emit(x.cols[0], "i32")
    prog = x.program()
    prog.ops[0].a = 5  # no such column
    import ctypes
    cols = (M.Col * 1)()
    cols[0].data, cols[0].type = a.ptr, M.TYPES["i32"]
    out = rf.DeviceBuffer(64)
    outs = (M.Out * 1)()
    outs[0].data, outs[0].type = out.ptr, M.TYPES["i32"]
    cell = rf.u64_cell()
    rc = rf.lib().rfx_mapfilter(mctx.handle, cols, 1, 10, ctypes.byref(prog), outs, 1, cell.ptr)
    assert rc == -1 and b"reads column 5" in rf.lib().rfx_last_error()


def test_long_program(rf, mctx, orc):
    """A program near the op limit (the compiled kernel keeps one copy of
    each phase for programs this long)."""
    n = 70_001
    a = (datagen.splitmix64(61, n) % np.uint64(1 << 20)).astype(np.int64)
    f = datagen.uniform01(62, n)

    def build(x, a_, f_):
        x.where(((a_ % 5) < 3) & (f_ < 0.9))
        acc = a_
        for k in range(1, 6):
            acc = acc * (k + 1) - (a_ % (k + 2)) + k
        x.emit(acc, "i64")
        g = f_
        for k in range(5):
            g = g * 1.5 + M.to_float(a_ % (k + 3))
        x.emit(g, "f64")

    x = M.MapFilter([(rf.DeviceBuffer(8), "i64"), (rf.DeviceBuffer(8), "f64")])
    build(x, *x.cols)
    assert x.program().n_ops > 32
    run(rf, mctx, orc, [a, f], ["i64", "f64"], build)


@pytest.mark.parametrize("compile_first", [False, True], ids=["uncompiled", "compiled"])
def test_graph_capture_and_replay(rf, orc, compile_first):
    """Captured into a HIP graph on the context's own stream: a program not
    yet compiled is recorded as the interpreter, a compiled one as its own
    kernel.  Every replay (eager work and synchronisation between them)
    rewrites the outputs and the kept count."""
    n = 123_457
    v = datagen.splitmix64(71 + int(compile_first), n).view(np.int64)
    stream = rf.Stream()
    c = rf.Context(0, stream.handle)
    try:
        col = rf.DeviceBuffer.from_numpy(v)
        x = M.MapFilter([(col, "i64")])
        a = x.cols[0]
        salt = 1000 + 17 * int(compile_first)  # a program no other test compiled
        x.where((a % 7) < 4)
        x.emit(a % salt, "i64")
        out = rf.DeviceBuffer(8 * n)
        cell = rf.u64_cell()
        x.run(c, n, [out], cell)  # sizes the workspace (and compiles, by default)
        c.synchronize()
        if not compile_first:
            x = M.MapFilter([(col, "i64")])
            a = x.cols[0]
            x.where((a % 7) < 4)
            x.emit(a % (salt + 1), "i64")
        ft, ot = x.trees()
        keep, want = orc.mapfilter(ft, ot, [v], n)
        c.graph_begin()
        x.run(c, n, [out], cell)
        g = c.graph_end()
        sentinel = np.array([7], dtype=np.uint64)
        for _ in range(3):
            out.zero()
            cell.upload(sentinel)
            assert rf.lib().rfx_device_synchronize() == 0
            g.launch()
            c.synchronize()
            assert rf.read_u64(cell) == int(keep.sum())
            assert np.array_equal(out.download(np.int64, int(keep.sum())), want[0])
        g.close()
    finally:
        c.close()
        stream.close()


def test_gather_rows(rf, mctx, orc):
    """LDX: columns of another length read at computed rows (the join's
    output gather), rows past the end reading 0, a packed tuple field read
    through a gather, and arithmetic on gathered values."""
    n, m = 60_001, 5_000
    idx = (datagen.splitmix64(81, n) % np.uint64(m + 50)).astype(np.int32)  # some past the end
    idx[:4] = [0, m - 1, m, -1]
    vals = datagen.uniform01(82, m) * 100.0
    small = (datagen.splitmix64(83, m) % np.uint64(200)).astype(np.uint8)
    t16 = np.stack([datagen.splitmix64(84, m), datagen.splitmix64(85, m)], axis=1).copy()
    bufs = [rf.DeviceBuffer.from_numpy(a) for a in (idx, vals, small, t16)]
    x = M.MapFilter([(bufs[0], "i32"), (bufs[1], "f64", m), (bufs[2], "u8", m),
                     (bufs[3], "tuple16", m)])
    i = x.cols[0]
    x.where(x.gather(2, i) < 150)
    x.emit(x.gather(1, i) * 2.0 + x.gather(2, i + 1), "f64")
    x.emit(x.gather(3, i, (60, 20, "unsigned")), "u32")
    x.emit(x.gather(2, i), "u8")
    outs = [rf.DeviceBuffer(8 * n), rf.DeviceBuffer(4 * n), rf.DeviceBuffer(n)]
    cell = rf.u64_cell()
    x.run(mctx, n, outs, cell)
    mctx.synchronize()
    kept = rf.read_u64(cell)
    ft, ot = x.trees()
    keep, want = orc.mapfilter(ft, ot, [idx, vals, small, t16], n)
    assert kept == int(keep.sum()) and 0 < kept < n
    got = [outs[0].download(np.float64, kept), outs[1].download(np.uint32, kept),
           outs[2].download(np.uint8, kept)]
    assert np.array_equal(got[0].view(np.uint64), want[0].view(np.uint64))
    assert np.array_equal(got[1], want[1]) and np.array_equal(got[2], want[2])
