"""String keys on the GPU: "str" fields of parsed text become addresses of
NUL-terminated strings in the text; rfx_join_find_bounds_string (sorted
search in compare_string2's order) against the oracle, and its row pairs
through rfx_join_main."""
import numpy as np
import pytest

import oracle
from redfox_amd import loaders

pytestmark = pytest.mark.gpu


def test_string_fields_point_into_the_text(rf, ctx):
    text = b"ab|cde|1\nx||2\r\n|q|\nlast|z"
    raw = np.frombuffer(text, dtype=np.uint8)
    buf = rf.DeviceBuffer(len(raw) + 16)
    buf.upload(raw)
    cols, n, short = loaders.parse_tbl(ctx, buf, ["str", "str", "i32"], "|", len(raw))
    want, wshort = oracle.parse_tbl(text, ["str", "str", "i32"])
    assert n == 4 and short == wshort
    after = buf.download(np.uint8, len(raw)).tobytes()
    for f in range(2):
        ptrs = cols[f].download(np.uint64, n)
        got = [None if p == 0 else int(p) - buf.ptr for p in ptrs]
        assert got == want[f]
        for off in got:
            if off is not None:  # the field now ends in a NUL
                end = after.index(b"\0", off)
                assert text[off:end].split(b"|")[0].rstrip(b"\r") == text[off:end]


def strings_on_device(rf, strings):
    """One buffer of NUL-terminated strings; their addresses (0 for None)."""
    blob, offs = b"", []
    for s in strings:
        if s is None:
            offs.append(None)
            continue
        offs.append(len(blob))
        blob += s + b"\0"
    buf = rf.DeviceBuffer(max(len(blob), 16))
    if blob:
        buf.upload(np.frombuffer(blob, dtype=np.uint8))
    ptrs = np.array([0 if o is None else buf.ptr + o for o in offs], dtype=np.uint64)
    return buf, rf.DeviceBuffer.from_numpy(ptrs) if len(ptrs) else rf.DeviceBuffer(8)


@pytest.mark.parametrize("nl,nr,seed", [(0, 5, 1), (5, 0, 2), (1, 1, 3), (300, 2000, 4),
                                        (5000, 700, 5), (20_000, 20_000, 6)])
def test_find_bounds_string(rf, ctx, nl, nr, seed):
    rng = np.random.default_rng(seed)
    alphabet = [b"a", b"b", b"ab", b"\x80", b"\xff", b"z", b"TRUCK", b"MAIL"]

    def draw(n):
        out = []
        for _ in range(n):
            k = int(rng.integers(0, 4))
            out.append(b"".join(alphabet[int(i)] for i in rng.integers(0, len(alphabet), k)))
        return sorted(out, key=oracle.string_key)
    left, right = draw(nl), draw(nr)
    # Empty strings on the left as null addresses (they read as "").
    left_dev = [None if s == b"" and i % 2 else s for i, s in enumerate(left)]
    lb_buf, lkeys = strings_on_device(rf, left_dev)
    rb_buf, rkeys = strings_on_device(rf, right)
    lb, lc = rf.DeviceBuffer(4 * max(nl, 1)), rf.DeviceBuffer(4 * max(nl, 1))
    tot = rf.u64_cell()
    ctx.join_find_bounds_string(lb, lc, tot, lkeys, nl, rkeys, nr)
    ctx.synchronize()
    want_lb, want_lc, want_total = oracle.find_bounds_string(left, right)
    total = rf.read_u64(tot)
    assert total == want_total
    if nl:
        assert np.array_equal(lb.download(np.int32, nl), want_lb)
        assert np.array_equal(lc.download(np.int32, nl), want_lc)
    if total:
        li, ri = rf.DeviceBuffer(4 * total), rf.DeviceBuffer(4 * total)
        ctx.join_main(li, ri, tot, total, lb, lc, nl)
        l_idx, r_idx = li.download(np.int32, total), ri.download(np.int32, total)
        assert all(left[a] == right[b] for a, b in zip(l_idx, r_idx))
        assert np.all(np.diff(l_idx) >= 0)


@pytest.mark.parametrize("n,seed", [(1, 1), (2, 2), (1000, 3), (50_000, 4)])
def test_sort_strings(rf, ctx, n, seed):
    """Stable sort of string addresses into compare_string2's order: bytes
    >= 0x80 before ASCII, prefixes first, null addresses as "", lengths up
    to 40 bytes (5 chunk passes)."""
    rng = np.random.default_rng(seed)
    alphabet = [b"a", b"b", b"ab", b"\x80", b"\xff", b"z", b"TRUCK#", b"0123456789"]
    strings = []
    for i in range(n):
        k = int(rng.integers(0, 6))
        s = b"".join(alphabet[int(j)] for j in rng.integers(0, len(alphabet), k))[:40]
        strings.append(None if (s == b"" and i % 3 == 0) else s)
    buf, addr = strings_on_device(rf, strings)
    scratch = rf.DeviceBuffer(8 * (n + 1))
    before = addr.download(np.uint64, n)
    ctx.sort_strings(addr, scratch, n)
    after = addr.download(np.uint64, n)
    order = sorted(range(n), key=lambda i: oracle.string_key(strings[i] or b""))  # stable
    assert after.tolist() == [int(before[i]) for i in order]


def test_parse_sort_join_strings(rf, ctx):
    """Text with string keys on both sides: parse, sort, sorted search,
    expansion; the pairs are exactly the equal-key row pairs."""
    rng = np.random.default_rng(12)
    modes = [b"TRUCK", b"MAIL", b"SHIP", b"AIR", b"REG AIR", b"FOB", b"RAIL", b"\xc3\xa9t\xc3\xa9"]
    lrows = [modes[int(i)] for i in rng.integers(0, len(modes) - 2, 300)]
    rrows = [modes[int(i)] for i in rng.integers(0, len(modes), 2000)]
    sides = []
    for rows in (lrows, rrows):
        text = b"".join(b"%d|%s|\n" % (i, s) for i, s in enumerate(rows))
        raw = np.frombuffer(text, dtype=np.uint8)
        buf = rf.DeviceBuffer(len(raw) + 16)
        buf.upload(raw)
        cols, n, _ = loaders.parse_tbl(ctx, buf, ["u32", "str"], "|", len(raw))
        ids, addr = cols
        scratch = rf.DeviceBuffer(8 * (n + 1))
        # Sort the addresses; the row ids follow through the address order.
        ptr_before = addr.download(np.uint64, n)
        ctx.sort_strings(addr, scratch, n)
        ptr_after = addr.download(np.uint64, n)
        row_of = {int(p): i for i, p in enumerate(ptr_before)}
        sides.append((buf, addr, n, [row_of[int(p)] for p in ptr_after]))
    (_, la, nl, lorder), (_, ra, nr, rorder) = sides
    lb, lc, tot = rf.DeviceBuffer(4 * nl), rf.DeviceBuffer(4 * nl), rf.u64_cell()
    ctx.join_find_bounds_string(lb, lc, tot, la, nl, ra, nr)
    total = rf.read_u64(tot)
    want = sorted((i, j) for i, a in enumerate(lrows) for j, b in enumerate(rrows) if a == b)
    assert total == len(want)
    li, ri = rf.DeviceBuffer(4 * total), rf.DeviceBuffer(4 * total)
    ctx.join_main(li, ri, tot, total, lb, lc, nl)
    got = sorted((lorder[a], rorder[b]) for a, b in zip(li.download(np.int32, total),
                                                       ri.download(np.int32, total)))
    assert got == want


@pytest.fixture(params=[1, 2], ids=["compiled", "interpreted"])
def mctx(ctx, request):
    ctx.tune("mapfilter.mode", request.param)
    yield ctx
    ctx.tune("mapfilter.mode", 0)


def test_mapfilter_string_predicates(rf, mctx):
    """MAPFILTER filters on string columns against constants (SCHUNK chunk
    comparisons), in both kernels, against the oracle on the same bytes."""
    from redfox_amd import mapfilter as M

    rng = np.random.default_rng(31)
    modes = [b"TRUCK", b"MAIL", b"SHIP", b"AIR", b"REG AIR", b"FOB", b"RAIL", b"AUTOMOBILE",
             b"AUTOMOBI", b"\xc3\xa9t\xc3\xa9", b""]
    rows = [modes[int(i)] for i in rng.integers(0, len(modes), 30_000)]
    text = b"".join(b"%d|%s|\n" % (i, s) for i, s in enumerate(rows))
    raw = np.frombuffer(text, dtype=np.uint8)
    buf = rf.DeviceBuffer(len(raw) + 16)
    buf.upload(raw)
    (ids, addr), n, _ = loaders.parse_tbl(mctx, buf, ["u32", "str"], "|", len(raw))
    col = np.empty(n, dtype=object)
    col[:] = rows
    for const, op in [(b"MAIL", "eq"), (b"AUTOMOBILE", "eq"), (b"AUTOMOBI", "ne"),
                      (b"M", "lt"), (b"REG AIR", "ge"), (b"\xc3", "gt"), (b"", "le")]:
        x = M.MapFilter([(addr, "u64"), (ids, "u32")])
        s, i = x.cols
        x.where(M.str_compare(s, const, op))
        x.emit(i, "u32")
        out, cell = rf.DeviceBuffer(4 * n), rf.u64_cell()
        x.run(mctx, n, [out], cell)
        mctx.synchronize()
        k = rf.read_u64(cell)
        keep, _ = oracle.mapfilter(x.trees()[0], [], [col, np.arange(n, dtype=np.uint32)], n)
        assert k == int(keep.sum()), (const, op)
        assert np.array_equal(out.download(np.uint32, k), np.flatnonzero(keep).astype(np.uint32))


def test_plan_string_filter(rf, ctx):
    """An RA plan filtering a STRING column against string constants
    (shipmode IN ('MAIL', 'SHIP') as two filters and a union)."""
    from redfox_amd import raplan

    rng = np.random.default_rng(33)
    modes = [b"TRUCK", b"MAIL", b"SHIP", b"AIR", b"REG AIR"]
    rows = [(modes[int(m)], int(v)) for m, v in zip(rng.integers(0, 5, 5000),
                                                     rng.integers(0, 10 ** 6, 5000))]
    text = b"".join(b"%s|%d|\n" % r for r in rows)
    raw = np.frombuffer(text, dtype=np.uint8)
    buf = rf.DeviceBuffer(len(raw) + 16)
    buf.upload(raw)
    (addr, val), n, _ = loaders.parse_tbl(ctx, buf, ["str", "i64"], "|", len(raw))
    b = raplan.Builder("modes")
    b.var("L", ["string"], ["int64"]).var("M", fields=["int64"]).var("S", fields=["int64"])
    b.var("U", fields=["int64"])
    s = b.sequence(0)
    s.mapfilter("M", "L", 2, b.comp("eq", b.idx(0, "string"), b.const_string("MAIL")),
                [b.idx(1, "int64")])
    s.mapfilter("S", "L", 2, b.comp("eq", b.const_string("SHIP"), b.idx(0, "string")),
                [b.idx(1, "int64")])
    s.setop("U", "union", "M", "S")
    s.halt()
    p = raplan.Plan(ctx, b.entry_exit(0, 0).bytes())
    p.set_columns("L", [(addr, "str"), (val, "i64")], n)
    p.run()
    assert p.get("U")[0].tolist() == sorted({v for m, v in rows if m in (b"MAIL", b"SHIP")})


def test_mapfilter_like(rf, mctx):
    """LIKE / NOT LIKE with '%' patterns in both kernels against the oracle
    (the reference's algorithm), over strings parsed from text."""
    from redfox_amd import mapfilter as M

    rng = np.random.default_rng(41)
    words = [b"special", b"requests", b"pending", b"deposits", b"carefully", b"ab", b"a", b""]
    rows = [b" ".join(words[int(i)] for i in rng.integers(0, len(words), int(rng.integers(0, 5))))
            for _ in range(20_000)]
    text = b"".join(b"%d|%s|\n" % (i, s) for i, s in enumerate(rows))
    raw = np.frombuffer(text, dtype=np.uint8)
    buf = rf.DeviceBuffer(len(raw) + 16)
    buf.upload(raw)
    (ids, addr), n, _ = loaders.parse_tbl(mctx, buf, ["u32", "str"], "|", len(raw))
    col = np.empty(n, dtype=object)
    col[:] = rows
    for pat, neg in [(b"%special%requests%", False), (b"special%", False), (b"%deposits", True),
                     (b"%a%b", False), (b"a", False), (b"%", True), (b"", False),
                     (b"carefully%ab%ab", False)]:
        pbuf = rf.DeviceBuffer(len(pat) + 16)
        pbuf.upload(np.frombuffer(pat + b"\0", dtype=np.uint8))
        pcol = rf.DeviceBuffer.from_numpy(np.array([pbuf.ptr], dtype=np.uint64))
        x = M.MapFilter([(addr, "u64"), (ids, "u32"), (pcol, "u64", 1)])
        s, i, _ = x.cols
        e = M.like(s, x.gather(2, 0), pat)
        x.where(~e if neg else e)
        x.emit(i, "u32")
        out, cell = rf.DeviceBuffer(4 * n), rf.u64_cell()
        x.run(mctx, n, [out], cell)
        mctx.synchronize()
        k = rf.read_u64(cell)
        keep, _ = oracle.mapfilter(x.trees()[0], [], [col], n)
        assert k == int(keep.sum()), (pat, neg)
        assert np.array_equal(out.download(np.uint32, k), np.flatnonzero(keep).astype(np.uint32))


def test_plan_like_tests(rf, ctx):
    from redfox_amd import raplan

    rng = np.random.default_rng(43)
    comments = [b"quickly special requests", b"final deposits", b"special packages",
                b"even requests sleep", b"blithely special accounts requests"]
    rows = [(comments[int(c)], int(v)) for c, v in zip(rng.integers(0, 5, 3000),
                                                        rng.integers(0, 10 ** 6, 3000))]
    text = b"".join(b"%s|%d|\n" % r for r in rows)
    raw = np.frombuffer(text, dtype=np.uint8)
    buf = rf.DeviceBuffer(len(raw) + 16)
    buf.upload(raw)
    (addr, val), n, _ = loaders.parse_tbl(ctx, buf, ["str", "i64"], "|", len(raw))
    b = raplan.Builder("like")
    b.var("C", ["string"], ["int64"]).var("L", fields=["int64"]).var("N", fields=["int64"])
    s = b.sequence(0)
    s.mapfilter("L", "C", 2, b.test("like", b.idx(0, "string"), "%special%requests%"),
                [b.idx(1, "int64")])
    s.mapfilter("N", "C", 2, b.and_(b.test("notlike", b.idx(0, "string"), "%requests%"),
                                    b.comp("lt", b.idx(1, "int64"), b.const_int(500_000))),
                [b.idx(1, "int64")])
    s.halt()
    p = raplan.Plan(ctx, b.entry_exit(0, 0).bytes())
    p.set_columns("C", [(addr, "str"), (val, "i64")], n)
    p.run()
    assert p.get("L")[0].tolist() == [v for c, v in rows if oracle.like(c, b"%special%requests%")]
    assert p.get("N")[0].tolist() == [v for c, v in rows
                                      if not oracle.like(c, b"%requests%") and v < 500_000]
