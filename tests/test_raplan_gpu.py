"""RA plans on the GPU: TPC-H Q1 as the reference's plan builds it
(q1.man.cpp), a transitive closure that loops until a fixpoint (MOVE,
COND, UNION, key join with reordering), and set operations, SINGLE,
PRODUCT, MIN/MAX and a signed filter; results against numpy and Python
sets computed from the same inputs."""
import numpy as np
import oracle
import pytest

from redfox_amd import datagen, raplan

pytestmark = pytest.mark.gpu
Q = "+$logicQ1:_"


@pytest.mark.parametrize("n_orders,delta_days", [(1, 90), (2_000, 90), (150_000, 90),
                                                 (40_000, 3000)])
def test_q1_plan(ctx, n_orders, delta_days):
    ln, od, ship, rf, ls, qty, price, disc, tax = datagen.lineitem_keyed(n_orders, n_orders)
    p = raplan.Plan(ctx, raplan.q1_plan())
    delta = delta_days * 86400
    p.set(Q + "dateDelta", [np.array([delta])])
    for name, col in (("L_SHIPDATE", ship), ("L_RETURNFLAG", rf), ("L_LINESTATUS", ls),
                      ("L_QUANTITY", qty), ("L_EXTENDEDPRICE", price), ("L_DISCOUNT", disc),
                      ("L_TAX", tax)):
        p.set(name, [ln, od, col])
    assert p.run() == 21
    keep = ship <= 912470400 - delta
    g = rf.astype(np.int64) * 2 + ls
    groups = sorted(np.unique(g[keep]).tolist())
    for out, vals in (("count_order", None), ("sum_qty", qty), ("sum_base_price", price),
                      ("avg_disc_adt", disc), ("sum_disc_price", price * (1 - disc)),
                      ("sum_charge", price * (1 - disc) * (1 + tax))):
        k0, k1, v = p.get(Q + out)
        assert (k0.astype(np.int64) * 2 + k1).tolist() == groups, out
        for k, got in zip(groups, v):
            sel = keep & (g == k)
            if vals is None:
                assert got == sel.sum()
            else:
                want = vals[sel].sum()
                assert abs(got - want) <= 1e-12 * abs(want), (out, k)


def closure_plan():
    b = raplan.Builder("closure")
    for v in ("edge", "path", "next", "new", "flip"):
        b.var(v, ["uint32"], ["uint32"])
    b.sequence(0).move("path", "edge").goto(1)
    s = b.sequence(1)
    s.mapfilter("flip", "path", 2, b.always(), [b.idx(1, "uint32"), b.idx(0, "uint32")])
    s.join("new", 1, "flip", "edge", [b.idx(1, "uint32"), b.idx(2, "uint32")])
    s.setop("next", "union", "path", "new")
    s.cond("next", "path", 2, 3)
    b.sequence(3).move("path", "next").goto(1)
    b.sequence(2).halt()
    return b.entry_exit(0, 2).bytes()


@pytest.mark.parametrize("nodes,edges", [(40, 80), (300, 450)])
def test_transitive_closure_loop(ctx, nodes, edges):
    rng = np.random.default_rng(nodes)
    E = rng.integers(0, nodes, (edges, 2)).astype(np.uint32)
    E[:, 0] += np.uint32(1 << 31)  # keys above 2^31 keep the unsigned order
    E[:, 1] += np.uint32(1 << 31)
    p = raplan.Plan(ctx, closure_plan())
    p.set("edge", [E[:, 0], E[:, 1]])
    p.run()
    got = list(zip(*[c.tolist() for c in p.get("path")]))
    succ = {}
    for x, y in E.tolist():
        succ.setdefault(x, set()).add(y)
    want = set()
    for s0 in succ:
        seen, stack = set(), list(succ[s0])
        while stack:
            y = stack.pop()
            if y not in seen:
                seen.add(y)
                stack.extend(succ.get(y, ()))
        want |= {(s0, y) for y in seen}
    assert got == sorted(want)


def sets_plan():
    b = raplan.Builder("sets")
    for v in ("a", "b", "u", "i", "d", "neg"):
        b.var(v, ["entity:NATION"], ["int32"])
    b.var("f", ["entity:NATION"], ["float64"])
    b.var("c", fields=["int64"])
    b.var("prod", ["entity:NATION"], ["int32", "int64"])
    b.var("lo", ["entity:NATION"], ["float64"])
    b.var("hi", ["entity:NATION"], ["float64"])
    s = b.sequence(0)
    s.setop("u", "union", "a", "b").setop("i", "intersection", "a", "b")
    s.setop("d", "difference", "a", "b")
    s.single("c", [b.const_int(-7)])
    s.product("prod", "a", "c")
    s.agg("lo", "f", 2, [0], "min", 1, 1).agg("hi", "f", 2, [0], "max", 1, 1)
    s.mapfilter("neg", "a", 2, b.and_(b.comp("lt", b.idx(1, "int32"), b.const_int(0)),
                                      b.comp("ne", b.idx(0, "entity:NATION"), b.const_int(3))),
                [b.idx(0, "entity:NATION"), b.idx(1, "int32")])
    s.halt()
    return b.entry_exit(0, 0).bytes()


def test_set_operations_single_product_minmax(ctx):
    rng = np.random.default_rng(11)
    A = [(int(k), int(v)) for k, v in zip(rng.integers(0, 25, 3000), rng.integers(-2**31, 2**31, 3000))]
    A += [(k, v) for k, v in A[:200]]  # duplicates: the set operations see sets
    B = [(int(k), int(v)) for k, v in zip(rng.integers(0, 25, 2000), rng.integers(-2**31, 2**31, 2000))]
    B += A[:700]
    p = raplan.Plan(ctx, sets_plan())
    p.set("a", [np.array([x for x, _ in A]), np.array([y for _, y in A])])
    p.set("b", [np.array([x for x, _ in B]), np.array([y for _, y in B])])
    fk, fv = rng.integers(0, 20, 5000), rng.normal(size=5000)
    p.set("f", [fk, fv])
    p.run()

    def rel(name):
        return list(zip(*[c.tolist() for c in p.get(name)]))
    assert rel("u") == sorted(set(A) | set(B))
    assert rel("i") == sorted(set(A) & set(B))
    assert rel("d") == sorted(set(A) - set(B))
    assert rel("prod") == [(x, y, -7) for x, y in A]
    ks = sorted(set(fk.tolist()))
    assert rel("lo") == [(k, fv[fk == k].min()) for k in ks]
    assert rel("hi") == [(k, fv[fk == k].max()) for k in ks]
    assert rel("neg") == [(x, y) for x, y in A if y < 0 and x != 3]


def test_string_key_join_plan(rf, ctx):
    """Two relations with a STRING key, loaded from text: a key join (rows
    of each side ordered by their strings, sorted search on the strings),
    then a projection of the joined fields; every equal-name pair once."""
    from redfox_amd import loaders

    rng = np.random.default_rng(21)
    names = [b"FRANCE", b"GERMANY", b"CHINA", b"PERU", b"EGYPT", b"\xc3\x89IRE", b"KENYA"]
    a_rows = [(names[int(i)], int(v)) for i, v in zip(rng.integers(0, 6, 400),
                                                       rng.integers(-1000, 1000, 400))]
    b_rows = [(names[int(i)], float(v)) for i, v in zip(rng.integers(1, 7, 900),
                                                         rng.integers(0, 10 ** 6, 900))]
    b = raplan.Builder("strings")
    b.var("A", ["string"], ["int32"]).var("B", ["string"], ["float64"])
    b.var("J", ["string"], ["int32", "float64"]).var("P", fields=["int32", "float64"])
    s = b.sequence(0)
    s.join("J", 1, "A", "B")
    s.mapfilter("P", "J", 3, b.always(), [b.idx(1, "int32"), b.idx(2, "float64")])
    s.halt()
    p = raplan.Plan(ctx, b.entry_exit(0, 0).bytes())
    keep = []
    for name, rows, typ in (("A", a_rows, "i32"), ("B", b_rows, "f64")):
        text = b"".join(b"%s|%s|\n" % (k, str(v).encode()) for k, v in rows)
        raw = np.frombuffer(text, dtype=np.uint8)
        buf = rf.DeviceBuffer(len(raw) + 16)
        buf.upload(raw)
        (addr, val), n, _ = loaders.parse_tbl(ctx, buf, ["str", typ], "|", len(raw))
        p.set_columns(name, [(addr, "str"), (val, "i32" if typ == "i32" else "f64")], n)
        keep.append(buf)  # the strings live in the text buffers
    p.run()
    x, y = p.get("P")
    want = sorted((va, vb) for ka, va in a_rows for kb, vb in b_rows if ka == kb)
    assert sorted(zip(x.tolist(), y.tolist())) == want
    assert p.rows("J") == len(want)


def test_string_group_by_plan(rf, ctx):
    """AGG grouped by a STRING column (COUNT, TOTAL, MIN): groups are the
    distinct strings, each once, with its aggregate."""
    from redfox_amd import loaders

    rng = np.random.default_rng(51)
    modes = [b"TRUCK", b"MAIL", b"SHIP", b"AIR", b"REG AIR", b"\xc3\xa9t\xc3\xa9", b"", b"MAILBOX"]
    rows = [(modes[int(m)], float(v) / 4) for m, v in zip(rng.integers(0, len(modes), 20_000),
                                                         rng.integers(0, 4000, 20_000))]
    text = b"".join(b"%s|%r|\n" % (m, v) for m, v in rows)
    raw = np.frombuffer(text, dtype=np.uint8)
    buf = rf.DeviceBuffer(len(raw) + 16)
    buf.upload(raw)
    (addr, val), n, _ = loaders.parse_tbl(ctx, buf, ["str", "f64"], "|", len(raw))
    consumed = buf.download(np.uint8, len(raw)).tobytes()

    def string_at(a):
        off = int(a) - buf.ptr
        return consumed[off:consumed.index(b"\0", off)]
    b = raplan.Builder("groupby")
    b.var("L", ["string"], ["float64"])
    b.var("C", ["string"], ["int64"]).var("T", ["string"], ["float64"])
    b.var("M", ["string"], ["float64"])
    s = b.sequence(0)
    s.agg("C", "L", 2, [0], "count").agg("T", "L", 2, [0], "total", 1, 1)
    s.agg("M", "L", 2, [0], "min", 1, 1)
    s.halt()
    p = raplan.Plan(ctx, b.entry_exit(0, 0).bytes())
    p.set_columns("L", [(addr, "str"), (val, "f64")], n)
    p.run()
    groups = {}
    for m, v in rows:
        groups.setdefault(m, []).append(v)
    for name, f in (("C", len), ("T", sum), ("M", min)):
        a, v = p.get(name)
        got = {string_at(x): y for x, y in zip(a, v)}
        assert len(got) == len(a) == len(groups), name
        for m, vs in groups.items():
            assert abs(got[m] - f(vs)) <= 1e-12 * max(1.0, abs(f(vs))), (name, m)


def test_reference_no_op_plan_moves_a_to_b(ctx):
    # The reference's own plan file (TPC-H/data/NO_OP.binaryGpu) over the
    # reference's own A (TPC-H/data/A: Tuple<32,32> rows, field 0 in the
    # high word): B must come out byte-equal to A.
    import os
    golden = os.path.join(os.path.dirname(__file__), "golden")
    with open(os.path.join(golden, "no_op.binaryGpu"), "rb") as f:
        p = raplan.Plan(ctx, f.read())
    a = np.fromfile(os.path.join(golden, "tpch_A.bin"), dtype="<u8")
    p.set("A", [(a >> 32).astype(np.uint32).view(np.int32),
                (a & 0xFFFFFFFF).astype(np.uint32).view(np.int32)])
    assert p.run() == 3
    f0, f1 = p.get("B")
    b = (f0.view(np.uint32).astype(np.uint64) << 32) | f1.view(np.uint32).astype(np.uint64)
    assert b.astype("<u8").tobytes() == a.tobytes()


def _load_text_relation(rf, ctx, loaders, rows, types):
    """rows of (bytes, value) as '|'-separated text in a buffer of their
    own, parsed into (string address, value) columns; the buffer and a
    host copy of it (for reading strings back) are returned too."""
    text = b"".join(b"%s|%s|\n" % (k, str(v).encode()) for k, v in rows)
    raw = np.frombuffer(text, dtype=np.uint8)
    buf = rf.DeviceBuffer(len(raw) + 16)
    buf.upload(raw)
    cols, n, _ = loaders.parse_tbl(ctx, buf, types, "|", len(raw))
    return cols, n, buf, buf.download(np.uint8, len(raw)).tobytes()


def _string_reader(bufs):
    def at(a):
        for buf, host in bufs:
            off = int(a) - buf.ptr
            if 0 <= off < len(host):
                return host[off:host.index(b"\0", off)]
        raise AssertionError(f"address {a:#x} outside the text buffers")
    return at


def test_string_set_operations_plan(rf, ctx):
    """UNION / INTERSECTION / DIFFERENCE of (string, int32) relations whose
    strings live in two different text buffers (equal strings, different
    addresses): tuples compare by their strings' content."""
    from redfox_amd import loaders

    rng = np.random.default_rng(71)
    words = [b"AIR", b"MAIL", b"RAIL", b"SHIP", b"TRUCK", b"", b"FOB", b"REG AIR", b"\xc3\xa9t\xc3\xa9"]

    def rows(n, lo, hi):
        return [(words[int(w)], int(v)) for w, v in zip(rng.integers(lo, hi, n),
                                                         rng.integers(-3, 4, n))]
    a_rows, b_rows = rows(3000, 0, 7), rows(2500, 2, 9)
    b = raplan.Builder("strsets")
    for v in ("A", "B", "U", "I", "D"):
        b.var(v, ["string"], ["int32"])
    s = b.sequence(0)
    s.setop("U", "union", "A", "B").setop("I", "intersection", "A", "B")
    s.setop("D", "difference", "A", "B")
    s.halt()
    p = raplan.Plan(ctx, b.entry_exit(0, 0).bytes())
    keep = []
    for name, r in (("A", a_rows), ("B", b_rows)):
        (addr, val), n, buf, host = _load_text_relation(rf, ctx, loaders, r, ["str", "i32"])
        p.set_columns(name, [(addr, "str"), (val, "i32")], n)
        keep.append((buf, host))
    p.run()
    at = _string_reader(keep)
    sa, sb = set(a_rows), set(b_rows)
    for name, want in (("U", sa | sb), ("I", sa & sb), ("D", sa - sb)):
        addr, val = p.get(name)
        got = [(at(x), int(y)) for x, y in zip(addr, val)]
        assert len(got) == len(set(got)), name  # a set: each tuple once
        assert sorted(got) == sorted(want), name
        # in the string order of compare_string2 (signed bytes), then by value
        assert got == sorted(got, key=lambda t: (oracle.string_key(t[0]), t[1])), name


def test_multi_column_string_key_join_plan(rf, ctx):
    """A join on two key columns, (string, int32), from relations in two
    text buffers: every pair equal in both keys once."""
    from redfox_amd import loaders

    rng = np.random.default_rng(72)
    names = [b"ALGERIA", b"ARGENTINA", b"BRAZIL", b"CANADA", b"EGYPT", b"ETHIOPIA"]
    a_rows = [(names[int(i)], int(v)) for i, v in zip(rng.integers(0, 5, 700), rng.integers(0, 9, 700))]
    b_rows = [(names[int(i)], int(v)) for i, v in zip(rng.integers(1, 6, 500), rng.integers(0, 9, 500))]
    b = raplan.Builder("strjoin2")
    b.var("A", ["string", "int32"]).var("B", ["string", "int32"])
    b.var("J", ["string", "int32"])
    s = b.sequence(0)
    s.join("J", 2, "A", "B")
    s.halt()
    p = raplan.Plan(ctx, b.entry_exit(0, 0).bytes())
    keep = []
    for name, r in (("A", a_rows), ("B", b_rows)):
        (addr, val), n, buf, host = _load_text_relation(rf, ctx, loaders, r, ["str", "i32"])
        p.set_columns(name, [(addr, "str"), (val, "i32")], n)
        keep.append((buf, host))
    p.run()
    at = _string_reader(keep)
    addr, val = p.get("J")
    got = sorted((at(x), int(y)) for x, y in zip(addr, val))
    want = sorted(ka for ka in a_rows for kb in b_rows if ka == kb)
    assert got == want


def test_group_by_string_and_int_plan(rf, ctx):
    """AGG over two domain columns, (string, int32): COUNT and TOTAL per
    distinct pair, the strings compared by content."""
    from redfox_amd import loaders

    rng = np.random.default_rng(73)
    flags = [b"A", b"N", b"R", b"NR", b"\xc3\xa9"]
    rows = [(flags[int(f)], int(s), float(v) / 8) for f, s, v in zip(
        rng.integers(0, 5, 20_000), rng.integers(0, 3, 20_000), rng.integers(0, 8000, 20_000))]
    text = b"".join(b"%s|%d|%r|\n" % r for r in rows)
    raw = np.frombuffer(text, dtype=np.uint8)
    buf = rf.DeviceBuffer(len(raw) + 16)
    buf.upload(raw)
    (addr, st, val), n, _ = loaders.parse_tbl(ctx, buf, ["str", "i32", "f64"], "|", len(raw))
    at = _string_reader([(buf, buf.download(np.uint8, len(raw)).tobytes())])
    b = raplan.Builder("groupby2")
    b.var("L", ["string", "int32"], ["float64"])
    b.var("C", ["string", "int32"], ["int64"]).var("T", ["string", "int32"], ["float64"])
    s = b.sequence(0)
    s.agg("C", "L", 3, [0, 1], "count").agg("T", "L", 3, [0, 1], "total", 2, 2)
    s.halt()
    p = raplan.Plan(ctx, b.entry_exit(0, 0).bytes())
    p.set_columns("L", [(addr, "str"), (st, "i32"), (val, "f64")], n)
    p.run()
    groups = {}
    for f, s_, v in rows:
        groups.setdefault((f, s_), []).append(v)
    for name, fn in (("C", len), ("T", sum)):
        a, k2, v = p.get(name)
        got = {(at(x), int(y)): z for x, y, z in zip(a, k2, v)}
        assert len(got) == len(a) == len(groups), name
        for key, vs in groups.items():
            assert abs(got[key] - fn(vs)) <= 1e-12 * max(1.0, abs(fn(vs))), (name, key)


def test_min_max_of_strings_plan(rf, ctx):
    """MIN/MAX whose value column is a STRING: the smallest/largest string
    of each group in the string order of joins and sorts (bytes as signed
    chars, oracle.string_key), grouped by an int32 key, by a string key and
    by (string, int32); TOTAL of a string column is refused."""
    from redfox_amd import loaders

    rng = np.random.default_rng(97)
    names = [b"Brand#13", b"Brand#1", b"brand#13", b"", b"\xc3\xa9t\xc3\xa9", b"Brand#13x",
             b"zz", b"Brand#2", b"\x7f", b"A"]
    mode = [b"AIR", b"MAIL", b"SHIP", b"\xc3\xa9"]
    rows = [(int(k), mode[int(m)], names[int(s)]) for k, m, s in zip(
        rng.integers(-3, 40, 30_000), rng.integers(0, 4, 30_000), rng.integers(0, 10, 30_000))]
    text = b"".join(b"%d|%s|%s|\n" % r for r in rows)
    raw = np.frombuffer(text, dtype=np.uint8)
    buf = rf.DeviceBuffer(len(raw) + 16)
    buf.upload(raw)
    (key, md, nm), n, _ = loaders.parse_tbl(ctx, buf, ["i32", "str", "str"], "|", len(raw))
    assert n == len(rows)
    at = _string_reader([(buf, buf.download(np.uint8, len(raw)).tobytes())])
    b = raplan.Builder("strminmax")
    b.var("L", ["int32", "string"], ["string"])
    for d in ("KLO", "KHI"):
        b.var(d, ["int32"], ["string"])
    for d in ("MLO", "MHI"):
        b.var(d, ["string"], ["string"])
    for d in ("PLO", "PHI"):
        b.var(d, ["string", "int32"], ["string"])
    s = b.sequence(0)
    s.agg("KLO", "L", 3, [0], "min", 2, 2).agg("KHI", "L", 3, [0], "max", 2, 2)
    s.agg("MLO", "L", 3, [1], "min", 2, 2).agg("MHI", "L", 3, [1], "max", 2, 2)
    s.agg("PLO", "L", 3, [1, 0], "min", 2, 2).agg("PHI", "L", 3, [1, 0], "max", 2, 2)
    s.halt()
    p = raplan.Plan(ctx, b.entry_exit(0, 0).bytes())
    p.set_columns("L", [(key, "i32"), (md, "str"), (nm, "str")], n)
    p.run()
    lo = lambda v: min(v, key=oracle.string_key)  # noqa: E731
    hi = lambda v: max(v, key=oracle.string_key)  # noqa: E731
    assert hi([b"\x7f", b"\xc3\xa9"]) == b"\x7f" and lo([b"", b"\x80"]) == b""
    for dest, group, fn in (("KLO", lambda r: (r[0],), lo), ("KHI", lambda r: (r[0],), hi),
                            ("MLO", lambda r: (r[1],), lo), ("MHI", lambda r: (r[1],), hi),
                            ("PLO", lambda r: (r[1], r[0]), lo),
                            ("PHI", lambda r: (r[1], r[0]), hi)):
        want = {}
        for r in rows:
            want.setdefault(group(r), []).append(r[2])
        cols = p.get(dest)
        got = {}
        for vals in zip(*[c.tolist() for c in cols]):
            g = tuple(at(v) if isinstance(t, bytes) else int(v)
                      for v, t in zip(vals[:-1], group(rows[0])))
            got[g] = at(vals[-1])
        assert len(got) == len(cols[0]) == len(want), dest
        assert got == {g: fn(v) for g, v in want.items()}, dest
        if dest.startswith("K"):
            assert cols[0].tolist() == sorted(k for (k,) in want)  # int keys in key order
    b2 = raplan.Builder("strtotal")
    b2.var("L", ["int32", "string"], ["string"]).var("BAD", ["int32"], ["float64"])
    b2.sequence(0).agg("BAD", "L", 3, [0], "total", 2, 2).halt()
    p2 = raplan.Plan(ctx, b2.entry_exit(0, 0).bytes())
    p2.set_columns("L", [(key, "i32"), (md, "str"), (nm, "str")], n)
    with pytest.raises(ValueError, match="TOTAL"):
        p2.run()
