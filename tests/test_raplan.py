"""RA plan front end without a GPU: the plan schema pinned by hand-encoded
wire bytes (field numbers and wire types of RelationalAlgebra.proto and
BloxCompiler.proto), the Q1 plan's shape, malformed input, and the packed
key layout checked through the oracle's evaluator (pack then unpack gives
the columns back, order is preserved) at edge values."""
import numpy as np
import pytest

import oracle
from redfox_amd import mapfilter as M
from redfox_amd import raplan


def varint(v):
    out = bytearray()
    v &= (1 << 64) - 1
    while True:
        b = v & 0x7F
        v >>= 7
        out.append(b | (0x80 if v else 0))
        if not v:
            return bytes(out)


def fld(num, v):
    """A field: varint for ints, length-delimited for bytes."""
    if isinstance(v, int):
        return varint(num << 3) + varint(v)
    return varint(num << 3 | 2) + varint(len(v)) + v


def test_schema_field_numbers_from_wire_bytes():
    # Type{kind UNARY(1), unary{name}} (BloxCompiler.proto:172-204)
    entity = fld(1, 1) + fld(3, fld(1, b"ORDER"))
    f64 = fld(1, 0) + fld(2, fld(1, 8) + fld(2, 64))
    var = fld(1, b"R") + fld(2, entity) + fld(3, f64) + fld(4, b"")   # GPUVariable
    # Exp INDEX (tag 1) offset 1 with its type: Exp.index = 3, Index{offset 1, etyp 2}
    exp = fld(1, 1) + fld(3, fld(1, 1) + fld(2, f64))
    # Compare ALWAYS = 3; GPUMapFilter{srcWidth 2, predicate 3, reordering 4, srcA 5}
    mf = fld(2, 2) + fld(3, fld(1, 3)) + fld(4, exp) + fld(5, b"R")
    op = fld(1, 10) + fld(10, mf)                                    # GPUOperator MAPFILTER
    assign = fld(1, b"S") + fld(2, op)                               # GPUAssign{dest, op}
    cmd = fld(1, 2) + fld(2, assign)                                 # GPUCommand ASSIGN
    join = fld(2, 1) + fld(3, b"R") + fld(4, b"R")                   # GPUJoin keyCount 1
    cmd2 = fld(1, 2) + fld(2, fld(1, b"T") + fld(2, fld(1, 7) + fld(7, join)))
    halt = fld(1, 6)
    seq = fld(1, 0) + fld(2, cmd) + fld(2, cmd2) + fld(2, halt)     # GPUSequence
    graph = fld(1, b"g") + fld(2, var) + fld(3, seq) + fld(4, 0) + fld(5, 0)
    g = raplan.parse(graph)
    assert g.graphName == "g" and g.entry == 0 and g.exit == 0
    v = g.variables[0]
    assert v.varName == "R" and v.keys[0].unary.name == "ORDER"
    assert v.fields[0].primitive.kind == 8 and v.fields[0].primitive.capacity == 64
    assert raplan.column_type(v.keys[0]) == "u32" and raplan.column_type(v.fields[0]) == "f64"
    ops = g.sequences[0].operators
    m = ops[0].assign.op.mapfilter
    assert ops[0].assign.dest == "S" and m.srcWidth == 2 and m.predicate.tag == 3
    assert m.reordering[0].index.offset == 1 and m.srcA == "R"
    assert ops[1].assign.op.join.keyCount == 1 and ops[2].tag == 6
    # The builder writes the same bytes for the same plan.
    b = raplan.Builder("g")
    b.var("R", ["entity:ORDER"], ["float64"])
    s = b.sequence(0)
    s.mapfilter("S", "R", 2, b.always(), [b.idx(1, "float64")])
    s.join("T", 1, "R", "R")
    s.halt()
    assert b.entry_exit(0, 0).bytes() == graph


def test_malformed_plans_are_refused():
    good = raplan.q1_plan()
    with pytest.raises(ValueError):
        raplan.parse(good[:len(good) // 2])
    with pytest.raises(ValueError):
        raplan.parse(b"\x0a\x05ab")                    # truncated string
    with pytest.raises(ValueError):
        raplan.parse(fld(1, b"g") + fld(4, 0))        # exit (required) missing
    for cut in range(0, len(good), 97):
        try:
            raplan.parse(good[:cut])
        except ValueError:
            pass


def test_q1_plan_shape():
    """The command sequence of q1.man.cpp:948-2329."""
    g = raplan.parse(raplan.q1_plan())
    assert len(g.variables) == 27
    ops = g.sequences[0].operators
    tags = [(c.assign.op.tag if c.tag == 2 else -c.tag) for c in ops]
    assert tags == [10, 7, 10, 7, 7, 11, 7, 11, 7, 11, 7, 11, 10, 10, 11, 7, 10, 10, 11, -4]
    assert [c.assign.dest for c in ops[:3]] == ["+$logicQ1:_startDate", "j_1", "m_1"]
    assert ops[1].assign.op.join.keyCount == 0 and ops[3].assign.op.join.keyCount == 2
    aggs = [c.assign.op.agg for c in ops if c.tag == 2 and c.assign.op.tag == 11]
    assert all(list(a.domains) == [2, 3] for a in aggs)
    assert [a.range[0].tag for a in aggs] == [0, 1, 1, 1, 1, 1]
    assert g.sequences[1].operators[0].tag == 6 and (g.entry, g.exit) == (0, 1)


class _Buf:
    ptr = 0


@pytest.mark.parametrize("types", [["u8", "u32"], ["u32", "u32"], ["i8", "u16", "i32"],
                                   ["i64"], ["u64"], ["i32", "i32"], ["u8"], []])
def test_packed_keys_round_trip_and_order(types):
    rng = np.random.default_rng(len(types) * 7 + 1)
    n = 400
    np_t = {"u8": np.uint8, "u16": np.uint16, "u32": np.uint32, "u64": np.uint64, "i8": np.int8,
            "i16": np.int16, "i32": np.int32, "i64": np.int64}
    cols = []
    for t in types:
        info = np.iinfo(np_t[t])
        c = rng.integers(info.min, info.max, n, dtype=np_t[t], endpoint=True)
        c[:4] = [info.min, info.max, 0, info.max - 1] if info.min == 0 else [info.min, info.max, -1, 0]
        cols.append(c)
    kt, bits, fields = raplan.key_layout(types)
    x = M.MapFilter([(_Buf(), t) for t in types])
    x.emit(raplan._pack_expr(x.cols, fields), kt)
    _, (keys,) = oracle.mapfilter(None, [(e.tree(), t) for e, t in x.outputs], cols, n)
    if not types:
        assert np.all(keys == 0)
        return
    # order: lexicographic order of the columns equals the keys' unsigned order
    lex = np.lexsort(tuple(c for c in reversed(cols)))
    assert np.all(np.diff(keys[lex].astype(np.uint64)) >= 0)
    y = M.MapFilter([(_Buf(), raplan._signed(kt))])
    for f, t in zip(fields, types):
        y.emit(raplan._unpack_expr(y.cols[0], f, bits), t)
    kcol = keys.view(np.int64) if kt == "u64" else keys
    _, back = oracle.mapfilter(None, [(e.tree(), t) for e, t in y.outputs], [kcol], n)
    for c, b in zip(cols, back):
        assert np.array_equal(c, b)


def test_wide_keys_and_unsupported_types_are_refused():
    with pytest.raises(ValueError):
        raplan.key_layout(["u64", "u8"])
    with pytest.raises(ValueError):
        raplan.key_layout(["f64"])
    m = raplan.messages()
    t = m["Type"](kind=0)
    t.primitive.kind = 10  # DECIMAL
    with pytest.raises(ValueError):
        raplan.column_type(t)
    t.primitive.kind = 9  # STRING: device addresses of NUL-terminated strings
    assert raplan.column_type(t) == "str"
    with pytest.raises(ValueError):
        raplan.key_layout(["str"])
    t = m["Type"](kind=1)
    t.unary.name = "PLANET"
    with pytest.raises(ValueError):
        raplan.column_type(t)


def test_q1_arithmetic_lowering_matches_the_expression():
    """m_2's reordering 4 * (1 - 5) lowered and evaluated by the oracle:
    one rounding per operation, as price * (1 - disc)."""
    b = raplan.Builder("x")
    e = b.arith("mul", b.idx(0, "float64"), b.arith("sub", b.const_float(1), b.idx(1, "float64")))
    x = M.MapFilter([(_Buf(), "f64"), (_Buf(), "f64")])
    v, t = raplan.lower_exp(e, lambda i: (x.cols[i], "f64"))
    assert t == "f64"
    rng = np.random.default_rng(4)
    p, d = rng.random(100) * 1e4, rng.integers(0, 11, 100) / 100.0
    got, _ = oracle.mapfilter_eval(v.tree(), [p, d])
    assert np.array_equal(got, p * (1.0 - d))
    sub = b.call("subtract", [b.const_datetime(912470400), b.idx(0, "int64"), b.const_string("days")])
    v, t = raplan.lower_exp(sub, lambda i: (M.Column(0, "i64"), "i64"))
    got, _ = oracle.mapfilter_eval(v.tree(), [np.array([7776000], dtype=np.int64)])
    assert t == "i64" and got.tolist() == [912470400 - 7776000]


def test_string_predicates_are_refused():
    b = raplan.Builder("x")
    x = M.MapFilter([(_Buf(), "u64")])
    col = lambda i: (x.cols[i], "str")  # noqa: E731
    with pytest.raises(ValueError, match="string"):
        raplan.lower_compare(b.comp("eq", b.idx(0, "string"), b.const_int(1)), col)
    with pytest.raises(ValueError, match="string"):
        raplan.lower_exp(b.arith("add", b.idx(0, "string"), b.const_int(1), "int"), col)


GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


def test_reference_no_op_plan_parses():
    # TPC-H/data/NO_OP.binaryGpu: the plan file the reference's front end
    # wrote for NO_OP.textGpu (B := A over two 32-bit INT fields).
    with open(f"{GOLDEN}/no_op.binaryGpu", "rb") as f:
        g = raplan.parse(f.read())
    assert g.graphName == "TPC-H/NO_OP.lbb"
    assert [v.varName for v in g.variables] == ["A", "B"]
    for v in g.variables:
        assert [raplan.column_type(t) for t in v.fields] == ["i32", "i32"]
        assert len(v.keys) == 0
    assert [s.uniqueIdentifier for s in g.sequences] == [0, 1]
    s0, s1 = g.sequences
    assert [c.tag for c in s0.operators] == [5, 4]  # MOVE, GOTO
    assert (s0.operators[0].move.dest, s0.operators[0].move.src1) == ("B", "A")
    assert s0.operators[1].jump.target == 1
    assert [c.tag for c in s1.operators] == [6]  # HALT
    assert (g.entry, g.exit) == (0, 1)
