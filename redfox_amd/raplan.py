"""RA plan front end: runs a Red Fox relational-algebra plan (a serialized
GPUGraph, RelationalAlgebra.proto:27-454) on the device operators.

The reference lowers a plan to HIR with generated CUDA per node
(RelationalAlgebraCompiler.cpp: addJoin :4466, addMapFilter :7220,
addMapJoin :7351, addAgg :7852, addUnion :3541, addDifference :4277,
addSingle :4126, addMove :8742, addBlock :8974).  Here every node runs
directly on the library's operators, columns kept as device arrays:

* a relation is a list of columns (one device array per key and field, in
  declaration order) and a row count; entity types are small unsigned
  integers (RETURNFLAG, LINESTATUS, LINENUMBER, REGION, NATION: u8;
  SUPPLIER: u16; ORDER, PART, CUSTOMER: u32; as the reference sizes them,
  RelationalAlgebraCompiler.cpp:1698-1733), DATETIME is i64 seconds,
  INT/UINT/FLOAT keep their capacity;
* MAPFILTER (predicate + reordering expressions) is one rfx_mapfilter
  pass;
* JOIN on the first keyCount columns packs those columns into one integer
  key per side, sorts (key, row) pairs, finds each left key's equal run on
  the right (rfx_join_find_bounds), expands it to row pairs
  (rfx_join_main), and gathers the output columns of the combined tuple
  (A's columns, then B's non-key columns; RelationalAlgebraCompiler.cpp:
  7366-7368) through the pairs in one MAPFILTER pass with gathers (LDX),
  which also evaluates computed join arguments.  keyCount 0 (and PRODUCT)
  is the same with a constant key: every left row meets every right row;
* AGG packs the domain columns into a key, sorts (key, value) pairs and
  reduces each run (TOTAL/MIN/MAX over FP64 values, COUNT), then unpacks
  the domain columns from the run keys: the result is sorted by its
  domain, as the reference's (:7852-8000);
* UNION / INTERSECTION / DIFFERENCE pack whole tuples, sort and dedupe
  both sides and run the merge-path set operation;
* SINGLE is a one-row relation of constants; MOVE copies; COND compares
  two relations' contents; GOTO and HALT steer between sequences.

Packed keys: each column maps order-preservingly to its unsigned width
(signed types biased by 2^(w-1)); the first column is the most
significant.  A key is one u32 (<= 32 bits) or one u64 (<= 64 bits).
Wider keys are refused.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import mapfilter as M
from . import rfx

# ---------------------------------------------------------------------------
# Schema: the messages of BloxCompiler.proto:115-204 (constants and types)
# and RelationalAlgebra.proto:27-454 that a plan uses, restated for the
# protobuf runtime.  (label, type, name, number); labels r/o/p = required,
# optional, repeated; a type naming a message or enum of this schema.
# ---------------------------------------------------------------------------
_COMMON = "blox.common.protocol"
_GPU = "blox.compiler.gpu"

_ENUMS = {
    (_COMMON, "Constant.Kind"): [("BOOL", 0), ("INT", 1), ("FLOAT", 2), ("STRING", 3),
                                 ("DATETIME", 4), ("UINT", 5)],
    (_COMMON, "Type.Kind"): [("PRIMITIVE", 0), ("UNARY", 1)],
    (_COMMON, "PrimitiveType.Kind"): [("BOOL", 0), ("COLOR", 1), ("SHAPE", 2), ("IMAGE", 3),
                                      ("BLOB", 4), ("DATETIME", 5), ("INT", 6), ("UINT", 7),
                                      ("FLOAT", 8), ("STRING", 9), ("DECIMAL", 10)],
    (_GPU, "OperatorTag"): [("UNION", 2), ("INTERSECTION", 3), ("PRODUCT", 4), ("SINGLE", 5),
                            ("DIFFERENCE", 6), ("JOIN", 7), ("MAPFILTER", 10), ("AGG", 11)],
    (_GPU, "CommandTag"): [("ASSIGN", 2), ("COND", 3), ("GOTO", 4), ("MOVE", 5), ("HALT", 6)],
    (_GPU, "Comparison"): [("Eq", 0), ("Ne", 1), ("Lt", 2), ("Le", 3), ("Gt", 4), ("Ge", 5)],
    (_GPU, "AddressTag"): [("VARADDR", 2), ("CONSTADDR", 3)],
    (_GPU, "ArithmeticOp"): [("Add", 0), ("Multiply", 1), ("Subtract", 2), ("Divide", 3),
                             ("Mod", 4)],
    (_GPU, "Numeric"): [("NInt", 0), ("NUInt", 1), ("NFloat", 2), ("NDecimal", 3)],
    (_GPU, "Fixed"): [("FString", 0), ("FBool", 1), ("FDatetime", 2), ("FColor", 3)],
    (_GPU, "UnivTag"): [("Ntag", 0), ("Ftag", 1)],
    (_GPU, "ExpTag"): [("CONSTEXP", 0), ("INDEX", 1), ("NAME", 2), ("ARITHEXP", 3),
                       ("CONVERT", 4), ("MIXED", 5), ("CALL", 6)],
    (_GPU, "CompareTag"): [("COMP", 0), ("TEST", 1), ("AND", 2), ("ALWAYS", 3)],
    (_GPU, "AggTag"): [("COUNT", 0), ("TOTAL", 1), ("MIN", 2), ("MAX", 3)],
}

_C = "." + _COMMON + "."
_G = "." + _GPU + "."
_MESSAGES = {
    (_COMMON, "Constant"): [("r", _C + "Constant.Kind", "kind", 1),
                            ("o", _C + "BoolConstant", "bool_constant", 2),
                            ("o", _C + "IntConstant", "int_constant", 3),
                            ("o", _C + "FloatConstant", "float_constant", 4),
                            ("o", _C + "StringConstant", "string_constant", 5),
                            ("o", _C + "DateTimeConstant", "date_time_constant", 6),
                            ("o", _C + "UnsignedIntConstant", "uint_constant", 7)],
    (_COMMON, "BoolConstant"): [("r", "bool", "value", 1)],
    (_COMMON, "IntConstant"): [("r", "int64", "value", 1), ("r", _C + "Type", "type", 2)],
    (_COMMON, "UnsignedIntConstant"): [("r", "uint64", "value", 1), ("r", _C + "Type", "type", 2)],
    (_COMMON, "FloatConstant"): [("r", "string", "value", 1), ("r", _C + "Type", "type", 2)],
    (_COMMON, "StringConstant"): [("r", "string", "value", 1)],
    (_COMMON, "DateTimeConstant"): [("o", "int64", "value", 1)],
    (_COMMON, "Type"): [("r", _C + "Type.Kind", "kind", 1),
                        ("o", _C + "PrimitiveType", "primitive", 2),
                        ("o", _C + "UnaryPredicateType", "unary", 3)],
    (_COMMON, "PrimitiveType"): [("r", _C + "PrimitiveType.Kind", "kind", 1),
                                 ("o", "uint32", "capacity", 2)],
    (_COMMON, "UnaryPredicateType"): [("r", "string", "name", 1)],
    (_GPU, "GPUGraph"): [("r", "string", "graphName", 1), ("p", _G + "GPUVariable", "variables", 2),
                         ("p", _G + "GPUSequence", "sequences", 3), ("r", "int32", "entry", 4),
                         ("r", "int32", "exit", 5)],
    (_GPU, "GPUVariable"): [("r", "string", "varName", 1), ("p", _C + "Type", "keys", 2),
                            ("p", _C + "Type", "fields", 3), ("r", "bytes", "initialData", 4)],
    (_GPU, "GPUUnion"): [("r", "string", "srcA", 2), ("r", "string", "srcB", 3)],
    (_GPU, "GPUIntersection"): [("r", "string", "srcA", 2), ("r", "string", "srcB", 3)],
    (_GPU, "GPUProduct"): [("r", "string", "srcA", 2), ("r", "string", "srcB", 3)],
    (_GPU, "GPUSingle"): [("p", _C + "Constant", "element", 1)],
    (_GPU, "GPUDifference"): [("r", "string", "srcA", 3), ("r", "string", "srcB", 4)],
    (_GPU, "GPUJoin"): [("r", "int32", "keyCount", 2), ("r", "string", "srcA", 3),
                        ("r", "string", "srcB", 4), ("p", _G + "Exp", "args", 5)],
    (_GPU, "GPUMapFilter"): [("r", "int32", "srcWidth", 2), ("r", _G + "Compare", "predicate", 3),
                             ("p", _G + "Exp", "reordering", 4), ("r", "string", "srcA", 5)],
    (_GPU, "GPUAgg"): [("r", "int32", "srcWidth", 2), ("p", "int32", "domains", 3),
                       ("p", _G + "Agg", "range", 4), ("r", "string", "srcA", 5)],
    (_GPU, "GPUOperator"): [("r", _G + "OperatorTag", "tag", 1),
                            ("o", _G + "GPUUnion", "unionop", 2),
                            ("o", _G + "GPUIntersection", "intersection", 3),
                            ("o", _G + "GPUProduct", "product", 4),
                            ("o", _G + "GPUSingle", "single", 5),
                            ("o", _G + "GPUDifference", "difference", 6),
                            ("o", _G + "GPUJoin", "join", 7),
                            ("o", _G + "GPUMapFilter", "mapfilter", 10),
                            ("o", _G + "GPUAgg", "agg", 11)],
    (_GPU, "GPUAssign"): [("r", "string", "dest", 1), ("r", _G + "GPUOperator", "op", 2)],
    (_GPU, "GPUCond"): [("r", "string", "src1", 1), ("r", "string", "src2", 2),
                        ("r", "int32", "yes", 3), ("r", "int32", "no", 4)],
    (_GPU, "GPUGoto"): [("r", "int32", "target", 1)],
    (_GPU, "GPUMove"): [("r", "string", "dest", 1), ("r", "string", "src1", 2)],
    (_GPU, "GPUCommand"): [("r", _G + "CommandTag", "tag", 1), ("o", _G + "GPUAssign", "assign", 2),
                           ("o", _G + "GPUCond", "cond", 3), ("o", _G + "GPUGoto", "jump", 4),
                           ("o", _G + "GPUMove", "move", 5), ("o", "int32", "halt", 6)],
    (_GPU, "Address"): [("r", _G + "AddressTag", "tag", 1), ("o", "int32", "varaddr", 2),
                        ("o", _C + "Constant", "constaddr", 3)],
    (_GPU, "GPUCompare"): [("r", _G + "Comparison", "comparison", 1),
                           ("r", _C + "PrimitiveType", "optype", 2),
                           ("r", _G + "Address", "addr1", 3), ("r", _G + "Address", "addr2", 4)],
    (_GPU, "GPUSequence"): [("r", "int32", "uniqueIdentifier", 1),
                            ("p", _G + "GPUCommand", "operators", 2)],
    (_GPU, "NumType"): [("r", _G + "Numeric", "typ", 1), ("r", "int32", "size", 2)],
    (_GPU, "Univ"): [("r", _G + "UnivTag", "tag", 1), ("o", _G + "NumType", "n", 2),
                     ("o", _G + "Fixed", "f", 3)],
    (_GPU, "ConstExp"): [("r", _C + "Constant", "literal", 1)],
    (_GPU, "Name"): [("r", "string", "ename", 1), ("r", _C + "Type", "etype", 2)],
    (_GPU, "Index"): [("r", "int32", "offset", 1), ("r", _C + "Type", "etyp", 2)],
    (_GPU, "ArithExp"): [("r", _G + "NumType", "domain", 1), ("r", _G + "ArithmeticOp", "op", 2),
                         ("r", _G + "Exp", "exp1", 3), ("r", _G + "Exp", "exp2", 4)],
    (_GPU, "Convert"): [("r", _G + "Univ", "convDom", 1), ("r", _G + "Univ", "convRng", 2),
                        ("r", _G + "Exp", "exp2", 3)],
    (_GPU, "Mixed"): [("r", _G + "NumType", "domain", 1), ("r", _G + "NumType", "range", 2),
                      ("r", _G + "Exp", "exp1", 3)],
    (_GPU, "Call"): [("r", _G + "Fixed", "calltype", 1), ("r", "string", "callname", 2),
                     ("p", _G + "Exp", "args", 3)],
    (_GPU, "Exp"): [("r", _G + "ExpTag", "tag", 1), ("o", _G + "ConstExp", "constExp", 2),
                    ("o", _G + "Index", "index", 3), ("o", _G + "Name", "name", 4),
                    ("o", _G + "ArithExp", "arithexp", 5), ("o", _G + "Convert", "convert", 6),
                    ("o", _G + "Mixed", "mixed", 7), ("o", _G + "Call", "call", 8)],
    (_GPU, "Comp"): [("r", _G + "Comparison", "comparison", 1), ("r", _G + "Exp", "op1", 2),
                     ("r", _G + "Exp", "op2", 3)],
    (_GPU, "Test"): [("r", _G + "Fixed", "testtype", 1), ("r", "string", "testname", 2),
                     ("p", _G + "Exp", "ops", 3)],
    (_GPU, "And"): [("r", _G + "Compare", "and1", 1), ("r", _G + "Compare", "and2", 2)],
    (_GPU, "Compare"): [("r", _G + "CompareTag", "tag", 1), ("o", _G + "Comp", "comp", 2),
                        ("o", _G + "Test", "test", 3), ("o", _G + "And", "andComp", 4)],
    (_GPU, "Count"): [("r", "int32", "aggrng", 1)],
    (_GPU, "Total"): [("r", "int32", "aggdom", 1), ("r", "int32", "aggrng", 2)],
    (_GPU, "Min"): [("r", "int32", "aggdom", 1), ("r", "int32", "aggrng", 2)],
    (_GPU, "Max"): [("r", "int32", "aggdom", 1), ("r", "int32", "aggrng", 2)],
    (_GPU, "Agg"): [("r", _G + "AggTag", "tag", 1), ("o", _G + "Count", "count", 2),
                    ("o", _G + "Total", "total", 3), ("o", _G + "Min", "min", 4),
                    ("o", _G + "Max", "max", 5)],
}

_classes = None


def messages():
    """Message classes of the plan schema, by name ("GPUGraph", "Exp", ...)."""
    global _classes
    if _classes is not None:
        return _classes
    from google.protobuf import descriptor_pb2, descriptor_pool, message_factory

    F = descriptor_pb2.FieldDescriptorProto
    scalar = {"string": F.TYPE_STRING, "bytes": F.TYPE_BYTES, "bool": F.TYPE_BOOL,
              "int32": F.TYPE_INT32, "int64": F.TYPE_INT64, "uint32": F.TYPE_UINT32,
              "uint64": F.TYPE_UINT64}
    labels = {"r": F.LABEL_REQUIRED, "o": F.LABEL_OPTIONAL, "p": F.LABEL_REPEATED}
    enum_names = {"." + pkg + "." + name for pkg, name in _ENUMS}
    pool = descriptor_pool.DescriptorPool()
    for pkg, fname in ((_COMMON, "rfx_blox_common.proto"), (_GPU, "rfx_ra_plan.proto")):
        fd = descriptor_pb2.FileDescriptorProto(name=fname, package=pkg, syntax="proto2")
        if pkg == _GPU:
            fd.dependency.append("rfx_blox_common.proto")
        nested: Dict[str, object] = {}
        for (p, name), fields in _MESSAGES.items():
            if p != pkg:
                continue
            m = fd.message_type.add(name=name)
            nested[name] = m
            for label, typ, fname_, num in fields:
                f = m.field.add(name=fname_, number=num, label=labels[label])
                if typ in scalar:
                    f.type = scalar[typ]
                else:
                    f.type = F.TYPE_ENUM if typ in enum_names else F.TYPE_MESSAGE
                    f.type_name = typ
        for (p, name), values in _ENUMS.items():
            if p != pkg:
                continue
            if "." in name:  # nested in its message
                outer, inner = name.split(".")
                e = nested[outer].enum_type.add(name=inner)
            else:
                e = fd.enum_type.add(name=name)
            for vname, v in values:
                e.value.add(name=vname, number=v)
        pool.Add(fd)
    _classes = {}
    for (pkg, name) in _MESSAGES:
        _classes[name] = message_factory.GetMessageClass(
            pool.FindMessageTypeByName(pkg + "." + name))
    return _classes


def parse(data: bytes):
    """A GPUGraph from its wire bytes (raises on malformed input or a
    missing required field)."""
    from google.protobuf.message import DecodeError

    g = messages()["GPUGraph"]()
    try:
        g.ParseFromString(bytes(data))
    except DecodeError as e:
        raise ValueError(f"not a GPUGraph: {e}") from None
    if not g.IsInitialized():
        raise ValueError("GPUGraph lacks required fields: " +
                         ", ".join(g.FindInitializationErrors()))
    return g


# ---------------------------------------------------------------------------
# Types
# ---------------------------------------------------------------------------
_ENTITY = {"RETURNFLAG": "u8", "LINESTATUS": "u8", "LINENUMBER": "u8", "REGION": "u8",
           "NATION": "u8", "SUPPLIER": "u16", "ORDER": "u32", "PART": "u32", "CUSTOMER": "u32"}
_NP = {"u8": np.uint8, "u16": np.uint16, "u32": np.uint32, "u64": np.uint64, "i8": np.int8,
       "i16": np.int16, "i32": np.int32, "i64": np.int64, "f32": np.float32, "f64": np.float64,
       "str": np.uint64}


def _mft(t: str) -> str:
    """The MAPFILTER element type of a relation column type (string columns
    are carried as their 64-bit addresses)."""
    return "u64" if t == "str" else t
_BITS = {"u8": 8, "u16": 16, "u32": 32, "u64": 64, "i8": 8, "i16": 16, "i32": 32, "i64": 64}


def column_type(t) -> str:
    """The device element type of a plan Type (RelationalAlgebraCompiler.cpp:
    1557-1800)."""
    if t.kind == 1:  # UNARY: an entity
        if t.unary.name not in _ENTITY:
            raise ValueError(f"entity type {t.unary.name!r} is not supported")
        return _ENTITY[t.unary.name]
    k, cap = t.primitive.kind, t.primitive.capacity or 64
    if k == 0:  # BOOL
        return "u8"
    if k == 5:  # DATETIME: seconds since 1970
        return "i64"
    if k in (6, 7):  # INT, UINT
        if cap not in (8, 16, 32, 64):
            raise ValueError(f"invalid integer size {cap}")
        return ("i" if k == 6 else "u") + str(cap)
    if k == 8:  # FLOAT
        if cap not in (32, 64):
            raise ValueError(f"invalid float size {cap}")
        return "f" + str(cap)
    if k == 9:  # STRING: the address of a NUL-terminated string in device memory
        return "str"
    raise ValueError(f"primitive type {k} is not supported (decimals and blobs are "
                     "outside the numeric operators)")


def _num_type(nt) -> str:
    size = nt.size or 64
    if nt.typ == 2:
        return "f64" if size >= 64 else "f32"
    if nt.typ in (0, 1):
        return ("i" if nt.typ == 0 else "u") + str(size if size in (8, 16, 32, 64) else 64)
    raise ValueError("decimal arithmetic is not supported")


class Relation:
    """Columns on the device (declaration order), a row count, and how many
    leading columns the rows are sorted by (lexicographically, by value),
    as the reference tracks isSorted (RelationalAlgebraCompiler.cpp:
    7376-7400)."""

    def __init__(self, cols: List[Tuple[object, str]], n: int, sorted_by: int = 0):
        self.cols, self.n, self.sorted_by = cols, n, sorted_by

    @property
    def types(self) -> List[str]:
        return [t for _, t in self.cols]


def sorted_prefix(columns: Sequence[np.ndarray]) -> int:
    """How many leading integer columns the rows are sorted by."""
    n = len(columns[0]) if columns else 0
    if n < 2:
        return sum(1 for c in columns if np.asarray(c).dtype.kind in "iu") if n else 0
    lt = np.zeros(n - 1, dtype=bool)
    eq = np.ones(n - 1, dtype=bool)
    s = 0
    for c in columns:
        c = np.asarray(c)
        if c.dtype.kind not in "iu":
            break
        a, b = c[:-1], c[1:]
        if not np.all(lt | (eq & (a <= b))):
            break
        s += 1
        lt |= eq & (a < b)
        eq &= a == b
    return s


def _identity_prefix(exps, limit: int) -> int:
    """Leading reordering expressions that are Index(i) at position i < limit."""
    k = 0
    for i, e in enumerate(exps):
        if i >= limit or e.tag != 1 or e.index.offset != i:
            break
        k += 1
    return k


# ---------------------------------------------------------------------------
# Expressions -> MAPFILTER expressions
# ---------------------------------------------------------------------------
def _constant(c):
    """(value, type) of a Constant."""
    if c.kind == 0:
        return int(c.bool_constant.value), "i64"
    if c.kind == 1:
        return int(c.int_constant.value), "i64"
    if c.kind == 5:
        return int(c.uint_constant.value), "u64"
    if c.kind == 2:
        return float(c.float_constant.value), "f64"
    if c.kind == 4:
        return int(c.date_time_constant.value), "i64"
    raise ValueError("string constants are not supported")


def _kind(t: str) -> str:
    return "f" if t.startswith("f") else "u" if t == "u64" else "i"


def _as(e, t: str):
    """e converted to the kind of type t."""
    if _kind(t) == "f":
        return M.to_float(e)
    return M.to_int(e)


def lower_exp(exp, col) -> Tuple[object, str]:
    """(expression, type) of an Exp; col(i) gives (expression, type) of
    source column i."""
    tag = exp.tag
    if tag == 1:  # INDEX
        return col(exp.index.offset)
    if tag == 0:  # CONSTEXP
        v, t = _constant(exp.constExp.literal)
        return M.Const(v), t
    if tag == 3:  # ARITHEXP
        t = _num_type(exp.arithexp.domain)
        a, ta = lower_exp(exp.arithexp.exp1, col)
        b, tb = lower_exp(exp.arithexp.exp2, col)
        if "str" in (ta, tb):
            raise ValueError("arithmetic on string columns is not supported")
        a, b = _as(a, t), _as(b, t)
        op = exp.arithexp.op
        if op == 4 and _kind(t) == "f":
            raise ValueError("mod of doubles is not supported")
        return {0: a + b, 1: a * b, 2: a - b, 3: a / b, 4: a % b}[op], t
    if tag == 5:  # MIXED: a numeric conversion
        t = _num_type(exp.mixed.range)
        a, _ = lower_exp(exp.mixed.exp1, col)
        return _as(a, t), t
    if tag == 4:  # CONVERT
        if exp.convert.convRng.tag != 0:
            raise ValueError("conversions to non-numeric types are not supported")
        t = _num_type(exp.convert.convRng.n)
        a, _ = lower_exp(exp.convert.exp2, col)
        return _as(a, t), t
    if tag == 6:  # CALL (RelationalAlgebraCompiler.cpp:730-760)
        c = exp.call
        if c.calltype != 2:
            raise ValueError(f"call of type {c.calltype} is not supported")
        if c.callname == "subtract":  # date - amount, as plain subtraction
            a, _ = lower_exp(c.args[0], col)
            b, _ = lower_exp(c.args[1], col)
            return M.to_int(a) - M.to_int(b), "i64"
        if c.callname == "part" and len(c.args) >= 2 and \
                c.args[1].constExp.literal.string_constant.value == "year":
            a, _ = lower_exp(c.args[0], col)
            return M.part_year(M.to_int(a)), "i64"
        if c.callname == "add":
            a, _ = lower_exp(c.args[0], col)
            unit = c.args[2].constExp.literal.string_constant.value if len(c.args) > 2 else ""
            if unit in ("months", "years"):
                if c.args[1].tag != 0:
                    raise ValueError(f"add {unit}: the amount must be a constant")
                k, _ = _constant(c.args[1].constExp.literal)
                f = M.add_month if unit == "months" else M.add_year
                return f(M.to_int(a), int(k)), "i64"
            b, _ = lower_exp(c.args[1], col)
            return M.to_int(a) + M.to_int(b), "i64"
        raise ValueError(f"call {c.callname!r} is not supported")
    raise ValueError(f"expression tag {tag} is not supported")


_CMP = {0: "eq", 1: "ne", 2: "lt", 3: "le", 4: "gt", 5: "ge"}


def test_patterns(cmp) -> List[bytes]:
    """The LIKE patterns of a predicate's string tests, in order."""
    if cmp.tag == 1:
        return [cmp.test.ops[1].constExp.literal.string_constant.value.encode()]
    if cmp.tag == 2:
        return test_patterns(cmp.andComp.and1) + test_patterns(cmp.andComp.and2)
    return []


def lower_compare(cmp, col, pattern=None):
    """A filter expression, or None for ALWAYS.  pattern(bytes) gives the
    expression of a LIKE pattern's device address."""
    if cmp.tag == 3:
        return None
    if cmp.tag == 1:  # string test: like / notlike (RelationalAlgebraCompiler.cpp:953-970)
        t = cmp.test
        if t.testtype != 0 or t.testname not in ("like", "notlike") or len(t.ops) != 2 \
                or t.ops[1].tag != 0 or t.ops[1].constExp.literal.kind != 3 or pattern is None:
            raise ValueError(f"string test {t.testname!r} is not supported")
        a, ta = lower_exp(t.ops[0], col)
        if ta != "str":
            raise ValueError("LIKE reads a string column")
        pat = t.ops[1].constExp.literal.string_constant.value.encode()
        e = M.like(a, pattern(pat), pat)
        return e if t.testname == "like" else ~e
    if cmp.tag == 0:
        op = _CMP[cmp.comp.comparison]
        e1, e2 = cmp.comp.op1, cmp.comp.op2
        s1 = e2.tag == 0 and e2.constExp.literal.kind == 3  # a string constant
        s2 = e1.tag == 0 and e1.constExp.literal.kind == 3
        if s1 or s2:
            # A string column against a string constant: chunk comparisons.
            ecol, econst = (e1, e2) if s1 else (e2, e1)
            a, ta = lower_exp(ecol, col)
            if ta != "str":
                raise ValueError("a string constant compares with a string column")
            if not s1:  # constant first: mirror the comparison
                op = {"lt": "gt", "le": "ge", "gt": "lt", "ge": "le"}.get(op, op)
            return M.str_compare(a, econst.constExp.literal.string_constant.value.encode(), op)
        a, ta = lower_exp(e1, col)
        b, tb = lower_exp(e2, col)
        if "str" in (ta, tb):
            raise ValueError("string columns compare only with string constants")
        op = _CMP[cmp.comp.comparison]
        return a.eq(b) if op == "eq" else a.ne(b) if op == "ne" else \
            {"lt": a < b, "le": a <= b, "gt": a > b, "ge": a >= b}[op]
    if cmp.tag == 2:
        x = lower_compare(cmp.andComp.and1, col, pattern)
        y = lower_compare(cmp.andComp.and2, col, pattern)
        return y if x is None else x if y is None else x & y
    raise ValueError(f"compare tag {cmp.tag} is not supported")


# ---------------------------------------------------------------------------
# Packed keys
# ---------------------------------------------------------------------------
def key_layout(types: Sequence[str]) -> Tuple[str, int, List[Tuple[int, int, int]]]:
    """(key type, key bits, [(shift, width, bias)] per column) of a packed
    key over columns of these types, first column most significant."""
    for t in types:
        if t not in _BITS:
            raise ValueError(f"a {t} column cannot be part of a key")
    widths = [_BITS[t] for t in types]
    total = sum(widths)
    if total == 0:
        return "u32", 1, []
    if total > 64:
        raise ValueError(f"a key of {total} bits is not supported")
    fields, shift = [], total
    for t, w in zip(types, widths):
        shift -= w
        fields.append((shift, w, (1 << (w - 1)) if t.startswith("i") else 0))
    return ("u32" if total <= 32 else "u64"), total, fields


def _pack_expr(cols: Sequence[object], fields) -> object:
    key = M.Const(0)
    for e, (shift, w, bias) in zip(cols, fields):
        u = e + bias if bias else e
        if w < 64:  # keep the column's own width (sign-extended values wrap into it)
            u = u % (1 << w) if bias else u
        key = key + u * (1 << shift) if shift else key + u
    return key


def _signed(kt: str) -> str:
    """The type a packed key column is read back as (unpacking works on
    signed 64-bit values)."""
    return "i64" if kt == "u64" else kt


def _unpack_expr(key, field, total: int):
    """Column `field` of a packed key read as a signed integer expression."""
    shift, w, bias = field
    if w == 64:
        v = key
    elif total < 64:
        v = key / (1 << shift) if shift else key
        v = v % (1 << w)
    else:
        # All 64 bits used: unsigned shifts from signed division.  h is the
        # key shifted right by one as unsigned (non-negative as int64).
        h = (key - (key % 2).ne(0)) / 2 + (key < 0) * (1 << 63)
        if shift:
            v = (h / (1 << (shift - 1))) % (1 << w)
        else:
            v = key - (h / (1 << (w - 1))) * (1 << w)
    return v - bias if bias else v


# ---------------------------------------------------------------------------
# The executor
# ---------------------------------------------------------------------------
class Plan:
    """A loaded plan bound to a context.  Relations of input variables are
    set with set(); run() executes the sequences from the entry; get()
    returns a relation's columns as numpy arrays."""

    def __init__(self, ctx, data: bytes, max_steps: int = 100_000):
        self.ctx, self.graph, self.max_steps = ctx, parse(data), max_steps
        self.decl: Dict[str, List[str]] = {}
        self.rel: Dict[str, Relation] = {}
        # Device buffers by power-of-two capacity, reused across nodes and
        # runs (freeing device memory synchronises the device).
        self._free: Dict[int, List[object]] = {}
        for v in self.graph.variables:
            types = [column_type(t) for t in list(v.keys) + list(v.fields)]
            self.decl[v.varName] = types
            self.rel[v.varName] = self._empty(types)
        self.steps = 0

    # -- relations --------------------------------------------------------
    def set(self, name: str, columns: Sequence[np.ndarray]) -> None:
        types = self.decl[name]
        if len(columns) != len(types):
            raise ValueError(f"{name} has {len(types)} columns")
        n = len(columns[0]) if columns else 0
        cols, host = [], []
        for a, t in zip(columns, types):
            a = np.ascontiguousarray(a, dtype=_NP[t])
            if len(a) != n:
                raise ValueError("columns of one relation have one length")
            b = self._buf(a.nbytes)
            if n:
                self._upload(b, a)
            cols.append((b, t))
            host.append(a)
        lead = types.index("str") if "str" in types else len(types)
        self._replace(name, Relation(cols, n, sorted_prefix(host[:lead])))

    def set_columns(self, name: str, columns: Sequence[Tuple[object, str]], n: int,
                    sorted_by: int = 0) -> None:
        """Set a relation from device columns (copied into the plan's own
        buffers); sorted_by: leading columns the rows are known sorted by."""
        types = self.decl[name]
        if [t for _, t in columns] != types:
            raise ValueError(f"{name} has columns {types}")
        cols = []
        for b, t in columns:
            nb = n * np.dtype(_NP[t]).itemsize
            c = self._buf(nb)
            if nb:
                self.ctx.copy(c, b, nb)
            cols.append((c, t))
        self._replace(name, Relation(cols, n, sorted_by))

    def get(self, name: str) -> List[np.ndarray]:
        r = self.rel[name]
        self.ctx.synchronize()
        return [b.download(_NP[t], r.n) for b, t in r.cols]

    def rows(self, name: str) -> int:
        return self.rel[name].n

    # -- device buffers -----------------------------------------------------
    def _buf(self, nbytes: int):
        cap = 256
        while cap < nbytes:
            cap *= 2
        free = self._free.get(cap)
        return free.pop() if free else rfx.DeviceBuffer(cap)

    def _alloc(self, t: str, n: int):
        return self._buf(n * np.dtype(_NP[t]).itemsize)

    def _cell(self, value: int = 0):
        c = self._buf(8)
        self._upload(c, np.array([value], dtype=np.uint64))
        return c

    def _upload(self, buf, arr: np.ndarray) -> None:
        """Host write into a (possibly pooled) buffer.  The copy is a
        blocking host-to-device copy, which is not ordered with the
        context's stream, so work queued there that may still read the
        buffer's previous contents is waited for first."""
        self.ctx.synchronize()
        buf.upload(arr)

    def _release(self, *bufs) -> None:
        """Return buffers to the pool (work already queued on the stream
        that reads them runs before any later reuse)."""
        for b in bufs:
            if b is not None:
                self._free.setdefault(b.nbytes, []).append(b)

    def _empty(self, types: Sequence[str]) -> Relation:
        return Relation([(self._buf(8), t) for t in types], 0, len(types))

    def _replace(self, name: str, rel: Relation) -> None:
        old = self.rel.get(name)
        self.rel[name] = rel
        if old is not None and old is not rel:
            keep = {id(b) for b, _ in rel.cols}
            self._release(*[b for b, _ in old.cols if id(b) not in keep])

    def _mapfilter(self, columns, n, build, out_types) -> Relation:
        """One MAPFILTER pass: build(x) sets the filter and outputs over
        MapFilter x of `columns`; returns the kept rows' outputs."""
        if len(columns) > M.MAX_COLS:
            raise ValueError(f"{len(columns)} columns in one pass (at most {M.MAX_COLS})")
        if len(out_types) > M.MAX_OUTS:
            return self._mapfilter_wide(columns, n, build, out_types)
        x = M.MapFilter([(c[0], _mft(c[1])) + tuple(c[2:]) for c in columns])
        build(x)
        x.outputs = [(e, _mft(t)) for e, t in x.outputs]
        outs = [self._alloc(t, n) for t in out_types]
        cell = self._buf(8)
        x.run(self.ctx, n, outs, cell)
        # Without a filter every row is kept: no need to wait for the count.
        kept = n if x.filter is None else rfx.read_u64(cell)
        self._release(cell)
        return Relation(list(zip(outs, out_types)), kept)

    def _mapfilter_wide(self, columns, n, build, out_types) -> Relation:
        # More outputs than one pass holds: the same filter, outputs in groups.
        rel_cols, kept = [], None
        for g in range(0, len(out_types), M.MAX_OUTS):
            def part(x, g=g):
                build(x)
                x.outputs = x.outputs[g:g + M.MAX_OUTS]
            r = self._mapfilter(columns, n, part, out_types[g:g + M.MAX_OUTS])
            rel_cols += r.cols
            kept = r.n if kept is None else kept
        return Relation(rel_cols, kept or 0)

    def _pack(self, rel: Relation, idx: Sequence[int]):
        """(sorted keys, row permutation or None, key type, key bits,
        layout) of the packed key over columns idx; no sort (and no
        permutation) when the rows are already sorted by those columns."""
        kt, bits, fields = key_layout([rel.types[i] for i in idx])
        n = rel.n
        src = [(rel.cols[i][0], rel.cols[i][1]) for i in idx]
        keys = self._mapfilter(src, n, lambda x: x.emit(_pack_expr(x.cols, fields), kt),
                               [kt]).cols[0][0]
        presorted = list(idx) == list(range(len(idx))) and rel.sorted_by >= len(idx)
        if presorted:
            return keys, None, kt, bits, fields
        perm = self._alloc("u32", n)
        self.ctx.iota_u32(perm, n)
        self.ctx.sort_pairs(keys, perm, n, 4 if kt == "u32" else 8, bits, 4)
        return keys, perm, kt, bits, fields

    def _result(self, dest: str, rel: Relation) -> None:
        """Store rel as dest, converting columns to dest's declared types."""
        types = self.decl.get(dest)
        if types is not None and types != rel.types:
            if len(types) != len(rel.types):
                raise ValueError(f"{dest}: {len(rel.types)} columns computed, "
                                 f"{len(types)} declared")
            conv = self._mapfilter(list(rel.cols), rel.n, lambda x: [
                x.emit(_as(c, t), t) for c, t in zip(x.cols, types)], types) if rel.n else \
                self._empty(types)
            conv.sorted_by = rel.sorted_by
            self._release(*[b for b, _ in rel.cols])
            rel = conv
        self._replace(dest, rel)
        self.decl.setdefault(dest, rel.types)

    # -- operators ----------------------------------------------------------
    def mapfilter(self, node) -> Relation:
        src = self.rel[node.srcA]
        if node.srcWidth and node.srcWidth != len(src.cols):
            raise ValueError(f"MAPFILTER over {node.srcA}: width {node.srcWidth}, "
                             f"relation has {len(src.cols)} columns")

        # LIKE patterns: NUL-terminated in device memory, each address in a
        # one-row column after the relation's columns.
        pats = test_patterns(node.predicate)
        pat_bufs, pat_cols = [], {}
        for k, pat in enumerate(pats):
            text = self._buf(len(pat) + 1)
            self._upload(text, np.frombuffer(pat + b"\0", dtype=np.uint8))
            cell = self._cell(text.ptr)
            pat_bufs += [text, cell]
            pat_cols.setdefault(pat, len(src.cols) + k)

        def build(x):
            col = lambda i: (x.cols[i], src.types[i])  # noqa: E731
            f = lower_compare(node.predicate, col, lambda pat: x.gather(pat_cols[pat], 0))
            if f is not None:
                x.where(f)
            for e in node.reordering:
                v, t = lower_exp(e, col)
                x.emit(_as(v, t), t)

        # The output types come from the expressions: lower once to learn them.
        probe = M.MapFilter([(b, _mft(t)) for b, t in src.cols])
        types = [lower_exp(e, lambda i: (probe.cols[i], src.types[i]))[1] for e in node.reordering]
        if src.n == 0:
            self._release(*pat_bufs)
            return self._empty(types)
        cols = list(src.cols) + [(pat_bufs[2 * k + 1], "u64", 1) for k in range(len(pats))]
        out = self._mapfilter(cols, src.n, build, types)
        self._release(*pat_bufs)
        out.sorted_by = _identity_prefix(node.reordering, src.sorted_by)  # a filter keeps order
        return out

    def join(self, key_count: int, a_name: str, b_name: str, args) -> Relation:
        A, B = self.rel[a_name], self.rel[b_name]
        k = key_count
        if k > min(len(A.cols), len(B.cols)):
            raise ValueError(f"join on {k} keys of {len(A.cols)}- and {len(B.cols)}-column inputs")
        if [A.types[i] for i in range(k)] != [B.types[i] for i in range(k)]:
            raise ValueError("joined key columns differ in type")
        comb = A.types + B.types[k:]
        if args:
            probe = M.MapFilter([(None, _mft(t)) for t in comb])
            out_types = [lower_exp(e, lambda i: (probe.cols[i], comb[i]))[1] for e in args]
        else:
            out_types = comb
        if A.n == 0 or B.n == 0:
            return self._empty(out_types)
        strings = "str" in A.types[:k]
        lb, lc, tot = self._alloc("i32", A.n), self._alloc("i32", A.n), self._buf(8)
        rtmp = []
        if strings and k > 1:
            # Several key columns, some of them strings: each string key
            # column compares as its ranks over both sides' strings, and the
            # packed integer keys join as usual (output columns come from
            # the original relations).
            (A2, B2), _, rtmp = self._ranked([A, B], [i for i in range(k) if A.types[i] == "str"])
            ka, pa, kt, _, _ = self._pack(A2, range(k))
            kb, pb, _, _, _ = self._pack(B2, range(k))
            self.ctx.join_find_bounds(lb, lc, tot, ka, A.n, kb, B.n, 4 if kt == "u32" else 8)
        elif strings:
            # Rows of each side in string order, their addresses in that
            # order, then the sorted search on the strings.
            ka, pa = self._string_order(A)
            kb, pb = self._string_order(B)
            self.ctx.join_find_bounds_string(lb, lc, tot, ka, A.n, kb, B.n)
        else:
            ka, pa, kt, _, _ = self._pack(A, range(k))
            kb, pb, _, _, _ = self._pack(B, range(k))
            self.ctx.join_find_bounds(lb, lc, tot, ka, A.n, kb, B.n, 4 if kt == "u32" else 8)
        total = rfx.read_u64(tot)
        if total >= 1 << 31:
            self._release(ka, kb, pa, pb, lb, lc, tot, *rtmp)
            raise ValueError(f"join of {total} rows exceeds the int32 row index range")
        if total == 0:
            self._release(ka, kb, pa, pb, lb, lc, tot, *rtmp)
            return self._empty(out_types)
        li, ri = self._alloc("i32", total), self._alloc("i32", total)
        self.ctx.join_main(li, ri, tot, total, lb, lc, A.n)
        # One pass: the combined tuple through the row pairs (through each
        # side's sort permutation where it was sorted here).
        cols = [(li, "i32"), (ri, "i32")]
        if pa is not None:
            cols.append((pa, "u32", A.n))
        if pb is not None:
            cols.append((pb, "u32", B.n))
        first = len(cols)
        cols += [(b, t, A.n) for b, t in A.cols] + [(b, t, B.n) for b, t in B.cols[k:]]
        na = len(A.cols)

        def build(x):
            ra = x.cols[0] if pa is None else x.gather(2, x.cols[0])
            rb = x.cols[1] if pb is None else x.gather(2 + (pa is not None), x.cols[1])

            def col(i):
                return x.gather(first + i, ra if i < na else rb), comb[i]
            if args:
                for e in args:
                    v, t = lower_exp(e, col)
                    x.emit(_as(v, t), t)
            else:
                for i in range(len(comb)):
                    v, t = col(i)
                    x.emit(v, t)
        out = self._mapfilter(cols, total, build, out_types)
        self._release(ka, kb, pa, pb, lb, lc, tot, li, ri, *rtmp)
        # Pairs come in key order, left rows in their order within a key:
        # sorted by A's sorted columns where A needed no sort, else by the key
        # (string order is not the addresses' order: no claim).
        s_out = 0 if strings else A.sorted_by if pa is None else k
        out.sorted_by = _identity_prefix(args, s_out) if args else s_out
        return out

    def _string_order(self, rel: Relation):
        """(addresses of column 0 in string order, the row permutation)."""
        n = rel.n
        perm, scratch = self._alloc("u32", n), self._alloc("u64", n + 1)
        self.ctx.iota_u32(perm, n)
        self.ctx.sort_strings_rows(rel.cols[0][0], perm, scratch, n)
        self._release(scratch)
        keys = self._mapfilter([(perm, "u32"), (rel.cols[0][0], "u64", n)], n,
                               lambda x: x.emit(x.gather(1, x.cols[0]), "u64"), ["u64"]).cols[0][0]
        return keys, perm

    def agg(self, node) -> Relation:
        src = self.rel[node.srcA]
        if len(node.range) != 1:
            raise ValueError("one aggregate per AGG node")
        a = node.range[0]
        dom = list(node.domains)
        which = {0: "count", 1: "total", 2: "min", 3: "max"}[a.tag]
        vcol = None if which == "count" else getattr(a, which).aggdom
        str_val = vcol is not None and src.types[vcol] == "str"
        out_t = "i64" if which == "count" else "str" if str_val else "f64"
        if str_val and which == "total":
            raise ValueError("TOTAL runs on floating-point columns, not strings")
        if vcol is not None and not str_val and not src.types[vcol].startswith("f"):
            raise ValueError("TOTAL/MIN/MAX run on floating-point columns (FP64 sums)")
        dom_types = [src.types[i] for i in dom]
        if src.n == 0:
            return self._empty(dom_types + [out_t])
        if str_val:
            # MIN/MAX of strings: the value column becomes its dense ranks in
            # string order (exact as doubles), the ranks are reduced, and each
            # group's rank gets back an address holding its string.
            (ranked,), vreps, vtmp = self._ranked([src], [vcol])
            ranked.sorted_by = min(src.sorted_by, vcol)
            out = self._agg_columns(ranked, dom, which, vcol, "f64")
            conv = self._rank_to_string(out, vreps[vcol], src.n)
            self._release(*[b for b, _ in out.cols], *vtmp)
            return conv
        return self._agg_columns(src, dom, which, vcol, out_t)

    def _rank_to_string(self, rel: Relation, rep, n_rep: int) -> Relation:
        """rel with its last column (string ranks as f64) replaced by the
        representative address of each rank (a STRING column)."""
        k = len(rel.cols) - 1
        cols = list(rel.cols) + [(rep, "f64", n_rep)]

        def build(x):
            for i, t in enumerate(rel.types[:k]):
                x.emit(x.cols[i], t)
            x.emit(M.to_int(x.gather(k + 1, M.to_int(x.cols[k]))), "u64")
        out = self._mapfilter(cols, rel.n, build, rel.types[:k] + ["str"])
        out.sorted_by = min(rel.sorted_by, k)
        return out

    def _agg_columns(self, src: Relation, dom, which: str, vcol, out_t: str) -> Relation:
        """AGG of a non-empty relation whose value column (if any) is
        numeric: groups by the domain columns, one aggregate per group."""
        dom_types = [src.types[i] for i in dom]
        if "str" in dom_types and len(dom) == 1:
            return self._agg_string(src, dom[0], which, vcol, out_t)
        reps, rtmp = {}, []
        if "str" in dom_types:
            # Several domain columns with strings among them: each string
            # column groups by its rank in string order, and each result
            # rank gets back an address holding its string.
            (src,), reps, rtmp = self._ranked([src], [i for i in dom if src.types[i] == "str"])
            dom_types = [src.types[i] for i in dom]
        kt, bits, fields = key_layout(dom_types)
        kbytes = 4 if kt == "u32" else 8
        n = src.n
        # keys (and values as f64) in one pass, then sort them together
        # unless the rows are already in domain order.
        srcc = [(src.cols[i][0], src.cols[i][1]) for i in dom]
        if vcol is not None:
            srcc.append(src.cols[vcol])

        def build(x):
            x.emit(_pack_expr(x.cols[:len(dom)], fields), kt)
            if vcol is not None:
                x.emit(M.to_float(x.cols[len(dom)]), "f64")
        kv = self._mapfilter(srcc, n, build, [kt] + (["f64"] if vcol is not None else []))
        keys = kv.cols[0][0]
        vals = kv.cols[1][0] if vcol is not None else None
        if not (dom == list(range(len(dom))) and src.sorted_by >= len(dom)):
            self.ctx.sort_pairs(keys, vals, n, kbytes, bits, 8 if vals is not None else 0)
        ncell, nout = self._cell(n), self._buf(8)
        okeys = self._alloc(kt, n)
        if which == "count":
            ovals, vt = self._alloc("u32", n), "u32"
            self.ctx.count_by_key(okeys, ovals, nout, keys, ncell, n, kbytes)
        else:
            ovals, vt = self._alloc("f64", n), "f64"
            op = {"total": 0, "min": 1, "max": 2}[which]
            self.ctx.reduce_by_key_f64(okeys, ovals, nout, keys, vals, ncell, n, kbytes, op)
        groups = rfx.read_u64(nout)

        slot = {d: 2 + k for k, d in enumerate(reps)}
        rep_cols = [(reps[d], "f64", src.n) for d in reps]

        def unpack(x):
            for d, f, t in zip(dom, fields, dom_types):
                v = _unpack_expr(x.cols[0], f, bits)
                if d in slot:
                    x.emit(M.to_int(x.gather(slot[d], v)), "u64")
                else:
                    x.emit(v, t)
            x.emit(_as(x.cols[1], out_t), out_t)
        out_types = ["str" if d in slot else t for d, t in zip(dom, dom_types)] + [out_t]
        out = self._mapfilter([(okeys, _signed(kt)), (ovals, vt)] + rep_cols, groups, unpack,
                              out_types)
        self._release(keys, vals, ncell, nout, okeys, ovals, *rtmp)
        out.sorted_by = 0 if reps else len(dom)
        return out

    def _agg_string(self, src: Relation, d: int, which: str, vcol, out_t: str) -> Relation:
        """AGG grouped by one string column: rows in string order, runs of
        equal strings numbered (rfx_string_groups), each run reduced; a
        group's string is its smallest address (addresses are below 2^53,
        exact as doubles)."""
        n = src.n
        perm, scratch = self._alloc("u32", n), self._alloc("u64", n + 1)
        self.ctx.iota_u32(perm, n)
        self.ctx.sort_strings_rows(src.cols[d][0], perm, scratch, n)
        cols = [(perm, "u32"), (src.cols[d][0], "u64", n)]
        if vcol is not None:
            cols.append((src.cols[vcol][0], src.cols[vcol][1], n))

        def gather(x):
            row = x.cols[0]
            a = x.gather(1, row)
            x.emit(a, "u64")
            x.emit(M.to_float(a), "f64")
            if vcol is not None:
                x.emit(M.to_float(x.gather(2, row)), "f64")
        g = self._mapfilter(cols, n, gather, ["u64", "f64"] + (["f64"] if vcol is not None else []))
        addr, addr_f = g.cols[0][0], g.cols[1][0]
        vals = g.cols[2][0] if vcol is not None else None
        group, heads, ngroups = self._alloc("u32", n), self._alloc("i32", n), self._buf(8)
        self.ctx.string_groups(addr, n, group, heads, ngroups)
        ncell, nout, nout2 = self._cell(n), self._buf(8), self._buf(8)
        okeys, okeys2, oaddr = self._alloc("u32", n), self._alloc("u32", n), self._alloc("f64", n)
        if which == "count":
            ovals, vt = self._alloc("u32", n), "u32"
            self.ctx.count_by_key(okeys, ovals, nout, group, ncell, n, 4)
        else:
            ovals, vt = self._alloc("f64", n), "f64"
            self.ctx.reduce_by_key_f64(okeys, ovals, nout, group, vals, ncell, n, 4,
                                       {"total": 0, "min": 1, "max": 2}[which])
        self.ctx.reduce_by_key_f64(okeys2, oaddr, nout2, group, addr_f, ncell, n, 4, 1)
        groups = rfx.read_u64(nout)
        out = self._mapfilter([(oaddr, "f64"), (ovals, vt)], groups, lambda x: (
            x.emit(M.to_int(x.cols[0]), "u64"), x.emit(_as(x.cols[1], out_t), out_t)),
            ["str", out_t])
        self._release(perm, scratch, addr, addr_f, vals, group, heads, ngroups, ncell, nout, nout2,
                      okeys, okeys2, oaddr, ovals)
        out.sorted_by = 0  # string order, not the addresses' order
        return out

    def _string_ranks(self, pieces):
        """Dense ranks, in string order, of the strings of several address
        columns taken together: pieces = [(address buffer, rows)].  Returns
        (a u32 rank column per piece, a representative address per rank as
        f64 (the smallest address of the rank's strings: exact, addresses are
        below 2^53), temporaries to release).  Equal strings get one rank
        wherever they live, so string columns compare as their ranks."""
        N = sum(n for _, n in pieces)
        allv = self._alloc("u64", N)
        off = 0
        for buf, n in pieces:
            if n:
                self.ctx.copy(allv, buf, 8 * n, dst_off=8 * off)
            off += n
        rows, scratch = self._alloc("u32", N), self._alloc("u64", N + 1)
        self.ctx.iota_u32(rows, N)
        self.ctx.sort_strings_rows(allv, rows, scratch, N)
        g = self._mapfilter([(rows, "u32"), (allv, "u64", N)], N, lambda x: (
            x.emit(x.gather(1, x.cols[0]), "u64"), x.emit(M.to_float(x.gather(1, x.cols[0])), "f64")),
            ["u64", "f64"])
        saddr, saddr_f = g.cols[0][0], g.cols[1][0]
        group, heads, ngroups = self._alloc("u32", N), self._alloc("i32", N), self._buf(8)
        self.ctx.string_groups(saddr, N, group, heads, ngroups)
        ncell, nout = self._cell(N), self._buf(8)
        okeys, rep = self._alloc("u32", N), self._alloc("f64", N)
        self.ctx.reduce_by_key_f64(okeys, rep, nout, group, saddr_f, ncell, N, 4, 1)
        # Ranks back in row order: sort the (row, rank) pairs by row.
        self.ctx.sort_pairs(rows, group, N, 4, max(1, int(N - 1).bit_length()), 4)
        ranks, off = [], 0
        for _, n in pieces:
            r = self._alloc("u32", max(n, 1))
            if n:
                self.ctx.copy(r, group, 4 * n, src_off=4 * off)
            ranks.append(r)
            off += n
        tmp = [allv, rows, scratch, saddr, saddr_f, group, heads, ngroups, ncell, nout, okeys]
        return ranks, rep, tmp

    def _ranked(self, rels, cols):
        """The relations with each string column in `cols` replaced by its
        ranks over all the relations' strings in that column (u32, string
        order); returns (relations, {column: representative addresses},
        temporaries)."""
        out = [list(r.cols) for r in rels]
        reps, tmp = {}, []
        for i in cols:
            ranks, rep, t = self._string_ranks([(r.cols[i][0], r.n) for r in rels])
            for o, rk in zip(out, ranks):
                o[i] = (rk, "u32")
            reps[i] = rep
            tmp += t + ranks + [rep]
        return [Relation(o, r.n, 0) for o, r in zip(out, rels)], reps, tmp

    def set_operation(self, op: int, a_name: str, b_name: str) -> Relation:
        """op: 0 union, 1 difference, 2 intersection, on whole tuples.
        String columns take part by their strings: they are replaced by
        their ranks in string order over both inputs, the set operation
        runs on the integer tuples, and each result rank gets back an
        address holding its string (the reference's Union.cu / Difference.cu
        compare packed tuples; strings compare by content here, as the join
        and GROUP BY on strings do)."""
        A, B = self.rel[a_name], self.rel[b_name]
        if A.types != B.types:
            raise ValueError("set operations need inputs of one type")
        si = [i for i, t in enumerate(A.types) if t == "str"]
        if not si:
            return self._set_op_rel(op, A, B)
        if A.n + B.n == 0:
            return self._empty(A.types)
        (A2, B2), reps, tmp = self._ranked([A, B], si)
        R = self._set_op_rel(op, A2, B2)
        if R.n == 0:
            res = self._empty(A.types)
        else:
            cols = list(R.cols) + [(reps[i], "f64", A.n + B.n) for i in si]
            slot = {i: len(R.cols) + k for k, i in enumerate(si)}
            res = self._mapfilter(cols, R.n, lambda x: [
                x.emit(M.to_int(x.gather(slot[i], x.cols[i])), "u64") if i in slot
                else x.emit(x.cols[i], t) for i, t in enumerate(A.types)], A.types)
        self._release(*[b for b, _ in R.cols], *tmp)
        res.sorted_by = 0  # string order, not the addresses' order
        return res

    def _set_op_rel(self, op: int, A: Relation, B: Relation) -> Relation:
        kt, bits, fields = key_layout(A.types)
        kbytes = 4 if kt == "u32" else 8
        sets, tmp = [], []
        for R in (A, B):
            if R.n == 0:
                sets.append((self._alloc(kt, 1), 0))
                tmp.append(sets[-1][0])
                continue
            keys, perm, _, _, _ = self._pack(R, range(len(R.cols)))
            uniq, cnt = self._alloc(kt, R.n), self._buf(8)
            self.ctx.unique(uniq, cnt, keys, R.n, kbytes)
            sets.append((uniq, rfx.read_u64(cnt)))
            tmp += [keys, perm, uniq, cnt]
        (a, na), (b, nb) = sets
        out, cnt = self._alloc(kt, na + nb + 1), self._buf(8)
        self.ctx.set_op(op, out, cnt, a, na, b, nb, kbytes)
        m = rfx.read_u64(cnt)
        if m == 0:
            res = self._empty(A.types)
        else:
            res = self._mapfilter([(out, _signed(kt))], m, lambda x: [
                x.emit(_unpack_expr(x.cols[0], f, bits), t) for f, t in zip(fields, A.types)],
                A.types)
        self._release(out, cnt, *tmp)
        res.sorted_by = len(A.types)
        return res

    def single(self, dest: str, elements) -> Relation:
        vals = [_constant(c) for c in elements]
        types = self.decl.get(dest) or [t for _, t in vals]
        if len(types) != len(vals):
            raise ValueError(f"SINGLE of {len(vals)} constants for {len(types)} columns")
        cols = []
        for (v, _), t in zip(vals, types):
            b = self._buf(8)
            self._upload(b, np.array([v], dtype=_NP[t]))
            cols.append((b, t))
        return Relation(cols, 1, len(types))

    def copy(self, name: str) -> Relation:
        r = self.rel[name]
        cols = []
        for b, t in r.cols:
            nb = r.n * np.dtype(_NP[t]).itemsize
            c = self._buf(nb)
            if nb:
                self.ctx.copy(c, b, nb)
            cols.append((c, t))
        return Relation(cols, r.n, r.sorted_by)

    def equal(self, a: str, b: str) -> bool:
        A, B = self.rel[a], self.rel[b]
        if A.n != B.n or A.types != B.types:
            return False
        return all(np.array_equal(x, y) for x, y in zip(self.get(a), self.get(b)))

    # -- control ------------------------------------------------------------
    def assign(self, cmd) -> None:
        op, dest = cmd.op, cmd.dest
        tag = op.tag
        if tag == 10:
            r = self.mapfilter(op.mapfilter)
        elif tag == 7:
            r = self.join(op.join.keyCount, op.join.srcA, op.join.srcB, list(op.join.args))
        elif tag == 4:
            r = self.join(0, op.product.srcA, op.product.srcB, [])
        elif tag == 11:
            r = self.agg(op.agg)
        elif tag == 2:
            r = self.set_operation(0, op.unionop.srcA, op.unionop.srcB)
        elif tag == 3:
            r = self.set_operation(2, op.intersection.srcA, op.intersection.srcB)
        elif tag == 6:
            r = self.set_operation(1, op.difference.srcA, op.difference.srcB)
        elif tag == 5:
            r = self.single(dest, list(op.single.element))
        else:
            raise ValueError(f"operator tag {tag} is not supported")
        self._result(dest, r)

    def run(self) -> int:
        """Execute from the entry sequence; returns the commands executed."""
        seqs = list(self.graph.sequences)
        where = {s.uniqueIdentifier: i for i, s in enumerate(seqs)}
        if self.graph.entry not in where:
            raise ValueError(f"entry {self.graph.entry} is not a sequence")
        i = where[self.graph.entry]
        count = 0
        while i is not None and i < len(seqs):
            nxt = i + 1
            for cmd in seqs[i].operators:
                count += 1
                if count > self.max_steps:
                    raise RuntimeError(f"plan ran more than {self.max_steps} commands")
                t = cmd.tag
                if t == 2:
                    self.assign(cmd.assign)
                    continue
                if t == 5:
                    self._replace(cmd.move.dest, self.copy(cmd.move.src1))
                    continue
                if t == 4:
                    target = cmd.jump.target
                elif t == 3:
                    target = cmd.cond.yes if self.equal(cmd.cond.src1, cmd.cond.src2) \
                        else cmd.cond.no
                elif t == 6:
                    target = None
                else:
                    raise ValueError(f"command tag {t} is not supported")
                if target is not None and target not in where:
                    raise ValueError(f"jump to missing sequence {target}")
                nxt = None if target is None else where[target]
                break
            if i == where.get(self.graph.exit) and nxt == i + 1:
                break  # the exit sequence ran to its end
            i = nxt
        self.ctx.synchronize()
        self.steps = count
        return count


# ---------------------------------------------------------------------------
# Building plans
# ---------------------------------------------------------------------------
class Builder:
    """Builds a GPUGraph with the plan schema's messages.

    Types are written "entity:NAME", "int64", "uint32", "float64",
    "datetime", "bool" or "string"; expressions with idx(), const_*(), arith(),
    call(); predicates with comp(), and_(), always()."""

    _PRIM = {"bool": 0, "datetime": 5, "int": 6, "uint": 7, "float": 8, "string": 9}

    def __init__(self, name: str):
        self.m = messages()
        self.g = self.m["GPUGraph"](graphName=name, entry=0, exit=0)

    # types and expressions
    def type(self, spec: str, into=None):
        t = into if into is not None else self.m["Type"]()
        if spec.startswith("entity:"):
            t.kind = 1
            t.unary.name = spec[7:]
            return t
        t.kind = 0
        base = spec.rstrip("0123456789")
        t.primitive.kind = self._PRIM[base]
        if spec != base:
            t.primitive.capacity = int(spec[len(base):])
        return t

    def idx(self, offset: int, spec: str):
        e = self.m["Exp"](tag=1)
        e.index.offset = offset
        self.type(spec, e.index.etyp)
        return e

    def _const(self, kind: int, setter):
        e = self.m["Exp"](tag=0)
        e.constExp.literal.kind = kind
        setter(e.constExp.literal)
        return e

    def const_int(self, v: int):
        def s(c):
            c.int_constant.value = v
            self.type("int64", c.int_constant.type)
        return self._const(1, s)

    def const_float(self, v):
        def s(c):
            c.float_constant.value = str(v)
            self.type("float64", c.float_constant.type)
        return self._const(2, s)

    def const_datetime(self, v: int):
        return self._const(4, lambda c: setattr(c.date_time_constant, "value", v))

    def const_string(self, v: str):
        return self._const(3, lambda c: setattr(c.string_constant, "value", v))

    def arith(self, op: str, a, b, typ: str = "float", size: int = 64):
        e = self.m["Exp"](tag=3)
        e.arithexp.domain.typ = {"int": 0, "uint": 1, "float": 2}[typ]
        e.arithexp.domain.size = size
        e.arithexp.op = {"add": 0, "mul": 1, "sub": 2, "div": 3, "mod": 4}[op]
        e.arithexp.exp1.CopyFrom(a)
        e.arithexp.exp2.CopyFrom(b)
        return e

    def call(self, name: str, args):
        e = self.m["Exp"](tag=6)
        e.call.calltype = 2  # FDatetime
        e.call.callname = name
        for a in args:
            e.call.args.add().CopyFrom(a)
        return e

    def always(self):
        return self.m["Compare"](tag=3)

    def comp(self, op: str, a, b):
        c = self.m["Compare"](tag=0)
        c.comp.comparison = {"eq": 0, "ne": 1, "lt": 2, "le": 3, "gt": 4, "ge": 5}[op]
        c.comp.op1.CopyFrom(a)
        c.comp.op2.CopyFrom(b)
        return c

    def test(self, name: str, a, pattern: str):
        """A string test (like / notlike) of expression a against a pattern."""
        c = self.m["Compare"](tag=1)
        c.test.testtype = 0  # FString
        c.test.testname = name
        c.test.ops.add().CopyFrom(a)
        c.test.ops.add().CopyFrom(self.const_string(pattern))
        return c

    def and_(self, x, y):
        c = self.m["Compare"](tag=2)
        c.andComp.and1.CopyFrom(x)
        c.andComp.and2.CopyFrom(y)
        return c

    # variables and commands
    def var(self, name: str, keys: Sequence[str] = (), fields: Sequence[str] = ()):
        v = self.g.variables.add(varName=name, initialData=b"")
        for k in keys:
            self.type(k, v.keys.add())
        for f in fields:
            self.type(f, v.fields.add())
        return self

    def sequence(self, ident: int):
        s = self.g.sequences.add(uniqueIdentifier=ident)
        return _Seq(self, s)

    def entry_exit(self, entry: int, exit_: int):
        self.g.entry, self.g.exit = entry, exit_
        return self

    def bytes(self) -> bytes:
        return self.g.SerializeToString()


class _Seq:
    def __init__(self, b: Builder, s):
        self.b, self.s = b, s

    def _assign(self, dest: str, tag: int):
        c = self.s.operators.add(tag=2)
        c.assign.dest = dest
        c.assign.op.tag = tag
        return c.assign.op

    def mapfilter(self, dest, src, width, predicate, reordering):
        op = self._assign(dest, 10).mapfilter
        op.srcA, op.srcWidth = src, width
        op.predicate.CopyFrom(predicate)
        for e in reordering:
            op.reordering.add().CopyFrom(e)
        return self

    def join(self, dest, key_count, a, b, args=()):
        op = self._assign(dest, 7).join
        op.keyCount, op.srcA, op.srcB = key_count, a, b
        for e in args:
            op.args.add().CopyFrom(e)
        return self

    def product(self, dest, a, b):
        op = self._assign(dest, 4).product
        op.srcA, op.srcB = a, b
        return self

    def agg(self, dest, src, width, domains, tag: str, dom: int = 0, rng: int = 0):
        op = self._assign(dest, 11).agg
        op.srcA, op.srcWidth = src, width
        op.domains.extend(domains)
        a = op.range.add(tag={"count": 0, "total": 1, "min": 2, "max": 3}[tag])
        if tag == "count":
            a.count.aggrng = rng
        else:
            getattr(a, tag).aggdom = dom
            getattr(a, tag).aggrng = rng
        return self

    def setop(self, dest, kind: str, a, b):
        tag = {"union": 2, "intersection": 3, "difference": 6}[kind]
        op = getattr(self._assign(dest, tag), {"union": "unionop"}.get(kind, kind))
        op.srcA, op.srcB = a, b
        return self

    def single(self, dest, constants):
        op = self._assign(dest, 5).single
        for c in constants:
            op.element.add().CopyFrom(c.constExp.literal)
        return self

    def move(self, dest, src):
        c = self.s.operators.add(tag=5)
        c.move.dest, c.move.src1 = dest, src
        return self

    def goto(self, target: int):
        self.s.operators.add(tag=4).jump.target = target
        return self

    def cond(self, a, b, yes: int, no: int):
        c = self.s.operators.add(tag=3)
        c.cond.src1, c.cond.src2, c.cond.yes, c.cond.no = a, b, yes, no
        return self

    def halt(self):
        self.s.operators.add(tag=6)
        return self


def q1_plan() -> bytes:
    """TPC-H Q1 as the reference's hand-written plan builds it
    (q1.man.cpp:31-2329): columns keyed by (LINENUMBER, ORDER), the date
    filter as a product with the start date, one key join per column, and
    one AGG per output grouped by (RETURNFLAG, LINESTATUS)."""
    b = Builder("TPC-H/logicQ1.lbb")
    Q = "+$logicQ1:_"
    key2 = ["entity:LINENUMBER", "entity:ORDER"]
    grp = ["entity:RETURNFLAG", "entity:LINESTATUS"]
    b.var(Q + "dateDelta", fields=["int64"])
    for out in ("sum_qty", "sum_base_price", "avg_disc_adt", "sum_disc_price", "sum_charge"):
        b.var(Q + out, grp, ["float64"])
    b.var(Q + "count_order", grp, ["int64"])
    b.var("L_SHIPDATE", key2, ["datetime"])
    b.var("L_RETURNFLAG", key2, ["entity:RETURNFLAG"])
    b.var("L_LINESTATUS", key2, ["entity:LINESTATUS"])
    for c in ("L_QUANTITY", "L_EXTENDEDPRICE", "L_DISCOUNT", "L_TAX"):
        b.var(c, key2, ["float64"])
    b.var(Q + "startDate", fields=["datetime"])
    b.var("j_1", key2, ["datetime", "datetime"])
    b.var("m_1", key2)
    b.var("j_2", key2, ["entity:RETURNFLAG"])
    b.var("j_3", key2, grp)
    for j in ("j_4", "j_5", "m_3", "m_5"):
        b.var(j, key2, grp + ["float64"])
    for j in ("j_6", "j_7"):
        b.var(j, key2, grp + ["float64", "float64"])
    for m in ("m_2", "m_4"):
        b.var(m, key2, grp + ["float64", "float64", "float64"])
    i = b.idx
    ln, od = i(0, "entity:LINENUMBER"), i(1, "entity:ORDER")
    rf, ls = i(2, "entity:RETURNFLAG"), i(3, "entity:LINESTATUS")
    s = b.sequence(0)
    s.mapfilter(Q + "startDate", Q + "dateDelta", 1, b.always(), [
        b.call("subtract", [b.const_datetime(912470400), i(0, "int64"), b.const_string("days")])])
    s.join("j_1", 0, "L_SHIPDATE", Q + "startDate",
           [ln, od, i(2, "datetime"), i(3, "datetime")])
    s.mapfilter("m_1", "j_1", 4, b.comp("le", i(2, "datetime"), i(3, "datetime")), [ln, od])
    s.join("j_2", 2, "m_1", "L_RETURNFLAG", [ln, od, rf])
    s.join("j_3", 2, "j_2", "L_LINESTATUS", [ln, od, rf, ls])
    s.agg(Q + "count_order", "j_3", 4, [2, 3], "count", rng=0)
    f4, f5, f6 = i(4, "float64"), i(5, "float64"), i(6, "float64")
    s.join("j_4", 2, "j_3", "L_QUANTITY", [ln, od, rf, ls, f4])
    s.agg(Q + "sum_qty", "j_4", 5, [2, 3], "total", 4, 4)
    s.join("j_5", 2, "j_3", "L_EXTENDEDPRICE", [ln, od, rf, ls, f4])
    s.agg(Q + "sum_base_price", "j_5", 5, [2, 3], "total", 4, 4)
    s.join("j_6", 2, "j_5", "L_DISCOUNT", [ln, od, rf, ls, f4, f5])
    s.agg(Q + "avg_disc_adt", "j_6", 6, [2, 3], "total", 5, 5)
    one = b.const_float(1)
    s.mapfilter("m_2", "j_6", 6, b.always(),
                [ln, od, rf, ls, f4, f5, b.arith("mul", f4, b.arith("sub", one, f5))])
    s.mapfilter("m_3", "m_2", 7, b.always(), [ln, od, rf, ls, f6])
    s.agg(Q + "sum_disc_price", "m_3", 5, [2, 3], "total", 4, 4)
    s.join("j_7", 2, "m_3", "L_TAX", [ln, od, rf, ls, f4, f5])
    s.mapfilter("m_4", "j_7", 6, b.always(),
                [ln, od, rf, ls, f4, f5, b.arith("mul", f4, b.arith("add", one, f5))])
    s.mapfilter("m_5", "m_4", 7, b.always(), [ln, od, rf, ls, f6])
    s.agg(Q + "sum_charge", "m_5", 5, [2, 3], "total", 4, 4)
    s.goto(1)
    b.sequence(1).halt()
    return b.entry_exit(0, 1).bytes()
