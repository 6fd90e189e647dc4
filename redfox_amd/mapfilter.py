"""MAPFILTER: expressions over columns, compiled to the register program
that rfx_mapfilter interprets (include/rfx_ra.h).

Expressions are built from column references and constants with Python
operators; each node has a type: "i" (int64), "u" (uint64: only 64-bit
unsigned columns) or "f" (double).  Arithmetic follows C's usual
conversions (a double operand makes the operation a double one), as the
C expressions the reference generated for set_arith do
(RelationalAlgebraKernel.cpp:1822-1939); constants stay exact doubles
rather than the 6-digit text the reference printed them as.

    x = MapFilter([(price, "f64"), (disc, "f64"), (ship, "i32")])
    p, d, s = x.cols
    x.where((s <= 2436) & (d > 0.05))
    x.emit(p * (1 - d), "f64")
    x.emit(s, "i32")
    n = x.run(ctx, n_rows, [out0, out1])
"""
import ctypes
from typing import List, Optional, Sequence, Tuple

from . import rfx

TYPES = {"u8": 1, "u16": 2, "u32": 3, "u64": 4, "i8": 5, "i16": 6, "i32": 7, "i64": 8,
         "f32": 9, "f64": 10, "tuple1": 11, "tuple2": 12, "tuple4": 13, "tuple8": 14,
         "tuple16": 15}
OP = dict(NOP=0, LDC=1, LDI=2, MOV=3, LDX=4, IADD=10, ISUB=11, IMUL=12, IDIV=13, IMOD=14, FADD=20,
          FSUB=21, FMUL=22, FDIV=23, I2F=30, U2F=31, F2I=32, ILT=40, ILE=41, IEQ=42, INE=43,
          ULT=44, ULE=45, FLT=46, FLE=47, FEQ=48, FNE=49, AND=50, OR=51, NOT=52, ADD_MONTH=60,
          ADD_YEAR=61, PART_YEAR=62, SCHUNK=63, LIKE=64, OUT=70)
MAX_OPS, MAX_COLS, MAX_OUTS, REGS = 96, 16, 8, 16


class Op(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("d", ctypes.c_uint8), ("a", ctypes.c_uint8),
                ("b", ctypes.c_uint8), ("reserved", ctypes.c_uint32), ("imm", ctypes.c_uint64)]


class Program(ctypes.Structure):
    _fields_ = [("n_filter", ctypes.c_uint32), ("n_ops", ctypes.c_uint32),
                ("keep_reg", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("ops", Op * MAX_OPS)]


class Col(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("type", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("rows", ctypes.c_uint64)]


class Out(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("type", ctypes.c_int32), ("reserved", ctypes.c_int32)]


def _f64_bits(v: float) -> int:
    return ctypes.c_uint64.from_buffer_copy(ctypes.c_double(float(v))).value


class Expr:
    type = "i"

    def tree(self):
        raise NotImplementedError

    # arithmetic
    def __add__(self, o): return Bin("add", self, lift(o))
    def __radd__(self, o): return Bin("add", lift(o), self)
    def __sub__(self, o): return Bin("sub", self, lift(o))
    def __rsub__(self, o): return Bin("sub", lift(o), self)
    def __mul__(self, o): return Bin("mul", self, lift(o))
    def __rmul__(self, o): return Bin("mul", lift(o), self)
    def __truediv__(self, o): return Bin("div", self, lift(o))
    def __rtruediv__(self, o): return Bin("div", lift(o), self)
    def __mod__(self, o): return Bin("mod", self, lift(o))
    def __rmod__(self, o): return Bin("mod", lift(o), self)
    # comparisons (results are 0/1 integers)
    def __lt__(self, o): return Cmp("lt", self, lift(o))
    def __le__(self, o): return Cmp("le", self, lift(o))
    def __gt__(self, o): return Cmp("gt", self, lift(o))
    def __ge__(self, o): return Cmp("ge", self, lift(o))
    def eq(self, o): return Cmp("eq", self, lift(o))
    def ne(self, o): return Cmp("ne", self, lift(o))
    # logic
    def __and__(self, o): return Logic("and", self, lift(o))
    def __or__(self, o): return Logic("or", self, lift(o))
    def __invert__(self): return Unary("not", self)


class Column(Expr):
    def __init__(self, index: int, ctype: str, field: Optional[Tuple[int, int, str]] = None):
        self.index, self.ctype, self.field = index, ctype, field
        if ctype.startswith("tuple"):
            # A packed tuple column is read through its fields (MapFilter.field).
            kind = field[2] if field else "unsigned"
            bits = field[1] if field else 64
            self.type = {"unsigned": "i" if bits < 64 else "u", "signed": "i", "double": "f"}[kind]
        else:
            self.type = "f" if ctype in ("f32", "f64") else "u" if ctype == "u64" else "i"

    def tree(self):
        return ("col", self.index) if self.field is None else ("field", self.index) + tuple(self.field)


class Gather(Column):
    """Column `index` at the row an integer expression gives (LDX); a row
    past the column's length reads 0."""

    def __init__(self, index: int, ctype: str, row: "Expr",
                 field: Optional[Tuple[int, int, str]] = None):
        super().__init__(index, ctype, field)
        if row.type == "f":
            raise ValueError("a gather's row must be an integer expression")
        self.row = row

    def tree(self):
        return ("gather", self.index, self.row.tree()) + ((tuple(self.field),) if self.field else ())


class Const(Expr):
    def __init__(self, v):
        if isinstance(v, bool):
            v = int(v)
        self.v = v
        self.type = "f" if isinstance(v, float) else "i"

    def tree(self):
        return ("float", float(self.v)) if self.type == "f" else ("int", int(self.v))


class Like(Expr):
    """The string at address s LIKE the pattern at address p (LIKE op);
    `pattern` is the pattern's bytes, for the expression tree."""
    type = "i"

    def __init__(self, s, p, pattern: bytes):
        if "f" in (s.type, p.type):
            raise ValueError("LIKE takes string addresses")
        self.s, self.p, self.pattern = s, p, bytes(pattern)

    def tree(self):
        return ("like", self.s.tree(), self.pattern)


def like(s, p, pattern: bytes) -> Expr:
    return Like(lift(s), lift(p), pattern)


def lift(v) -> Expr:
    return v if isinstance(v, Expr) else Const(v)


class Bin(Expr):
    def __init__(self, op, x, y):
        self.op, self.x, self.y = op, x, y
        self.type = "f" if "f" in (x.type, y.type) else "u" if "u" in (x.type, y.type) else "i"
        if self.type == "u" and op in ("div", "mod"):
            raise ValueError("unsigned 64-bit division is not supported")

    def tree(self):
        return (self.op, self.x.tree(), self.y.tree())


class Cmp(Expr):
    type = "i"

    def __init__(self, op, x, y):
        self.op, self.x, self.y = op, x, y

    def tree(self):
        return (self.op, self.x.tree(), self.y.tree())


class Logic(Expr):
    type = "i"

    def __init__(self, op, x, y):
        if "f" in (x.type, y.type):
            raise ValueError("logical operands must be integers")
        self.op, self.x, self.y = op, x, y

    def tree(self):
        return (self.op, self.x.tree(), self.y.tree())


class Unary(Expr):
    def __init__(self, op, x, k=0):
        self.op, self.x, self.k = op, x, k
        self.type = {"not": "i", "i2f": "f", "f2i": "i", "add_month": "i", "add_year": "i",
                     "part_year": "i", "schunk": "u"}[op]
        if op in ("add_month", "add_year", "part_year", "not", "schunk") and x.type == "f":
            raise ValueError(f"{op} needs an integer operand")

    def tree(self):
        if self.op in ("add_month", "add_year", "schunk"):
            return (self.op, self.x.tree(), self.k)
        return (self.op, self.x.tree())


def to_float(e) -> Expr: return Unary("i2f", lift(e)) if lift(e).type != "f" else lift(e)
def to_int(e) -> Expr: return Unary("f2i", lift(e)) if lift(e).type == "f" else lift(e)
def add_month(e, k: int) -> Expr: return Unary("add_month", lift(e), int(k))
def add_year(e, k: int) -> Expr: return Unary("add_year", lift(e), int(k))
def part_year(e) -> Expr: return Unary("part_year", lift(e))
def str_chunk(e, k: int) -> Expr: return Unary("schunk", lift(e), int(k))


def _const_chunks(value: bytes, n: int):
    """The first n chunks of a constant string as SCHUNK computes them."""
    codes = [0 if b == 0 else (b ^ 0x80) + (1 if b >= 0x80 else 0)
             for b in value.split(b"\0", 1)[0]]
    codes += [0] * (8 * n - len(codes))
    return [int.from_bytes(bytes(codes[8 * k:8 * k + 8]), "big") for k in range(n)]


def str_compare(e, value: bytes, op: str) -> Expr:
    """Compare the string at address e with a constant string (op: eq, ne,
    lt, le, gt, ge) in compare_string2's order, chunk by chunk: chunks up to
    the one holding the constant's terminator decide."""
    value = value.split(b"\0", 1)[0]
    n = len(value) // 8 + 1
    c = _const_chunks(value, n)
    x = [str_chunk(e, k) for k in range(n)]
    if op in ("eq", "ne"):
        r = x[0].eq(c[0])
        for k in range(1, n):
            r = r & x[k].eq(c[k])
        return r if op == "eq" else ~r
    strict = op in ("lt", "gt")
    less = op in ("lt", "le")
    # Lexicographic: decided at the first chunk that differs.
    r = None
    for k in reversed(range(n)):
        here = (x[k] < c[k]) if less else (x[k] > c[k])
        if k == n - 1:
            tail = here if strict else (here | x[k].eq(c[k]))
        else:
            tail = here | (x[k].eq(c[k]) & r)
        r = tail
    return r


class _Gen:
    """Register code: gen(e, r) leaves e in register r and uses only r and
    above."""

    def __init__(self, cols):
        self.ops: List[Tuple[int, int, int, int, int]] = []
        self.cols = cols

    def emit(self, code, d=0, a=0, b=0, imm=0):
        self.ops.append((OP[code], d, a, b, imm & ((1 << 64) - 1)))

    def conv(self, r, have, want):
        if want == "f" and have != "f":
            self.emit("U2F" if have == "u" else "I2F", r, r)
        elif want != "f" and have == "f":
            self.emit("F2I", r, r)

    def gen(self, e: Expr, r: int):
        if r >= REGS:
            raise ValueError(f"expression too deep for {REGS} registers")
        if isinstance(e, Column):
            if e.ctype.startswith("tuple") and e.field is None:
                raise ValueError("a tuple column is read through a field (MapFilter.field)")
            imm = 0
            if e.field is not None:
                shift, bits, kind = e.field
                imm = shift | (bits << 8) | ({"unsigned": 0, "signed": 1, "double": 2}[kind] << 16)
            if isinstance(e, Gather):
                self.gen(e.row, r)
                self.emit("LDX", r, e.index, r, imm)
            else:
                self.emit("LDC", r, e.index, 0, imm)
        elif isinstance(e, Const):
            self.emit("LDI", r, 0, 0, _f64_bits(e.v) if e.type == "f" else int(e.v))
        elif isinstance(e, Bin):
            self.gen(e.x, r)
            self.conv(r, e.x.type, e.type)
            self.gen(e.y, r + 1)
            self.conv(r + 1, e.y.type, e.type)
            f = e.type == "f"
            code = {"add": "FADD" if f else "IADD", "sub": "FSUB" if f else "ISUB",
                    "mul": "FMUL" if f else "IMUL", "div": "FDIV" if f else "IDIV",
                    "mod": "IMOD"}[e.op]
            if f and e.op == "mod":
                raise ValueError("% needs integer operands")
            self.emit(code, r, r, r + 1)
        elif isinstance(e, Cmp):
            t = "f" if "f" in (e.x.type, e.y.type) else "u" if "u" in (e.x.type, e.y.type) else "i"
            self.gen(e.x, r)
            self.conv(r, e.x.type, t if t == "f" else e.x.type)
            self.gen(e.y, r + 1)
            self.conv(r + 1, e.y.type, t if t == "f" else e.y.type)
            a, b, op = r, r + 1, e.op
            if op in ("gt", "ge"):  # x > y  <=>  y < x
                a, b, op = b, a, {"gt": "lt", "ge": "le"}[op]
            pre = {"f": "F", "u": "U", "i": "I"}[t]
            if op in ("eq", "ne"):
                pre = "F" if t == "f" else "I"
            self.emit(pre + op.upper(), r, a, b)
        elif isinstance(e, Like):
            self.gen(e.s, r)
            self.gen(e.p, r + 1)
            self.emit("LIKE", r, r, r + 1)
        elif isinstance(e, Logic):
            self.gen(e.x, r)
            self.gen(e.y, r + 1)
            self.emit(e.op.upper(), r, r, r + 1)
        elif isinstance(e, Unary):
            self.gen(e.x, r)
            if e.op == "not":
                self.emit("NOT", r, r)
            elif e.op == "i2f":
                self.conv(r, e.x.type, "f")
            elif e.op == "f2i":
                self.conv(r, "f", "i")
            elif e.op == "schunk":
                self.emit("SCHUNK", r, r, 0, e.k)
            else:
                self.emit(e.op.upper(), r, r, 0, e.k)
        else:
            raise TypeError(type(e))


class MapFilter:
    """A MAPFILTER over device columns: a filter and output expressions."""

    def __init__(self, columns: Sequence[Tuple]):
        """columns: (buffer, type) or (buffer, type, rows); rows bounds the
        rows a gather may read (default: the pass's row count)."""
        if len(columns) > MAX_COLS:
            raise ValueError("at most 16 input columns")
        self.columns = [tuple(c[:2]) for c in columns]
        self.rows = [int(c[2]) if len(c) > 2 else 0 for c in columns]
        self.cols = [Column(i, t) for i, (_, t) in enumerate(self.columns)]
        self.filter: Optional[Expr] = None
        self.outputs: List[Tuple[Expr, str]] = []

    def field(self, index: int, shift: int, bits: int, kind: str = "unsigned") -> Column:
        return Column(index, self.columns[index][1], (shift, bits, kind))

    def gather(self, index: int, row, field: Optional[Tuple[int, int, str]] = None) -> Column:
        """Column `index` at row `row` (an integer expression)."""
        return Gather(index, self.columns[index][1], lift(row), field)

    def where(self, cond: Expr) -> "MapFilter":
        self.filter = cond
        return self

    def emit(self, e, out_type: str) -> "MapFilter":
        if len(self.outputs) >= MAX_OUTS:
            raise ValueError("at most 8 outputs")
        self.outputs.append((lift(e), out_type))
        return self

    def program(self) -> Program:
        g = _Gen(self.columns)
        if self.filter is not None:
            g.gen(self.filter, 0)
            if self.filter.type == "f":
                raise ValueError("the filter must be an integer (0 or 1) expression")
        n_filter = len(g.ops)
        for k, (e, t) in enumerate(self.outputs):
            g.gen(e, 0)
            g.conv(0, e.type, "f" if t in ("f32", "f64") else "i")
            g.emit("OUT", 0, 0, k)
        if len(g.ops) > MAX_OPS:
            raise ValueError(f"{len(g.ops)} ops, more than {MAX_OPS}")
        p = Program()
        p.n_filter, p.n_ops, p.keep_reg = n_filter, len(g.ops), 0
        for i, (code, d, a, b, imm) in enumerate(g.ops):
            p.ops[i].code, p.ops[i].d, p.ops[i].a, p.ops[i].b, p.ops[i].imm = code, d, a, b, imm
        return p

    def run(self, ctx, n_rows: int, outs: Sequence[object], n_kept) -> None:
        """Enqueue on ctx; *n_kept (a device u64 cell) gets the kept count."""
        if len(outs) != len(self.outputs):
            raise ValueError("one output buffer per emitted expression")
        prog = self.program()
        cols = (Col * max(len(self.columns), 1))()
        for i, (buf, t) in enumerate(self.columns):
            cols[i].data, cols[i].type, cols[i].rows = buf.ptr, TYPES[t], self.rows[i]
        outa = (Out * max(len(outs), 1))()
        for k, (buf, (_, t)) in enumerate(zip(outs, self.outputs)):
            outa[k].data, outa[k].type = buf.ptr, TYPES[t]
        rfx._check("rfx_mapfilter", rfx.lib().rfx_mapfilter(
            ctx.handle, cols, len(self.columns), n_rows, ctypes.byref(prog), outa, len(outs),
            n_kept.ptr))

    def trees(self):
        """The filter and output expressions as nested tuples."""
        return (self.filter.tree() if self.filter is not None else None,
                [(e.tree(), t) for e, t in self.outputs])
