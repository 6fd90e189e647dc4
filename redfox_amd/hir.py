"""Harmony IR programs: an encoder for the HIR file format and a binding to
the executor in librfx_hip.so (rfx_hir_*).

The format is HarmonyIR.proto's KernelControlFlowGraph behind a u64 byte
count, followed by `testCount` (u64 byte count, Test) records
(RelationalAlgebraCompiler.cpp:9429-9437,9502-9520).  The encoder writes
the protobuf wire format directly: field numbers and types are those of
HarmonyIR.proto; no protobuf library is needed.
"""
import ctypes
import struct
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import rfx
from .rfx import RfxError, _check

# DataType, AccessMode and KernelType codes of HarmonyIR.proto.
I8, I16, I32, I64, I128, F32, F64 = 1, 2, 3, 4, 5, 6, 7
IN, OUT, INOUT = 1, 2, 3
COMPUTE, BINARY, CONTROL, BRANCH, GET_SIZE, RESIZE, UPDATE_SIZE, EXIT = range(1, 9)


def _varint(v: int) -> bytes:
    out = bytearray()
    v &= (1 << 64) - 1
    while True:
        b = v & 0x7F
        v >>= 7
        if v:
            out.append(b | 0x80)
        else:
            out.append(b)
            return bytes(out)


def _vfield(num: int, v: int) -> bytes:
    return _varint(num << 3) + _varint(int(v))


def _bfield(num: int, b: bytes) -> bytes:
    return _varint((num << 3) | 2) + _varint(len(b)) + b


@dataclass
class Variable:
    name: int
    type: int = I8
    size: int = 0            # bytes, as Harmony's runtime reads it
    output: bool = False
    input: bool = False
    data: Optional[bytes] = None
    filename: Optional[str] = None

    def encode(self) -> bytes:
        out = (_vfield(1, self.name) + _vfield(2, self.type) + _vfield(3, self.size)
               + _vfield(4, self.output) + _vfield(5, self.input))
        if self.data is not None:
            out += _bfield(6, bytes(self.data))
        if self.filename is not None:
            out += _bfield(7, self.filename.encode())
        return out


@dataclass
class Kernel:
    name: str
    type: int
    operands: Sequence[Tuple[int, int]] = ()   # (mode, variable)
    targets: Sequence[int] = ()
    code: bytes = b""
    packed_targets: bool = False

    def encode(self) -> bytes:
        out = _bfield(1, self.name.encode()) + _bfield(2, self.code) + _vfield(3, self.type)
        for mode, var in self.operands:
            out += _bfield(4, _vfield(1, mode) + _vfield(2, var))
        if self.packed_targets and self.targets:
            out += _bfield(5, b"".join(_varint(t) for t in self.targets))
        else:
            for t in self.targets:
                out += _vfield(5, t)
        return out


@dataclass
class Block:
    id: int
    kernels: List[Kernel] = field(default_factory=list)
    control: Optional[Kernel] = None

    def encode(self) -> bytes:
        ctl = self.control or Kernel("exit", EXIT)
        return (_vfield(1, self.id) + b"".join(_bfield(2, k.encode()) for k in self.kernels)
                + _bfield(3, ctl.encode()))


@dataclass
class Test:
    name: str
    program: str
    inputs: Sequence[Tuple[int, bytes]] = ()
    outputs: Sequence[Tuple[int, bytes]] = ()
    features: Sequence[int] = ()

    def encode(self) -> bytes:
        def var(name, data):
            return Variable(name, I8, len(data), False, False, bytes(data)).encode()
        out = b"".join(_bfield(1, var(n, d)) for n, d in self.inputs)
        out += b"".join(_bfield(2, var(n, d)) for n, d in self.outputs)
        out += _bfield(3, self.program.encode()) + _bfield(4, self.name.encode())
        for f in self.features:
            out += _vfield(5, f)
        return out


def encode(name: str, blocks: Sequence[Block], variables: Sequence[Variable], entry: int,
           exit: int, tests: Sequence[Test] = ()) -> bytes:
    """A HIR file image: u64 length + program message + test records."""
    cfg = (_bfield(1, name.encode()) + b"".join(_bfield(2, b.encode()) for b in blocks)
           + b"".join(_bfield(3, v.encode()) for v in variables)
           + _vfield(4, entry) + _vfield(5, exit) + _vfield(6, len(tests)))
    out = struct.pack("<Q", len(cfg)) + cfg
    for t in tests:
        body = t.encode()
        out += struct.pack("<Q", len(body)) + body
    return out


class Program:
    """A builder: numbered variables and blocks in order, the last block the
    exit (each block's control branches to the next)."""

    def __init__(self, name: str):
        self.name = name
        self.variables: List[Variable] = []
        self.blocks: List[Block] = []
        self.tests: List[Test] = []

    def var(self, type: int = I8, size: int = 0, data=None, input: bool = False,
            output: bool = False) -> int:
        if data is not None:
            data = data.tobytes() if isinstance(data, np.ndarray) else bytes(data)
            if not size:
                size = len(data)
        v = Variable(len(self.variables) + 1, type, size, output, input, data)
        self.variables.append(v)
        return v.name

    def u64(self, value: int, input: bool = True) -> int:
        return self.var(I64, 8, struct.pack("<Q", value & ((1 << 64) - 1)), input=input)

    def block(self, *kernels: Kernel) -> int:
        bid = 100 + len(self.blocks)
        self.blocks.append(Block(bid, list(kernels)))
        return bid

    def test(self, name: str, inputs=(), outputs=()) -> None:
        conv = lambda d: d.tobytes() if isinstance(d, np.ndarray) else bytes(d)
        self.tests.append(Test(name, self.name, [(n, conv(d)) for n, d in inputs],
                               [(n, conv(d)) for n, d in outputs]))

    def encode(self) -> bytes:
        blocks = []
        for i, b in enumerate(self.blocks):
            nxt = self.blocks[i + 1].id if i + 1 < len(self.blocks) else None
            ctl = Kernel("branch", BRANCH, targets=[nxt]) if nxt is not None else Kernel("exit", EXIT)
            blocks.append(Block(b.id, b.kernels, ctl))
        return encode(self.name, blocks, self.variables, blocks[0].id, blocks[-1].id, self.tests)


def binary(name: str, *operands: Tuple[int, int]) -> Kernel:
    return Kernel(name, BINARY, operands)


def get_size(target: int, source: int) -> Kernel:
    return Kernel("get_size", GET_SIZE, [(OUT, target), (IN, source)])


def resize(variable: int, size: int) -> Kernel:
    return Kernel("resize", RESIZE, [(OUT, variable), (IN, size)])


def update_size(variable: int, size: int) -> Kernel:
    return Kernel("update_size", UPDATE_SIZE, [(OUT, variable), (IN, size)])


def tuple_text(fields) -> str:
    """A tuple for an rfx1 descriptor: field widths in bits, each an int or
    (bits, kind) with kind "i", "f" (float) or "p" (string address)."""
    if not fields:
        return "-"
    out = []
    for f in fields:
        if isinstance(f, tuple):
            bits, kind = f
            out.append(f"{bits}{'' if kind == 'i' else kind}")
        else:
            out.append(str(int(f)))
    return ",".join(out)


def descriptor(family: str, **kw) -> bytes:
    """The rfx1 descriptor of a kernel instance: what Red Fox's generated
    source bakes in (RelationalAlgebraKernel::cudaSourceRepresentation,
    RelationalAlgebraKernel.cpp:849-909), written into the kernel's code
    bytes.  Tuple members (d, d1, a, b, c, e) take field lists; dom, parts
    and nt ints; map a list of field indices; agg and cmp strings."""
    parts = ["rfx1", family]
    for k in ("d", "d1", "a", "b", "c", "e"):
        if k in kw:
            parts.append(f"{k}={tuple_text(kw[k])}")
    for k in ("dom", "parts", "nt"):
        if k in kw:
            parts.append(f"{k}={int(kw[k])}")
    if "map" in kw:
        parts.append("map=" + ",".join(str(int(i)) for i in kw["map"]))
    for k in ("agg", "cmp"):
        if k in kw:
            parts.append(f"{k}={kw[k]}")
    return " ".join(parts).encode()


def compute(name: str, dims: Sequence[int], *operands: Tuple[int, int], code: bytes = b"") -> Kernel:
    """A ComputeKernel record: dims are the ctaCount, threadCount and
    sharedBytes variables (operands 0-2, read by value)."""
    return Kernel(name, COMPUTE, [(IN, d) for d in dims] + list(operands), code=code)


def _bits(fields) -> int:
    return sum(f[0] if isinstance(f, tuple) else int(f) for f in fields)


def _data_type(fields) -> int:
    """getTupleDataType (RelationalAlgebraCompiler.cpp:494-512): I8..I128
    by the tuple's total bits."""
    b = _bits(fields)
    for code, lim in ((1, 8), (2, 16), (3, 32), (4, 64), (5, 128)):
        if b <= lim:
            return code
    raise ValueError(f"{b}-bit tuple")


class Lowering:
    """Red Fox's lowering of RA nodes into Harmony kernel sequences, written
    into a Program with rfx1 descriptors in the kernels' code bytes: the
    same kernels, names, operands and access modes, in the same order, as
    RelationalAlgebraCompiler.cpp emits them.  Relations are packed tuple
    buffers (I8 variables holding byte sizes); tuples are field lists as
    in descriptor()."""

    def __init__(self, prog: Program, first_id: int = 1):
        self.p = prog
        self.next_id = first_id
        self.consts = {}

    def _id(self, family: str) -> str:
        name = f"{family}{self.next_id}"
        self.next_id += 1
        return name

    def const(self, v: int) -> int:
        """getConstantIntVariable: one I64 variable per value."""
        if v not in self.consts:
            self.consts[v] = self.p.u64(v)
        return self.consts[v]

    def temp_buffer(self) -> int:
        return self.p.var(I8, 0)

    def temp_int(self) -> int:
        return self.p.var(I64, 8)

    def _dims(self, cta: int, threads: int):
        return (self.const(cta), self.const(threads), self.const(0))

    def sort_key(self, ks: List[Kernel], key: int, key_t, size: int) -> None:
        """addSortKey (:3410-3460)."""
        fam = "set_mgpusortkey" if _bits(key_t) > 37 else "set_b40csortkey"
        arg = self.const(_data_type(key_t) if fam == "set_mgpusortkey" else _bits(key_t))
        ks.append(Kernel(self._id(fam), BINARY, [(INOUT, key), (IN, size), (IN, arg)],
                         code=descriptor(fam, d=key_t)))

    def sort_pair(self, ks: List[Kernel], key: int, key_t, val: int, val_t, size: int) -> None:
        """addSortPair (:3462-3539): mergesort past 33 key bits or for wide
        values, else the LSD radix sort."""
        kb, vbytes = _bits(key_t), (_bits(val_t) + 7) // 8
        mgpu = kb > 33 or vbytes > 16 or (kb >= 23 and vbytes > 8) or (kb >= 32 and vbytes > 4)
        if mgpu:
            fam, a0 = "set_mgpusortpair", self.const(_data_type(key_t))
        else:
            fam, a0 = "set_b40csortpair", self.const(kb)
        ks.append(Kernel(self._id(fam), BINARY,
                         [(INOUT, key), (INOUT, val), (IN, size), (IN, a0),
                          (IN, self.const(_data_type(val_t)))],
                         code=descriptor(fam, d=key_t, d1=val_t)))

    def _split(self, ks: List[Kernel], src: int, src_t, size: int, keycount: int):
        key_t, val_t = list(src_t[:keycount]), list(src_t[keycount:])
        key, val = self.temp_buffer(), self.temp_buffer()
        ksize, vsize = self.temp_int(), self.temp_int()
        ks.append(compute(self._id("split_get_result_size"), self._dims(1, 1), (IN, size),
                          (OUT, ksize), (OUT, vsize),
                          code=descriptor("split_get_result_size", d=key_t, d1=val_t, a=src_t)))
        ks.append(resize(key, ksize))
        ks.append(resize(val, vsize))
        ks.append(compute(self._id("set_split"), self._dims(350, 256), (INOUT, key), (INOUT, val),
                          (IN, src), (IN, size),
                          code=descriptor("set_split", d=key_t, d1=val_t, a=src_t, dom=keycount)))
        return key, key_t, ksize, val, val_t, vsize

    def join(self, ks: List[Kernel], dest: int, dest_t, a: int, a_t, b: int, b_t, keycount: int,
             a_sorted: bool = False, b_sorted: bool = False) -> None:
        """addJoin (:4466-5226)."""
        lsize, rsize = self.temp_int(), self.temp_int()
        ks.append(get_size(lsize, a))
        ks.append(get_size(rsize, b))
        if keycount < len(a_t):
            lk, lk_t, lks, lv, lv_t, _ = self._split(ks, a, a_t, lsize, keycount)
        else:
            lk, lk_t, lks, lv, lv_t = a, list(a_t), lsize, None, []
        if keycount < len(b_t):
            rk, rk_t, rks, rv, rv_t, _ = self._split(ks, b, b_t, rsize, keycount)
        else:
            rk, rk_t, rks, rv, rv_t = b, list(b_t), rsize, None, []
        if not a_sorted:
            if lv is not None:
                self.sort_pair(ks, lk, lk_t, lv, lv_t, lks)
            else:
                self.sort_key(ks, lk, lk_t, lks)
        if not b_sorted:
            if rv is not None:
                self.sort_pair(ks, rk, rk_t, rv, rv_t, rks)
            else:
                self.sort_key(ks, rk, rk_t, rks)
        lb, lc, lbs = self.temp_buffer(), self.temp_buffer(), self.temp_int()
        ks.append(compute(self._id("join_get_temp_size"), self._dims(1, 1), (OUT, lbs), (IN, lsize),
                          code=descriptor("join_get_temp_size", d=a_t)))
        ks.append(resize(lb, lbs))
        ks.append(resize(lc, lbs))
        total = self.temp_int()
        ks.append(Kernel(self._id("mgpu_join_find_bounds"), BINARY,
                         [(INOUT, lb), (INOUT, lc), (OUT, total), (IN, lk), (IN, lks), (IN, rk),
                          (IN, rks)],
                         code=descriptor("mgpu_join_find_bounds", a=lk_t, b=rk_t)))
        isz = self.temp_int()
        ks.append(compute(self._id("join_get_result_size"), self._dims(1, 1), (OUT, isz),
                          (INOUT, total), code=descriptor("join_get_result_size", d=dest_t)))
        li, ri = self.temp_buffer(), self.temp_buffer()
        ks.append(resize(li, isz))
        ks.append(resize(ri, isz))
        ks.append(Kernel(self._id("mgpu_join_main"), BINARY,
                         [(INOUT, li), (INOUT, ri), (IN, isz), (IN, lb), (IN, lc), (IN, lbs),
                          (IN, rsize)],
                         code=descriptor("mgpu_join_main")))
        ks.append(resize(dest, total))
        ops = [(INOUT, dest), (IN, lk), (IN, li), (IN, rk), (IN, ri), (IN, isz)]
        if lv is not None:
            ops.append((IN, lv))
        if rv is not None:
            ops.append((IN, rv))
        ks.append(compute(self._id("mgpu_join_gather"), self._dims(350, 256), *ops,
                          code=descriptor("mgpu_join_gather", d=dest_t, a=lk_t, b=rk_t, c=lv_t,
                                          e=rv_t, dom=keycount)))

    def count(self, ks: List[Kernel], dest: int, dest_t, src: int, src_t, domains: Sequence[int],
              src_sorted: bool = False) -> None:
        """addAgg's COUNT path (:8335-8641): project the grouping fields,
        sort them, generate ones, set_count, then merge keys and counts into
        the destination tuples."""
        size = self.temp_int()
        ks.append(get_size(size, src))
        proj_t = [src_t[i] for i in domains]
        proj = self.temp_buffer()
        if (_bits(proj_t) + 7) // 8 != (_bits(src_t) + 7) // 8:
            ks.append(compute(self._id("project_get_result_size"), self._dims(1, 1), (INOUT, size),
                              code=descriptor("project_get_result_size", d=proj_t, a=src_t)))
        ks.append(resize(proj, size))
        ks.append(compute(self._id("set_project"), self._dims(350, 256), (INOUT, proj), (IN, src),
                          (IN, size), code=descriptor("set_project", d=proj_t, a=src_t, map=domains)))
        if not src_sorted:
            self.sort_key(ks, proj, proj_t, size)
        gen_t, vsize, gen = [32], self.temp_int(), self.temp_buffer()
        ks.append(compute(self._id("generate_get_result_size"), self._dims(1, 1), (OUT, vsize),
                          (IN, size), code=descriptor("generate_get_result_size", d=gen_t, a=proj_t)))
        ks.append(resize(gen, vsize))
        ks.append(compute(self._id("set_generate"), self._dims(350, 256), (INOUT, gen), (IN, vsize),
                          code=descriptor("set_generate", d=gen_t)))
        rkey, rval = self.temp_buffer(), self.temp_buffer()
        ks.append(resize(rkey, size))
        ks.append(resize(rval, vsize))
        ks.append(Kernel(self._id("set_count"), BINARY,
                         [(INOUT, rkey), (INOUT, rval), (IN, proj), (INOUT, size), (IN, gen),
                          (INOUT, vsize)],
                         code=descriptor("set_count", a=proj_t, b=gen_t, agg="count")))
        ks.append(compute(self._id("merge_get_result_size"), self._dims(1, 1), (INOUT, size),
                          (IN, vsize), code=descriptor("merge_get_result_size", d=dest_t, a=gen_t)))
        ks.append(resize(dest, size))
        ks.append(compute(self._id("set_merge"), self._dims(350, 256), (INOUT, dest), (IN, rkey),
                          (IN, size), (IN, rval), (IN, vsize),
                          code=descriptor("set_merge", d=dest_t, a=proj_t, b=gen_t,
                                          dom=len(proj_t))))


class RfxHirSummary(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 128), ("blocks", ctypes.c_uint32),
                ("variables", ctypes.c_uint32), ("kernels", ctypes.c_uint32),
                ("tests", ctypes.c_uint32), ("entry", ctypes.c_uint32), ("exit", ctypes.c_uint32)]


class Hir:
    """A loaded HIR program (parsing needs no GPU; running does)."""

    def __init__(self, image: bytes):
        self._h = ctypes.c_void_p()
        buf = ctypes.create_string_buffer(bytes(image), len(image))
        _check("rfx_hir_load", rfx.lib().rfx_hir_load(ctypes.byref(self._h), buf, len(image)))

    @classmethod
    def from_file(cls, path: str) -> "Hir":
        with open(path, "rb") as f:
            return cls(f.read())

    def close(self) -> None:
        if self._h:
            rfx.lib().rfx_hir_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        s = RfxHirSummary()
        _check("rfx_hir_info", rfx.lib().rfx_hir_info(self._h, ctypes.byref(s)))
        return {"name": s.name.decode(errors="replace"), "blocks": s.blocks, "variables": s.variables,
                "kernels": s.kernels, "tests": s.tests, "entry": s.entry, "exit": s.exit}

    def schedule(self, block: int) -> List[str]:
        buf = ctypes.create_string_buffer(1 << 16)
        _check("rfx_hir_schedule", rfx.lib().rfx_hir_schedule(self._h, block, buf, len(buf)))
        text = buf.value.decode(errors="replace")
        return text.split("; ") if text else []

    def set_input(self, variable: int, data) -> None:
        b = data.tobytes() if isinstance(data, np.ndarray) else bytes(data)
        buf = ctypes.create_string_buffer(b, max(len(b), 1))
        _check("rfx_hir_set_input", rfx.lib().rfx_hir_set_input(self._h, variable, buf, len(b)))

    def run(self) -> None:
        _check("rfx_hir_run", rfx.lib().rfx_hir_run(self._h))

    def read(self, variable: int, dtype=np.uint8) -> np.ndarray:
        addr, n = ctypes.c_void_p(), ctypes.c_uint64()
        _check("rfx_hir_variable", rfx.lib().rfx_hir_variable(self._h, variable, ctypes.byref(addr), ctypes.byref(n)))
        out = np.empty(n.value, dtype=np.uint8)
        got = ctypes.c_uint64()
        _check("rfx_hir_read", rfx.lib().rfx_hir_read(self._h, variable, out.ctypes.data_as(ctypes.c_void_p),
                                      n.value, ctypes.byref(got)))
        return out.view(dtype)

    def run_tests(self) -> Tuple[int, int, str]:
        p, f = ctypes.c_uint32(), ctypes.c_uint32()
        rep = ctypes.create_string_buffer(1 << 16)
        _check("rfx_hir_run_tests", rfx.lib().rfx_hir_run_tests(self._h, ctypes.byref(p), ctypes.byref(f), rep, len(rep)))
        return p.value, f.value, rep.value.decode(errors="replace")


__all__ = ["Program", "Hir", "Variable", "Kernel", "Block", "Test", "encode", "binary",
           "compute", "descriptor", "tuple_text", "Lowering", "get_size", "resize", "update_size",
           "RfxError"]
