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
           "get_size", "resize", "update_size", "RfxError"]
