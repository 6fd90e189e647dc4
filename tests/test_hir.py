"""HIR file reader and schedule, without a GPU: parsing, the
liveness-driven allocate/free schedule of Runtime::_scheduleBasicBlock
(Runtime.cpp:303-445) on hand-worked programs, and rejection of malformed
files (every truncation and random byte flips of a valid file)."""
import ctypes
import struct

import numpy as np
import pytest

from redfox_amd import hir as H
from redfox_amd import rfx


def straight_program(packed=False):
    """Variables: 1 input A, 2 and 3 temporaries, 4 output, 5 input u64.
    Block 100: X(Out 2, In 1); Y(Out 3, In 2); Z(InOut 3, In 5).
    Block 101: W(Out 4, In 3, In 2).  Block 102: empty, the exit."""
    v = [H.Variable(1, H.I32, 16, input=True, data=bytes(16)), H.Variable(2, H.I8, 32),
         H.Variable(3, H.I8, 32), H.Variable(4, H.I8, 8, output=True),
         H.Variable(5, H.I64, 8, input=True, data=bytes(8))]
    b0 = H.Block(100, [H.binary("X", (H.OUT, 2), (H.IN, 1)), H.binary("Y", (H.OUT, 3), (H.IN, 2)),
                       H.binary("Z", (H.INOUT, 3), (H.IN, 5))],
                 H.Kernel("br", H.BRANCH, targets=[101], packed_targets=packed))
    b1 = H.Block(101, [H.binary("W", (H.OUT, 4), (H.IN, 3), (H.IN, 2))],
                 H.Kernel("br", H.BRANCH, targets=[102], packed_targets=packed))
    b2 = H.Block(102, [], H.Kernel("exit", H.EXIT))
    tests = [H.Test("t0", "straight", [(1, bytes(16))], [(4, bytes(8))]),
             H.Test("t1", "straight", [], [(4, b"\x01" * 8)], features=[3, 7])]
    return H.encode("straight", [b0, b1, b2], v, 100, 102, tests)


def test_info():
    h = H.Hir(straight_program())
    assert h.info() == {"name": "straight", "blocks": 3, "variables": 5, "kernels": 4,
                        "tests": 2, "entry": 100, "exit": 102}


@pytest.mark.parametrize("packed", [False, True])
def test_schedule_follows_liveness(packed):
    """Worked by hand from the rules: a pure Out operand of a non-input is
    allocated before its kernel; a pure In operand of a non-input,
    non-output variable is freed after its last use in the walk that
    starts from the block's live-out set; InOut operands do neither.
    Variable 2 is live out of block 100 (W reads it), so it is freed in
    block 101, after W, with 3."""
    h = H.Hir(straight_program(packed))
    assert h.schedule(0) == ["alloc 2", "run X", "alloc 3", "run Y", "run Z"]
    assert h.schedule(1) == ["alloc 4", "run W", "free 3", "free 2"]
    assert h.schedule(2) == []


def test_loop_back_edge_keeps_values_live():
    """Block 101 branches back to 100, which reads variable 3 (written in
    101).  Without the back edge 3 would be freed after its last use in
    101; with it, 3 is live out of 101 and is not freed there."""
    v = [H.Variable(1, H.I8, 8, input=True, data=bytes(8)), H.Variable(2, H.I8, 8),
         H.Variable(3, H.I8, 8)]

    def prog(back):
        b0 = H.Block(100, [H.binary("P", (H.OUT, 2), (H.IN, 1), (H.INOUT, 3))],
                     H.Kernel("br", H.BRANCH, targets=[101]))
        b1 = H.Block(101, [H.binary("Q", (H.OUT, 3), (H.IN, 2)), H.binary("R", (H.IN, 3))],
                     H.Kernel("c", H.CONTROL, targets=[100, 102] if back else [102]))
        b2 = H.Block(102, [], H.Kernel("exit", H.EXIT))
        return H.encode("loop", [b0, b1, b2], v, 100, 102)

    assert H.Hir(prog(False)).schedule(1) == ["alloc 3", "run Q", "free 2", "run R", "free 3"]
    assert H.Hir(prog(True)).schedule(1) == ["alloc 3", "run Q", "free 2", "run R"]


def test_size_kernels_allocate_nothing_themselves():
    """Resize and UpdateSize targets are not allocated by the schedule
    (Resize allocates when it runs); GetSize targets are."""
    p = H.Program("sizes")
    a = p.var(H.I32, data=np.arange(4, dtype=np.uint32), input=True)
    s = p.var(H.I64, 8)
    r = p.var(H.I8, 0, output=True)
    p.block(H.get_size(s, a), H.resize(r, s), H.update_size(r, s))
    h = H.Hir(p.encode())
    assert h.schedule(0) == ["alloc 2", "run get_size", "run resize", "run update_size", "free 2"]


def malformed():
    good = straight_program()
    n = struct.unpack("<Q", good[:8])[0]
    v_ok = H.Variable(1, H.I8, 8, input=True, data=bytes(8))
    blk = lambda *ks: H.Block(1, list(ks), H.Kernel("exit", H.EXIT))
    yield "short", good[:5]
    yield "length_past_end", struct.pack("<Q", len(good)) + good[8:]
    yield "cut_message", struct.pack("<Q", n - 3) + good[8:8 + n - 3]
    yield "missing_test", good[:8 + n]
    yield "undeclared_variable", H.encode("p", [blk(H.binary("X", (H.IN, 9)))], [v_ok], 1, 1)
    yield "duplicate_variable", H.encode("p", [blk()], [v_ok, v_ok], 1, 1)
    yield "unknown_entry", H.encode("p", [blk()], [v_ok], 7, 1)
    yield "unknown_target", H.encode(
        "p", [H.Block(1, [], H.Kernel("br", H.BRANCH, targets=[5]))], [v_ok], 1, 1)
    yield "bad_type", H.encode("p", [blk()], [H.Variable(1, 9, 8)], 1, 1)
    yield "bad_mode", H.encode("p", [blk(H.Kernel("X", H.BINARY, [(4, 1)]))], [v_ok], 1, 1)
    yield "size_kernel_one_operand", H.encode(
        "p", [blk(H.Kernel("g", H.GET_SIZE, [(H.OUT, 1)]))], [v_ok], 1, 1)
    # A variable without its required size field.
    body = H._vfield(1, 1) + H._vfield(2, 1) + H._vfield(4, 0) + H._vfield(5, 0)
    cfg = (H._bfield(1, b"p") + H._bfield(2, blk().encode()) + H._bfield(3, body)
           + H._vfield(4, 1) + H._vfield(5, 1) + H._vfield(6, 0))
    yield "missing_required", struct.pack("<Q", len(cfg)) + cfg
    # A group (wire type 3) where a field should be.
    cfg = H._bfield(1, b"p") + H._varint((9 << 3) | 3)
    yield "group_wire_type", struct.pack("<Q", len(cfg)) + cfg


@pytest.mark.parametrize("name,image", list(malformed()), ids=[n for n, _ in malformed()])
def test_malformed_files_are_rejected(name, image):
    with pytest.raises(H.RfxError) as e:
        H.Hir(image)
    assert e.value.status == -1


def test_every_truncation_is_rejected_cleanly():
    good = straight_program()
    H.Hir(good)
    for cut in range(len(good)):
        with pytest.raises(H.RfxError):
            H.Hir(good[:cut])


def test_random_byte_flips_never_crash():
    good = bytearray(straight_program())
    rng = np.random.default_rng(1234)
    loaded = 0
    for _ in range(3000):
        img = bytearray(good)
        for pos in rng.integers(8, len(img), size=int(rng.integers(1, 4))):
            img[pos] = int(rng.integers(0, 256))
        try:
            h = H.Hir(bytes(img))
            h.info()
            for b in range(h.info()["blocks"]):
                h.schedule(b)
            loaded += 1
        except H.RfxError:
            pass
    assert loaded > 0  # some flips only change payload bytes


def test_descriptor_text():
    assert H.descriptor("set_split", d=[32], d1=[32], a=[32, 32], dom=1) == \
        b"rfx1 set_split d=32 d1=32 a=32,32 dom=1"
    assert H.descriptor("mgpu_join_gather", d=[40, 24, 24], a=[40], b=[40], c=[24], e=[],
                        dom=1) == b"rfx1 mgpu_join_gather d=40,24,24 a=40 b=40 c=24 e=- dom=1"
    assert H.descriptor("set_count", a=[(64, "p")], b=[(64, "f")], agg="count") == \
        b"rfx1 set_count a=64p b=64f agg=count"
    assert H.descriptor("set_select", d=[16, 48], cmp="0:lt:300", parts=128, nt=128) == \
        b"rfx1 set_select d=16,48 parts=128 nt=128 cmp=0:lt:300"


def launch(name, kind, code, slots):
    L = rfx.lib()
    arr = (ctypes.c_uint64 * max(len(slots), 1))(*slots)
    st = L.rfx_harmony_launch(name.encode(), kind, code, len(code), arr, len(slots))
    return st, L.rfx_last_error().decode()


def test_launch_refusals_before_any_device_work():
    # Each is refused while checking the record, before the default context
    # (and so any device) is touched.
    st, msg = launch("copy3", 1, b"", [1, 1, 0])
    assert st == -5 and "compute kernel family 'copy'" in msg
    st, msg = launch("set_split4", 2, b"", [0] * 4)  # set_split is a compute family
    assert st == -5 and "binary kernel family 'set_split'" in msg
    st, msg = launch("set_split4", 1, b"", [350, 256, 0, 1, 2, 3, 4])
    assert st == -1 and "no rfx1 descriptor" in msg
    st, msg = launch("set_split4", 1, b"rfx1 set_split d=32 d1=33x a=32,32 dom=1", [1] * 7)
    assert st == -1 and "bad 'd1=33x'" in msg
    st, msg = launch("set_split4", 1, b"rfx1 set_merge d=32", [1] * 7)
    assert st == -1 and "descriptor is for 'set_merge'" in msg
    st, msg = launch("set_split4", 1, b"rfx1 set_split d=32 d1=32 a=32,32 dom=1", [1, 1])
    assert st == -1 and "3 or more" in msg
    st, msg = launch("mgpu_join_find_bounds9", 2, b"", [0] * 7)
    assert st == -1 and "no rfx1 descriptor" in msg
    st, msg = launch("set_select2", 1, b"rfx1 set_select d=16 cmp=0:between:3", [1] * 7)
    assert st == -1 and "cmp=0:between:3" in msg
