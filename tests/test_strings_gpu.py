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
