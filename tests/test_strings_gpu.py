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
