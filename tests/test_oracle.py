"""Pins the CPU oracle before it is trusted as the GPU checker.

Each oracle function is checked against (a) the reference's own fixtures and
known answers (TPC-H/data/A from TC.tbl, TestJoin's identity join), and (b)
independent formulations that a plausible slip (dropped term, wrong tie
rule, off-by-one, swapped operand) would fail: Python's stable sort, nested
loop joins, collections.Counter multisets, math.fsum, the tuple as a bit
string.
"""
import collections
import math
import os

import numpy as np
import pytest

from redfox_amd import datagen

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64, 16: np.uint64}


def ints(arr, width):
    """Element values as Python ints (16-byte elements: lo | hi << 64)."""
    if width == 16:
        return [int(arr[2 * i]) | (int(arr[2 * i + 1]) << 64) for i in range(arr.size // 2)]
    return [int(x) for x in arr]


def from_ints(vals, width):
    if width == 16:
        out = np.empty(2 * len(vals), dtype=np.uint64)
        out[0::2] = [v & ((1 << 64) - 1) for v in vals]
        out[1::2] = [v >> 64 for v in vals]
        return out
    return np.array(vals, dtype=DT[width])


def rand_elems(width, n, seed, mod=None):
    raw = datagen.splitmix64(seed, (n * width + 7) // 8 + 1).view(np.uint8)[: n * width]
    a = raw.view(DT[width]).copy()
    if mod is not None:
        v = [x % mod for x in ints(a, width)]
        return from_ints(v, width)
    return a


# ---------------------------------------------------------------------------
# Tuple codec
# ---------------------------------------------------------------------------
def test_fixture_A_matches_TC_tbl(orc):
    """TPC-H/data/A is TC.tbl packed as Tuple<32,32> (TC_data_process.cu:80-90)."""
    words = np.fromfile(os.path.join(GOLDEN, "tpch_A.bin"), dtype=np.uint64)
    rows = [tuple(int(x) for x in line.split()) for line in open(os.path.join(GOLDEN, "TC.tbl"))]
    assert len(words) == len(rows) == 5
    bits = np.array([32, 32], dtype=np.uint32)
    for word, (c1, c2) in zip(words, rows):
        w = np.zeros(2, dtype=np.uint64)
        orc.lib().rfxo_tuple_insert(w.ctypes.data, bits.ctypes.data, 2, 0, c1)
        orc.lib().rfxo_tuple_insert(w.ctypes.data, bits.ctypes.data, 2, 1, c2)
        assert int(w[0]) == int(word) and int(w[1]) == 0
        got = np.array([int(word), 0], dtype=np.uint64)
        assert orc.lib().rfxo_tuple_extract(got.ctypes.data, bits.ctypes.data, 2, 0) == c1
        assert orc.lib().rfxo_tuple_extract(got.ctypes.data, bits.ctypes.data, 2, 1) == c2
    assert int(words[0]) == 0x0000000100000006


def test_tuple_closed_forms(orc):
    # field 0 in the most significant bits of the used width (BitsAfter)
    bits = np.array([32, 32, 32], dtype=np.uint32)
    L = orc.lib()
    assert L.rfxo_bits_after(bits.ctypes.data, 3, 0) == 64
    assert L.rfxo_bits_after(bits.ctypes.data, 3, 2) == 0
    assert L.rfxo_tuple_bytes(bits.ctypes.data, 3) == 16
    b2 = np.array([2, 1, 64], dtype=np.uint32)
    assert L.rfxo_bits_after(b2.ctypes.data, 3, 0) == 65
    assert L.rfxo_tuple_bytes(b2.ctypes.data, 3) == 16
    b3 = np.array([3, 2], dtype=np.uint32)
    assert L.rfxo_tuple_bytes(b3.ctypes.data, 2) == 1


# Tuple shapes: config-1 rows, join output, Q1 aggregate output, and words
# of 8, 32, 64 and 128 bits with odd and word-crossing field widths.
SHAPES = [(32, 32), (32, 32, 32), (2, 1, 64), (16, 8, 8), (3, 23, 38), (64, 64), (5, 3),
          (1, 40, 40), (7,)]


def layout_bits(bits):
    """The tuple as a bit string, most significant bit first: padding up
    to the word (the smallest of 8..128 bits holding the fields,
    Tuple.h:30-37), then field 0, field 1, ... (Tuple.h:718-790).  A
    formulation by string slicing, independent of the oracle's shifts."""
    used = sum(bits)
    word = next(w for w in (8, 16, 32, 64, 128) if used <= w)
    spans, at = [], word - used
    for b in bits:
        spans.append((at, at + b))
        at += b
    return word, spans


def test_codec_matches_bit_string_layout(orc):
    """extract / insert of every field, and the split (stripValues /
    stripKeys) and merge (restoreValues) forms, against slices of the
    tuple's bit string."""
    L = orc.lib()
    words = datagen.splitmix64(77, 2 * 48).reshape(48, 2)
    for bits in SHAPES:
        word, spans = layout_bits(bits)
        b = np.array(bits, dtype=np.uint32)
        nf = len(bits)
        assert L.rfxo_tuple_bytes(b.ctypes.data, nf) == word // 8
        for w0 in words:
            val = (int(w0[0]) | (int(w0[1]) << 64)) & ((1 << word) - 1)
            s = format(val, f"0{word}b")
            w = from_ints([val], 16)
            for f, (lo, hi) in enumerate(spans):
                assert L.rfxo_tuple_extract(w.ctypes.data, b.ctypes.data, nf, f) == \
                    int(s[lo:hi], 2) % (1 << 64), (bits, f)
                if bits[f] <= 64:
                    v = int(datagen.splitmix64(val & 0xFFFF, 1)[0])
                    a = w.copy()
                    L.rfxo_tuple_insert(a.ctypes.data, b.ctypes.data, nf, f, v)
                    fb = format(v % (1 << bits[f]), f"0{bits[f]}b")
                    assert ints(a, 16)[0] == int(s[:lo] + fb + s[hi:], 2), (bits, f)
            for keys in range(1, nf):
                cut = spans[keys - 1][1]  # the key fields end here
                shift = int(L.rfxo_bits_after(b.ctypes.data, nf, keys - 1))
                assert shift == word - cut
                # split of the tuple (stripValues, stripKeys) and merge
                # (restoreValues | values) by the oracle, against slices.
                w8 = from_ints([val], 16)
                kb, vb = (cut + 7) // 8, (word - cut + 7) // 8
                kb = next(x for x in (1, 2, 4, 8, 16) if x >= kb)
                vb = next(x for x in (1, 2, 4, 8, 16) if x >= vb)
                if word == 128:
                    k, v = orc.split(w8, 16, kb, vb, shift)
                else:
                    k, v = orc.split(w8[:1].view(np.uint64).astype(DT[word // 8]), word // 8,
                                     kb, vb, shift)
                assert ints(k, kb)[0] == int(s[:cut], 2)
                assert ints(v, vb)[0] == int(s[cut:], 2)
                back = orc.merge(word // 8, k, kb, v, vb, shift)
                assert ints(back, word // 8)[0] == val


def test_split_merge_roundtrip(orc):
    """split (stripValues/stripKeys) then merge (restoreValues | value)
    restores every Tuple<32,32> row."""
    rows = datagen.splitmix64(5, 1000)
    k, v = orc.split(rows, 8, 4, 4, 32)
    assert np.array_equal(k, (rows >> np.uint64(32)).astype(np.uint32))
    assert np.array_equal(v, (rows & np.uint64(0xFFFFFFFF)).astype(np.uint32))
    back = orc.merge(8, k, 4, v, 4, 32)
    assert np.array_equal(back, rows)


def test_project_swaps_fields(orc):
    rows = datagen.splitmix64(6, 100)
    out = orc.project(8, rows, 8, [0, 32], [32, 32], [32, 0])
    want = (rows << np.uint64(32)) | (rows >> np.uint64(32))
    assert np.array_equal(out, want)


# ---------------------------------------------------------------------------
# Sort
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kb,bits", [(1, 3), (1, 8), (2, 14), (4, 23), (4, 32), (8, 37),
                                     (8, 64), (16, 100), (16, 128)])
def test_sort_is_stable_sort_by_low_bits(orc, kb, bits):
    n = 2000
    keys = rand_elems(kb, n, 3 + bits)
    vals = np.arange(n, dtype=np.uint32)
    k, v = orc.sort_pairs(keys, vals, kb, bits, 4)
    iv = ints(keys, kb)
    mask = (1 << bits) - 1
    order = sorted(range(n), key=lambda i: iv[i] & mask)
    assert list(v) == order
    assert ints(k, kb) == [iv[i] for i in order]


def test_sort_value_widths_follow_keys(orc):
    n = 500
    keys = rand_elems(4, n, 1, mod=37)
    for vb in (1, 2, 8, 16):
        vals = rand_elems(vb, n, 2)
        k, v = orc.sort_pairs(keys, vals, 4, 32, vb)
        order = sorted(range(n), key=lambda i: int(keys[i]))
        assert ints(v, vb) == [ints(vals, vb)[i] for i in order]


# ---------------------------------------------------------------------------
# Join
# ---------------------------------------------------------------------------
def nested_loop_join(a, b):
    return [(i, j) for i, x in enumerate(a) for j, y in enumerate(b) if x == y]


@pytest.mark.parametrize("kb", [1, 2, 4, 8, 16])
def test_join_matches_nested_loops(orc, kb):
    a = from_ints(sorted(ints(rand_elems(kb, 300, 10 + kb, mod=40), kb)), kb)
    b = from_ints(sorted(ints(rand_elems(kb, 400, 20 + kb, mod=40), kb)), kb)
    lb, lc, total = orc.join_find_bounds(a, b, kb)
    li, ri = orc.join_main(lb, lc, total)
    want = nested_loop_join(ints(a, kb), ints(b, kb))
    assert list(zip(li.tolist(), ri.tolist())) == want
    ai, bi = orc.cpujoin(a, b, kb)
    assert np.array_equal(ai, li) and np.array_equal(bi, ri)
    av, bv = ints(a, kb), ints(b, kb)
    for i, x in enumerate(av):
        assert lb[i] == sum(1 for y in bv if y < x)


def test_join_identity_known_answer(orc):
    """TestJoin.cu:142-156: identity join of 2^15 u16 keys gives
    (i << 16) | i."""
    n = 1 << 15
    keys = np.arange(n, dtype=np.uint16)
    lb, lc, total = orc.join_find_bounds(keys, keys, 2)
    assert total == n
    li, ri = orc.join_main(lb, lc, total)
    out = orc.join_gather(4, keys, 2, None, 0, keys, 2, li, ri, 16, 0)
    i = np.arange(n, dtype=np.uint32)
    assert np.array_equal(out, (i << np.uint32(16)) | i)


def test_join_gather_closed_form(orc):
    """gather_value_value: restoreValues(key) | lval << bits(rval) | rval."""
    lk = np.array([1, 2], dtype=np.uint32)
    lv = np.array([10, 20], dtype=np.uint32)
    rv = np.array([100, 200, 300], dtype=np.uint32)
    li = np.array([0, 1, 1], dtype=np.int32)
    ri = np.array([0, 1, 2], dtype=np.int32)
    out = orc.join_gather(16, lk, 4, lv, 4, rv, 4, li, ri, 64, 32).reshape(-1, 2)
    for g in range(3):
        val = int(out[g, 0]) | (int(out[g, 1]) << 64)
        assert val == (int(lk[li[g]]) << 64) | (int(lv[li[g]]) << 32) | int(rv[ri[g]])


def test_join_empty_sides(orc):
    e = np.zeros(0, dtype=np.uint32)
    a = np.arange(5, dtype=np.uint32)
    assert orc.join_find_bounds(e, a, 4)[2] == 0
    lb, lc, t = orc.join_find_bounds(a, e, 4)
    assert t == 0 and not lb.any() and not lc.any()


# ---------------------------------------------------------------------------
# Reduce-by-key
# ---------------------------------------------------------------------------
def test_reduce_by_key_matches_groupby(orc):
    keys = np.sort(rand_elems(4, 5000, 9, mod=97))
    vals = datagen.uniform01(9, 5000) * 1000.0
    ok, ov = orc.reduce_by_key_f64(keys, vals, 4, 0)
    groups = collections.OrderedDict()
    for k, v in zip(keys.tolist(), vals.tolist()):
        groups.setdefault(k, []).append(v)
    assert ok.tolist() == list(groups)
    for got, vs in zip(ov.tolist(), groups.values()):
        assert got == math.fsum(vs)  # compensated sum: correctly rounded here
    omin = orc.reduce_by_key_f64(keys, vals, 4, 1)[1]
    omax = orc.reduce_by_key_f64(keys, vals, 4, 2)[1]
    assert omin.tolist() == [min(v) for v in groups.values()]
    assert omax.tolist() == [max(v) for v in groups.values()]
    cnt = orc.reduce_by_key_u32(keys, np.ones(5000, dtype=np.uint32), 4)[1]
    assert cnt.tolist() == [len(v) for v in groups.values()]


def test_reduce_by_key_runs_not_sets(orc):
    """Runs of equal adjacent keys, as thrust::reduce_by_key: a key that
    reappears later starts a new group."""
    keys = np.array([1, 1, 2, 1, 1], dtype=np.uint8)
    vals = np.array([1.0, 2.0, 3.0, 4.0, 5.0])
    ok, ov = orc.reduce_by_key_f64(keys, vals, 1, 0)
    assert ok.tolist() == [1, 2, 1] and ov.tolist() == [3.0, 3.0, 9.0]


def test_single_reductions(orc):
    v = datagen.uniform01(3, 10000) * 1e6
    assert orc.total_double(v) == math.fsum(v.tolist())
    u = (datagen.u32(4, 1000) >> np.uint32(8))
    assert orc.count_u32(u) == int(u.astype(np.uint64).sum()) % (1 << 32)
    bits = datagen.splitmix64(5, 1000)
    assert orc.max_double_bits(bits) == int(bits.max())


# ---------------------------------------------------------------------------
# Set operations
# ---------------------------------------------------------------------------
def multiset(op, a, b):
    ca, cb = collections.Counter(a), collections.Counter(b)
    if op == 0:
        c = ca | cb   # max
    elif op == 1:
        c = ca - cb   # max(m - n, 0)
    else:
        c = ca & cb   # min
    return sorted(c.elements())


@pytest.mark.parametrize("kb", [1, 2, 4, 8, 16])
@pytest.mark.parametrize("op", [0, 1, 2])
def test_set_ops_multiset_semantics(orc, kb, op):
    a = sorted(ints(rand_elems(kb, 700, 50 + kb, mod=300), kb))
    b = sorted(ints(rand_elems(kb, 500, 60 + kb, mod=300), kb))
    got = orc.set_op(op, from_ints(a, kb), from_ints(b, kb), kb)
    assert ints(got, kb) == multiset(op, a, b)


def test_set_ops_compare_unsigned(orc):
    """int64 sets compare as u64 (Union.cu:106-118): negatives sort last."""
    a = np.array([1, 5, (1 << 63), (1 << 64) - 1], dtype=np.uint64)
    b = np.array([5, 7, (1 << 63)], dtype=np.uint64)
    assert orc.set_op(0, a, b, 8).tolist() == [1, 5, 7, 1 << 63, (1 << 64) - 1]
    assert orc.set_op(1, a, b, 8).tolist() == [1, (1 << 64) - 1]
    assert orc.set_op(2, a, b, 8).tolist() == [5, 1 << 63]


def test_unique(orc):
    a = np.array([1, 1, 2, 3, 3, 3, 9], dtype=np.uint32)
    assert orc.unique(a, 4).tolist() == [1, 2, 3, 9]
    assert orc.unique(np.zeros(0, dtype=np.uint32), 4).size == 0


# ---------------------------------------------------------------------------
# Select
# ---------------------------------------------------------------------------
def test_select_is_stable_filter(orc):
    rows = datagen.splitmix64(8, 3000)
    f = (rows >> np.uint64(40)) & np.uint64(0xFFFF)
    got = orc.select(rows, 8, 40, 16, 3, 30000)  # le
    assert np.array_equal(got, rows[f <= 30000])
    got = orc.select(rows, 8, 40, 16, 0, int(f[17]))  # eq
    assert np.array_equal(got, rows[f == f[17]])


def test_select_signed_and_double(orc):
    vals = np.array([-5, -1, 0, 3, 7], dtype=np.int32).view(np.uint32).astype(np.uint64)
    got = orc.select(vals, 8, 0, 32, 6, 0)  # slt 0
    assert got.tolist() == vals[:2].tolist()
    d = np.array([1.0, 2.0, 2.0000001, 3.0]).view(np.uint64)
    two = int(np.array([2.0]).view(np.uint64)[0])
    # ltdouble: (two - one) > 1e-6 ; ledouble: (two - one) > -1e-6
    assert orc.select(d, 8, 0, 64, 10, two).view(np.float64).tolist() == [1.0]
    assert orc.select(d, 8, 0, 64, 11, two).view(np.float64).tolist() == [1.0, 2.0, 2.0000001]


# ---------------------------------------------------------------------------
# Q1
# ---------------------------------------------------------------------------
def test_q1_matches_numpy_formulation(orc):
    n = 200_000
    cols = datagen.lineitem(datagen.BASE_SEED + 2, n)
    sd, fl, st, qty, price, disc, tax = cols
    got = orc.q1(sd, fl, st, qty, price, disc, tax, 2436)
    keep = sd <= 2436
    g = (fl.astype(np.int64) * 2 + st)[keep]
    dp = price * (1.0 - disc)
    ch = dp * (1.0 + tax)
    for grp in range(6):
        m = g == grp
        for col, arr in enumerate([qty.astype(np.float64), price, dp, ch, disc]):
            want = math.fsum(arr[keep][m].tolist())
            assert got[grp, col] == pytest.approx(want, rel=1e-15)
        assert got[grp, 5] == m.sum()
    assert got[:, 5].sum() == keep.sum()
