"""Text loading without a GPU: the oracle's parser pinned against the
reference's own loader output (TPC-H/data/A is TC.tbl parsed with atoi and
packed as Tuple<32,32>, TC_data_process.cu:40-90), Python's correctly
rounded float() for decimals, datetime for dates, C atoi known answers,
and the dbgen-style generator's values."""
import datetime
import os

import numpy as np
import pytest

import oracle
from redfox_amd import datagen, loaders

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_tc_tbl_parses_to_the_reference_fixture():
    text = open(os.path.join(GOLDEN, "TC.tbl"), "rb").read()
    (c1, c2), short = oracle.parse_tbl(text, ["i32", "i32"], b"\t")
    assert short == 0
    words = np.fromfile(os.path.join(GOLDEN, "tpch_A.bin"), dtype=np.uint64)
    bits = np.array([32, 32], dtype=np.uint32)
    for a, b, word in zip(c1, c2, words):
        w = np.zeros(2, dtype=np.uint64)
        oracle.lib().rfxo_tuple_insert(w.ctypes.data, bits.ctypes.data, 2, 0, int(a))
        oracle.lib().rfxo_tuple_insert(w.ctypes.data, bits.ctypes.data, 2, 1, int(b))
        assert int(w[0]) == int(word)


def test_atoi_and_decimal_known_answers():
    assert [oracle._atoi(s) for s in (b"12", b"  -7x", b"+3", b"", b"-", b"9" * 3, b" 4 5")] == \
        [12, -7, 3, 0, 0, 999, 4]
    rng = np.random.default_rng(2)
    for _ in range(3000):
        digits = int(rng.integers(1, 16))
        frac = int(rng.integers(0, min(digits, 8) + 1))
        m = int(rng.integers(0, 10 ** digits))
        s = str(m).rjust(digits, "0")
        txt = (s[:digits - frac] or "0") + ("." + s[digits - frac:] if frac else "")
        if rng.random() < 0.3:
            txt = "-" + txt
        assert oracle._decimal(txt.encode()) == float(txt), txt
        assert oracle._decimal(txt.encode() + b"|junk") == float(txt)


def test_dates_match_datetime():
    day = datetime.date(1992, 1, 1)
    rows = []
    while day.year < 1999:
        rows.append(day)
        day += datetime.timedelta(days=3)
    text = b"".join(f"{d.isoformat()}|x\n".encode() for d in rows)
    (sec, ch), short = oracle.parse_tbl(text, ["date", "char"])
    assert short == 0 and np.all(ch == ord("x"))
    assert sec.tolist() == [(d - datetime.date(1970, 1, 1)).days * 86400 for d in rows]


def test_rows_fields_and_line_ends():
    text = b"1|a|2.5|\r\n-3|b\n\n7|c|0.125|extra|\n8"
    cols, short = oracle.parse_tbl(text, ["i64", "char", "f64"])
    assert [c.tolist() for c in cols] == [[1, -3, 0, 7, 8], [97, 98, 0, 99, 0],
                                          [2.5, 0.0, 0.0, 0.125, 0.0]]
    assert short == 3  # "-3|b", the empty row and "8"
    assert oracle.parse_tbl(b"", ["i32"])[0][0].size == 0


def test_generated_lineitem_text_encodes_its_values():
    text, *vals = datagen.lineitem_tbl(500, 4)
    cols, short = oracle.parse_tbl(text, loaders.LINEITEM_SPEC)
    assert short == 0
    for c, v in zip(cols, vals):
        assert np.array_equal(c, v)


def test_unknown_field_type_is_refused():
    with pytest.raises(ValueError):
        loaders.parse_tbl(None, b"1\n", ["int"])


def _random_strings(rng, n):
    alphabet = [b"a", b"b", b"ab", b"\x80", b"\xff", b"z", b""]
    out = []
    for _ in range(n):
        k = int(rng.integers(0, 4))
        out.append(b"".join(alphabet[int(i)] for i in rng.integers(0, len(alphabet), k)))
    return out


def test_string_order_two_ways():
    """string_key's order (signed bytes, prefixes first) equals sorting with
    the compare_string2 transliteration as the comparator."""
    import functools
    rng = np.random.default_rng(7)
    xs = _random_strings(rng, 300) + [b"a\0zz", b"a"]
    by_key = sorted(xs, key=oracle.string_key)
    by_cmp = sorted(xs, key=functools.cmp_to_key(
        lambda a, b: -1 if oracle.string_less(a, b) else 1 if oracle.string_less(b, a) else 0))
    assert [oracle.string_key(x) for x in by_key] == [oracle.string_key(x) for x in by_cmp]
    assert oracle.string_less(b"\xff", b"a") and oracle.string_less(b"", b"\x80")
    assert not oracle.string_less(b"a\0b", b"a") and not oracle.string_less(b"a", b"a\0b")


def test_find_bounds_string_against_counting():
    rng = np.random.default_rng(8)
    left = sorted(_random_strings(rng, 80), key=oracle.string_key)
    right = sorted(_random_strings(rng, 120), key=oracle.string_key)
    lb, lc, total = oracle.find_bounds_string(left, right)
    run = 0
    for i, x in enumerate(left):
        below = sum(oracle.string_less(r, x) for r in right)
        equal = sum(not oracle.string_less(r, x) and not oracle.string_less(x, r) for r in right)
        assert lb[i] == below and lc[i] == run
        run += equal
    assert total == run


def test_string_fields_offsets():
    cols, short = oracle.parse_tbl(b"ab|cde|1\nx||2\nlast|z", ["str", "str", "i32"])
    assert cols[0] == [0, 9, 14] and cols[1] == [3, 11, None] and short == 1
