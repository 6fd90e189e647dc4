"""Seeded synthetic inputs for the RA workloads (numpy, host side).

Every generator is a splitmix64 stream (Steele, Lea & Flood 2014): value i of
stream `seed` is mix(seed + (i + 1) * 0x9E3779B97F4A7C15).  The device
generators rfx_gen_u32 / rfx_gen_u64 compute the same values, so data made on
the GPU and on the CPU agree bit for bit.
"""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
BASE_SEED = 0x52464F58000000CC  # + workload index


def splitmix64(seed: int, n: int, start: int = 0) -> np.ndarray:
    """Values start .. start+n-1 of splitmix64 stream `seed` as uint64."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def u32(seed: int, n: int) -> np.ndarray:
    """High 32 bits of each stream value (what rfx_gen_u32 writes)."""
    return (splitmix64(seed, n) >> np.uint64(32)).astype(np.uint32)


def uniform01(seed: int, n: int) -> np.ndarray:
    """Doubles in [0, 1) from the top 53 bits."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def zipf_ranks(seed: int, n: int, s: float, n_max: int) -> np.ndarray:
    """Zipf(s) ranks in [1, n_max] by rejection-inversion (Hormann &
    Derflinger 1996), vectorised; rejected draws are redrawn from the next
    block of the stream."""
    out = np.empty(n, dtype=np.uint64)
    filled = 0
    q = 1.0 - s

    def H(x):
        return (np.exp(q * np.log(x)) - 1.0) / q

    def Hinv(y):
        return np.exp(np.log1p(q * y) / q)

    h_x1 = H(1.5) - 1.0
    h_n = H(n_max + 0.5)
    s_const = 2.0 - Hinv(H(2.5) - np.exp(-s * np.log(2.0)))
    block = 0
    while filled < n:
        need = n - filled
        m = max(1024, int(need * 1.25))
        u = uniform01(seed + 0x1000003 * block, m)
        block += 1
        ur = h_n + u * (h_x1 - h_n)
        x = Hinv(ur)
        k = np.floor(x + 0.5)
        k = np.clip(k, 1, n_max)
        accept = (k - x <= s_const) | (ur >= H(k + 0.5) - np.exp(-s * np.log(k)))
        got = k[accept].astype(np.uint64)[:need]
        out[filled:filled + got.size] = got
        filled += got.size
    return out


def mix64(x: np.ndarray) -> np.ndarray:
    """The splitmix64 finaliser applied to each element (a bijection)."""
    with np.errstate(over="ignore"):
        z = x.astype(np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def join_tables(seed: int, n_r: int, n_s: int, key_bits: int = 16):
    """Join inputs: R and S rows (key in [0, 2^key_bits), u32 payload)."""
    r = splitmix64(seed, n_r)
    s = splitmix64(seed + 1, n_s)
    rk = (r >> np.uint64(64 - key_bits)).astype(np.uint32)
    rv = (r & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    sk = (s >> np.uint64(64 - key_bits)).astype(np.uint32)
    sv = (s & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    return rk, rv, sk, sv


def lineitem(seed: int, n: int):
    """Q1 columns: shipdate i32 U[0,2526), returnflag u8 U{0,1,2},
    linestatus u8 U{0,1}, qty i32 U[1,50], price f64 = qty*U{90000..200000}/100,
    disc f64 U{0..10}/100, tax f64 U{0..8}/100."""
    a = splitmix64(seed, n)
    b = splitmix64(seed + 1, n)
    shipdate = ((a >> np.uint64(32)) % np.uint64(2526)).astype(np.int32)
    flag = ((a >> np.uint64(16)) & np.uint64(0xFFFF)) % np.uint64(3)
    status = ((a >> np.uint64(8)) & np.uint64(0xFF)) % np.uint64(2)
    qty = ((a & np.uint64(0xFF)) % np.uint64(50) + np.uint64(1)).astype(np.int32)
    cents = ((b >> np.uint64(32)) % np.uint64(110001)) + np.uint64(90000)
    price = qty.astype(np.float64) * (cents.astype(np.float64) / 100.0)
    disc = (((b >> np.uint64(16)) & np.uint64(0xFFFF)) % np.uint64(11)).astype(np.float64) / 100.0
    tax = ((b & np.uint64(0xFFFF)) % np.uint64(9)).astype(np.float64) / 100.0
    return (shipdate, flag.astype(np.uint8), status.astype(np.uint8), qty, price, disc, tax)


def lineitem_keyed(n_orders: int, seed: int):
    """Lineitem as column relations keyed by (LINENUMBER, ORDER), the form
    the reference's Q1 plan reads (q1.man.cpp): 1-7 lines per order,
    sorted by the keys.  Returns (linenumber u8, order u32, shipdate i64
    seconds over 1992-1998, returnflag u8 0-2, linestatus u8 0-1, quantity,
    extendedprice, discount, tax as f64)."""
    rng = np.random.default_rng(seed)
    lines = rng.integers(1, 8, n_orders)
    order = np.repeat(np.arange(1, n_orders + 1, dtype=np.uint32), lines)
    ln = (np.arange(len(order)) - np.repeat(np.cumsum(lines) - lines, lines) + 1).astype(np.uint8)
    n = len(order)
    ship = (rng.integers(8035, 8035 + 2557, n) * 86400).astype(np.int64)
    rf = rng.integers(0, 3, n).astype(np.uint8)
    ls = rng.integers(0, 2, n).astype(np.uint8)
    qty = rng.integers(1, 51, n).astype(np.float64)
    price = rng.random(n) * 1e4
    disc = rng.integers(0, 11, n) / 100.0
    tax = rng.integers(0, 9, n) / 100.0
    o = np.lexsort((order, ln))
    return [a[o] for a in (ln, order, ship, rf, ls, qty, price, disc, tax)]


def _ascii_digits(v: np.ndarray, width: int) -> np.ndarray:
    """Zero-padded decimal digits of non-negative integers, (n, width) ASCII."""
    out = np.empty((len(v), width), dtype=np.uint8)
    x = v.astype(np.uint64).copy()
    for i in range(width - 1, -1, -1):
        out[:, i] = 48 + (x % np.uint64(10)).astype(np.uint8)
        x //= np.uint64(10)
    return out


def lineitem_tbl(n_orders: int, seed: int):
    """dbgen-style lineitem.tbl text ('|' after every field, rows in
    (orderkey, linenumber) order, fixed-width numbers) and the values it
    encodes: (text, order u32, linenumber u8, quantity, extendedprice,
    discount, tax f64, returnflag, linestatus u8 chars, shipdate i64
    seconds).  Decimal values are the correctly rounded value of their text
    (cents / 100)."""
    rng = np.random.default_rng(seed)
    lines = rng.integers(1, 8, n_orders)
    order = np.repeat(np.arange(1, n_orders + 1, dtype=np.uint32), lines)
    ln = (np.arange(len(order)) - np.repeat(np.cumsum(lines) - lines, lines) + 1).astype(np.uint8)
    n = len(order)
    qty = rng.integers(1, 51, n)
    cents = rng.integers(90_000, 10_500_000, n)
    disc = rng.integers(0, 11, n)
    tax = rng.integers(0, 9, n)
    rf = np.frombuffer(b"ANR", dtype=np.uint8)[rng.integers(0, 3, n)]
    ls = np.frombuffer(b"FO", dtype=np.uint8)[rng.integers(0, 2, n)]
    day = rng.integers(8035, 8035 + 2557, n)
    dates = np.datetime_as_string(np.datetime64("1970-01-01") + day.astype("timedelta64[D]"))
    date_bytes = np.char.encode(dates).view(np.uint8).reshape(n, 10)
    bar = np.full((n, 1), ord("|"), dtype=np.uint8)
    dot = np.full((n, 1), ord("."), dtype=np.uint8)
    zero = np.full((n, 1), ord("0"), dtype=np.uint8)
    tail = np.frombuffer(b"DELIVER IN PERSON|TRUCK|carefully final deposits|", np.uint8)
    parts = [_ascii_digits(order, 8), bar, _ascii_digits(rng.integers(1, 200_000, n), 6), bar,
             _ascii_digits(rng.integers(1, 10_000, n), 5), bar, _ascii_digits(ln, 1), bar,
             _ascii_digits(qty, 2), dot, zero, zero, bar,
             _ascii_digits(cents // 100, 6), dot, _ascii_digits(cents % 100, 2), bar,
             zero, dot, _ascii_digits(disc, 2), bar, zero, dot, _ascii_digits(tax, 2), bar,
             rf[:, None], bar, ls[:, None], bar, date_bytes, bar, date_bytes, bar, date_bytes, bar,
             np.broadcast_to(tail, (n, len(tail))), np.full((n, 1), 10, dtype=np.uint8)]
    text = np.concatenate(parts, axis=1).tobytes()
    return (text, order, ln, qty.astype(np.float64), cents / 100.0, disc / 100.0, tax / 100.0,
            rf, ls, (day * 86400).astype(np.int64))
