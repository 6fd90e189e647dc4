"""ctypes binding of librfx_cpu.so, the reported CPU baseline: the
reference's operator algorithms restated in C++ (rfx_cpu.cpp) and run with
OpenMP on the host's cores.

TEST INFRASTRUCTURE ONLY: used by bench.py's cpu_baseline leg and by
tests/test_cpu_baseline.py, which checks it against the serial oracle.
Each timing call returns the median of `reps` runs after one warm-up, in
seconds (std::chrono::steady_clock), with input copies outside the timed
region.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import c_double, c_int, c_int32, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "librfx_cpu.so")
P, U64, I = c_void_p, c_uint64, c_int

_SIG = {
    "rfxc_threads": (I, []),
    "rfxc_set_threads": (None, [I]),
    "rfxc_sort_pairs": (c_double, [P, P, P, P, U64, I, I]),
    "rfxc_join": (c_double, [P, U64, P, P, P, U64, P, P, U64, I]),
    "rfxc_reduce_sum_count": (c_double, [P, P, P, P, P, P, U64, I]),
    "rfxc_q1": (c_double, [P, P, P, P, P, P, P, P, U64, c_int32, I]),
    "rfxc_select16": (c_double, [P, P, P, U64, c_uint32, I]),
    "rfxc_mapfilter": (c_double, [P, P, P, P, P, P, P, U64, c_int32, I]),
    "rfxc_set_op": (c_double, [P, P, P, U64, P, U64, I, I]),
}
_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} missing: run `make -C oracle`")
        _lib = ctypes.CDLL(LIB)
        for name, (res, args) in _SIG.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


class threads:
    """Context manager running the baseline on `n` threads (0: all the
    OpenMP default allows)."""

    def __init__(self, n: int):
        self.n = n

    def __enter__(self):
        self.saved = lib().rfxc_threads()
        if self.n:
            lib().rfxc_set_threads(self.n)
        return lib().rfxc_threads()

    def __exit__(self, *exc):
        lib().rfxc_set_threads(self.saved)


def sort_pairs(keys: np.ndarray, vals: np.ndarray, reps: int = 5):
    k, v = np.empty_like(keys), np.empty_like(vals)
    dt = lib().rfxc_sort_pairs(_p(k), _p(v), _p(keys), _p(vals), len(keys), keys.itemsize, reps)
    return dt, k, v


def join(rk, rv, sk, sv, reps: int = 5):
    """Join core over sorted u32 columns -> 16-byte rows (2 u64 each)."""
    lo = np.searchsorted(sk, rk, "left")
    hi = np.searchsorted(sk, rk, "right")
    cap = int((hi - lo).sum())
    out = np.empty(2 * max(cap, 1), dtype=np.uint64)
    tot = np.zeros(1, dtype=np.uint64)
    dt = lib().rfxc_join(_p(out), cap, _p(tot), _p(rk), _p(rv), len(rk), _p(sk), _p(sv), len(sk),
                         reps)
    return dt, out[: 2 * int(tot[0])]


def reduce_sum_count(keys, vals, reps: int = 5):
    n = len(keys)
    ok, os_, oc = np.empty(n, np.uint32), np.empty(n, np.float64), np.empty(n, np.uint32)
    runs = np.zeros(1, dtype=np.uint64)
    dt = lib().rfxc_reduce_sum_count(_p(ok), _p(os_), _p(oc), _p(runs), _p(keys), _p(vals), n,
                                     reps)
    r = int(runs[0])
    return dt, ok[:r], os_[:r], oc[:r]


def q1(shipdate, flag, status, qty, price, disc, tax, cutoff, reps: int = 5):
    out = np.zeros(48, dtype=np.float64)
    dt = lib().rfxc_q1(_p(out), _p(shipdate), _p(flag), _p(status), _p(qty), _p(price), _p(disc),
                       _p(tax), len(shipdate), cutoff, reps)
    return dt, out.reshape(6, 8)


def select16(tuples: np.ndarray, k: int, reps: int = 5):
    n = len(tuples) // 2
    out = np.empty_like(tuples)
    kept = np.zeros(1, dtype=np.uint64)
    dt = lib().rfxc_select16(_p(out), _p(kept), _p(tuples), n, k, reps)
    return dt, out[: 2 * int(kept[0])]


def mapfilter(shipdate, price, disc, tax, cutoff, reps: int = 5):
    n = len(shipdate)
    dp, ch = np.empty(n), np.empty(n)
    kept = np.zeros(1, dtype=np.uint64)
    dt = lib().rfxc_mapfilter(_p(dp), _p(ch), _p(kept), _p(shipdate), _p(price), _p(disc),
                              _p(tax), n, cutoff, reps)
    m = int(kept[0])
    return dt, dp[:m], ch[:m]


def set_op(op: int, a: np.ndarray, b: np.ndarray, reps: int = 5):
    out = np.empty(len(a) + len(b), dtype=np.uint64)
    nout = np.zeros(1, dtype=np.uint64)
    dt = lib().rfxc_set_op(_p(out), _p(nout), _p(a), len(a), _p(b), len(b), op, reps)
    return dt, out[: int(nout[0])]
