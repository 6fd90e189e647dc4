"""The served form: mixed requests over device-resident relations.

rfx_server (serve.hip) does the serving: admission by HBM budget, one
stream per slot, whole plans enqueued without host synchronisation, and
completion polling; this module builds the resident relations on the
device, draws request streams and summarises latencies.

Workload: relations of floor(10^U[5,8]) rows (the range is a parameter),
requests 30% select-project, 30% join-aggregate, 20% three-way join chain,
20% set operations (union / intersection / difference, equally likely) on
sorted unique u64 sets, arriving as a Poisson process of rate lambda.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_uint32, c_uint64, c_void_p
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import rfx as R

PLAN_SELECT_PROJECT, PLAN_JOIN_AGG, PLAN_JOIN3, PLAN_SET_OP = 0, 1, 2, 3
PLAN_NAMES = {0: "select_project", 1: "join_agg", 2: "join3", 3: "set_op"}
MODE_PLAIN, MODE_CONTINUOUS = 0, 1
SET_UNION, SET_DIFFERENCE, SET_INTERSECTION = 0, 1, 2
CMP_LT = 2
MIX = (0.3, 0.3, 0.2, 0.2)   # select-project, join-agg, join3, set ops
GRP_BITS = 6
GOLDEN = 0x9E3779B97F4A7C15
M64 = (1 << 64) - 1


class Relation(ctypes.Structure):
    _fields_ = [("col", c_void_p * 2), ("rows", c_uint64), ("key_bits", c_uint32),
                ("val_bits", c_uint32)]


class Request(ctypes.Structure):
    _fields_ = [("schema", c_uint32), ("param", c_uint32), ("constant", c_uint64),
                ("arrival_ns", c_uint64), ("sink", c_void_p), ("sink_bytes", c_uint64),
                ("admit_ns", c_uint64), ("done_ns", c_uint64), ("mem_bytes", c_uint64),
                ("rows", c_uint64), ("digest", c_uint64 * 3), ("status", ctypes.c_int32),
                ("reserved", c_uint32)]


class ServeStats(ctypes.Structure):
    _fields_ = [("wall_ns", c_uint64), ("peak_bytes", c_uint64), ("budget_bytes", c_uint64),
                ("blocked_on_memory", c_uint64), ("peak_inflight", c_uint32),
                ("slots", c_uint32)]


class Server:
    """An rfx_server: `slots` streams and a budget of HBM for requests."""

    def __init__(self, device: int = 0, slots: int = 4, budget_bytes: int = 8 << 30):
        h = c_void_p()
        R._check("rfx_server_create", R.lib().rfx_server_create(byref(h), device, slots,
                                                                budget_bytes))
        self.handle = h.value
        self.slots = slots

    def add_schema(self, plan: int, rels: Sequence[Relation]) -> int:
        arr = (Relation * len(rels))(*rels)
        sid = c_uint32()
        R._check("rfx_server_add_schema", R.lib().rfx_server_add_schema(
            self.handle, plan, arr, len(rels), byref(sid)))
        return int(sid.value)

    def request_bytes(self, req: Request) -> int:
        b = c_uint64()
        R._check("rfx_server_request_bytes",
                 R.lib().rfx_server_request_bytes(self.handle, byref(req), byref(b)))
        return int(b.value)

    def run(self, reqs, mode: int = MODE_CONTINUOUS) -> ServeStats:
        st = ServeStats()
        R._check("rfx_server_run", R.lib().rfx_server_run(
            self.handle, reqs, len(reqs), mode, byref(st)))
        return st

    def close(self) -> None:
        if getattr(self, "handle", None):
            R.lib().rfx_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bits_for(n: int) -> int:
    """Radix bits covering keys 0..n-1."""
    return max(1, int(n - 1).bit_length())


def permutation_u32(ctx: R.Context, n: int, seed: int) -> R.DeviceBuffer:
    """A random permutation of 0..n-1 (row indices sorted by random keys)."""
    keys = R.DeviceBuffer(8 * n)
    perm = R.DeviceBuffer(4 * n)
    ctx.gen_u64(keys, n, seed)
    ctx.iota_u32(perm, n)
    ctx.sort_pairs(keys, perm, n, 8, 64, 4)
    ctx.synchronize()
    keys.close()
    return perm


def masked_u32(ctx: R.Context, n: int, seed: int, bits: int) -> R.DeviceBuffer:
    raw = R.DeviceBuffer(4 * n)
    out = R.DeviceBuffer(4 * n)
    ctx.gen_u32(raw, n, seed)
    ctx.project(out, 4, raw, 4, n, [0], [bits], [0])
    ctx.synchronize()
    raw.close()
    return out


def zipf_fk(ctx: R.Context, n: int, target_keys: R.DeviceBuffer, nd: int, seed: int,
            s: float = 1.1, measure: Optional[R.DeviceBuffer] = None):
    """Foreign keys drawn Zipf(s) over the rows of a key column, and a
    measure U{0..100000}/100 (the batched workload's fact generator)."""
    fk = R.DeviceBuffer(4 * n)
    m = measure if measure is not None else R.DeviceBuffer(8 * n)
    row_off = R.DeviceBuffer.from_numpy(np.array([0, n], dtype=np.uint64))
    dim_off = R.DeviceBuffer.from_numpy(np.array([0, nd], dtype=np.uint64))
    ctx.gen_fk_batch(fk, m, row_off, dim_off, target_keys, 1, n, seed, s)
    ctx.synchronize()
    return fk, m


def sorted_set_u64(ctx: R.Context, n: int, seed: int, first: int = 0):
    """Sorted unique u64: splitmix64 stream values first .. first+n-1."""
    buf = R.DeviceBuffer(8 * n)
    ctx.gen_u64(buf, n, (seed + first * GOLDEN) & M64)
    ctx.sort_pairs(buf, None, n, 8, 64, 0)
    cnt = R.u64_cell()
    ctx.unique(buf, cnt, buf, n, 8)
    ctx.synchronize()
    return buf, R.read_u64(cnt)


class Catalog:
    """Device-resident relations registered as schemas with a server.

    per_kind schemas of each plan; each relation has floor(10^U[lo,hi])
    rows.  self.schemas[plan] lists dicts with the schema id, the relation
    sizes and the buffers (kept alive here)."""

    def __init__(self, ctx: R.Context, server: Server, seed: int = 0x52464F58000000D0,
                 per_kind: int = 4, exp=(5.0, 8.0)):
        self.ctx, self.server = ctx, server
        self.schemas: Dict[int, List[dict]] = {p: [] for p in range(4)}
        self.bytes = 0
        self.order = []
        rng = np.random.default_rng(seed & 0xFFFFFFFF)

        def size():
            return int(np.floor(10.0 ** rng.uniform(exp[0], exp[1])))

        sd = seed
        for k in range(per_kind):
            sd += 0x100
            self._select_project(size(), sd + 1)
            self._join_agg(size(), size(), sd + 2)
            self._join3(size(), size(), size(), sd + 3)
            self._set_op(size(), size(), sd + 4)

    def _rel(self, col0, col1, rows, key_bits=0, val_bits=0) -> Relation:
        r = Relation()
        r.col[0] = col0.ptr
        r.col[1] = col1.ptr if col1 is not None else None
        r.rows, r.key_bits, r.val_bits = rows, key_bits, val_bits
        return r

    def register(self, server: Server) -> None:
        """Register the same relations with another (fresh) server; the
        schema ids come out the same."""
        for sid, plan, rels in self.order:
            if server.add_schema(plan, rels) != sid:
                raise ValueError("register: the server already has schemas")

    def _add(self, plan, rels, bufs, sizes):
        sid = self.server.add_schema(plan, rels)
        self.order.append((sid, plan, list(rels)))
        self.bytes += sum(b.nbytes for b in bufs)
        self.schemas[plan].append({"id": sid, "sizes": sizes, "bufs": bufs, "rels": list(rels)})

    def _select_project(self, n, seed):
        c = self.ctx
        ident = R.DeviceBuffer(4 * n)
        c.iota_u32(ident, n)
        qty = masked_u32(c, n, seed, 6)
        val = R.DeviceBuffer(8 * n)
        c.gen_u64(val, n, seed + 1)
        hi = R.DeviceBuffer(8 * n)
        c.merge(hi, 8, ident, 4, qty, 4, n, 32)
        tup = R.DeviceBuffer(16 * n)
        c.merge(tup, 16, hi, 8, val, 8, n, 64)
        c.synchronize()
        for b in (ident, qty, val, hi):
            b.close()
        self._add(PLAN_SELECT_PROJECT, [self._rel(tup, None, n)], [tup], (n,))

    def _join_agg(self, nf, nd, seed):
        c = self.ctx
        dkey = permutation_u32(c, nd, seed)
        grp = masked_u32(c, nd, seed + 1, GRP_BITS)
        fk, m = zipf_fk(c, nf, dkey, nd, seed + 2)
        kb = bits_for(nd)
        self._add(PLAN_JOIN_AGG, [self._rel(fk, m, nf, kb), self._rel(dkey, grp, nd, kb, GRP_BITS)],
                  [fk, m, dkey, grp], (nf, nd))

    def _join3(self, nf, ns, nt, seed):
        c = self.ctx
        tkey = permutation_u32(c, nt, seed)
        tval = R.DeviceBuffer(4 * nt)
        c.gen_u32(tval, nt, seed + 1)
        skey = permutation_u32(c, ns, seed + 2)
        fk2, scratch = zipf_fk(c, ns, tkey, nt, seed + 3)
        scratch.close()
        fk1, m = zipf_fk(c, nf, skey, ns, seed + 4)
        self._add(PLAN_JOIN3, [self._rel(fk1, m, nf, bits_for(ns)),
                               self._rel(skey, fk2, ns, bits_for(ns), bits_for(nt)),
                               self._rel(tkey, tval, nt, bits_for(nt))],
                  [fk1, m, skey, fk2, tkey, tval], (nf, ns, nt))

    def _set_op(self, na, nb, seed):
        a, na_u = sorted_set_u64(self.ctx, na, seed)
        b, nb_u = sorted_set_u64(self.ctx, nb, seed, first=na // 2)
        self._add(PLAN_SET_OP, [self._rel(a, None, na_u), self._rel(b, None, nb_u)], [a, b],
                  (na_u, nb_u))


def make_requests(cat: Catalog, n: int, rate: Optional[float], seed: int = 0x52464F58000000D1,
                  mix=MIX):
    """n requests drawn from the catalog's schemas; arrivals Poisson of
    `rate` per second (None: all arrive at time 0).  The requests depend on
    seed alone, so one seed at several rates gives the same requests."""
    rng = np.random.default_rng(seed & 0xFFFFFFFF)
    arr = np.random.default_rng((seed + 1) & 0xFFFFFFFF)
    kinds = [p for p in range(4) if cat.schemas[p]]
    w = np.array([mix[p] for p in kinds], dtype=np.float64)
    w /= w.sum()
    reqs = (Request * n)()
    t = 0.0
    for i in range(n):
        p = kinds[int(rng.choice(len(kinds), p=w))]
        sch = cat.schemas[p][int(rng.integers(len(cat.schemas[p])))]
        r = reqs[i]
        r.schema = sch["id"]
        if p == PLAN_SELECT_PROJECT:
            r.param, r.constant = CMP_LT, int(rng.integers(1, 64))
        elif p == PLAN_SET_OP:
            r.param = int(rng.integers(0, 3))
        if rate:
            t += arr.exponential(1.0 / rate)
        r.arrival_ns = int(t * 1e9)
    return reqs


def plan_of(cat: Catalog, schema_id: int) -> int:
    for p, lst in cat.schemas.items():
        if any(s["id"] == schema_id for s in lst):
            return p
    raise KeyError(schema_id)


def copy_requests(reqs):
    """Fresh request array with the same inputs and no outputs."""
    out = (Request * len(reqs))()
    for i, r in enumerate(reqs):
        o = out[i]
        o.schema, o.param, o.constant = r.schema, r.param, r.constant
        o.arrival_ns, o.sink, o.sink_bytes = r.arrival_ns, r.sink, r.sink_bytes
    return out


def outcome(r: Request):
    """What must agree between runs: status, rows and column digests."""
    return (int(r.status), int(r.rows), tuple(int(x) for x in r.digest))


def latency_summary(reqs, stats: ServeStats) -> dict:
    lat = np.array([(r.done_ns - r.arrival_ns) for r in reqs], dtype=np.float64) / 1e6
    first = min(r.arrival_ns for r in reqs)
    last = max(r.done_ns for r in reqs)
    span = (last - first) / 1e9
    return {"requests": len(reqs), "qps": len(reqs) / span if span > 0 else float("inf"),
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "peak_inflight": int(stats.peak_inflight),
            "peak_gib": stats.peak_bytes / 2**30, "blocked_on_memory": int(stats.blocked_on_memory)}


def fmix64(z: np.ndarray) -> np.ndarray:
    z = z.astype(np.uint64)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def digest_host(words: np.ndarray, c: int) -> int:
    """The digest of one output column (u64 row words), computed on the
    host: sum_i fmix64(w_i + fmix64(0x9E3779B97F4A7C15 * (c + 1) + i))."""
    with np.errstate(over="ignore"):
        i = np.arange(len(words), dtype=np.uint64)
        salt = np.uint64((GOLDEN * (c + 1)) & M64)
        h = fmix64(words.astype(np.uint64) + fmix64(salt + i))
        return int(h.sum(dtype=np.uint64))


__all__ = ["Server", "Catalog", "Relation", "Request", "ServeStats", "make_requests",
           "latency_summary", "digest_host", "outcome", "copy_requests", "plan_of"]
