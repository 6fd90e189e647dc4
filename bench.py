#!/usr/bin/env python3
"""Benchmark of the RA hot path on one MI355X.

Headline (`value`): stable LSD radix sort of 2^28 (u32 key, u32 value) pairs,
keys uniform over the full 32-bit range, values = row index - the sort the
reference runs as set_b40csortpair (b40cSort.cu:752-780).  One step sorts one
such batch that is already resident in HBM (each step has its own input, all
generated before the timed region).  `roofline` covers the dominant kernel,
the onesweep digit pass, timed per launch with HIP events on its stream
during the timed steps.  `cpu_baseline` is the oracle's serial LSD sort on a
bounded sample, on this host.  `ops` adds the other workloads: the u64
Zipf(1.1) sort, the config-1 join and an HBM-scale join, TPC-H Q1 over 60M
rows, and 512 batched join-aggregate instances (plain per-operator launches
vs one captured graph).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "rows/s + % HBM roofline per op (join, sort, agg); batched qps, p99 latency"
SEED = 0x52464F58000000CC


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def median(xs):
    return float(np.median(np.array(xs)))


COPY_GBPS = None  # measured device-to-device copy rate (read + write bytes)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sort_onesweep_2p28_u32_traffic.txt")
AGG_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "reduce_by_key_2p28_zipf_traffic.txt")


def roofline(bytes_per_launch, launches, total_ms, traffic=None):
    avg_ms = total_ms / max(launches, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    r = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
         "launch_ms": round(avg_ms, 4), "launches": launches,
         "bytes_per_launch": int(bytes_per_launch)}
    if COPY_GBPS:
        r["frac_of_copy"] = round(achieved / COPY_GBPS, 4)
    return r


def copy_bandwidth(rf, ctx, nbytes=1 << 30, reps=5):
    """Device copy of 1 GiB (the reference kernel of known size for the
    traffic counters): bytes read + written per second."""
    a, b = rf.DeviceBuffer(nbytes), rf.DeviceBuffer(nbytes)
    ctx.copy(b, a, nbytes)
    ctx.synchronize()
    t = rf.Timer(ctx)
    times = []
    for _ in range(reps):
        t.begin()
        ctx.copy(b, a, nbytes)
        times.append(t.end())
    t.close()
    a.close()
    b.close()
    return 2 * nbytes / (median(times) * 1e-3) / 1e9


def measured_traffic(path=TRAFFIC_FILE):
    """HBM bytes per launch of a kernel (the onesweep pass by default) from
    its committed rocprofv3 PMC summary (FETCH_SIZE and WRITE_SIZE, scaled by
    the same counters on the 1 GiB copy), or None."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def kernel_traffic(profile, kernel):
    """HBM bytes per dispatch of one kernel from a committed rocprofv3
    summary under profiles/ (the `traffic_bytes_per_dispatch` line of its
    `## <kernel>` section: FETCH_SIZE and WRITE_SIZE scaled by the same
    counters on the 1 GiB copy), or None.  The summary was taken on the
    same workload as the bench line that quotes it."""
    try:
        with open(os.path.join(ROOT, "profiles", profile)) as f:
            inside = False
            for line in f:
                if line.startswith("## "):
                    inside = line[3:].startswith(kernel)
                elif inside and line.startswith("traffic_bytes_per_dispatch:"):
                    return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


# ---------------------------------------------------------------------------
def bench_sort_u32(rf, ctx, steps, warmup):
    n = 1 << 28
    total = steps + warmup
    if total * 8 * n > 200 << 30:
        raise SystemExit("bench.py: too many steps for distinct resident inputs (max 25)")
    keys, vals = [], []
    for i in range(total):
        k = rf.DeviceBuffer(4 * n)
        v = rf.DeviceBuffer(4 * n)
        ctx.gen_u32(k, n, SEED + 1 + i)
        ctx.iota_u32(v, n)
        keys.append(k)
        vals.append(v)
    ctx.reserve(rf.lib().rfx_sort_workspace(n, 4, 32, 4))
    for i in range(warmup):
        ctx.sort_pairs(keys[i], vals[i], n, 4, 32, 4)
    ctx.synchronize()
    ctx.profile(True)
    timer = rf.Timer(ctx)
    ctx.synchronize()
    t0 = time.perf_counter()
    timer.begin()
    for i in range(warmup, total):
        ctx.sort_pairs(keys[i], vals[i], n, 4, 32, 4)
    ev_ms = timer.end()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    launches, sweep_ms = ctx.profile_read("sort.onesweep")
    hist = ctx.profile_read("sort.hist")
    ctx.profile(False)
    # spot check: the last output is sorted and stable
    k = keys[-1].download(np.uint32, n)
    v = vals[-1].download(np.uint32, n)
    ok = bool(np.all(k[1:] >= k[:-1]) and np.all(v[1:][k[1:] == k[:-1]] > v[:-1][k[1:] == k[:-1]]))
    for b in keys + vals:
        b.close()
    # Each onesweep launch reads and writes every (key, value): 16 B/row.
    rl = roofline(16 * n, launches, sweep_ms, measured_traffic())
    return {"n": n, "wall_s": wall, "event_ms": ev_ms, "ms_per_step": 1e3 * wall / steps,
            "rows_per_s": n * steps / wall, "roofline": rl, "verified_sorted": ok,
            "hist_ms": hist[1] / max(hist[0], 1),
            "pass_model_GBps": 68 * n * steps / wall / 1e9}


def bench_sort_u64_zipf(rf, ctx, iters=3):
    n = 1 << 27
    k0 = rf.DeviceBuffer(8 * n)
    v0 = rf.DeviceBuffer(4 * n)
    ctx.gen_zipf_u64(k0, n, SEED + 2, 1.1, 1 << 32)
    k = rf.DeviceBuffer(8 * n)
    v = rf.DeviceBuffer(4 * n)
    L = rf.lib()
    times = []
    for i in range(iters + 1):
        if i == 1:
            ctx.profile(True)  # per-launch figures over the timed calls only
        L.rfx_memcpy_dtod(k.ptr, k0.ptr, 8 * n)
        ctx.iota_u32(v, n)
        ctx.synchronize()
        t = rf.Timer(ctx)
        t.begin()
        ctx.sort_pairs(k, v, n, 8, 64, 4)
        times.append(t.end())
        t.close()
    launches, ms = ctx.profile_read("sort.onesweep")
    ctx.profile(False)
    ms_sort = median(times[1:])
    kk = k.download(np.uint64, n)
    ok = bool(np.all(kk[1:] >= kk[:-1]))
    for b in (k0, v0, k, v):
        b.close()
    return {"n": n, "ms": ms_sort, "rows_per_s": n / (ms_sort * 1e-3),
            "pass_model_GBps": 200 * n / (ms_sort * 1e-3) / 1e9,
            "roofline": roofline(24 * n, launches, ms, kernel_traffic(
                "sort_onesweep_2p27_u64_zipf.txt", "k_onesweep<8, 4, 512, 14, 8,")
                if n == 1 << 27 else None), "verified_sorted": ok}


def bench_join(rf, ctx, n_r, n_s, key_bits, iters):
    """Sort-merge join R x S -> (key, r_val, s_val) 16-byte rows.  'core'
    times find_bounds + expand/gather on sorted inputs; 'plan' adds the two
    sorts."""
    from redfox_amd import datagen

    rk, rv, sk, sv = datagen.join_tables(SEED, n_r, n_s, key_bits)
    src = {k: rf.DeviceBuffer.from_numpy(a) for k, a in dict(rk=rk, rv=rv, sk=sk, sv=sv).items()}
    work = {k: rf.DeviceBuffer(b.nbytes) for k, b in src.items()}
    lb = rf.DeviceBuffer(4 * n_r)
    lc = rf.DeviceBuffer(4 * n_r)
    tot = rf.u64_cell()
    L = rf.lib()

    def reset():
        for k in src:
            L.rfx_memcpy_dtod(work[k].ptr, src[k].ptr, src[k].nbytes)

    reset()
    ctx.sort_pairs(work["rk"], work["rv"], n_r, 4, 32, 4)
    ctx.sort_pairs(work["sk"], work["sv"], n_s, 4, 32, 4)
    ctx.join_find_bounds(lb, lc, tot, work["rk"], n_r, work["sk"], n_s, 4)
    ctx.synchronize()
    total = rf.read_u64(tot)
    out = rf.DeviceBuffer(16 * max(total, 1))
    plan, core = [], []
    ctx.profile(True)
    for _ in range(iters + 2):
        reset()
        ctx.synchronize()
        t = rf.Timer(ctx)
        t.begin()
        ctx.sort_pairs(work["rk"], work["rv"], n_r, 4, 32, 4)
        ctx.sort_pairs(work["sk"], work["sv"], n_s, 4, 32, 4)
        mid = rf.Timer(ctx)
        mid.begin()
        ctx.join_find_bounds(lb, lc, tot, work["rk"], n_r, work["sk"], n_s, 4)
        ctx.join_expand_gather(out, 16, work["rk"], 4, work["rv"], 4, work["sv"], 4, lb, lc,
                               n_r, tot, total, 64, 32)
        core.append(mid.end())
        plan.append(t.end())
        t.close()
        mid.close()
    fb = ctx.profile_read("join.find_bounds")
    eg = ctx.profile_read("join.expand_gather")
    ctx.profile(False)
    core_ms, plan_ms = median(core[2:]), median(plan[2:])
    core_bytes = 8 * n_r + 4 * n_s + 16 * total  # keys+lb/count, right keys, output
    res = {"R": n_r, "S": n_s, "out_rows": total, "core_ms": core_ms, "plan_ms": plan_ms,
           "core_rows_per_s": (n_r + n_s) / (core_ms * 1e-3),
           "plan_rows_per_s": (n_r + n_s) / (plan_ms * 1e-3),
           "core_GBps": core_bytes / (core_ms * 1e-3) / 1e9,
           "find_bounds_ms": fb[1] / max(fb[0], 1), "expand_gather_ms": eg[1] / max(eg[0], 1)}
    # expand+gather algorithmic bytes: lower bound + count per left row,
    # r_val per left row, s_val per output (gathered), 16 B out per output.
    eg_bytes = 12 * n_r + 4 * total + 16 * total
    prof = "join_find_bounds_expand_2p24x2p26.txt" if (n_r, n_s) == (1 << 24, 1 << 26) else ""
    res["expand_gather_roofline"] = roofline(eg_bytes, eg[0], eg[1], kernel_traffic(
        prof, "k_expand<1, 512, 4096,") if prof else None)
    # find_bounds reads both key columns and writes a lower bound and a
    # scanned count per left row.
    res["find_bounds_roofline"] = roofline(4 * (n_r + n_s) + 8 * n_r, fb[0], fb[1], kernel_traffic(
        prof, "k_find_bounds<unsigned int, 256, 4096>") if prof else None)
    # The same plan (two sorts, find_bounds, expansion) captured once into
    # a HIP graph and replayed: one launch per plan.  The inputs are
    # restored before each replay, outside the timed region as above, and
    # the output of every replay must equal the eager one.
    want = out.download(np.uint8, 16 * total)
    stream = rf.Stream()
    gctx = rf.Context(0, stream.handle)  # capture needs a stream of its own

    def restore():
        for k in src:
            gctx.copy(work[k], src[k], src[k].nbytes)

    def enqueue():
        gctx.sort_pairs(work["rk"], work["rv"], n_r, 4, 32, 4)
        gctx.sort_pairs(work["sk"], work["sv"], n_s, 4, 32, 4)
        gctx.join_find_bounds(lb, lc, tot, work["rk"], n_r, work["sk"], n_s, 4)
        gctx.join_expand_gather(out, 16, work["rk"], 4, work["rv"], 4, work["sv"], 4, lb, lc,
                                n_r, tot, total, 64, 32)
    restore()
    enqueue()  # sizes gctx's workspace before the capture
    gctx.synchronize()
    gctx.graph_begin()
    enqueue()
    graph = gctx.graph_end()
    graph_ms = []
    for i in range(iters + 2):
        out.zero()
        restore()
        gctx.synchronize()
        t = rf.Timer(gctx)
        t.begin()
        graph.launch()
        graph_ms.append(t.end())
        t.close()
        if i < 2 and not np.array_equal(out.download(np.uint8, 16 * total), want):
            raise RuntimeError("join plan: a graph replay differs from the eager plan")
    graph.close()
    gctx.close()
    stream.close()
    res["plan_graph_ms"] = median(graph_ms[2:])
    for b in list(src.values()) + list(work.values()) + [lb, lc, tot, out]:
        b.close()
    return res


def bench_q1(rf, ctx, iters=5):
    n = 60_000_000
    cols = [rf.DeviceBuffer(4 * n), rf.DeviceBuffer(n), rf.DeviceBuffer(n), rf.DeviceBuffer(4 * n),
            rf.DeviceBuffer(8 * n), rf.DeviceBuffer(8 * n), rf.DeviceBuffer(8 * n)]
    ctx.gen_lineitem(cols, n, SEED + 2)
    out = rf.DeviceBuffer(48 * 8)
    ctx.q1(out, *cols, n, 2436)
    ctx.synchronize()
    ctx.profile(True)
    times = []
    for _ in range(iters):
        t = rf.Timer(ctx)
        t.begin()
        ctx.q1(out, *cols, n, 2436)
        times.append(t.end())
        t.close()
    launches, ms = ctx.profile_read("q1.partial")
    ctx.profile(False)
    res = out.download(np.float64, 48).reshape(6, 8)
    for b in cols + [out]:
        b.close()
    m = median(times)
    return {"n": n, "ms": m, "rows_per_s": n / (m * 1e-3),
            "roofline": roofline(34 * n, launches, ms,
                                 kernel_traffic("q1_partial_60M.txt", "k_q1_partial")),
            "kept_rows": int(res[:, 5].sum())}


def bench_mapfilter(rf, ctx, iters=5):
    """MAPFILTER over 2^27 lineitem rows: WHERE shipdate <= 2436, emitting
    disc_price = price * (1 - disc) and charge = disc_price * (1 + tax)."""
    from redfox_amd import mapfilter as M

    n = 1 << 27
    cols = [rf.DeviceBuffer(4 * n), rf.DeviceBuffer(n), rf.DeviceBuffer(n), rf.DeviceBuffer(4 * n),
            rf.DeviceBuffer(8 * n), rf.DeviceBuffer(8 * n), rf.DeviceBuffer(8 * n)]
    ctx.gen_lineitem(cols, n, SEED + 9)
    ship, price, disc, tax = cols[0], cols[4], cols[5], cols[6]
    x = M.MapFilter([(ship, "i32"), (price, "f64"), (disc, "f64"), (tax, "f64")])
    s, p, d, t = x.cols
    x.where(s <= 2436)
    x.emit(p * (1 - d), "f64")
    x.emit(p * (1 - d) * (1 + t), "f64")
    outs = [rf.DeviceBuffer(8 * n), rf.DeviceBuffer(8 * n)]
    kept_cell = rf.u64_cell()
    t0 = time.perf_counter()
    x.run(ctx, n, outs, kept_cell)  # compiles the program on its first use
    ctx.synchronize()
    first_ms = (time.perf_counter() - t0) * 1e3
    ctx.profile(True)
    times = []
    for _ in range(iters):
        tm = rf.Timer(ctx)
        tm.begin()
        x.run(ctx, n, outs, kept_cell)
        times.append(tm.end())
        tm.close()
    launches, ms = ctx.profile_read("mapfilter")
    ctx.profile(False)
    kept = rf.read_u64(kept_cell)
    for b in cols + outs + [kept_cell]:
        b.close()
    m = median(times)
    # Every row's four columns are read once; kept rows write two doubles.
    return {"n": n, "kept": kept, "ms": m, "rows_per_s": n / (m * 1e-3), "first_call_ms": first_ms,
            "roofline": roofline(28 * n + 16 * kept, launches, ms, kernel_traffic(
                "mapfilter_compiled_2p27_lineitem.txt", "rfx_mf_jit"))}


PROF_TAGS = ("join.find_bounds", "join.expand", "join.expand_gather", "join.expand_project",
             "join.gather", "mapfilter", "q1.partial", "rbk.pass", "reduce.single",
             "select", "setop", "sort.hist", "sort.onesweep", "unique")


def bench_q1_plan(rf, ctx, iters=3, n_orders=1 << 22):
    """The reference's TPC-H Q1 plan (q1.man.cpp) run by the plan front end
    over lineitem column relations keyed by (LINENUMBER, ORDER), about
    2^24 rows (4 per order): a product with the start date, a date
    filter, five key joins, two arithmetic projections, six AGGs.  Inputs
    resident in HBM; the timed region is Plan.run() (every node, with the
    host reading each result size as Harmony does)."""
    from redfox_amd import datagen, raplan

    cols = datagen.lineitem_keyed(n_orders, SEED + 30)
    ln, od = cols[0], cols[1]
    n = len(ln)
    p = raplan.Plan(ctx, raplan.q1_plan())
    p.set("+$logicQ1:_dateDelta", [np.array([90 * 86400])])
    for name, c in zip(("L_SHIPDATE", "L_RETURNFLAG", "L_LINESTATUS", "L_QUANTITY",
                        "L_EXTENDEDPRICE", "L_DISCOUNT", "L_TAX"), cols[2:]):
        p.set(name, [ln, od, c])
    t0 = time.perf_counter()
    p.run()  # compiles the plan's MAPFILTER programs
    first_ms = (time.perf_counter() - t0) * 1e3
    times, kernel_ms = [], []
    for _ in range(iters):
        ctx.synchronize()
        ctx.profile(True)
        t0 = time.perf_counter()
        p.run()
        times.append((time.perf_counter() - t0) * 1e3)
        kernel_ms.append(sum(ctx.profile_read(t)[1] for t in PROF_TAGS))
        ctx.profile(False)
    m = median(times)
    count = p.get("+$logicQ1:_count_order")[2]
    return {"n": n, "ms": m, "rows_per_s": n / (m * 1e-3), "first_run_ms": first_ms,
            "kernel_ms": median(kernel_ms), "commands": p.steps,
            "kept_rows": int(count.sum()), "groups": int(len(count))}


def bench_tbl(rf, ctx, iters=3, n_orders=1 << 20):
    """lineitem.tbl text (about 4M dbgen-style rows, 147 bytes each)
    resident in HBM to the Q1 plan's typed columns: a row-count pass, then
    the row-ends pass and the field pass into columns of that size."""
    from redfox_amd import datagen, loaders

    text = datagen.lineitem_tbl(n_orders, SEED + 40)[0]
    raw = np.frombuffer(text, dtype=np.uint8)
    buf = rf.DeviceBuffer.from_numpy(raw)
    spec = loaders.LINEITEM_SPEC
    loaders.parse_tbl(ctx, buf, spec, "|", len(raw))
    ctx.synchronize()
    ctx.profile(True)
    times = []
    for _ in range(iters):
        t0 = time.perf_counter()
        cols, n, _ = loaders.parse_tbl(ctx, buf, spec, "|", len(raw))
        times.append((time.perf_counter() - t0) * 1e3)
        for c in cols:
            if c is not None:
                c.close()
    rows_l, rows_ms = ctx.profile_read("tbl.rows")
    fld_l, fld_ms = ctx.profile_read("tbl.fields")
    ctx.profile(False)
    buf.close()
    m = median(times)
    out_bytes = sum(loaders.WIDTH[s] for s in spec)
    return {"bytes": len(raw), "rows": n, "ms": m, "GBps_text": len(raw) / (m * 1e-3) / 1e9,
            "rows_per_s": n / (m * 1e-3),
            # parse_tbl runs the rows pass twice per call: a count (capacity
            # 0, no row ends stored), then the parse (8 B of row end per row).
            "rows_pass": roofline(len(raw) + 4 * n, rows_l, rows_ms, kernel_traffic(
                "loader_stringjoin_q1plan_kernels.txt", "k_tbl_rows")),
            "fields_pass": roofline(len(raw) + (out_bytes + 16) * n, fld_l, fld_ms, kernel_traffic(
                "loader_stringjoin_q1plan_kernels.txt", "k_tbl_fields"))}


def bench_join_string(rf, ctx, iters=5, nl=1 << 18, nr=1 << 22):
    """String-key join (find_bounds_string + expansion): 2^18 distinct left
    keys 'SUPPLIER#%09d' against 2^22 right keys drawn from them (16 per
    left key), both sorted, as NUL-terminated strings in HBM."""
    keys = np.array([b"SUPPLIER#%09d" % i for i in range(nl)])
    width = keys.dtype.itemsize + 1
    rng = np.random.default_rng(SEED + 50)
    right_idx = np.sort(rng.integers(0, nl, nr))

    def blob_and_ptrs(idx):
        mat = np.zeros((len(idx), width), dtype=np.uint8)
        mat[:, :width - 1] = keys[idx].view(np.uint8).reshape(len(idx), width - 1)
        buf = rf.DeviceBuffer.from_numpy(mat.reshape(-1))
        ptrs = rf.DeviceBuffer.from_numpy(buf.ptr + width * np.arange(len(idx), dtype=np.uint64))
        return buf, ptrs
    lbuf, lp = blob_and_ptrs(np.arange(nl))
    rbuf, rp = blob_and_ptrs(right_idx)
    lb, lc, tot = rf.DeviceBuffer(4 * nl), rf.DeviceBuffer(4 * nl), rf.u64_cell()
    ctx.join_find_bounds_string(lb, lc, tot, lp, nl, rp, nr)
    ctx.synchronize()
    total = rf.read_u64(tot)
    li, ri = rf.DeviceBuffer(4 * total), rf.DeviceBuffer(4 * total)
    times = []
    ctx.profile(True)
    for _ in range(iters):
        t = rf.Timer(ctx)
        t.begin()
        ctx.join_find_bounds_string(lb, lc, tot, lp, nl, rp, nr)
        ctx.join_main(li, ri, tot, total, lb, lc, nl)
        times.append(t.end())
        t.close()
    _, fb_ms = ctx.profile_read("join.find_bounds_string")
    ctx.profile(False)
    m = median(times)
    for b in (lbuf, lp, rbuf, rp, lb, lc, tot, li, ri):
        b.close()
    return {"nl": nl, "nr": nr, "out_rows": total, "ms": m, "rows_per_s": (nl + nr) / (m * 1e-3),
            "search_ms": fb_ms / iters}


def bench_agg(rf, ctx, iters=5):
    """GROUP BY key: SUM(v), COUNT(*) over 2^28 rows sorted by key (u32
    keys Zipf(1.1) over 2^20 values, f64 values) in one reduce-by-key pass."""
    n, groups = 1 << 28, 1 << 20
    keys, vals = rf.DeviceBuffer(4 * n), rf.DeviceBuffer(8 * n)
    ident = rf.DeviceBuffer(4 * groups)
    ctx.iota_u32(ident, groups)
    row_off = rf.DeviceBuffer.from_numpy(np.array([0, n], dtype=np.uint64))
    dim_off = rf.DeviceBuffer.from_numpy(np.array([0, groups], dtype=np.uint64))
    ctx.gen_fk_batch(keys, vals, row_off, dim_off, ident, 1, n, SEED + 5, 1.1)
    ctx.sort_pairs(keys, vals, n, 4, 20, 8)
    ok, ov, oc = rf.DeviceBuffer(4 * groups), rf.DeviceBuffer(8 * groups), rf.DeviceBuffer(4 * groups)
    nout, ncell = rf.u64_cell(), rf.u64_cell(n)
    # One untimed call first (code-object load, workspace), so that the
    # timed calls and the per-launch figures cover the same calls.
    ctx.reduce_count_by_key_f64(ok, ov, oc, nout, keys, vals, ncell, n, 4, 0, out_cap=groups)
    ctx.synchronize()
    ctx.profile(True)
    times = []
    for _ in range(iters):
        t = rf.Timer(ctx)
        t.begin()
        ctx.reduce_count_by_key_f64(ok, ov, oc, nout, keys, vals, ncell, n, 4, 0, out_cap=groups)
        times.append(t.end())
        t.close()
    # The roofline charges the algorithmic bytes (each row read once, each
    # group written once) to the reduce kernel's launches.
    lo, mo = ctx.profile_read("rbk.pass")
    ctx.profile(False)
    g = rf.read_u64(nout)
    cnt = oc.download(np.uint32, g)
    res = {"n": n, "groups": g, "ms": median(times), "rows_per_s": n / (median(times) * 1e-3),
           "counts_sum_to_n": bool(int(cnt.sum(dtype=np.uint64)) == n),
           "roofline": roofline(12 * n + 16 * g, lo, mo, measured_traffic(AGG_TRAFFIC_FILE))}
    for b in (keys, vals, ident, row_off, dim_off, ok, ov, oc, nout, ncell):
        b.close()
    return res


def bench_select(rf, ctx, iters=5):
    """SELECT on 2^27 16-byte tuples, a 32-bit field < 2^31 (keeps half)."""
    n = 1 << 27
    hi, lo = rf.DeviceBuffer(8 * n), rf.DeviceBuffer(8 * n)
    ctx.gen_u64(hi, n, SEED + 6)
    ctx.gen_u64(lo, n, SEED + 7)
    tup, out = rf.DeviceBuffer(16 * n), rf.DeviceBuffer(16 * n)
    ctx.merge(tup, 16, hi, 8, lo, 8, n, 64)
    nout = rf.u64_cell()
    ctx.synchronize()
    hi.close()
    lo.close()
    times = []
    for i in range(iters + 1):
        if i == 1:
            ctx.profile(True)  # per-launch figures over the timed calls only
        t = rf.Timer(ctx)
        t.begin()
        ctx.select(out, nout, tup, n, 16, 64, 32, 2, 1 << 31)
        times.append(t.end())
        t.close()
    launches, ms = ctx.profile_read("select")
    ctx.profile(False)
    kept = rf.read_u64(nout)
    res = {"n": n, "kept": kept, "ms": median(times[1:]), "rows_per_s": n / (median(times[1:]) * 1e-3),
           "roofline": roofline(16 * n + 16 * kept, launches, ms, kernel_traffic(
               "select_2p27_16B.txt", "k_select<u128,") if n == 1 << 27 else None)}
    # The autotuner's choice of rows per thread for 16-, 8- and 4-byte
    # tuples of the same bytes (a 32-bit field < 2^31: half kept).
    from redfox_amd.autotune import tune

    res["autotune"] = {}
    for w in (16, 8, 4):
        m = 16 * n // w
        sh = 64 if w == 16 else 0
        t = tune(ctx, lambda: ctx.select(out, nout, tup, m, w, sh, 32, 2, 1 << 31),
                 knobs={"select.vt": (0, 8, 16, 24, 32)}, reps=3)
        res["autotune"][f"{w}B"] = {"choice": t["choice"]["select.vt"],
                                    "ms": t["timings_ms"]["select.vt"]}
    ctx.tune("select.vt", 0)
    for b in (tup, out, nout):
        b.close()
    return res


def bench_setop(rf, ctx, iters=5):
    """Union, intersection and difference of two sorted unique u64 sets of
    2^27 elements each (half of B inside A)."""
    from redfox_amd import serving

    n = 1 << 27
    a, na = serving.sorted_set_u64(ctx, n, SEED + 8)
    b, nb = serving.sorted_set_u64(ctx, n, SEED + 8, first=n // 2)
    out = rf.DeviceBuffer(8 * (na + nb))
    nout = rf.u64_cell()
    res = {"na": na, "nb": nb}
    for op, name in ((0, "union"), (2, "intersection"), (1, "difference")):
        times = []
        for i in range(iters + 1):
            if i == 1:
                ctx.profile(True)  # per-launch figures over the timed calls only
            t = rf.Timer(ctx)
            t.begin()
            ctx.set_op(op, out, nout, a, na, b, nb, 8)
            times.append(t.end())
            t.close()
        launches, ms = ctx.profile_read("setop")
        ctx.profile(False)
        m = rf.read_u64(nout)
        res[name] = {"out": m, "ms": median(times[1:]),
                     "rows_per_s": (na + nb) / (median(times[1:]) * 1e-3),
                     "roofline": roofline(8 * (na + nb + m), launches, ms, kernel_traffic(
                         "set_ops_2p27_u64.txt", f"k_set_op<unsigned long, {op}, 512, 6144>")
                         if n == 1 << 27 else None)}
    for x in (a, b, out, nout):
        x.close()
    return res


def bench_batched(rf, ctx, reps=5):
    """512 join + group-aggregate instances (fact 10^U[3,7] rows, Zipf(1.1)
    foreign keys into dimensions of 10^U[2,5] rows): the plain per-operator
    form (one instance and one operator launch at a time, sizes read back
    as Harmony does) against the same plan batched into one HIP graph, its
    tile shapes chosen first by the autotuner (not timed)."""
    from redfox_amd import batched

    w = batched.Workload(ctx, n_inst=512)
    plain = batched.PlainRunner(w, ctx)
    plain.run()
    ctx.synchronize()
    t0 = time.perf_counter()
    plain.run()
    ctx.synchronize()
    t_plain = time.perf_counter() - t0
    br = batched.BatchedRunner(w)
    tuned = br.autotune()
    br.capture()
    times = []
    for _ in range(reps):
        br.synchronize()
        t0 = time.perf_counter()
        br.run()
        br.synchronize()
        times.append(time.perf_counter() - t0)
    t_b = median(times)
    a, b = plain.results(), br.results()
    match = all(np.array_equal(x[0], y[0]) and np.array_equal(x[2], y[2]) and
                np.allclose(x[1], y[1], rtol=1e-12, atol=0) for x, y in zip(a, b))
    br.graph.close()
    # The same batching with the plain form's (unfused) operators: how much
    # of the speedup is batching and how much is fusion.
    bu = batched.BatchedRunner(w, fused=False)
    bu.capture()
    times = []
    for _ in range(reps):
        bu.synchronize()
        t0 = time.perf_counter()
        bu.run()
        bu.synchronize()
        times.append(time.perf_counter() - t0)
    t_u = median(times)
    match = match and all(np.array_equal(x[0], y[0]) and np.array_equal(x[2], y[2]) and
                          np.allclose(x[1], y[1], rtol=1e-12, atol=0)
                          for x, y in zip(a, bu.results()))
    bu.graph.close()
    res = {"instances": w.n_inst, "fact_rows": w.N, "dim_rows": w.D,
           "plain_s": t_plain, "batched_ms": 1e3 * t_b,
           "qps_plain": w.n_inst / t_plain, "qps_batched": w.n_inst / t_b,
           "speedup": t_plain / t_b, "rows_per_s_batched": w.N / t_b,
           "batched_unfused_ms": 1e3 * t_u, "speedup_unfused": t_plain / t_u,
           "input_GBps_batched": w.bytes_per_batch() / t_b / 1e9,
           "results_match": bool(match), "autotuned": tuned["choice"],
           "autotune_timings_ms": tuned["timings_ms"], "autotune_s": tuned.get("tuning_s")}
    return res


def bench_served(rf, ctx, n_req=200, slots=8, rates=(0.25, 0.5, 0.75, 0.9), exp=(5.0, 8.0)):
    """Mixed requests (30% select-project, 30% join-aggregate, 20% three-way
    join chain, 20% u64 set operations) over device-resident relations of
    floor(10^U[exp]) rows: saturation throughput of the continuous server
    (all requests arriving at once) against the plain one-request-at-a-time
    form with host synchronisation after every operator, then Poisson
    arrivals at fractions of the saturation rate for p50/p99 latency."""
    from redfox_amd import serving as S

    free, _ = mem_info(rf)
    srv0 = S.Server(0, slots=slots, budget_bytes=1 << 30)
    cat = S.Catalog(ctx, srv0, per_kind=4, exp=exp)
    free, _ = mem_info(rf)
    budget = int(min(200 << 30, free - (16 << 30)))
    srv = S.Server(0, slots=slots, budget_bytes=budget)
    cat.register(srv)
    srv0.close()
    warm = S.make_requests(cat, 32, rate=None, seed=1)
    srv.run(warm)
    reqs = S.make_requests(cat, n_req, rate=None, seed=2)
    st = srv.run(reqs)
    sat = S.latency_summary(reqs, st)
    plain = S.copy_requests(reqs)
    stp = srv.run(plain, S.MODE_PLAIN)
    pl = S.latency_summary(plain, stp)
    same = [S.outcome(r) for r in reqs] == [S.outcome(r) for r in plain]
    ok = all(r.status == 0 for r in reqs)
    rows = sum(_request_rows(cat, r) for r in reqs)
    sweep = []
    for f in rates:
        # The same requests, arriving as a Poisson process of rate lam.
        lam = f * sat["qps"]
        rq = S.make_requests(cat, n_req, rate=lam, seed=2)
        st2 = srv.run(rq)
        d = S.latency_summary(rq, st2)
        d["lambda"] = lam
        same = same and [S.outcome(r) for r in rq] == [S.outcome(r) for r in plain]
        sweep.append(d)
    srv.close()
    return {"catalog_relations_GiB": cat.bytes / 2**30, "budget_GiB": budget / 2**30,
            "slots": slots, "requests": n_req, "input_rows_per_batch": rows,
            "saturation": sat, "plain": pl, "speedup_qps": sat["qps"] / pl["qps"],
            "results_identical_in_every_run": bool(same and ok), "poisson": sweep}


def _request_rows(cat, r):
    from redfox_amd import serving as S

    p = S.plan_of(cat, r.schema)
    sch = next(s for s in cat.schemas[p] if s["id"] == r.schema)
    return int(sum(sch["sizes"]))


def mem_info(rf):
    import ctypes

    f, t = ctypes.c_uint64(), ctypes.c_uint64()
    rf.lib().rfx_mem_info(ctypes.byref(f), ctypes.byref(t))
    return int(f.value), int(t.value)


def summary(name, res):
    """One stderr line per op: its times and every roofline fraction in it
    (the JSON line itself is too long for a log tail)."""
    parts = []

    def walk(prefix, d):
        for k, v in d.items():
            if not isinstance(v, dict):
                if k in ("ms", "core_ms", "plan_ms", "plan_graph_ms", "batched_ms", "speedup",
                         "speedup_qps", "batched_unfused_ms", "speedup_unfused"):
                    parts.append(f"{prefix}{k}={v:.4g}")
                elif k == "choice":
                    parts.append(f"{prefix}{k}={v}")
                continue
            if k == "roofline" or k.endswith("_roofline") or k.endswith("_pass"):
                tag = prefix + (k[: -len("_roofline")] if k.endswith("_roofline") else
                                k if k != "roofline" else "")
                tag = tag.rstrip(".")
                parts.append(f"{tag or 'kernel'}: {v['launch_ms']:.4f} ms/launch "
                             f"{v['achieved']:.0f} GB/s frac {v['frac']:.3f}")
            elif k in ("saturation", "plain"):
                parts.append(f"{prefix}{k} qps={v['qps']:.0f} p99={v['p99_ms']:.1f}ms")
            elif k not in ("cpu_baseline", "autotune_timings_ms", "autotuned"):
                walk(f"{prefix}{k}.", v)

    walk("", res)
    return f"  {name}: " + "; ".join(parts)


def _host_cpu():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_line(run, rows, sample):
    """A CPU baseline figure: `run(reps)` times the restated algorithm
    (oracle/rfx_cpu.cpp) on the sample and returns the median seconds of
    `reps` runs after one warm-up.  Timed on all the OpenMP threads the host
    gives this process (OMP_NUM_THREADS), then on one thread."""
    from oracle import cpu_baseline as cb

    with cb.threads(0) as cores:
        dt = run(5)
    with cb.threads(1):
        dt1 = run(3)
    return {"value": round(rows / dt, 1), "unit": "rows/s", "cores": cores, "kind": "port",
            "value_1_thread": round(rows / dt1, 1),
            "sample": f"{sample}; median of 5 runs after a warm-up ({dt:.4f} s) on {cores} "
                      f"threads, median of 3 ({dt1:.4f} s) on 1 thread; host {_host_cpu()}, "
                      f"oracle/rfx_cpu.cpp (g++ -O3 -march=x86-64-v3 -fopenmp)"}


def cpu_baseline_sort():
    """The restated stable LSD radix sort (8-bit digits, as the reference's
    b40c sort) on 2^25 of the 2^28 pairs."""
    from oracle import cpu_baseline as cb
    from redfox_amd import datagen

    n = 1 << 25
    keys = datagen.u32(SEED + 1, n)
    vals = np.arange(n, dtype=np.uint32)
    return cpu_line(lambda reps: cb.sort_pairs(keys, vals, reps)[0], n,
                    "2^25 (u32 key, u32 value) pairs of the same generator, stable LSD radix "
                    "sort, 4 passes of 8-bit digits")


def cpu_ops_baseline(names):
    """The restated algorithms (OpenMP, all host threads and one) on bounded
    samples of each op line's workload, as rows/s.  Lines whose work has no
    restatement here (the text loader, the string join, the Q1 plan, the
    batched and served forms) carry no CPU figure."""
    from oracle import cpu_baseline as cb
    from redfox_amd import datagen

    out = {}
    n = 1 << 24
    if "sort_u64_zipf_2^27" in names:
        k = datagen.mix64(datagen.zipf_ranks(SEED + 2, n, 1.1, 1 << 32))
        v = np.arange(n, dtype=np.uint32)
        out["sort_u64_zipf_2^27"] = cpu_line(
            lambda reps: cb.sort_pairs(k, v, reps)[0], n,
            "2^24 (u64 Zipf(1.1) key, u32 value) pairs, stable LSD radix sort, 8 passes")
    for name, nr, ns, kb in (("join_65536x262144", 65536, 262144, 16),
                             ("join_2^24x2^26", 1 << 22, 1 << 24, 22)):
        if name not in names:
            continue
        rk, rv, sk, sv = datagen.join_tables(SEED, nr, ns, kb)
        o1, o2 = np.argsort(rk, kind="stable"), np.argsort(sk, kind="stable")
        rk, rv, sk, sv = rk[o1], rv[o1], sk[o2], sv[o2]
        sample = (f"the same {nr} x {ns} join" if nr == 65536 else
                  f"{nr} x {ns} rows, keys in [0, 2^{kb})")
        out[name] = cpu_line(lambda reps: cb.join(rk, rv, sk, sv, reps)[0], nr + ns,
                             sample + ", sorted inputs: sorted search, scan, expansion and "
                             "gather of (key, r_val, s_val)")
    if "q1_60M" in names or "mapfilter_2^27" in names:
        cols = datagen.lineitem(SEED + 2, n)
        if "q1_60M" in names:
            out["q1_60M"] = cpu_line(lambda reps: cb.q1(*cols, 2436, reps)[0], n,
                                     "2^24 lineitem rows of the same generator, Q1 filter and "
                                     "6-group sums (Neumaier) and counts")
        if "mapfilter_2^27" in names:
            out["mapfilter_2^27"] = cpu_line(
                lambda reps: cb.mapfilter(cols[0], cols[4], cols[5], cols[6], 2436, reps)[0], n,
                "2^24 lineitem rows, the same filter and two computed columns")
    if "agg_2^28" in names:
        g = np.sort((datagen.mix64(datagen.zipf_ranks(SEED + 21, n, 1.1, 1 << 20)) %
                     np.uint64(1 << 20)).astype(np.uint32))
        val = datagen.uniform01(SEED + 22, n) * 1e3
        out["agg_2^28"] = cpu_line(lambda reps: cb.reduce_sum_count(g, val, reps)[0], n,
                                   "2^24 sorted rows, Zipf(1.1) keys over 2^20 values, "
                                   "SUM (Neumaier) and COUNT per run")
    if "select_2^27" in names:
        t = datagen.splitmix64(SEED + 23, 2 * n)
        out["select_2^27"] = cpu_line(lambda reps: cb.select16(t, 1 << 31, reps)[0], n,
                                      "2^24 16-byte tuples, half kept, stable compaction")
    if "setop_2^27" in names:
        a = np.unique(datagen.splitmix64(SEED + 24, n))
        b = np.unique(np.concatenate([a[len(a) // 2:], datagen.splitmix64(SEED + 25, n // 2)]))
        out["setop_2^27"] = cpu_line(lambda reps: cb.set_op(0, a, b, reps)[0], len(a) + len(b),
                                     "union of two 2^24 sorted u64 sets, half of B in A")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-ops", action="store_true", help="headline sort only")
    ap.add_argument("--only", default="", help="comma-separated operator lines to run "
                    "besides the headline sort (e.g. mapfilter_2^27)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: measures one GPU; multi-GPU runs are not supported "
              "(the kernels run on one GPU's shard of a relation)", file=sys.stderr)
        sys.exit(2)
    import redfox_amd as rf

    if rf.device_count() < 1:
        raise SystemExit("bench.py: no HIP device visible")
    ctx = rf.Context(0)
    global COPY_GBPS
    COPY_GBPS = copy_bandwidth(rf, ctx)
    log(f"device copy {COPY_GBPS:.0f} GB/s (read + write)")
    log("sort 2^28 u32 pairs ...")
    s = bench_sort_u32(rf, ctx, args.steps, args.warmup)
    log(f"  {s['ms_per_step']:.3f} ms/step, {s['rows_per_s'] / 1e9:.2f} G rows/s, "
        f"onesweep {s['roofline']['launch_ms']:.3f} ms/launch frac {s['roofline']['frac']}")
    ops = {"sort_u32_2^28": s}
    lines = [
        ("sort_u64_zipf_2^27", "sort 2^27 u64 Zipf(1.1) pairs", lambda: bench_sort_u64_zipf(rf, ctx)),
        ("join_65536x262144", "join config 1", lambda: bench_join(rf, ctx, 65536, 262144, 16, 20)),
        ("join_2^24x2^26", "join 2^24 x 2^26", lambda: bench_join(rf, ctx, 1 << 24, 1 << 26, 24, 5)),
        ("q1_60M", "q1 60M", lambda: bench_q1(rf, ctx)),
        ("agg_2^28", "group-by sum+count 2^28", lambda: bench_agg(rf, ctx)),
        ("select_2^27", "select 2^27 16-byte tuples", lambda: bench_select(rf, ctx)),
        ("mapfilter_2^27", "mapfilter 2^27 lineitem rows", lambda: bench_mapfilter(rf, ctx)),
        ("q1_plan_2^24", "Q1 plan over 2^24 keyed lineitem rows", lambda: bench_q1_plan(rf, ctx)),
        ("tbl_lineitem_4M", "lineitem.tbl text to columns", lambda: bench_tbl(rf, ctx)),
        ("join_string_2^18x2^22", "string-key join", lambda: bench_join_string(rf, ctx)),
        ("setop_2^27", "set operations 2^27 + 2^27 u64", lambda: bench_setop(rf, ctx)),
        ("batched_512", "batched 512 join-aggregate instances", lambda: bench_batched(rf, ctx)),
        ("served_mix", "served mixed requests", lambda: bench_served(rf, ctx)),
        ("served_small_mix", "served mixed small requests",
         lambda: bench_served(rf, ctx, n_req=2000, exp=(3.0, 5.0))),
    ]
    only = set(args.only.split(",")) if args.only else None
    if only and not only <= {name for name, _, _ in lines}:
        raise SystemExit(f"bench.py: unknown --only lines {sorted(only - {n for n, _, _ in lines})}")
    if not args.no_ops:
        for name, what, fn in lines:
            if only is None or name in only:
                log(what + " ...")
                ops[name] = fn()
                log(summary(name, ops[name]))
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline_sort()
        if not args.no_ops:
            for name, base in cpu_ops_baseline(set(ops)).items():
                ops[name]["cpu_baseline"] = base
    line = {
        "metric": METRIC,
        "value": round(s["rows_per_s"], 1),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (splitmix64, device-generated, resident in HBM)",
        "config": {"workload": "stable LSD radix sort of 2^28 (u32 key uniform over the full "
                               "range, u32 value = row index) pairs",
                   "rows": 1 << 28, "key_bits": 32, "value_bytes": 4},
        "roofline": s["roofline"],
        "device_copy_GBps": round(COPY_GBPS, 1),
        "cpu_baseline": cpu,
        "ops": ops,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
