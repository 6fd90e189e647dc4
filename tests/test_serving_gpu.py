"""The served form: every request's result is the same alone, in a mix of
concurrent requests, under a budget tight enough to defer admissions, and
in the plain (one operator at a time, sizes read back) execution; and the
outputs themselves match a numpy evaluation of each plan (integers exact,
FP64 sums within 1e-12 relative)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def served(rf, ctx):
    from redfox_amd import serving as S

    srv = S.Server(0, slots=4, budget_bytes=1 << 30)
    cat = S.Catalog(ctx, srv, seed=0x5EED, per_kind=2, exp=(3.0, 4.7))
    yield S, srv, cat
    srv.close()


def run_alone(S, srv, reqs, mode):
    outs = []
    for i in range(len(reqs)):
        one = S.copy_requests(reqs[i:i + 1])
        one[0].arrival_ns = 0
        srv.run(one, mode)
        outs.append(S.outcome(one[0]))
    return outs


def test_alone_equals_mix_equals_plain(served):
    S, srv, cat = served
    reqs = S.make_requests(cat, 48, rate=20000.0, seed=7)
    st = srv.run(reqs, S.MODE_CONTINUOUS)
    mix = [S.outcome(r) for r in reqs]
    assert all(m[0] == 0 for m in mix), mix
    assert st.peak_inflight > 1
    assert run_alone(S, srv, reqs, S.MODE_CONTINUOUS) == mix
    plain = S.copy_requests(reqs)
    srv.run(plain, S.MODE_PLAIN)
    assert [S.outcome(r) for r in plain] == mix
    for r in reqs:
        assert r.admit_ns >= r.arrival_ns and r.done_ns >= r.admit_ns


def test_budget_defers_admission_without_changing_results(served, rf):
    S, srv, cat = served
    reqs = S.make_requests(cat, 32, rate=None, seed=8)
    need = max(srv.request_bytes(r) for r in reqs)
    # The same resident relations behind a budget of about one request.
    small = S.Server(0, slots=8, budget_bytes=need + (1 << 20))
    cat.register(small)
    mixed = S.copy_requests(reqs)
    st = small.run(mixed, S.MODE_CONTINUOUS)
    assert st.blocked_on_memory > 0
    assert st.peak_bytes <= st.budget_bytes
    ref = S.copy_requests(reqs)
    srv.run(ref, S.MODE_CONTINUOUS)
    assert [S.outcome(r) for r in mixed] == [S.outcome(r) for r in ref]
    small.close()


def expected_select_project(sch, constant):
    t = sch["bufs"][0].download(np.uint64, 2 * sch["sizes"][0]).reshape(-1, 2)
    lo, hi = t[:, 0], t[:, 1]
    qty = hi & np.uint64(0xFFFFFFFF)
    keep = qty < np.uint64(constant)
    return [((hi[keep] >> np.uint64(32)) << np.uint64(32)) | (lo[keep] & np.uint64(0xFFFFFFFF))]


def join_pairs(lk, rk):
    """Row pairs of a sort-merge join of sorted key arrays, in the order the
    expansion writes them: left rows in order, each left row's matches in
    right order."""
    lo = np.searchsorted(rk, lk, "left")
    cnt = np.searchsorted(rk, lk, "right") - lo
    li = np.repeat(np.arange(len(lk)), cnt)
    start = np.cumsum(cnt) - cnt
    ri = lo[li] + (np.arange(len(li)) - start[li])
    return li, ri


def expected_join_agg_cols(fk, m, dk, g):
    of, od = np.argsort(fk, kind="stable"), np.argsort(dk, kind="stable")
    li, ri = join_pairs(fk[of], dk[od])
    gf, mf = g[od][ri], m[of][li]
    keys = np.unique(gf)
    sums = np.array([math.fsum(mf[gf == k].tolist()) for k in keys])
    counts = np.array([(gf == k).sum() for k in keys], dtype=np.uint32)
    return [keys, sums, counts]


def expected_join_agg(sch):
    fk_b, m_b, dk_b, g_b = sch["bufs"]
    nf, nd = sch["sizes"]
    return expected_join_agg_cols(fk_b.download(np.uint32, nf), m_b.download(np.float64, nf),
                                  dk_b.download(np.uint32, nd), g_b.download(np.uint32, nd))


def expected_join3_cols(fk1, m, sk, fk2, tk, tv):
    o1, os_, ot = (np.argsort(x, kind="stable") for x in (fk1, sk, tk))
    li, ri = join_pairs(fk1[o1], sk[os_])
    k2, v2 = fk2[os_][ri], m[o1][li]
    o2 = np.argsort(k2, kind="stable")
    k2, v2 = k2[o2], v2[o2]
    li, ri = join_pairs(k2, tk[ot])
    return [tv[ot][ri], v2[li]]


def expected_join3(sch):
    fk1_b, m_b, sk_b, fk2_b, tk_b, tv_b = sch["bufs"]
    nf, ns, nt = sch["sizes"]
    return expected_join3_cols(fk1_b.download(np.uint32, nf), m_b.download(np.float64, nf),
                               sk_b.download(np.uint32, ns), fk2_b.download(np.uint32, ns),
                               tk_b.download(np.uint32, nt), tv_b.download(np.uint32, nt))


def expected_set(sch, op):
    a = sch["bufs"][0].download(np.uint64, sch["sizes"][0])
    b = sch["bufs"][1].download(np.uint64, sch["sizes"][1])
    f = {0: np.union1d, 1: np.setdiff1d, 2: np.intersect1d}[op]
    return [f(a, b)]


def test_outputs_match_numpy(served, rf):
    S, srv, cat = served
    widths = {0: [8], 1: [4, 8, 4], 2: [4, 8], 3: [8]}
    dtypes = {4: np.uint32, 8: np.uint64}
    cases = []
    for p in range(4):
        for sch in cat.schemas[p]:
            params = [(0, 0)]
            if p == 0:
                params = [(2, 1), (2, 40), (2, 64)]
            if p == 3:
                params = [(0, 0), (1, 0), (2, 0)]
            for prm, const in params:
                cases.append((p, sch, prm, const))
    reqs = (S.Request * len(cases))()
    sinks = []
    for i, (p, sch, prm, const) in enumerate(cases):
        sizes = sch["sizes"]
        cap = {0: sizes[0], 1: 64, 2: sizes[0],
               3: (sizes[0] + sizes[1]) if prm == 0 else sizes[0]}[p]
        nbytes = cap * sum(widths[p])
        sink = rf.DeviceBuffer(max(nbytes, 8))
        sinks.append((sink, cap))
        reqs[i].schema, reqs[i].param, reqs[i].constant = sch["id"], prm, const
        reqs[i].sink, reqs[i].sink_bytes = sink.ptr, nbytes
    srv.run(reqs, S.MODE_CONTINUOUS)
    for i, (p, sch, prm, const) in enumerate(cases):
        r = reqs[i]
        assert r.status == 0
        sink, cap = sinks[i]
        exp = {0: lambda: expected_select_project(sch, const), 1: lambda: expected_join_agg(sch),
               2: lambda: expected_join3(sch), 3: lambda: expected_set(sch, prm)}[p]()
        assert r.rows == len(exp[0]), (p, r.rows, len(exp[0]))
        off = 0
        for c, (w, e) in enumerate(zip(widths[p], exp)):
            is_f64 = e.dtype == np.float64
            got = sink.download(np.float64 if is_f64 else dtypes[w], r.rows, offset=off)
            if is_f64 and p == 1:
                np.testing.assert_allclose(got, e, rtol=1e-12, atol=0)
            else:
                assert np.array_equal(got, e), (p, c)
            words = got.view(np.uint64) if is_f64 else got.astype(np.uint64)
            assert S.digest_host(words, c) == r.digest[c], (p, c)
            off += cap * w


def test_request_errors(served, rf):
    S, srv, cat = served
    bad = (S.Request * 1)()
    bad[0].schema = 999
    with pytest.raises(rf.RfxError):
        srv.run(bad)
    # A union of two 10^5-row sets needs more than a 1 MiB budget holds.
    n = 100000
    a, b = rf.DeviceBuffer(8 * n), rf.DeviceBuffer(8 * n)
    a.zero()
    b.zero()
    rels = [S.Relation(), S.Relation()]
    for r, buf in zip(rels, (a, b)):
        r.col[0], r.rows = buf.ptr, n
    tiny = S.Server(0, slots=1, budget_bytes=1 << 20)
    big = (S.Request * 1)()
    big[0].schema = tiny.add_schema(3, rels)
    big[0].param = 0
    assert tiny.request_bytes(big[0]) > (1 << 20)
    with pytest.raises(rf.RfxError, match="budget"):
        tiny.run(big)
    tiny.close()


def check_sink(sink, rows, exp, widths, cap, rtol):
    off = 0
    for w, e in zip(widths, exp):
        is_f64 = e.dtype == np.float64
        got = sink.download(np.float64 if is_f64 else {4: np.uint32, 8: np.uint64}[w], rows,
                            offset=off)
        if is_f64 and rtol:
            np.testing.assert_allclose(got, e, rtol=rtol, atol=0)
        else:
            assert np.array_equal(got, e)
        off += cap * w


def test_many_to_many_joins_match_numpy(served, rf, ctx):
    """Dimension keys that repeat (a quarter of them twice) or are missing
    (the most frequent Zipf rank): each fact row joins every matching row.
    The join sizes are measured when the relations are registered, and
    every form of the request gives the numpy result."""
    S, srv, cat = served
    sch = cat.schemas[1][0]
    nf, nd = sch["sizes"]
    fk_b, m_b, dk_b, g_b = sch["bufs"]
    fk, m = fk_b.download(np.uint32, nf), m_b.download(np.float64, nf)
    dk, g = dk_b.download(np.uint32, nd), g_b.download(np.uint32, nd)
    q = nd // 4
    dk2, g2 = np.concatenate([dk[1:], dk[:q]]), np.concatenate([g[1:], g[:q]])
    bd, bg = rf.DeviceBuffer.from_numpy(dk2), rf.DeviceBuffer.from_numpy(g2)
    r1 = S.Relation()
    r1.col[0], r1.col[1] = bd.ptr, bg.ptr
    r1.rows, r1.key_bits, r1.val_bits = len(dk2), sch["rels"][1].key_bits, sch["rels"][1].val_bits
    sid_a = srv.add_schema(1, [sch["rels"][0], r1])
    exp_a = expected_join_agg_cols(fk, m, dk2, g2)

    s3 = cat.schemas[2][0]
    nf3, ns, nt = s3["sizes"]
    fk1_b, m3_b, sk_b, fk2_b, tk_b, tv_b = s3["bufs"]
    sk, fk2 = sk_b.download(np.uint32, ns), fk2_b.download(np.uint32, ns)
    tk, tv = tk_b.download(np.uint32, nt), tv_b.download(np.uint32, nt)
    qs, qt = ns // 3, nt // 5
    sk2, fk22 = np.concatenate([sk[1:], sk[:qs]]), np.concatenate([fk2[1:], fk2[:qs]])
    tk2, tv2 = np.concatenate([tk, tk[:qt]]), np.concatenate([tv, tv[:qt] ^ np.uint32(0xA5A5)])
    bufs3 = [rf.DeviceBuffer.from_numpy(x) for x in (sk2, fk22, tk2, tv2)]
    rs, rt = S.Relation(), S.Relation()
    rs.col[0], rs.col[1] = bufs3[0].ptr, bufs3[1].ptr
    rs.rows, rs.key_bits, rs.val_bits = len(sk2), s3["rels"][1].key_bits, s3["rels"][1].val_bits
    rt.col[0], rt.col[1] = bufs3[2].ptr, bufs3[3].ptr
    rt.rows, rt.key_bits = len(tk2), s3["rels"][2].key_bits
    sid_3 = srv.add_schema(2, [s3["rels"][0], rs, rt])
    exp_3 = expected_join3_cols(fk1_b.download(np.uint32, nf3), m3_b.download(np.float64, nf3),
                                sk2, fk22, tk2, tv2)
    assert len(exp_3[0]) != nf3  # not a 1:1 chain

    widths_a, widths_3 = [4, 8, 4], [4, 8]
    cap_a, cap_3 = 64, len(exp_3[0])
    for mode in (S.MODE_CONTINUOUS, S.MODE_PLAIN):
        sinks = [rf.DeviceBuffer(cap_a * 16), rf.DeviceBuffer(max(cap_3 * 12, 8))]
        req = (S.Request * 2)()
        for r, sid, sink in zip(req, (sid_a, sid_3), sinks):
            r.schema, r.sink, r.sink_bytes = sid, sink.ptr, sink.nbytes
        srv.run(req, mode)
        assert req[0].status == 0 and req[1].status == 0, (req[0].status, req[1].status)
        assert req[0].rows == len(exp_a[0]) and req[1].rows == cap_3
        check_sink(sinks[0], req[0].rows, exp_a, widths_a, cap_a, 1e-12)
        check_sink(sinks[1], req[1].rows, exp_3, widths_3, cap_3, 0)
        for sink in sinks:
            sink.close()
    for b in [bd, bg] + bufs3:
        b.close()
