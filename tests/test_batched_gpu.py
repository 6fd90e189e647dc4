"""Batched join-aggregate instances: the captured one-graph plan gives the
same per-instance results as the plain per-operator execution and as a
direct numpy evaluation of the query (keys and counts exact, FP64 sums
within 1e-12 relative)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def expected(w, fk, measure, which=None):
    """Per instance: the sort-merge join's row pairs (each fact row with
    every dimension row of its key), then the groups' fsum and count."""
    out = []
    for b in range(w.n_inst) if which is None else which:
        r0, r1 = int(w.row_off[b]), int(w.row_off[b + 1])
        d0, d1 = int(w.dim_off[b]), int(w.dim_off[b + 1])
        lk, rk = fk[r0:r1], w.dkey_host[d0:d1]
        ro = np.argsort(rk, kind="stable")
        rks = rk[ro]
        lo = np.searchsorted(rks, lk, "left")
        cnt = np.searchsorted(rks, lk, "right") - lo
        li = np.repeat(np.arange(len(lk)), cnt)
        ri = lo[li] + (np.arange(len(li)) - (np.cumsum(cnt) - cnt)[li])
        g = w.grp_host[d0:d1][ro][ri]
        m = measure[r0:r1][li]
        o = np.argsort(g, kind="stable")
        keys, first, counts = np.unique(g[o], return_index=True, return_counts=True)
        ms = m[o]
        ends = np.append(first[1:], len(o))
        sums = np.array([math.fsum(ms[a:e].tolist()) for a, e in zip(first, ends)])
        out.append((keys, sums, counts.astype(np.uint32)))
    return out


def check(a, b):
    assert len(a) == len(b)
    for (ka, sa, ca), (kb, sb, cb) in zip(a, b):
        assert np.array_equal(ka, kb)
        assert np.array_equal(ca, cb)
        np.testing.assert_allclose(sa, sb, rtol=1e-12, atol=0)


@pytest.mark.parametrize("dup", [0.0, 0.4])
def test_batched_equals_plain_equals_numpy(rf, ctx, dup):
    """dup = 0: every fact row joins one dimension row; dup = 0.4: 40% of
    each dimension's keys appear twice, so the join is many-to-many."""
    from redfox_amd import batched

    w = batched.Workload(ctx, n_inst=48, seed=11, fact_exp=(2.0, 5.0), dim_exp=(1.0, 3.5),
                         dup=dup)
    fk = w.fk.download(np.uint32, w.N)
    measure = w.measure.download(np.float64, w.N)
    # every fk references a dimension row
    for b in range(w.n_inst):
        d = int(w.d[b])
        assert int(fk[int(w.row_off[b]):int(w.row_off[b + 1])].max(initial=0)) < d
    want = expected(w, fk, measure)
    plain = batched.PlainRunner(w, ctx)
    plain.run()
    check(plain.results(), want)
    br = batched.BatchedRunner(w)
    assert (br.J == w.N) == (dup == 0.0)
    br.capture()
    for _ in range(3):  # replays give the same results
        br.run()
        br.synchronize()
        check(br.results(), want)
    br.run_eager()
    br.synchronize()
    check(br.results(), want)
    bu = batched.BatchedRunner(w, fused=False)  # the plain form's operators, batched
    bu.capture()
    bu.run()
    bu.synchronize()
    check(bu.results(), want)


def test_zipf_foreign_keys_are_skewed(rf, ctx):
    from redfox_amd import batched

    w = batched.Workload(ctx, n_inst=4, seed=3, fact_exp=(5.0, 5.0), dim_exp=(3.0, 3.0))
    fk = w.fk.download(np.uint32, w.N)
    b0 = fk[: int(w.n[0])]
    counts = np.bincount(b0, minlength=int(w.d[0]))
    top = np.sort(counts)[::-1]
    assert top[0] > 0.05 * len(b0)  # Zipf(1.1) over 1000 keys: rank 1 ~ 10%
    assert top[0] / top[1] > 1.5
    # the hottest key is the dimension's first key (rank 1 through dkey)
    assert int(np.argmax(counts)) == int(w.dkey_host[0])


def test_batched_512_instances_sample_against_numpy(rf, ctx):
    """The measured workload (512 instances, 5.6e8 fact rows): the graph
    replay's results for the largest and smallest instance and six more
    chosen at random equal numpy's."""
    from redfox_amd import batched

    w = batched.Workload(ctx, n_inst=512)
    br = batched.BatchedRunner(w)
    br.capture()
    br.run()
    br.synchronize()
    got = br.results()
    rng = np.random.default_rng(512)
    which = sorted({int(np.argmax(w.n)), int(np.argmin(w.n))} |
                   set(int(x) for x in rng.choice(w.n_inst, 6, replace=False)))
    fk = w.fk.download(np.uint32, w.N)
    measure = w.measure.download(np.float64, w.N)
    check([got[b] for b in which], expected(w, fk, measure, which))
