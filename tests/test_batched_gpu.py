"""Batched join-aggregate instances: the captured one-graph plan gives the
same per-instance results as the plain per-operator execution and as a
direct numpy evaluation of the query (keys and counts exact, FP64 sums
within 1e-12 relative)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def expected(w, fk, measure):
    out = []
    for b in range(w.n_inst):
        r0, r1 = int(w.row_off[b]), int(w.row_off[b + 1])
        d0, d1 = int(w.dim_off[b]), int(w.dim_off[b + 1])
        grp_of = np.empty(d1 - d0, dtype=np.uint32)
        grp_of[w.dkey_host[d0:d1]] = w.grp_host[d0:d1]
        g = grp_of[fk[r0:r1]]
        keys = np.unique(g)
        sums = np.array([math.fsum(measure[r0:r1][g == k].tolist()) for k in keys])
        counts = np.array([(g == k).sum() for k in keys], dtype=np.uint32)
        out.append((keys, sums, counts))
    return out


def check(a, b):
    assert len(a) == len(b)
    for (ka, sa, ca), (kb, sb, cb) in zip(a, b):
        assert np.array_equal(ka, kb)
        assert np.array_equal(ca, cb)
        np.testing.assert_allclose(sa, sb, rtol=1e-12, atol=0)


def test_batched_equals_plain_equals_numpy(rf, ctx):
    from redfox_amd import batched

    w = batched.Workload(ctx, n_inst=48, seed=11, fact_exp=(2.0, 5.0), dim_exp=(1.0, 3.5))
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
    br.capture()
    for _ in range(3):  # replays give the same results
        br.run()
        br.synchronize()
        check(br.results(), want)
    br.run_eager()
    br.synchronize()
    check(br.results(), want)


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
