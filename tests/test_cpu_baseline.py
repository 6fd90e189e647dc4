"""The reported CPU baseline (oracle/rfx_cpu.cpp, OpenMP) computes what the
serial oracle computes: integer and byte results bit-exact, FP64 sums
within 1e-12 relative, on one thread and on several.  CPU only."""
import numpy as np
import pytest

from redfox_amd import datagen

REL = 1e-12


@pytest.fixture(scope="module")
def cb():
    from oracle import cpu_baseline

    cpu_baseline.lib()
    return cpu_baseline


@pytest.fixture(params=[1, 4])
def nthreads(cb, request):
    with cb.threads(request.param) as got:
        assert got == request.param
        yield got


def test_sort_pairs_u32_and_u64(cb, orc, nthreads):
    n = 100_003
    k32 = datagen.u32(11, n)
    v = np.arange(n, dtype=np.uint32)
    _, k, vv = cb.sort_pairs(k32, v, reps=1)
    ok, ov = orc.sort_pairs(k32, v, 4, 32, 4)
    assert np.array_equal(k, ok) and np.array_equal(vv, ov)
    k64 = datagen.splitmix64(12, n) % np.uint64(5000) * np.uint64(0x9E3779B97F4A7C15)
    _, k, vv = cb.sort_pairs(k64, v, reps=1)
    ok, ov = orc.sort_pairs(k64, v, 8, 64, 4)
    assert np.array_equal(k, ok) and np.array_equal(vv, ov)


def test_join_matches_oracle(cb, orc, nthreads):
    rk, rv, sk, sv = datagen.join_tables(13, 20_000, 80_000, key_bits=14)
    rk, rv = orc.sort_pairs(rk, rv, 4, 32, 4)
    sk, sv = orc.sort_pairs(sk, sv, 4, 32, 4)
    _, got = cb.join(rk, rv, sk, sv, reps=1)
    lb, lc, tot = orc.join_find_bounds(rk, sk, 4)
    li, ri = orc.join_main(lb, lc, tot)
    want = orc.join_gather(16, rk, 4, rv, 4, sv, 4, li, ri, 64, 32)
    assert np.array_equal(got, want)


def test_reduce_sum_count(cb, orc, nthreads):
    n = 200_001
    keys = np.sort((datagen.splitmix64(14, n) % np.uint64(3000)).astype(np.uint32))
    keys[50_000:120_000] = keys[50_000]  # one run across several chunks
    keys = np.sort(keys)
    vals = datagen.uniform01(15, n) * 1e3 - 200.0
    _, k, s, c = cb.reduce_sum_count(keys, vals, reps=1)
    ok, ov = orc.reduce_by_key_f64(keys, vals, 4, 0)
    _, oc = orc.reduce_by_key_u32(keys, np.ones(n, np.uint32), 4)
    assert np.array_equal(k, ok) and np.array_equal(c, oc)
    np.testing.assert_allclose(s, ov, rtol=REL, atol=1e-9)


def test_q1(cb, orc, nthreads):
    cols = datagen.lineitem(16, 300_000)
    _, got = cb.q1(*cols, 2436, reps=1)
    want = orc.q1(*cols, 2436)
    np.testing.assert_allclose(got[:, :5], want[:, :5], rtol=REL, atol=0)
    assert np.array_equal(got[:, 5], want[:, 5])


def test_select16(cb, orc, nthreads):
    n = 100_000
    t = datagen.splitmix64(17, 2 * n)
    _, got = cb.select16(t, 1 << 31, reps=1)
    assert np.array_equal(got, orc.select(t, 16, 64, 32, 2, 1 << 31))


def test_mapfilter(cb, nthreads):
    cols = datagen.lineitem(18, 100_000)
    ship, price, disc, tax = cols[0], cols[4], cols[5], cols[6]
    _, dp, ch = cb.mapfilter(ship, price, disc, tax, 2436, reps=1)
    keep = ship <= 2436
    want_dp = price[keep] * (1 - disc[keep])
    assert np.array_equal(dp, want_dp) and np.array_equal(ch, want_dp * (1 + tax[keep]))


@pytest.mark.parametrize("op", [0, 1, 2])
def test_set_ops_with_duplicates(cb, orc, nthreads, op):
    a = np.sort(datagen.splitmix64(19, 50_000) % np.uint64(20_000))
    b = np.sort(datagen.splitmix64(20, 40_000) % np.uint64(20_000))
    _, got = cb.set_op(op, a, b, reps=1)
    assert np.array_equal(got, orc.set_op(op, a, b, 8))
