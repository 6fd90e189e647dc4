"""Device data generators agree with the host generators the tests use, and
the per-launch profiler records the operators' kernels."""
import numpy as np
import pytest

from redfox_amd import datagen

pytestmark = pytest.mark.gpu


def test_u32_u64_streams_match_host(rf, ctx):
    n = 100_003
    a = rf.DeviceBuffer(4 * n)
    b = rf.DeviceBuffer(8 * n)
    ctx.gen_u32(a, n, 12345)
    ctx.gen_u64(b, n, 999)
    ctx.synchronize()
    assert np.array_equal(a.download(np.uint32, n), datagen.u32(12345, n))
    assert np.array_equal(b.download(np.uint64, n), datagen.splitmix64(999, n))


def test_lineitem_matches_host_bit_exact(rf, ctx):
    n = 250_001
    cols = [rf.DeviceBuffer(4 * n), rf.DeviceBuffer(n), rf.DeviceBuffer(n), rf.DeviceBuffer(4 * n),
            rf.DeviceBuffer(8 * n), rf.DeviceBuffer(8 * n), rf.DeviceBuffer(8 * n)]
    ctx.gen_lineitem(cols, n, 77)
    ctx.synchronize()
    host = datagen.lineitem(77, n)
    dts = [np.int32, np.uint8, np.uint8, np.int32, np.float64, np.float64, np.float64]
    for buf, want, dt in zip(cols, host, dts):
        got = buf.download(dt, n)
        assert np.array_equal(got.view(np.uint8), want.astype(dt).view(np.uint8))


def test_zipf_keys_are_skewed(rf, ctx):
    n = 1 << 20
    b = rf.DeviceBuffer(8 * n)
    ctx.gen_zipf_u64(b, n, 5, 1.1, 1 << 32)
    ctx.synchronize()
    k = b.download(np.uint64, n)
    vals, counts = np.unique(k, return_counts=True)
    order = np.argsort(counts)[::-1]
    # rank 1 of Zipf(1.1) over 2^32 holds a few percent of the mass and
    # outnumbers rank 2 by about 2^1.1; keys are mix64(rank)
    assert counts[order[0]] > 0.02 * n
    assert 1.6 < counts[order[0]] / counts[order[1]] < 2.7
    assert vals[order[0]] == datagen.mix64(np.array([1], dtype=np.uint64))[0]
    assert vals[order[1]] == datagen.mix64(np.array([2], dtype=np.uint64))[0]


def test_profiler_counts_launches(rf, ctx):
    n = 1 << 20
    k = rf.DeviceBuffer(4 * n)
    v = rf.DeviceBuffer(4 * n)
    ctx.gen_u32(k, n, 1)
    ctx.iota_u32(v, n)
    ctx.profile(True)
    ctx.sort_pairs(k, v, n, 4, 32, 4)
    launches, ms = ctx.profile_read("sort.onesweep")
    hist = ctx.profile_read("sort.hist")
    ctx.profile(False)
    assert launches == 4 and ms > 0
    assert hist[0] == 1
