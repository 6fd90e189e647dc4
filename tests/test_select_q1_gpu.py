"""Select / project / split / merge and the fused Q1 aggregate on the GPU vs
the oracle.  Integer and byte outputs bit-exact; Q1 FP64 sums within 1e-12
relative of the serial compensated oracle."""
import numpy as np
import pytest

from redfox_amd import datagen

pytestmark = pytest.mark.gpu

DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64, 16: np.uint64}
REL = 1e-12


def rows_of(tb, n, seed):
    raw = datagen.splitmix64(seed, (n * tb + 7) // 8 + 1).view(np.uint8)[: n * tb]
    return raw.view(DT[tb]).copy()


def gpu_select(rf, ctx, rows, tb, shift, bits, op, k):
    n = rows.nbytes // tb
    I = rf.DeviceBuffer.from_numpy(rows)
    O = rf.DeviceBuffer(max(rows.nbytes, 1))
    cnt = rf.u64_cell(99)
    ctx.select(O, cnt, I, n, tb, shift, bits, op, k)
    ctx.synchronize()
    m = rf.read_u64(cnt)
    return O.download(rows.dtype, m * tb // rows.itemsize)


@pytest.mark.parametrize("tb", [1, 2, 4, 8, 16])
@pytest.mark.parametrize("op", [0, 1, 2, 3, 4, 5, 6, 7, 8, 9])
def test_select_integer_ops(rf, ctx, orc, tb, op):
    n = 70_001
    rows = rows_of(tb, n, 3 + tb)
    bits = min(8 * tb, 16)
    shift = 8 * tb - bits
    k = int(datagen.splitmix64(tb + op, 1)[0]) & ((1 << bits) - 1)
    if op in (0, 1):  # eq/ne: a value that occurs (row 5's field)
        row5 = int(rows[10]) | (int(rows[11]) << 64) if tb == 16 else int(rows[5])
        k = (row5 >> shift) & ((1 << bits) - 1)
    got = gpu_select(rf, ctx, rows, tb, shift, bits, op, k)
    want = orc.select(rows, tb, shift, bits, op, k)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("op", [10, 11, 12, 13])
def test_select_double_ops(rf, ctx, orc, op):
    n = 50_000
    d = np.round(datagen.uniform01(7, n) * 10.0, 6)
    rows = d.view(np.uint64).copy()
    k = int(np.array([5.0]).view(np.uint64)[0])
    got = gpu_select(rf, ctx, rows, 8, 0, 64, op, k)
    assert np.array_equal(got, orc.select(rows, 8, 0, 64, op, k))


# Tiles hold 512 x 24 rows of 8 and 16 bytes and 512 x 32 rows of narrower
# tuples: sizes around one and two tiles, and a prime size spanning many.
@pytest.mark.parametrize("tb,n", [(8, 0), (8, 1), (8, 12287), (8, 12288), (8, 12289),
                                  (8, 24577), (8, 1_000_003), (16, 1), (16, 12287), (16, 12288),
                                  (16, 12289), (16, 24577), (16, 1_000_003), (4, 16383),
                                  (4, 16384), (4, 16385), (4, 32769), (2, 16385), (1, 16383)])
def test_select_sizes_keep_order(rf, ctx, orc, tb, n):
    rows = rows_of(tb, n, 17)
    shift = 8 * tb - 24  # a 16-bit field below the top byte
    got = gpu_select(rf, ctx, rows, tb, shift, 16, 3, 30_000)
    assert np.array_equal(got, orc.select(rows, tb, shift, 16, 3, 30_000))


@pytest.mark.parametrize("tb", [4, 8, 16])
@pytest.mark.parametrize("n", [12288, 16384, 100_003])
def test_select_keeps_all_or_none(rf, ctx, orc, tb, n):
    rows = rows_of(tb, n, 18)
    shift = 8 * tb - 16
    for k, kept in ((0xFFFF, n), (0, 0)):  # field <= 0xFFFF always, field < 0 never
        op = 3 if kept else 2
        got = gpu_select(rf, ctx, rows, tb, shift, 16, op, k)
        want = orc.select(rows, tb, shift, 16, op, k)
        assert len(want) * want.itemsize // tb == kept
        assert np.array_equal(got, want)



@pytest.mark.parametrize("vt", [8, 16, 24, 32])
@pytest.mark.parametrize("tb", [1, 4, 8, 16])
def test_select_rows_per_thread_knob(rf, ctx, orc, tb, vt):
    """Every select.vt tile shape keeps the oracle's rows in order: a
    ragged last tile, many tiles, and a field in the tuple's top bits."""
    n = 300_007
    rows = rows_of(tb, n, 40 + tb)
    bits = min(8 * tb, 16)
    shift = 8 * tb - bits
    try:
        ctx.tune("select.vt", vt)
        got = gpu_select(rf, ctx, rows, tb, shift, bits, 2, 1 << (bits - 1))
    finally:
        ctx.tune("select.vt", 0)
    assert np.array_equal(got, orc.select(rows, tb, shift, bits, 2, 1 << (bits - 1)))

def test_project_split_merge(rf, ctx, orc):
    n = 100_000
    rows = rows_of(8, n, 23)
    I = rf.DeviceBuffer.from_numpy(rows)
    # project: swap the two 32-bit fields
    O = rf.DeviceBuffer(8 * n)
    ctx.project(O, 8, I, 8, n, [0, 32], [32, 32], [32, 0])
    # split a 96-bit (key 32 | r 32 | s 32) tuple: key = top field
    t16 = rows_of(16, n, 24)
    t16[1::2] &= np.uint64(0xFFFFFFFF)
    T = rf.DeviceBuffer.from_numpy(t16)
    K = rf.DeviceBuffer(4 * n)
    V = rf.DeviceBuffer(8 * n)
    ctx.split(K, 4, V, 8, T, 16, n, 64)
    M = rf.DeviceBuffer(16 * n)
    ctx.merge(M, 16, K, 4, V, 8, n, 64)
    ctx.synchronize()
    assert np.array_equal(O.download(np.uint64, n), orc.project(8, rows, 8, [0, 32], [32, 32], [32, 0]))
    ok, ov = orc.split(t16, 16, 4, 8, 64)
    assert np.array_equal(K.download(np.uint32, n), ok)
    assert np.array_equal(V.download(np.uint64, n), ov)
    assert np.array_equal(M.download(np.uint64, 2 * n), t16)
    assert np.array_equal(M.download(np.uint64, 2 * n), orc.merge(16, ok, 4, ov, 8, 64))


def run_q1(rf, ctx, cols, cutoff):
    n = len(cols[0])
    bufs = [rf.DeviceBuffer.from_numpy(c) for c in cols]
    out = rf.DeviceBuffer(48 * 8)
    ctx.q1(out, *bufs, n, cutoff)
    ctx.synchronize()
    return out.download(np.float64, 48).reshape(6, 8)


@pytest.mark.parametrize("n", [1, 3, 1000, 1_000_003])
def test_q1_matches_oracle(rf, ctx, orc, n):
    cols = datagen.lineitem(datagen.BASE_SEED + 2, n)
    got = run_q1(rf, ctx, cols, 2436)
    want = orc.q1(*cols, 2436)
    np.testing.assert_allclose(got[:, :5], want[:, :5], rtol=REL, atol=0)
    assert np.array_equal(got[:, 5], want[:, 5])
    c = np.where(want[:, 5] > 0, want[:, 5], 1)
    np.testing.assert_allclose(got[:, 6], np.where(want[:, 5] > 0, want[:, 0] / c, 0), rtol=REL)
    np.testing.assert_allclose(got[:, 7], np.where(want[:, 5] > 0, want[:, 1] / c, 0), rtol=REL)


def test_q1_full_size_properties(rf, ctx):
    """60M rows: counts per group sum to the rows kept by the filter, and a
    second run gives identical bits."""
    n = 60_000_000
    cols = datagen.lineitem(datagen.BASE_SEED + 2, n)
    got = run_q1(rf, ctx, cols, 2436)
    again = run_q1(rf, ctx, cols, 2436)
    assert np.array_equal(got.view(np.uint64), again.view(np.uint64))
    assert int(got[:, 5].sum()) == int((cols[0] <= 2436).sum())
    # sum_qty is exact in FP64 (integers below 2^53)
    keep = cols[0] <= 2436
    g = (cols[1].astype(np.int64) * 2 + cols[2])[keep]
    q = cols[3][keep].astype(np.int64)
    for grp in range(6):
        assert got[grp, 0] == float(q[g == grp].sum())
