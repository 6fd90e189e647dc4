"""Set operations and unique on the GPU vs the oracle (bit-exact), with
multiset duplicates, every element width, unsigned int64 compares and
empty inputs."""
import numpy as np
import pytest

from redfox_amd import datagen

pytestmark = pytest.mark.gpu

DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64, 16: np.uint64}
OPS = [0, 1, 2]


def sorted_elems(orc, eb, n, seed, rng):
    v = datagen.splitmix64(seed, n) % np.uint64(max(rng, 1))
    if eb == 16:
        e = np.empty(2 * n, dtype=np.uint64)
        e[0::2] = datagen.mix64(v) & np.uint64(3)
        e[1::2] = v
    else:
        e = v.astype(DT[eb])
    return orc.sort_pairs(e, None, eb, 8 * eb, 0)[0]


def gpu_set(rf, ctx, op, a, b, eb):
    na, nb = a.nbytes // eb, b.nbytes // eb
    A = rf.DeviceBuffer.from_numpy(a)
    B = rf.DeviceBuffer.from_numpy(b)
    O = rf.DeviceBuffer(max((na + nb) * eb, 1))
    cnt = rf.u64_cell(12345)
    ctx.set_op(op, O, cnt, A, na, B, nb, eb)
    ctx.synchronize()
    m = rf.read_u64(cnt)
    return O.download(a.dtype, m * eb // a.itemsize)


@pytest.mark.parametrize("eb", [1, 2, 4, 8, 16])
@pytest.mark.parametrize("op", OPS)
def test_multisets(rf, ctx, orc, op, eb):
    rng = {1: 100, 2: 5000, 4: 50_000, 8: 50_000, 16: 50_000}[eb]
    a = sorted_elems(orc, eb, 60_000, 1 + eb, rng)
    b = sorted_elems(orc, eb, 45_000, 2 + eb, rng)
    assert np.array_equal(gpu_set(rf, ctx, op, a, b, eb), orc.set_op(op, a, b, eb))


@pytest.mark.parametrize("op", OPS)
def test_sorted_unique_int64_sets(rf, ctx, orc, op):
    """Config-5 form: sorted unique int64 sets compared as u64."""
    n = 1 << 20
    a = np.unique(datagen.splitmix64(101, n))
    b = np.unique(np.concatenate([a[::3], datagen.splitmix64(102, n // 2)]))
    assert np.array_equal(gpu_set(rf, ctx, op, a, b, 8), orc.set_op(op, a, b, 8))


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("na,nb", [(0, 0), (0, 10), (10, 0), (1, 1), (3000, 5), (5, 3000)])
def test_edge_sizes(rf, ctx, orc, op, na, nb):
    a = sorted_elems(orc, 4, na, 7, 20)
    b = sorted_elems(orc, 4, nb, 8, 20)
    assert np.array_equal(gpu_set(rf, ctx, op, a, b, 4), orc.set_op(op, a, b, 4))


def repeated(eb, values, copies):
    v = np.repeat(np.asarray(values, dtype=np.uint64), copies)
    if eb == 16:  # high word holds the value, low word 0
        e = np.zeros(2 * v.size, dtype=np.uint64)
        e[1::2] = v
        return e
    return v.astype(DT[eb])


# Runs longer than a tile (3072-6144 items) on both inputs: ranks and
# counts continue across tile edges in both directions.
@pytest.mark.parametrize("eb", [4, 8, 16])
@pytest.mark.parametrize("op", OPS)
def test_heavy_duplicates_across_tiles(rf, ctx, orc, op, eb):
    a = repeated(eb, range(5), 7000)
    b = repeated(eb, range(2, 8), 4000)
    assert np.array_equal(gpu_set(rf, ctx, op, a, b, eb), orc.set_op(op, a, b, eb))
    a = repeated(eb, [1, 2, 3], [20000, 3, 15000])
    b = repeated(eb, [1, 3, 4], [9000, 30000, 2])
    assert np.array_equal(gpu_set(rf, ctx, op, a, b, eb), orc.set_op(op, a, b, eb))


@pytest.mark.parametrize("eb", [1, 4, 8, 16])
def test_unique(rf, ctx, orc, eb):
    a = sorted_elems(orc, eb, 200_000, 40 + eb, 3000)
    n = a.nbytes // eb
    A = rf.DeviceBuffer.from_numpy(a)
    O = rf.DeviceBuffer(a.nbytes)
    cnt = rf.u64_cell()
    ctx.unique(O, cnt, A, n, eb)
    ctx.synchronize()
    m = rf.read_u64(cnt)
    assert np.array_equal(O.download(a.dtype, m * eb // a.itemsize), orc.unique(a, eb))
    # in place
    ctx.unique(A, cnt, A, n, eb)
    ctx.synchronize()
    assert np.array_equal(A.download(a.dtype, m * eb // a.itemsize), orc.unique(a, eb))
