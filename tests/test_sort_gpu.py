"""Radix sort parity on the GPU: rfx_sort_pairs vs the oracle's stable LSD
sort (bit-exact keys and values), plus full-size properties."""
import numpy as np
import pytest

from redfox_amd import datagen

pytestmark = pytest.mark.gpu

DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64, 16: np.uint64}


def make(width, n, seed):
    raw = datagen.splitmix64(seed, (n * width + 7) // 8 + 1).view(np.uint8)
    return np.ascontiguousarray(raw[: n * width]).view(DT[width]).copy()


def run_sort(rf, ctx, keys, vals, kb, bits, vb):
    n = keys.nbytes // kb
    kbuf = rf.DeviceBuffer.from_numpy(keys)
    vbuf = rf.DeviceBuffer.from_numpy(vals) if vb else None
    ctx.sort_pairs(kbuf, vbuf, n, kb, bits, vb)
    ctx.synchronize()
    k = kbuf.download(keys.dtype, keys.size)
    v = vbuf.download(vals.dtype, vals.size) if vb else None
    return k, v


WIDTHS = [(kb, vb) for kb in (1, 2, 4, 8, 16) for vb in (0, 1, 2, 4, 8, 16)]


@pytest.mark.parametrize("kb,vb", WIDTHS)
def test_widths_match_oracle(rf, ctx, orc, kb, vb):
    n = 50_003
    keys = make(kb, n, 11 + kb)
    vals = make(vb, n, 97 + vb) if vb else None
    bits = 8 * kb
    gk, gv = run_sort(rf, ctx, keys, vals, kb, bits, vb)
    ok, ov = orc.sort_pairs(keys, vals, kb, bits, vb)
    assert np.array_equal(gk, ok)
    if vb:
        assert np.array_equal(gv, ov)


# key_bits values the reference dispatches on (b40cSort.cu:85-826) and the
# widths it keeps them in; high bits above key_bits are random and ignored.
@pytest.mark.parametrize("kb,bits", [(1, 2), (1, 3), (1, 5), (2, 14), (4, 18), (4, 23),
                                     (4, 26), (4, 32), (8, 37), (8, 64), (16, 100)])
def test_key_bits_low_bits_only(rf, ctx, orc, kb, bits):
    n = 70_001
    keys = make(kb, n, 5 + bits)
    vals = np.arange(n, dtype=np.uint32)
    gk, gv = run_sort(rf, ctx, keys, vals, kb, bits, 4)
    ok, ov = orc.sort_pairs(keys, vals, kb, bits, 4)
    assert np.array_equal(gk, ok)
    assert np.array_equal(gv, ov)


@pytest.mark.parametrize("n", [0, 1, 2, 3, 63, 64, 65, 4095, 4096, 4097, 8191, 8193,
                               100_003, (1 << 20) + 7])
def test_sizes(rf, ctx, orc, n):
    keys = make(4, n, 1000 + n)
    vals = np.arange(n, dtype=np.uint32)
    gk, gv = run_sort(rf, ctx, keys, vals, 4, 32, 4)
    ok, ov = orc.sort_pairs(keys, vals, 4, 32, 4)
    assert np.array_equal(gk, ok)
    assert np.array_equal(gv, ov)


def test_all_equal_keys_skip_every_pass(rf, ctx, orc):
    n = 30_000
    keys = np.full(n, 0xDEADBEEF, dtype=np.uint32)
    vals = np.arange(n, dtype=np.uint32)[::-1].copy()
    gk, gv = run_sort(rf, ctx, keys, vals, 4, 32, 4)
    assert np.array_equal(gk, keys) and np.array_equal(gv, vals)


def test_odd_number_of_passes_copies_back(rf, ctx, orc):
    # keys < 2^16 in u32 with 24 key bits: pass 2 is an identity and is
    # skipped, leaving an odd number of passes.
    n = 123_457
    keys = (make(4, n, 3) & np.uint32(0xFFFF)).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    gk, gv = run_sort(rf, ctx, keys, vals, 4, 24, 4)
    ok, ov = orc.sort_pairs(keys, vals, 4, 24, 4)
    assert np.array_equal(gk, ok) and np.array_equal(gv, ov)


def test_zipf_u64_keys(rf, ctx, orc):
    n = 200_000
    ranks = datagen.zipf_ranks(datagen.BASE_SEED + 2, n, 1.1, 1 << 32)
    keys = datagen.mix64(ranks)
    vals = np.arange(n, dtype=np.uint32)
    gk, gv = run_sort(rf, ctx, keys, vals, 8, 64, 4)
    ok, ov = orc.sort_pairs(keys, vals, 8, 64, 4)
    assert np.array_equal(gk, ok) and np.array_equal(gv, ov)


def test_repeat_calls_reuse_workspace(rf, ctx, orc):
    for n in (300_000, 1000, 77_777):
        keys = make(4, n, n)
        vals = np.arange(n, dtype=np.uint32)
        gk, gv = run_sort(rf, ctx, keys, vals, 4, 32, 4)
        ok, ov = orc.sort_pairs(keys, vals, 4, 32, 4)
        assert np.array_equal(gk, ok) and np.array_equal(gv, ov)


def test_full_size_u32_pairs_properties(rf, ctx):
    """2^28 (u32 key, u32 row index) pairs: the result is sorted, stable
    (indices ascend within equal keys) and a permutation of the input
    (each key equals the input key at its index)."""
    n = 1 << 28
    seed = datagen.BASE_SEED + 1
    kbuf = rf.DeviceBuffer(4 * n)
    vbuf = rf.DeviceBuffer(4 * n)
    ctx.gen_u32(kbuf, n, seed)
    ctx.iota_u32(vbuf, n)
    ctx.sort_pairs(kbuf, vbuf, n, 4, 32, 4)
    ctx.synchronize()
    k = kbuf.download(np.uint32, n)
    v = vbuf.download(np.uint32, n)
    kbuf.close()
    vbuf.close()
    assert np.all(k[1:] >= k[:-1])
    same = k[1:] == k[:-1]
    assert np.all(v[1:][same] > v[:-1][same])
    src = datagen.u32(seed, n)
    assert np.array_equal(src[v], k)


def test_full_size_u64_zipf_pairs_properties(rf, ctx):
    """2^27 (u64 Zipf(1.1) key, u32 row index) pairs, the size the sort is
    measured at: sorted, stable and a permutation of the input (the keys
    are generated on the device and copied before the sort)."""
    n = 1 << 27
    seed = datagen.BASE_SEED + 2
    k0 = rf.DeviceBuffer(8 * n)
    kbuf = rf.DeviceBuffer(8 * n)
    vbuf = rf.DeviceBuffer(4 * n)
    ctx.gen_zipf_u64(k0, n, seed, 1.1, 1 << 32)
    ctx.copy(kbuf, k0, 8 * n)
    ctx.iota_u32(vbuf, n)
    ctx.sort_pairs(kbuf, vbuf, n, 8, 64, 4)
    ctx.synchronize()
    src = k0.download(np.uint64, n)
    k = kbuf.download(np.uint64, n)
    v = vbuf.download(np.uint32, n)
    for b in (k0, kbuf, vbuf):
        b.close()
    assert np.all(k[1:] >= k[:-1])
    same = k[1:] == k[:-1]
    assert same.mean() > 0.3  # Zipf(1.1): many rows repeat a key
    assert np.all(v[1:][same] > v[:-1][same])
    assert np.array_equal(src[v], k)
    assert np.array_equal(np.sort(v), np.arange(n, dtype=np.uint32))


@pytest.mark.parametrize("digit", [8, 9, 10])
@pytest.mark.parametrize("ipt", [0, 12])
@pytest.mark.parametrize("kb,vb,bits", [(4, 0, 32), (4, 4, 18), (4, 4, 20), (4, 4, 26),
                                        (4, 8, 31), (8, 4, 64), (8, 0, 37), (8, 8, 45)])
def test_wider_digits_match_oracle(rf, ctx, orc, digit, ipt, kb, vb, bits):
    """sort.digit 8, 9 and 10 forced over the width chosen from the key bits
    (9 for 17-18 bits, 10 for 19-20): digits that straddle the 32-bit halves
    of 64-bit keys, a narrow last digit, a skipped pass, a ragged last tile;
    keys and values bit-exact against the oracle."""
    n = 300_007
    keys = make(kb, n, 700 + bits)
    if bits == 26:
        keys &= DT[kb](0x3FFFC00)  # bits 0-9 zero: the first pass is skipped
    vals = make(vb, n, 800 + vb) if vb else None
    try:
        ctx.tune("sort.digit", digit)
        ctx.tune("sort.ipt", ipt)
        gk, gv = run_sort(rf, ctx, keys, vals, kb, bits, vb)
    finally:
        ctx.tune("sort.digit", 0)
        ctx.tune("sort.ipt", 0)
    ok, ov = orc.sort_pairs(keys, vals, kb, bits, vb)
    assert np.array_equal(gk, ok)
    if vb:
        assert np.array_equal(gv, ov)


@pytest.mark.parametrize("kb,vb,bits", [(4, 4, 32), (4, 0, 20), (8, 4, 64), (16, 8, 100), (2, 2, 16)])
def test_wide_status_words_match_oracle(rf, ctx, orc, kb, vb, bits):
    """The 64-bit look-back status words that sorts of more than 2^30 - 1
    rows take, forced on a small sort (knob sort.status = 64): keys and
    values bit-exact against the oracle."""
    n = 250_007
    keys = make(kb, n, 900 + bits)
    vals = make(vb, n, 901) if vb else None
    try:
        ctx.tune("sort.status", 64)
        gk, gv = run_sort(rf, ctx, keys, vals, kb, bits, vb)
    finally:
        ctx.tune("sort.status", 0)
    ok, ov = orc.sort_pairs(keys, vals, kb, bits, vb)
    assert np.array_equal(gk, ok)
    if vb:
        assert np.array_equal(gv, ov)


def test_more_than_2p30_rows_properties(rf, ctx):
    """2^30 + 12,345 (u32 key, u32 row index) pairs, past the 32-bit status
    words' 2^30 - 1 rows: sorted, stable and a permutation of the input."""
    n = (1 << 30) + 12_345
    k0 = rf.DeviceBuffer(4 * n)
    kbuf, vbuf = rf.DeviceBuffer(4 * n), rf.DeviceBuffer(4 * n)
    ctx.gen_u32(k0, n, datagen.BASE_SEED + 77)
    ctx.copy(kbuf, k0, 4 * n)
    ctx.iota_u32(vbuf, n)
    ctx.sort_pairs(kbuf, vbuf, n, 4, 32, 4)
    ctx.synchronize()
    src = k0.download(np.uint32, n)
    k0.close()
    k = kbuf.download(np.uint32, n)
    kbuf.close()
    v = vbuf.download(np.uint32, n)
    vbuf.close()
    assert np.all(k[1:] >= k[:-1])
    same = k[1:] == k[:-1]
    assert np.all(v[1:][same] > v[:-1][same])
    assert np.array_equal(src[v], k)
    seen = np.zeros(n, dtype=bool)
    seen[v] = True
    assert seen.all()
