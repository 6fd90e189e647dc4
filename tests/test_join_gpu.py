"""Sort-merge join parity on the GPU: find_bounds, expansion and gather vs
the oracle (bit-exact), on the reference's config-1 shape and on edge cases
(duplicates, empty sides, no matches, skew, every key width)."""
import numpy as np
import pytest

from redfox_amd import datagen

pytestmark = pytest.mark.gpu

DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64, 16: np.uint64}


def sorted_keys(orc, width, n, seed, key_range):
    """n sorted keys of `width` bytes drawn from [0, key_range)."""
    v = datagen.splitmix64(seed, n) % np.uint64(max(1, key_range))
    if width == 16:
        # spread values into both words; the high word carries the order
        k = np.empty(2 * n, dtype=np.uint64)
        k[0::2] = datagen.mix64(v) & np.uint64(0xFF)
        k[1::2] = v
    else:
        k = v.astype(DT[width])
    k, _ = orc.sort_pairs(k, None, width, 8 * width, 0)
    return k


def gpu_join(rf, ctx, left, right, kb):
    nl, nr = left.nbytes // kb, right.nbytes // kb
    L = rf.DeviceBuffer.from_numpy(left)
    R = rf.DeviceBuffer.from_numpy(right)
    lb = rf.DeviceBuffer(4 * nl)
    lc = rf.DeviceBuffer(4 * nl)
    tot = rf.u64_cell(12345)
    ctx.join_find_bounds(lb, lc, tot, L, nl, R, nr, kb)
    ctx.synchronize()
    t = rf.read_u64(tot)
    li = rf.DeviceBuffer(4 * max(t, 1))
    ri = rf.DeviceBuffer(4 * max(t, 1))
    ctx.join_main(li, ri, tot, t, lb, lc, nl)
    ctx.synchronize()
    return (lb.download(np.int32, nl), lc.download(np.int32, nl), t,
            li.download(np.int32, t), ri.download(np.int32, t))


def check_join(rf, ctx, orc, left, right, kb):
    glb, glc, gt, gli, gri = gpu_join(rf, ctx, left, right, kb)
    olb, olc, ot = orc.join_find_bounds(left, right, kb)
    assert gt == ot
    assert np.array_equal(glb, olb)
    assert np.array_equal(glc, olc)
    oli, ori = orc.join_main(olb, olc, ot)
    assert np.array_equal(gli, oli)
    assert np.array_equal(gri, ori)
    return gt


def test_config1_shape(rf, ctx, orc):
    """R 65,536 x S 262,144 rows, int32 keys uniform in [0, 2^16): sort
    both sides on the GPU, join, gather (key, r_val, s_val) 16-byte rows."""
    rk, rv, sk, sv = datagen.join_tables(datagen.BASE_SEED, 65536, 262144)
    bufs = {k: rf.DeviceBuffer.from_numpy(v) for k, v in dict(rk=rk, rv=rv, sk=sk, sv=sv).items()}
    ctx.sort_pairs(bufs["rk"], bufs["rv"], 65536, 4, 32, 4)
    ctx.sort_pairs(bufs["sk"], bufs["sv"], 262144, 4, 32, 4)
    lb = rf.DeviceBuffer(4 * 65536)
    lc = rf.DeviceBuffer(4 * 65536)
    tot = rf.u64_cell()
    ctx.join_find_bounds(lb, lc, tot, bufs["rk"], 65536, bufs["sk"], 262144, 4)
    ctx.synchronize()
    t = rf.read_u64(tot)
    li = rf.DeviceBuffer(4 * t)
    ri = rf.DeviceBuffer(4 * t)
    ctx.join_main(li, ri, tot, t, lb, lc, 65536)
    out = rf.DeviceBuffer(16 * t)
    ctx.join_gather(out, 16, bufs["rk"], 4, bufs["rv"], 4, bufs["sv"], 4, li, ri, tot, t, 64, 32)
    fused = rf.DeviceBuffer(16 * t)
    ctx.join_expand_gather(fused, 16, bufs["rk"], 4, bufs["rv"], 4, bufs["sv"], 4, lb, lc, 65536,
                           tot, t, 64, 32)
    ctx.synchronize()

    ork, orv = orc.sort_pairs(rk, rv, 4, 32, 4)
    osk, osv = orc.sort_pairs(sk, sv, 4, 32, 4)
    olb, olc, ot = orc.join_find_bounds(ork, osk, 4)
    oli, ori = orc.join_main(olb, olc, ot)
    want = orc.join_gather(16, ork, 4, orv, 4, osv, 4, oli, ori, 64, 32)
    assert t == ot
    assert np.array_equal(li.download(np.int32, t), oli)
    assert np.array_equal(ri.download(np.int32, t), ori)
    assert np.array_equal(out.download(np.uint64, 2 * t), want)
    assert np.array_equal(fused.download(np.uint64, 2 * t), want)
    # (key, left position, right position) order
    words = want.reshape(-1, 2)
    key = words[:, 1] & np.uint64(0xFFFFFFFF)
    assert np.all(key[1:] >= key[:-1])


@pytest.mark.parametrize("kb", [1, 2, 4, 8, 16])
def test_key_widths(rf, ctx, orc, kb):
    rng = {1: 200, 2: 3000, 4: 5000, 8: 5000, 16: 5000}[kb]
    left = sorted_keys(orc, kb, 20_000, 100 + kb, rng)
    right = sorted_keys(orc, kb, 30_000, 200 + kb, rng)
    check_join(rf, ctx, orc, left, right, kb)


@pytest.mark.parametrize("nl,nr", [(0, 100), (100, 0), (0, 0), (1, 1), (1, 5000),
                                   (5000, 1), (2047, 2049), (4096, 4096)])
def test_sizes(rf, ctx, orc, nl, nr):
    left = sorted_keys(orc, 4, nl, 7, 50)
    right = sorted_keys(orc, 4, nr, 8, 50)
    check_join(rf, ctx, orc, left, right, 4)


def test_no_matches(rf, ctx, orc):
    left = (np.arange(10_000, dtype=np.uint32) * 2)
    right = (np.arange(10_000, dtype=np.uint32) * 2 + 1)
    assert check_join(rf, ctx, orc, left, right, 4) == 0


def test_all_one_key_cross_product(rf, ctx, orc):
    left = np.full(300, 7, dtype=np.uint32)
    right = np.full(1000, 7, dtype=np.uint32)
    assert check_join(rf, ctx, orc, left, right, 4) == 300_000


def test_identity_join_known_answer(rf, ctx, orc):
    """TestJoin.cu:91-156: joining 0..2^15-1 with itself gives row i =
    (i << 16) | i for 16-bit keys."""
    n = 1 << 15
    keys = np.arange(n, dtype=np.uint16)
    glb, glc, gt, gli, gri = gpu_join(rf, ctx, keys, keys, 2)
    assert gt == n
    assert np.array_equal(gli, np.arange(n)) and np.array_equal(gri, np.arange(n))
    K = rf.DeviceBuffer.from_numpy(keys)
    LI = rf.DeviceBuffer.from_numpy(gli)
    RI = rf.DeviceBuffer.from_numpy(gri)
    tot = rf.u64_cell(n)
    out = rf.DeviceBuffer(4 * n)
    ctx.join_gather(out, 4, K, 2, None, 0, K, 2, LI, RI, tot, n, 16, 0)
    ctx.synchronize()
    got = out.download(np.uint32, n)
    i = np.arange(n, dtype=np.uint32)
    assert np.array_equal(got, (i << np.uint32(16)) | i)


def test_zipf_skewed_foreign_keys(rf, ctx, orc):
    """Dimension keys 0..d-1 against Zipf(1.1) fact keys: one key matches a
    large share of the facts (load balance across tiles)."""
    d = 1000
    dim = np.arange(d, dtype=np.uint32)
    fk = (datagen.zipf_ranks(datagen.BASE_SEED + 4, 300_000, 1.1, d) - np.uint64(1)).astype(np.uint32)
    fact = np.sort(fk, kind="stable")
    check_join(rf, ctx, orc, dim, fact, 4)   # dim-major output, fan-out up to ~10^5
    check_join(rf, ctx, orc, fact, dim, 4)   # fact-major output, fan-out 1


@pytest.mark.parametrize("shape", [(4, 4, 0, 4, 32, 0), (8, 4, 4, 0, 32, 0), (16, 4, 4, 4, 64, 32),
                                   (8, 2, 2, 2, 32, 16), (16, 8, 0, 8, 64, 0)])
def test_gather_variants(rf, ctx, orc, shape):
    """gather_key_value / value_key / value_value / key_key shapes
    (ModernGPUJoin.h:43-200)."""
    ob, kb, lvb, rvb, kshift, lvshift = shape
    left = sorted_keys(orc, kb, 3000, 31, 400)
    right = sorted_keys(orc, kb, 5000, 32, 400)
    olb, olc, ot = orc.join_find_bounds(left, right, kb)
    oli, ori = orc.join_main(olb, olc, ot)
    lval = datagen.splitmix64(41, 3000 * 2).view(np.uint8)[: 3000 * lvb].view(DT[lvb]).copy() if lvb else None
    rval = datagen.splitmix64(42, 5000 * 2).view(np.uint8)[: 5000 * rvb].view(DT[rvb]).copy() if rvb else None
    want = orc.join_gather(ob, left, kb, lval, lvb, rval, rvb, oli, ori, kshift, lvshift)
    held = {}  # device buffers stay referenced until the kernels have run

    def up(name, a):
        held[name] = rf.DeviceBuffer.from_numpy(a) if a is not None else None
        return held[name]

    out = rf.DeviceBuffer(ob * max(ot, 1))
    tot = rf.u64_cell(ot)
    ctx.join_gather(out, ob, up("l", left), kb, up("lv", lval), lvb, up("rv", rval), rvb,
                    up("li", oli), up("ri", ori), tot, ot, kshift, lvshift)
    fused = rf.DeviceBuffer(ob * max(ot, 1))
    ctx.join_expand_gather(fused, ob, held["l"], kb, held["lv"], lvb, held["rv"], rvb,
                           up("lb", olb), up("lc", olc), 3000, tot, ot, kshift, lvshift)
    ctx.synchronize()
    assert np.array_equal(out.download(want.dtype, want.size), want)
    assert np.array_equal(fused.download(want.dtype, want.size), want)


def test_capacity_clamps_expansion(rf, ctx, orc):
    """A device total larger than the caller's capacity never writes past
    the capacity."""
    left = np.full(100, 3, dtype=np.uint32)
    right = np.full(100, 3, dtype=np.uint32)
    olb, olc, ot = orc.join_find_bounds(left, right, 4)
    cap = 1000
    li = rf.DeviceBuffer(4 * (cap + 64))
    ri = rf.DeviceBuffer(4 * (cap + 64))
    li.zero()
    ri.zero()
    tot = rf.u64_cell(ot)
    d_lb, d_lc = rf.DeviceBuffer.from_numpy(olb), rf.DeviceBuffer.from_numpy(olc)
    ctx.join_main(li, ri, tot, cap, d_lb, d_lc, 100)
    ctx.synchronize()
    oli, ori = orc.join_main(olb, olc, ot)
    assert np.array_equal(li.download(np.int32, cap), oli[:cap])
    assert not li.download(np.int32, cap + 64)[cap:].any()


def test_expand_project_matches_gather_then_project(rf, ctx, orc):
    """Fused expansion + projection equals gather followed by project."""
    left = sorted_keys(orc, 4, 20_000, 61, 3000)
    right = sorted_keys(orc, 4, 8_000, 62, 3000)
    lval = datagen.splitmix64(63, 20_000)
    rval = datagen.u32(64, 8_000)
    olb, olc, ot = orc.join_find_bounds(left, right, 4)
    oli, ori = orc.join_main(olb, olc, ot)
    rows = orc.join_gather(16, left, 4, lval, 8, rval, 4, oli, ori, 96, 32)
    want_k = orc.project(4, rows, 16, [96 + 3, 0], [9, 6], [6, 0])
    want_v = orc.project(8, rows, 16, [32], [64], [0])
    up = rf.DeviceBuffer.from_numpy
    K = rf.DeviceBuffer(4 * ot)
    V = rf.DeviceBuffer(8 * ot)
    tot = rf.u64_cell(ot)
    d_lb, d_lc, d_l, d_lv, d_rv = up(olb), up(olc), up(left), up(lval), up(rval)
    ctx.join_expand_project(d_lb, d_lc, 20_000, tot, ot, d_l, 4, d_lv, 8, d_rv, 4,
                            K, 4, V, 8, [(0, 0, 3, 9, 6), (0, 2, 0, 6, 0), (1, 1, 0, 64, 0)])
    ctx.synchronize()
    assert np.array_equal(K.download(np.uint32, ot), want_k)
    assert np.array_equal(V.download(np.uint64, ot), want_v)


@pytest.mark.parametrize("max_run", [3, 300, 700])
def test_equal_runs_crossing_tiles(rf, ctx, orc, max_run):
    """Right-side runs of every length from 0 to max_run, so that runs end
    exactly at, just before and just past a tile's end and past the right
    keys a tile keeps beyond its end (the global-memory continuation)."""
    m = 30_000
    runs = (datagen.splitmix64(97 + max_run, m) % np.uint64(max_run + 1)).astype(np.int64)
    right = np.repeat(np.arange(m, dtype=np.uint32), runs)
    left = np.repeat(np.arange(m, dtype=np.uint32), (np.arange(m) % 3) + 0)  # 0, 1 or 2 copies
    for kb, dt in ((4, np.uint32), (8, np.uint64)):
        check_join(rf, ctx, orc, left.astype(dt), right.astype(dt), kb)


def test_hbm_scale_shape_against_oracle(rf, ctx, orc):
    """The bench's HBM-scale join shape scaled down 16x (2^20 x 2^22 keys in
    [0, 2^20)), every lower bound, count and index pair against the oracle."""
    left = sorted_keys(orc, 4, 1 << 20, 301, 1 << 20)
    right = sorted_keys(orc, 4, 1 << 22, 302, 1 << 20)
    assert check_join(rf, ctx, orc, left, right, 4) > (1 << 21)



def test_config1_plan_replayed_from_a_graph(rf, ctx, orc):
    """The config-1 plan (restore the unsorted inputs, sort both sides,
    find_bounds, expansion with gather) captured once into a HIP graph:
    every replay, with eager work and synchronisations between replays,
    writes the tuples the eager plan wrote."""
    n_r, n_s = 65536, 262144
    rk, rv, sk, sv = datagen.join_tables(datagen.BASE_SEED + 40, n_r, n_s, 16)
    src = [rf.DeviceBuffer.from_numpy(x) for x in (rk, rv, sk, sv)]
    work = [rf.DeviceBuffer(x.nbytes) for x in src]
    lb, lc, tot = rf.DeviceBuffer(4 * n_r), rf.DeviceBuffer(4 * n_r), rf.u64_cell()
    stream = rf.Stream()
    gctx = rf.Context(0, stream.handle)  # capture needs a stream of its own

    def enqueue(total):
        for w, s in zip(work, src):
            gctx.copy(w, s, s.nbytes)
        gctx.sort_pairs(work[0], work[1], n_r, 4, 32, 4)
        gctx.sort_pairs(work[2], work[3], n_s, 4, 32, 4)
        gctx.join_find_bounds(lb, lc, tot, work[0], n_r, work[2], n_s, 4)
        if total:
            gctx.join_expand_gather(out, 16, work[0], 4, work[1], 4, work[3], 4, lb, lc, n_r,
                                    tot, total, 64, 32)

    enqueue(0)
    gctx.synchronize()
    total = rf.read_u64(tot)
    out = rf.DeviceBuffer(16 * total)
    enqueue(total)  # also sizes the context's workspace before the capture
    gctx.synchronize()
    want = out.download(np.uint64, 2 * total)
    # the eager tuples against the oracle's join of the same tables
    ok, ov = orc.sort_pairs(rk, rv, 4, 32, 4)
    sk_s, sv_s = orc.sort_pairs(sk, sv, 4, 32, 4)
    olb, olc, ot = orc.join_find_bounds(ok, sk_s, 4)
    oli, ori = orc.join_main(olb, olc, ot)
    assert ot == total
    tup = np.empty(2 * total, dtype=np.uint64)
    tup[0::2] = (ov[oli].astype(np.uint64) << np.uint64(32)) | sv_s[ori].astype(np.uint64)
    tup[1::2] = ok[oli].astype(np.uint64)
    assert np.array_equal(want, tup)
    gctx.graph_begin()
    enqueue(total)
    g = gctx.graph_end()
    for _ in range(3):
        out.zero()
        for w in work:
            w.zero()
        gctx.synchronize()
        g.launch()
        gctx.synchronize()
        assert np.array_equal(out.download(np.uint64, 2 * total), want)
    g.close()
    gctx.close()
    stream.close()
