"""HIR programs run by the executor on the GPU, compared with the oracle:
a join plan in addJoin's kernel order with Harmony's size kernels
(GetSize, Resize) and liveness frees, a set plan whose result size comes
back through UpdateSize, an aggregate plan, re-runs with new inputs, and
the self-test records (byte-exact expected outputs)."""
import numpy as np
import pytest

from redfox_amd import datagen
from redfox_amd import hir as H

pytestmark = pytest.mark.gpu

RFX_I32, RFX_I64, RFX_F64 = 3, 4, 10


def join_program(rk, rv, sk, sv):
    """addJoin's order (RelationalAlgebraCompiler.cpp:4495-5226): sizes,
    two sorts, find bounds, result size, resized index and result
    buffers, join main, gather."""
    p = H.Program("join")
    RK = p.var(H.I32, data=rk, input=True)
    RV = p.var(H.I32, data=rv, input=True)
    SK = p.var(H.I32, data=sk, input=True)
    SV = p.var(H.I32, data=sv, input=True)
    r_bytes, s_bytes = p.var(H.I64, 8), p.var(H.I64, 8)
    bits, vtype, ktype = p.u64(32), p.u64(RFX_I32), p.u64(RFX_I32)
    lb, lc = p.var(H.I8, 0), p.var(H.I8, 0)
    result, index_size = p.var(H.I64, 8), p.var(H.I64, 8)
    out_bytes = p.u64(16)
    li, ri = p.var(H.I8, 0), p.var(H.I8, 0)
    out = p.var(H.I8, 0, output=True)
    desc = p.u64(16 | (4 << 8) | (4 << 16) | (4 << 24) | (64 << 32) | (32 << 40))
    p.block(H.get_size(r_bytes, RK), H.get_size(s_bytes, SK),
            H.binary("set_b40csortpair17", (H.INOUT, RK), (H.INOUT, RV), (H.IN, r_bytes),
                     (H.IN, bits), (H.IN, vtype)),
            H.binary("set_b40csortpair18", (H.INOUT, SK), (H.INOUT, SV), (H.IN, s_bytes),
                     (H.IN, bits), (H.IN, vtype)))
    p.block(H.resize(lb, r_bytes), H.resize(lc, r_bytes),
            H.binary("mgpu_join_find_bounds19", (H.INOUT, lb), (H.INOUT, lc), (H.OUT, result),
                     (H.IN, RK), (H.IN, r_bytes), (H.IN, SK), (H.IN, s_bytes), (H.IN, ktype)),
            H.binary("join_get_result_size20", (H.OUT, index_size), (H.INOUT, result),
                     (H.IN, out_bytes)),
            H.resize(li, index_size), H.resize(ri, index_size), H.resize(out, result))
    p.block(H.binary("mgpu_join_main21", (H.INOUT, li), (H.INOUT, ri), (H.IN, index_size),
                     (H.IN, lb), (H.IN, lc), (H.IN, r_bytes), (H.IN, s_bytes)),
            H.binary("mgpu_join_gather22", (H.INOUT, out), (H.IN, RK), (H.IN, li), (H.IN, SK),
                     (H.IN, ri), (H.IN, index_size), (H.IN, RV), (H.IN, SV), (H.IN, desc)))
    return p, dict(RK=RK, RV=RV, SK=SK, SV=SV, out=out, lb=lb, li=li)


def oracle_join(orc, rk, rv, sk, sv):
    ork, orv = orc.sort_pairs(rk, rv, 4, 32, 4)
    osk, osv = orc.sort_pairs(sk, sv, 4, 32, 4)
    olb, olc, ot = orc.join_find_bounds(ork, osk, 4)
    oli, ori = orc.join_main(olb, olc, ot)
    out = orc.join_gather(16, ork, 4, orv, 4, osv, 4, oli, ori, 64, 32)
    return ork, orv, osk, osv, out


def test_join_plan(rf, orc):
    rk, rv, sk, sv = datagen.join_tables(datagen.BASE_SEED, 65536, 262144)
    p, v = join_program(rk, rv, sk, sv)
    h = H.Hir(p.encode())
    # Liveness frees: lower bounds and counts after join main, the index
    # buffers after the gather (their last In operands).
    sched = h.schedule(2)
    assert sched.index(f"free {v['lb']}") > sched.index("run mgpu_join_main21")
    assert f"free {v['li']}" in sched and f"free {v['out']}" not in sched
    h.run()
    ork, orv, osk, osv, want = oracle_join(orc, rk, rv, sk, sv)
    assert np.array_equal(h.read(v["RK"], np.uint32), ork)
    assert np.array_equal(h.read(v["RV"], np.uint32), orv)
    assert np.array_equal(h.read(v["SV"], np.uint32), osv)
    assert np.array_equal(h.read(v["out"], np.uint64), want)
    with pytest.raises(H.RfxError):  # freed by the schedule
        h.read(v["lb"])
    h.close()


def test_self_tests_byte_exact(rf, orc):
    rk, rv, sk, sv = datagen.join_tables(datagen.BASE_SEED + 3, 4096, 20000)
    p, v = join_program(rk, rv, sk, sv)
    ork, orv, osk, osv, want = oracle_join(orc, rk, rv, sk, sv)
    p.test("declared_inputs", outputs=[(v["out"], want), (v["RK"], ork)])
    # Different inputs through the test record itself.
    rk2, rv2, sk2, sv2 = datagen.join_tables(datagen.BASE_SEED + 4, 3000, 9000)
    _, _, _, _, want2 = oracle_join(orc, rk2, rv2, sk2, sv2)
    p.test("other_inputs", inputs=[(v["RK"], rk2), (v["RV"], rv2), (v["SK"], sk2), (v["SV"], sv2)],
           outputs=[(v["out"], want2)])
    bad = bytearray(want.tobytes())
    bad[100] ^= 1
    p.test("corrupt_reference", outputs=[(v["out"], bytes(bad))])
    p.test("short_reference", outputs=[(v["out"], want.tobytes()[:-16])])
    h = H.Hir(p.encode())
    passed, failed, report = h.run_tests()
    assert (passed, failed) == (2, 2), report
    assert f"variable {v['out']} differs from the reference at byte 100" in report
    assert "corrupt_reference" in report and "short_reference" in report
    # The declared inputs are back after the tests.
    h.run()
    assert np.array_equal(h.read(v["out"], np.uint64), want)
    h.close()


def set_program(a, b, kernel):
    p = H.Program("sets")
    A = p.var(H.I64, data=a, input=True)
    B = p.var(H.I64, data=b, input=True)
    ab, bb, size = p.var(H.I64, 8), p.var(H.I64, 8), p.var(H.I64, 8)
    typ = p.u64(RFX_I64)
    out = p.var(H.I8, a.nbytes + b.nbytes, output=True)
    p.block(H.get_size(ab, A), H.get_size(bb, B),
            H.binary(kernel, (H.OUT, out), (H.OUT, size), (H.IN, A), (H.IN, ab), (H.IN, B),
                     (H.IN, bb), (H.IN, typ)),
            H.update_size(out, size))
    return p, A, B, out


@pytest.mark.parametrize("kernel,op", [("set_union3", 0), ("set_difference4", 1),
                                       ("set_intersection9", 2)])
def test_set_plan_size_through_update_size(rf, orc, kernel, op):
    a = np.unique(datagen.splitmix64(21, 40_000))
    b = np.unique(np.concatenate([a[::3], datagen.splitmix64(22, 15_000)]))
    p, A, B, out = set_program(a, b, kernel)
    h = H.Hir(p.encode())
    h.run()
    assert np.array_equal(h.read(out, np.uint64), orc.set_op(op, a, b, 8))
    # New inputs of other sizes, same program.
    a2 = np.unique(datagen.splitmix64(23, 9_000))
    b2 = np.unique(datagen.splitmix64(24, 30_000))
    h.set_input(A, a2)
    h.set_input(B, b2)
    with pytest.raises(H.RfxError):  # out is declared for the first inputs' sizes
        h.set_input(out, a2)
    h.run()
    assert np.array_equal(h.read(out, np.uint64), orc.set_op(op, a2, b2, 8))
    h.close()


def test_aggregate_plan(rf, orc):
    """set_reduce then set_count over the same sorted keys, result sizes
    taken back into the result variables with UpdateSize."""
    n = 150_000
    keys = np.sort((datagen.splitmix64(31, n) % np.uint64(1000)).astype(np.uint32))
    vals = datagen.uniform01(32, n) * 1e3
    p = H.Program("agg")
    K = p.var(H.I32, data=keys, input=True)
    V = p.var(H.F64, data=vals, input=True)
    ones = p.var(H.I32, data=np.ones(n, dtype=np.uint32), input=True)
    ks, vs = p.var(H.I64, 8), p.var(H.I64, 8)
    ks2, vs2 = p.var(H.I64, 8), p.var(H.I64, 8)
    rk = p.var(H.I8, keys.nbytes, output=True)
    rv = p.var(H.I8, vals.nbytes, output=True)
    rk2 = p.var(H.I8, keys.nbytes, output=True)
    rc = p.var(H.I8, keys.nbytes, output=True)
    desc = p.u64(4 | (0 << 8) | (RFX_F64 << 16))
    kbytes = p.u64(4)
    p.block(H.get_size(ks, K), H.get_size(vs, V),
            H.binary("set_reduce5", (H.OUT, rk), (H.OUT, rv), (H.IN, K), (H.INOUT, ks), (H.IN, V),
                     (H.INOUT, vs), (H.IN, desc)),
            H.update_size(rk, ks), H.update_size(rv, vs))
    p.block(H.get_size(ks2, K), H.get_size(vs2, ones),
            H.binary("set_count6", (H.OUT, rk2), (H.OUT, rc), (H.IN, K), (H.INOUT, ks2),
                     (H.IN, ones), (H.INOUT, vs2), (H.IN, kbytes)),
            H.update_size(rk2, ks2), H.update_size(rc, vs2))
    h = H.Hir(p.encode())
    h.run()
    ok, ov = orc.reduce_by_key_f64(keys, vals, 4, 0)
    assert np.array_equal(h.read(rk, np.uint32), ok)
    np.testing.assert_allclose(h.read(rv, np.float64), ov, rtol=1e-12, atol=0)
    assert np.array_equal(h.read(rk2, np.uint32), ok)
    assert h.read(rc, np.uint32).tolist() == np.bincount(keys)[ok].tolist()
    h.close()


def test_unknown_compute_family_is_refused(rf):
    p = H.Program("ptx")
    one, zero = p.u64(1), p.u64(0)
    a = p.var(H.I32, data=np.arange(8, dtype=np.uint32), input=True)
    p.block(H.compute("copy3", (one, one, zero), (H.IN, a), code=b"ptx text is not kept"))
    h = H.Hir(p.encode())
    with pytest.raises(H.RfxError, match="compute kernel family 'copy'") as e:
        h.run()
    assert e.value.status == -5
    h.close()


# ---------------------------------------------------------------------------
# Red Fox's own kernel sequences (RelationalAlgebraCompiler.cpp), PTX
# compute kernels included, each kernel's instance types in an rfx1
# descriptor in its code bytes.
# ---------------------------------------------------------------------------
def tbytes(t):
    b = sum(t)
    return 1 if b <= 8 else 2 if b <= 16 else 4 if b <= 32 else 8 if b <= 64 else 16


def after(t, i):
    return sum(t[i + 1:])


def pack(cols, t):
    """Rows of tuple t (field 0 in the high bits) from per-field columns."""
    out = [0] * len(cols[0])
    for f, c in enumerate(cols):
        sh = after(t, f)
        out = [o | (int(x) << sh) for o, x in zip(out, c)]
    w = tbytes(t)
    if w <= 8:
        return np.array(out, dtype={1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[w])
    return np.array([[o & (2**64 - 1), o >> 64] for o in out], dtype=np.uint64).ravel()


def oracle_lowered_join(orc, a, a_t, b, b_t, d_t, keycount):
    def side(x, t):
        kt, vt = t[:keycount], t[keycount:]
        kb, vb = tbytes(kt), tbytes(vt) if vt else 0
        if vt:
            k, v = orc.split(x, tbytes(t), kb, vb, after(t, keycount - 1))
        else:
            k, v = x, None
        k, v = orc.sort_pairs(k, v, kb, 8 * kb, vb)
        return k, kb, v, vb, vt
    lk, kb, lv, lvb, _ = side(a, a_t)
    rk, _, rv, rvb, rvt = side(b, b_t)
    lb, lc, tot = orc.join_find_bounds(lk, rk, kb)
    li, ri = orc.join_main(lb, lc, tot)
    return orc.join_gather(tbytes(d_t), lk, kb, lv, lvb, rv, rvb, li, ri, after(d_t, keycount - 1),
                           sum(rvt)), tot


JOIN_SHAPES = [
    # (A tuple, B tuple, D tuple): b40c sorts on 32-bit keys, split both sides
    ([32, 32], [32, 32], [32, 32, 32]),
    # B is all key (no split on the right): gather_value_key
    ([32, 32], [32], [32, 32]),
    # 40-bit keys: mergesort pairs, 8-byte key columns, 16-byte output
    ([40, 24], [40, 24], [40, 24, 24]),
]


@pytest.mark.parametrize("a_t,b_t,d_t", JOIN_SHAPES)
def test_reference_join_sequence(rf, orc, a_t, b_t, d_t):
    """addJoin's complete kernel sequence (:4466-5226): get_size, the split
    size/resize/split kernels, the sorts, join_get_temp_size, resizes,
    mgpu_join_find_bounds, join_get_result_size, mgpu_join_main and the
    PTX mgpu_join_gather, against the oracle bit for bit."""
    na, nb = 30_011, 90_007
    spread = 1 << (a_t[0] - 12)

    def rel(seed, n, t):
        keys = (datagen.splitmix64(seed, n) % np.uint64(4096)) * np.uint64(spread)
        cols = [keys] + [datagen.splitmix64(seed + 1 + i, n) >> np.uint64(64 - w)
                         for i, w in enumerate(t[1:])]
        return pack(cols, t)

    a, b = rel(61, na, a_t), rel(71, nb, b_t)
    p = H.Program("join_lowered")
    A = p.var(H.I8, data=a, input=True)
    B = p.var(H.I8, data=b, input=True)
    D = p.var(H.I8, 0, output=True)
    ks = []
    H.Lowering(p, first_id=40).join(ks, D, d_t, A, a_t, B, b_t, 1)
    p.block(*ks)
    fams = [k.name.rstrip("0123456789") for k in ks if k.type in (H.COMPUTE, H.BINARY)]
    assert "set_split" in fams and "mgpu_join_gather" in fams and "join_get_temp_size" in fams
    h = H.Hir(p.encode())
    h.run()
    want, tot = oracle_lowered_join(orc, a, a_t, b, b_t, d_t, 1)
    assert tot > na
    got = h.read(D, want.dtype)
    assert len(got) == len(want) and np.array_equal(got, want)
    h.close()


def test_reference_count_sequence(rf, orc):
    """addAgg's COUNT path (:8335-8641): set_project, the key sort,
    generate_get_result_size + set_generate, set_count,
    merge_get_result_size + set_merge; D = (key, count) rows."""
    n = 200_003
    src_t, d_t = [8, 24], [8, 32]
    key = (datagen.splitmix64(81, n) % np.uint64(200)).astype(np.uint64)
    other = datagen.splitmix64(82, n) >> np.uint64(40)
    src = pack([key, other], src_t)
    p = H.Program("count_lowered")
    S = p.var(H.I8, data=src, input=True)
    D = p.var(H.I8, 0, output=True)
    ks = []
    H.Lowering(p, first_id=7).count(ks, D, d_t, S, src_t, [0])
    p.block(*ks)
    h = H.Hir(p.encode())
    h.run()
    u, c = np.unique(key, return_counts=True)
    want = (u.astype(np.uint64) << np.uint64(32)) | c.astype(np.uint64)
    assert np.array_equal(h.read(D, np.uint64), want)
    h.close()


def test_reference_select_sequence(rf, orc):
    """addSelectField's kernels (:6234-6530): set_select (field against a
    constant), scan over ctaCount + 1 entries, select_get_result_size,
    resize, select_gather; kept rows in input order."""
    n = 100_019
    t = [16, 48]
    f0 = datagen.splitmix64(91, n) % np.uint64(1000)
    f1 = datagen.splitmix64(92, n) >> np.uint64(16)
    src = pack([f0, f1], t)
    p = H.Program("select_lowered")
    S = p.var(H.I8, data=src, input=True)
    D = p.var(H.I8, 0, output=True)
    in_size, out_size = p.var(H.I64, 8), p.var(H.I64, 8)
    temp, hist = p.var(H.I8, 0), p.var(H.I8, 512 * 4)
    c1, c0, c128, c129, c256 = p.u64(1), p.u64(0), p.u64(128), p.u64(129), p.u64(256)
    d = dict(d=t, parts=128, nt=128)
    p.block(H.get_size(in_size, S), H.resize(temp, in_size),
            H.compute("set_select3", (c128, c128, c0), (H.INOUT, temp), (H.OUT, hist), (H.IN, S),
                      (H.IN, in_size), code=H.descriptor("set_select", cmp="0:lt:300", **d)),
            H.compute("scan4", (c1, c129, c0), (H.INOUT, hist), code=H.descriptor("scan")),
            H.compute("select_get_result_size5", (c1, c1, c0), (H.OUT, out_size), (H.IN, hist),
                      code=H.descriptor("select_get_result_size", **d)),
            H.resize(D, out_size),
            H.compute("select_gather6", (c128, c256, c0), (H.INOUT, D), (H.IN, temp),
                      (H.IN, in_size), (H.IN, hist), code=H.descriptor("select_gather", **d)))
    h = H.Hir(p.encode())
    h.run()
    assert np.array_equal(h.read(D, np.uint64), src[f0 < 300])
    h.close()


def test_harmony_string_families(rf, orc):
    """String (Pointer) keys through the Harmony families: set_mgpusortpair
    (sort_string_pair), mgpu_join_find_bounds (find_bounds_string) and
    set_count (count_string), against Python's sorted(), bisect and
    Counter."""
    import collections
    import ctypes
    words = [b"TRUCK", b"AIR", b"MAIL", b"SHIP", b"RAIL", b"REG AIR", b"FOB", b"AIR REG",
             b"\xe9t\xe9", b"", b"A", b"TRUCKS"]
    rng = np.random.default_rng(5)
    n = 5000
    pick = rng.integers(0, len(words), n)
    blob, offs = bytearray(), []
    for i in pick:  # every row its own copy: equal strings at different addresses
        offs.append(len(blob))
        blob += words[i] + b"\0"
    dev = rf.DeviceBuffer.from_numpy(np.frombuffer(bytes(blob) + b"\0" * 8, dtype=np.uint8))
    addr = (np.array(offs, dtype=np.uint64) + np.uint64(dev.ptr))
    strs = [words[i] for i in pick]
    L = rf.lib()
    L.rfx_harmony_launch.restype = ctypes.c_int
    L.rfx_harmony_launch.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                                     ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                     ctypes.c_uint32]

    def launch(name, kind, desc, *slots):
        arr = (ctypes.c_uint64 * len(slots))(*[int(x) for x in slots])
        st = L.rfx_harmony_launch(name.encode(), kind, desc, len(desc), arr, len(slots))
        assert st == 0, L.rfx_last_error()

    K = rf.DeviceBuffer.from_numpy(addr)
    V = rf.DeviceBuffer.from_numpy(np.arange(n, dtype=np.uint32))
    launch("set_mgpusortpair11", 2, H.descriptor("set_mgpusortpair", d=[(64, "p")], d1=[32]),
           K.ptr, V.ptr, 8 * n, 4, 3)
    order = sorted(range(n), key=lambda i: orc.string_key(strs[i]))  # stable
    got_rows = V.download(np.uint32, n)
    assert got_rows.tolist() == order
    got_addr = K.download(np.uint64, n)
    assert np.array_equal(got_addr, addr[order])
    sorted_strs = [strs[i] for i in order]

    # find bounds of the distinct words (sorted) in the sorted rows
    left = sorted(set(words[:8]), key=orc.string_key)
    lblob, loffs = bytearray(), []
    for w in left:
        loffs.append(len(lblob))
        lblob += w + b"\0"
    ldev = rf.DeviceBuffer.from_numpy(np.frombuffer(bytes(lblob) + b"\0" * 8, dtype=np.uint8))
    LK = rf.DeviceBuffer.from_numpy(np.array(loffs, dtype=np.uint64) + np.uint64(ldev.ptr))
    nl = len(left)
    lb, lc, tot = rf.DeviceBuffer(4 * nl), rf.DeviceBuffer(4 * nl), rf.u64_cell(0)
    launch("mgpu_join_find_bounds12", 2,
           H.descriptor("mgpu_join_find_bounds", a=[(64, "p")], b=[(64, "p")]),
           lb.ptr, lc.ptr, tot.ptr, LK.ptr, 8 * nl, K.ptr, 8 * n)
    olb, olc, otot = orc.find_bounds_string(left, sorted_strs)
    assert rf.read_u64(tot) == otot
    assert lb.download(np.int32, nl).tolist() == olb.tolist()
    assert lc.download(np.int32, nl).tolist() == olc.tolist()

    # count over runs of equal strings: the first row's address keys each run
    ones = rf.DeviceBuffer.from_numpy(np.ones(n, dtype=np.uint32))
    RK, RC = rf.DeviceBuffer(8 * n), rf.DeviceBuffer(4 * n)
    ksz, vsz = rf.u64_cell(8 * n), rf.u64_cell(4 * n)
    launch("set_count13", 2, H.descriptor("set_count", a=[(64, "p")], b=[32], agg="count"),
           RK.ptr, RC.ptr, K.ptr, ksz.ptr, ones.ptr, vsz.ptr)
    runs = []
    for s_ in sorted_strs:
        if not runs or runs[-1][0] != s_:
            runs.append([s_, 0])
        runs[-1][1] += 1
    g = len(runs)
    assert rf.read_u64(ksz) == 8 * g and rf.read_u64(vsz) == 4 * g
    assert RC.download(np.uint32, g).tolist() == [c for _, c in runs]
    firsts, i = [], 0
    for _, c in runs:
        firsts.append(int(got_addr[i]))
        i += c
    assert RK.download(np.uint64, g).tolist() == firsts
    assert collections.Counter(strs) == {s_: c for s_, c in runs}
