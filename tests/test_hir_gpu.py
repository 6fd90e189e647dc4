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


def test_ptx_kernel_is_refused(rf):
    p = H.Program("ptx")
    a = p.var(H.I32, data=np.arange(8, dtype=np.uint32), input=True)
    p.block(H.Kernel("copy", H.COMPUTE, [(H.IN, a)], code=b"ptx text is not kept"))
    h = H.Hir(p.encode())
    with pytest.raises(H.RfxError, match="PTX kernel 'copy'") as e:
        h.run()
    assert e.value.status == -5
    h.close()
