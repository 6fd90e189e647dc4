"""Tile shapes chosen by the autotuner change time, never results: the
batched plan under every knob value gives bit-identical outputs, and the
tuner leaves the context on the values it reports."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def outputs(br):
    br.run_eager()
    br.synchronize()
    g = int(br.nout.download(np.uint64, 1)[0])
    return (br.okeys.download(np.uint32, g), br.osum.download(np.uint64, g),
            br.ocnt.download(np.uint32, g))


def test_knobs_do_not_change_results(rf, ctx):
    from redfox_amd import autotune, batched

    w = batched.Workload(ctx, n_inst=24, seed=5, fact_exp=(2.0, 4.5), dim_exp=(1.0, 3.0))
    br = batched.BatchedRunner(w)
    ref = outputs(br)
    for knob, values in autotune.KNOBS.items():
        for v in values:
            br.ctx.tune(knob, v)
            got = outputs(br)
            assert all(np.array_equal(a, b) for a, b in zip(got, ref)), (knob, v)
        br.ctx.tune(knob, 0)
    res = br.autotune(reps=1)
    for knob, v in res["choice"].items():
        assert br.ctx.tuning(knob) == v
    got = outputs(br)
    assert all(np.array_equal(a, b) for a, b in zip(got, ref))


def test_bad_knob_is_an_error(rf, ctx):
    with pytest.raises(rf.RfxError):
        ctx.tune("sort.ipt", 7)
    with pytest.raises(rf.RfxError):
        ctx.tune("no.such.knob", 0)
    assert ctx.tuning("sort.ipt") == 0
