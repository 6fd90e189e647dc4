"""Many foreign-key join + group-aggregate query instances on one GPU.

The query of each instance i is

    SELECT grp, SUM(measure), COUNT(*) FROM fact_i JOIN dim_i ON fk = dkey
    GROUP BY grp

with dim_i = (dkey: a permutation of 0..d_i-1, grp: U[0, 64)) and fact_i =
(fk: Zipf(1.1) rank over [1, d_i] mapped through dkey, measure: U{0..100000}
/ 100).  With dup > 0 the first floor(dup * d_i) keys of the permutation
appear a second time (with their own grp), so the join is many-to-many.
Two executions of the same operator plan are provided:

* plain: the way Harmony runs a Red Fox plan - one instance after another,
  one operator launch at a time, the host reading each size variable before
  the next operator (Runtime.cpp:99-118, 1054-1314);
* batched: all instances as one plan.  Keys are prefixed with the instance
  number (rfx_segment_keys), so one radix sort, one find_bounds, one
  expansion, one regrouping sort and one reduce-by-key (sum and count
  together) cover every instance.  The whole sequence is captured once into a HIP graph and
  replayed with a single launch.  Stable sorts keep each instance's rows in
  the order the plain form sees them, so keys and counts are identical and
  FP64 sums agree to rounding.
"""
from __future__ import annotations

import numpy as np

from . import rfx as R
from .datagen import uniform01

FK_BITS = 17     # dimensions of up to 10^5 rows
GRP_BITS = 6     # 64 groups
GROUPS = 1 << GRP_BITS


class Workload:
    """Instance sizes, dimension tables and fact tables, resident on the
    device.  fact_exp / dim_exp give the log10 ranges of the row counts."""

    def __init__(self, ctx: R.Context, n_inst: int = 512, seed: int = 0x52464F58000000CF,
                 fact_exp=(3.0, 7.0), dim_exp=(2.0, 5.0), zipf_s: float = 1.1, dup: float = 0.0):
        self.ctx = ctx
        self.n_inst = n_inst
        u = uniform01(seed, 2 * n_inst)
        self.n = np.floor(10.0 ** (fact_exp[0] + (fact_exp[1] - fact_exp[0]) * u[:n_inst])).astype(np.int64)
        keys = np.floor(10.0 ** (dim_exp[0] + (dim_exp[1] - dim_exp[0]) * u[n_inst:])).astype(np.int64)
        assert int(keys.max()) <= (1 << FK_BITS)
        extra = np.floor(dup * keys).astype(np.int64)
        self.d = keys + extra  # dimension rows
        self.row_off = np.concatenate([[0], np.cumsum(self.n)]).astype(np.uint64)
        self.dim_off = np.concatenate([[0], np.cumsum(self.d)]).astype(np.uint64)
        self.N = int(self.row_off[-1])
        self.D = int(self.dim_off[-1])
        rng = np.random.default_rng(seed & 0xFFFFFFFF)
        perms = [rng.permutation(int(x)) for x in keys]
        self.dkey_host = np.concatenate(
            [np.concatenate([p, p[:int(e)]]) for p, e in zip(perms, extra)]).astype(np.uint32)
        self.grp_host = rng.integers(0, GROUPS, self.D, dtype=np.uint32)
        self.dkey = R.DeviceBuffer.from_numpy(self.dkey_host)
        self.grp = R.DeviceBuffer.from_numpy(self.grp_host)
        self.row_off_dev = R.DeviceBuffer.from_numpy(self.row_off)
        self.dim_off_dev = R.DeviceBuffer.from_numpy(self.dim_off)
        self.fk = R.DeviceBuffer(4 * self.N)
        self.measure = R.DeviceBuffer(8 * self.N)
        ctx.gen_fk_batch(self.fk, self.measure, self.row_off_dev, self.dim_off_dev, self.dkey,
                         n_inst, self.N, seed + 7, zipf_s)
        ctx.synchronize()

    def bytes_per_batch(self) -> int:
        """Algorithmic input bytes: 12 per fact row, 8 per dimension row."""
        return 12 * self.N + 8 * self.D


def _results(keys, sums, counts, n_inst):
    """Split (instance << 6 | grp) runs into per-instance (grp, sum, count)."""
    inst = keys >> np.uint32(GRP_BITS)
    out = []
    for b in range(n_inst):
        m = inst == b
        out.append((keys[m] & np.uint32(GROUPS - 1), sums[m], counts[m]))
    return out


class PlainRunner:
    """One instance at a time, one operator at a time, sizes read back."""

    def __init__(self, w: Workload, ctx: R.Context):
        self.w, self.ctx = w, ctx
        nmax, dmax = int(w.n.max()), int(w.d.max())
        B = R.DeviceBuffer
        self.fkw, self.mw = B(4 * nmax), B(8 * nmax)
        self.dkw, self.dgw = B(4 * dmax), B(4 * dmax)
        self.lb, self.lc = B(4 * nmax), B(4 * nmax)
        self.tot = R.u64_cell()
        self.jcap = 0
        self._resize(nmax)
        self.okeys = B(4 * GROUPS * w.n_inst)
        self.osum = B(8 * GROUPS * w.n_inst)
        self.ocnt = B(4 * GROUPS * w.n_inst)
        self.okeys2 = B(4 * GROUPS)
        self.nout = R.DeviceBuffer(8 * w.n_inst)
        self.nout2 = R.u64_cell()

    def _resize(self, rows: int) -> None:
        """Joined-row buffers of at least `rows` rows (Harmony resizes a
        variable to the size it read back)."""
        if rows <= self.jcap:
            return
        B = R.DeviceBuffer
        self.jcap = rows
        self.out16 = B(16 * rows)
        self.key2, self.val = B(4 * rows), B(8 * rows)

    def run(self) -> None:
        w, c = self.w, self.ctx
        for b in range(w.n_inst):
            n, d = int(w.n[b]), int(w.d[b])
            r0, d0 = int(w.row_off[b]), int(w.dim_off[b])
            # split the relations into working key / value columns
            c.copy(self.fkw, w.fk, 4 * n, src_off=4 * r0)
            c.copy(self.mw, w.measure, 8 * n, src_off=8 * r0)
            c.copy(self.dkw, w.dkey, 4 * d, src_off=4 * d0)
            c.copy(self.dgw, w.grp, 4 * d, src_off=4 * d0)
            c.synchronize()
            c.sort_pairs(self.fkw, self.mw, n, 4, 32, 8)
            c.synchronize()
            c.sort_pairs(self.dkw, self.dgw, d, 4, 32, 4)
            c.synchronize()
            c.join_find_bounds(self.lb, self.lc, self.tot, self.fkw, n, self.dkw, d, 4)
            c.synchronize()
            t = R.read_u64(self.tot)  # join_get_result_size + Resize
            self._resize(t)
            c.join_expand_gather(self.out16, 16, self.fkw, 4, self.mw, 8, self.dgw, 4, self.lb,
                                 self.lc, n, self.tot, t, 96, 32)
            c.synchronize()
            c.project(self.key2, 4, self.out16, 16, t, [0], [GRP_BITS], [0])
            c.synchronize()
            c.project(self.val, 8, self.out16, 16, t, [32], [64], [0])
            c.synchronize()
            c.sort_pairs(self.key2, self.val, t, 4, 8, 8)
            c.synchronize()
            _reduce_into(c, self, b, t)
            c.synchronize()
            R.read_u64(self.nout, b)  # Harmony reads the output size

    def results(self):
        out = []
        nouts = self.nout.download(np.uint64, self.w.n_inst)
        for b in range(self.w.n_inst):
            g = int(nouts[b])
            k = self.okeys.download(np.uint32, g, offset=4 * GROUPS * b)
            s = self.osum.download(np.float64, g, offset=8 * GROUPS * b)
            cnt = self.ocnt.download(np.uint32, g, offset=4 * GROUPS * b)
            out.append((k, s, cnt))
        return out


def _reduce_into(c: R.Context, p: PlainRunner, b: int, t: int) -> None:
    """Reduce-by-key and count of instance b into its slice of the result
    buffers (the instance's output variables)."""
    L = R.lib()
    ok = p.okeys.ptr + 4 * GROUPS * b
    os_ = p.osum.ptr + 8 * GROUPS * b
    oc = p.ocnt.ptr + 4 * GROUPS * b
    nout = p.nout.ptr + 8 * b
    R._check("rfx_reduce_by_key_f64", L.rfx_reduce_by_key_f64(
        c.handle, ok, os_, nout, p.key2.ptr, p.val.ptr, p.tot.ptr, t, 4, 0))
    R._check("rfx_count_by_key", L.rfx_count_by_key(
        c.handle, p.okeys2.ptr, oc, p.nout2.ptr, p.key2.ptr, p.tot.ptr, t, 4))


class BatchedRunner:
    """All instances as one plan, captured into one HIP graph.  fused=False
    runs the plain form's operators instead of the fused ones (expansion
    and gather into 16-byte rows, two projections, then reduce-by-key and
    count-by-key), so that the gain of batching alone can be told from
    that of fusion."""

    def __init__(self, w: Workload, fused: bool = True):
        self.w = w
        self.fused = fused
        self.stream = R.Stream()
        self.ctx = R.Context(0, self.stream.handle)
        N, D = w.N, w.D
        B = R.DeviceBuffer
        self.ck, self.mw = B(4 * N), B(8 * N)
        self.dck, self.dgw = B(4 * D), B(4 * D)
        self.lb, self.lc = B(4 * N), B(4 * N)
        self.tot = R.u64_cell()
        cap = GROUPS * w.n_inst
        self.okeys, self.osum, self.ocnt = B(4 * cap), B(8 * cap), B(4 * cap)
        self.nout = R.u64_cell()
        self.seg_bits = max(1, int(np.ceil(np.log2(w.n_inst))))
        self.graph = None
        # The join's size, measured once: the plan's joined columns are
        # sized from it (the relations are resident and do not change).
        self._enqueue_bounds()
        self.ctx.synchronize()
        self.J = R.read_u64(self.tot)
        self.key2, self.val = B(4 * max(self.J, 1)), B(8 * max(self.J, 1))
        if not fused:
            self.out16 = B(16 * max(self.J, 1))
            self.okeys2, self.nout2 = B(4 * cap), R.u64_cell()

    def _enqueue_bounds(self) -> None:
        w, c, N, D = self.w, self.ctx, self.w.N, self.w.D
        kb = FK_BITS + self.seg_bits
        c.segment_keys(self.ck, w.fk, w.row_off_dev, w.n_inst, N, FK_BITS)
        c.copy(self.mw, w.measure, 8 * N)
        c.segment_keys(self.dck, w.dkey, w.dim_off_dev, w.n_inst, D, FK_BITS)
        c.copy(self.dgw, w.grp, 4 * D)
        c.sort_pairs(self.ck, self.mw, N, 4, kb, 8)
        c.sort_pairs(self.dck, self.dgw, D, 4, kb, 4)
        c.join_find_bounds(self.lb, self.lc, self.tot, self.ck, N, self.dck, D, 4)

    def _enqueue(self) -> None:
        w, c, J = self.w, self.ctx, self.J
        self._enqueue_bounds()
        if not self.fused:
            # The plain form's operators over all instances at once: the
            # 16-byte row (instance-prefixed key, grp, measure), the two
            # projections, then sum and count as two reductions.
            c.join_expand_gather(self.out16, 16, self.ck, 4, self.mw, 8, self.dgw, 4, self.lb,
                                 self.lc, w.N, self.tot, J, 96, 32)
            c.project(self.key2, 4, self.out16, 16, J, [96 + FK_BITS, 0],
                      [self.seg_bits, GRP_BITS], [GRP_BITS, 0])
            c.project(self.val, 8, self.out16, 16, J, [32], [64], [0])
            c.sort_pairs(self.key2, self.val, J, 4, self.seg_bits + GRP_BITS, 8)
            L = R.lib()
            R._check("rfx_reduce_by_key_f64", L.rfx_reduce_by_key_f64(
                c.handle, self.okeys.ptr, self.osum.ptr, self.nout.ptr, self.key2.ptr,
                self.val.ptr, self.tot.ptr, J, 4, 0))
            R._check("rfx_count_by_key", L.rfx_count_by_key(
                c.handle, self.okeys2.ptr, self.ocnt.ptr, self.nout2.ptr, self.key2.ptr,
                self.tot.ptr, J, 4))
            return
        # The joined row goes straight to the (instance, grp) key column and
        # the measure column: gather + project + split in one pass.
        c.join_expand_project(self.lb, self.lc, w.N, self.tot, J, self.ck, 4, self.mw, 8, self.dgw,
                              4, self.key2, 4, self.val, 8,
                              [(0, 0, FK_BITS, self.seg_bits, GRP_BITS),  # instance
                               (0, 2, 0, GRP_BITS, 0),                    # grp
                               (1, 1, 0, 64, 0)])                         # measure
        c.sort_pairs(self.key2, self.val, J, 4, self.seg_bits + GRP_BITS, 8)
        c.reduce_count_by_key_f64(self.okeys, self.osum, self.ocnt, self.nout, self.key2, self.val,
                                  self.tot, J, 4, 0, out_cap=GROUPS * w.n_inst)

    def autotune(self, reps: int = 2) -> dict:
        """Choose the operators' tile shapes for this workload by timing the
        whole plan (redfox_amd.autotune); capture() afterwards records them."""
        from .autotune import KNOBS, tune

        # The plan has no select: its knob would only time noise.
        knobs = {k: v for k, v in KNOBS.items() if not k.startswith("select.")}
        return tune(self.ctx, self._enqueue, knobs=knobs, reps=reps)

    def capture(self) -> None:
        """Run once eagerly (sizes the workspace), then record the graph."""
        self._enqueue()
        self.ctx.synchronize()
        self.ctx.graph_begin()
        self._enqueue()
        self.graph = self.ctx.graph_end()

    def run(self) -> None:
        self.graph.launch()

    def run_eager(self) -> None:
        self._enqueue()

    def synchronize(self) -> None:
        self.ctx.synchronize()

    def results(self):
        if R.read_u64(self.tot) != self.J:
            raise RuntimeError("batched join: the join's size changed since the plan was sized "
                               "(the relations must stay unchanged)")
        g = R.read_u64(self.nout)
        keys = self.okeys.download(np.uint32, g)
        return _results(keys, self.osum.download(np.float64, g), self.ocnt.download(np.uint32, g),
                        self.w.n_inst)
