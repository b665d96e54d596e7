"""The entry point that allocates no device memory
(poyg_pairs_cost_linear_noalloc) against the plan path it varies
(POYG_MODE_LINEAR_COST through poyg_plan_pairs / poyg_plan_run): identical
costs on two calls in a row, the second with a smaller batch (the scratch
keeps the first call's contents, which must not leak into the second), and
the size limits refused with a clear error."""
import ctypes as C

import numpy as np
import pytest

from poyd_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from poyd_amd import engine
    try:
        return engine.Context(0)
    except engine.PoygError as e:
        pytest.skip(f"no GPU: {e}")


def plan_costs(ctx, cm, w, dws):
    from poyd_amd import engine
    return engine.run_pairs(ctx, cm, engine.POYG_MODE_LINEAR_COST, w.a, w.b, dws).cost


def noalloc_costs(ctx, cm, w, dws):
    from poyd_amd import engine
    L = engine.load()
    n = len(w.a)
    seq = np.concatenate([w.a.seq, w.b.seq])
    d_seq = engine.DeviceBuffer.from_array(ctx, seq)
    d_cost = engine.DeviceBuffer(ctx, 4 * n)
    d_cost.upload(np.full(n, -7, np.int32))
    oa = np.ascontiguousarray(w.a.off, np.int64)
    ob = np.ascontiguousarray(w.b.off + len(w.a.seq), np.int64)
    la = np.ascontiguousarray(w.a.ln, np.int32)
    lb = np.ascontiguousarray(w.b.ln, np.int32)
    dw = np.ascontiguousarray(dws, np.int32)
    rc = L.poyg_pairs_cost_linear_noalloc(ctx.h, cm.h, n, oa.ctypes.data, la.ctypes.data, ob.ctypes.data,
                                          lb.ctypes.data, dw.ctypes.data, d_seq.ptr, d_cost.ptr, None)
    assert rc == 0, L.poyg_last_error()
    return d_cost.download(np.int32, n)


def test_noalloc_equals_plan_path_twice(ctx):
    from poyd_amd import engine
    print("[case] no-alloc linear costs, 1500 then 300 pairs", flush=True)
    cm = engine.CostMatrix.of_sub_indel(1, 2)
    big = synth.pairs_uniform(1500, 100, 1200, 0.12, 0.03, 71, "noalloc-a")
    small = synth.pairs_uniform(300, 20, 400, 0.2, 0.05, 72, "noalloc-b")
    for w in (big, small):
        dws = engine.batch_deltaw(w.a, w.b)
        got = noalloc_costs(ctx, cm, w, dws)
        want = plan_costs(ctx, cm, w, dws)
        assert np.array_equal(got, want)
        # the distance band of DOS.distance (?deltaw = max(8, |la - lb|))
        dws2 = engine.batch_deltaw(w.a, w.b, np.maximum(8, np.abs(w.a.ln.astype(np.int64) - w.b.ln)))
        assert np.array_equal(noalloc_costs(ctx, cm, w, dws2), plan_costs(ctx, cm, w, dws2))


def test_noalloc_refuses_what_does_not_fit(ctx):
    from poyd_amd import engine
    print("[case] no-alloc limits", flush=True)
    L = engine.load()
    cm = engine.CostMatrix.of_sub_indel(1, 2)
    cap = L.poyg_noalloc_scratch_bytes()
    assert cap == 32 << 20

    def call(n, la, lb, m=cm):
        oa = np.zeros(n, np.int64)
        lens_a = np.full(n, la, np.int32)
        lens_b = np.full(n, lb, np.int32)
        dw = np.full(n, 10, np.int32)
        return L.poyg_pairs_cost_linear_noalloc(ctx.h, m.h, n, oa.ctypes.data, lens_a.ctypes.data, oa.ctypes.data,
                                                lens_b.ctypes.data, dw.ctypes.data, 16, 16, None)

    # a shorter sequence past the systolic fill's 2,048 columns
    assert call(1, 3000, 2100) == -4
    assert b"2048" in L.poyg_last_error()
    # more pairs than the scratch holds (2 x 2,048 prefix sums each)
    assert call(4000, 2048, 2048) == -4
    assert b"scratch" in L.poyg_last_error()
    # an affine matrix
    assert call(1, 10, 10, engine.CostMatrix.of_sub_indel(1, 1, 3)) == -4
