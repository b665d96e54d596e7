"""Repeated runs of large batches give identical outputs, byte for byte.

Cross-lane moves in the anti-diagonal fills (DPP) follow EXEC changes only
after a guard; a missing guard showed up as wrong tie masks in a few pairs
per batch, different ones per run.  These tests run the same large batches
five times each and compare every cost and every output sequence (the bytes
the output lengths cover) of every run with the first run's."""
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


def outputs(plan, d_cost, d_out, d_len):
    cost = d_cost.download(np.int32, plan.n)
    out = d_out.download(np.uint8, plan.out_bytes)
    lens = d_len.download(np.int32, plan.n * plan.slots).reshape(plan.n, plan.slots)
    seqs = []
    for k in range(plan.n):
        cap = int(plan.cap[k])
        for s in range(plan.slots):
            base = int(plan.out_off[k]) + s * cap
            seqs.append(out[base + cap - int(lens[k, s]): base + cap].tobytes())
    return cost, lens, seqs


@pytest.mark.parametrize("kind,n", [("affine", 16384), ("linear", 4096)])
def test_repeated_runs_identical(ctx, kind, n):
    from poyd_amd import engine
    print(f"[case] {kind} {n} pairs x 5 runs", flush=True)
    if kind == "affine":
        w = synth.pairs_loguniform(n, 50, 5000, 0.05, 0.30, 811, "repeat-affine")
        cm = engine.CostMatrix.of_sub_indel(1, 1, 3)
        plan = engine.PairPlan(ctx, cm, engine.POYG_MODE_AFFINE_ALIGN, w.a.off, w.a.ln, w.b.off + len(w.a.seq),
                               w.b.ln)
    else:
        w = synth.pairs_uniform(n, 200, 1500, 0.12, 0.03, 812, "repeat-linear")
        cm = engine.CostMatrix.of_sub_indel(1, 2)
        plan = engine.PairPlan(ctx, cm, engine.POYG_MODE_LINEAR_ALIGN, w.a.off, w.a.ln, w.b.off + len(w.a.seq),
                               w.b.ln, engine.batch_deltaw(w.a, w.b))
    d_seq = engine.DeviceBuffer.from_array(ctx, np.concatenate([w.a.seq, w.b.seq]))
    d_cost = engine.DeviceBuffer(ctx, 4 * plan.n)
    d_out = engine.DeviceBuffer(ctx, plan.out_bytes)
    d_len = engine.DeviceBuffer(ctx, 4 * plan.n * plan.slots)
    first = None
    for r in range(5):
        plan.run(d_seq.ptr, d_cost.ptr, d_out.ptr, d_len.ptr)
        ctx.sync()
        assert plan.status() == 0
        got = outputs(plan, d_cost, d_out, d_len)
        if first is None:
            first = got
            continue
        assert np.array_equal(got[0], first[0]), r
        assert np.array_equal(got[1], first[1]), r
        bad = [k for k, (a, b) in enumerate(zip(got[2], first[2])) if a != b]
        assert not bad, (r, bad[:10])
    plan.close()
