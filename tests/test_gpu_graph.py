"""Recorded passes (poyg_graph_*): a tree's Fitch down and up passes and the
run of an affine and a linear alignment plan, each recorded once as a HIP
graph and replayed three times, with other work queued and synchronized
between the replays; every replay's outputs equal the eager run's, byte for
byte (the counters and flags the passes reset are reset by kernels: a graph
holds no memset node)."""
import numpy as np
import pytest

from test_gpu_parity import batch_of, mixed_pairs, note

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from poyd_amd import engine
    try:
        return engine.Context(0)
    except engine.PoygError as e:
        pytest.skip(f"no GPU: {e}")


def fitch_tree(rng, n_taxa, L):
    from poyd_amd import synth
    parent, _, leaves = synth.yule_tree(rng, n_taxa, 0.05)
    nn = len(parent)
    kids = [[] for _ in range(nn)]
    for v in range(1, nn):
        kids[parent[v]].append(v)
    internal = [v for v in range(nn) if kids[v]]
    height = np.zeros(nn, np.int64)
    for v in range(nn - 1, -1, -1):
        if kids[v]:
            height[v] = 1 + max(height[c] for c in kids[v])
    depth = np.zeros(nn, np.int64)
    for v in range(1, nn):
        depth[v] = depth[parent[v]] + 1
    fin = {v: nn + q for q, v in enumerate(internal)}
    down = [np.array([[kids[v][0], kids[v][1], v] for v in internal if height[v] == h], np.int32)
            for h in range(1, int(height.max()) + 1)]
    up = []
    for dpt in range(1, int(depth.max()) + 1):
        vs = [v for v in internal if depth[v] == dpt]
        if vs:
            up.append(np.array([[fin[parent[v]], v, kids[v][0], kids[v][1], fin[v]] for v in vs], np.int32))
    return nn, leaves, fin, down, up


def test_fitch_passes_graph_replays_equal_eager(ctx):
    from poyd_amd import engine
    note("graph: Fitch down + up passes")
    rng = np.random.default_rng(61)
    L = 4096
    nn, leaves, fin, down, up = fitch_tree(rng, 96, L)
    rows = nn + len(fin)
    lib = engine.load()
    chars = engine.DeviceBuffer(ctx, rows * L)
    chars.upload(np.zeros(rows * L, np.uint8))
    states = rng.integers(2, 9, size=L).astype(np.uint8)
    for v in leaves:
        row = np.left_shift(np.uint8(1), rng.integers(0, 8, size=L, dtype=np.uint8) % states).astype(np.uint8)
        lib.poyg_dev_upload(ctx.h, chars.ptr + int(v) * L, row.ctypes.data, L)
    dl = [[engine.DeviceBuffer.from_array(ctx, np.ascontiguousarray(a[:, c])) for c in range(3)] for a in down]
    du = [[engine.DeviceBuffer.from_array(ctx, np.ascontiguousarray(a[:, c])) for c in range(5)] for a in up]
    added = engine.DeviceBuffer(ctx, 8 * nn)

    def passes(stream, up_too=True):
        engine.dev_zero(ctx, added.ptr, 8 * nn, stream)
        for a, (ia, ib, ir) in zip(down, dl):
            engine._check(lib.poyg_fitch_down(ctx.h, chars.ptr, L, len(a), ia.ptr, ib.ptr, ir.ptr, added.ptr,
                                              stream), "poyg_fitch_down")
        if up_too:
            for a, (iA, iN, i1, i2, iF) in zip(up, du):
                engine._check(lib.poyg_fitch_up(ctx.h, chars.ptr, L, len(a), iA.ptr, iN.ptr, i1.ptr, i2.ptr,
                                                iF.ptr, stream), "poyg_fitch_up")

    # the root's final sets are its preliminary ones (the leaves never change)
    passes(None, up_too=False)
    ctx.sync()
    root = chars.download(np.uint8, L)
    lib.poyg_dev_upload(ctx.h, chars.ptr + fin[0] * L, root.ctypes.data, L)
    passes(None)
    ctx.sync()
    want_chars = chars.download(np.uint8, rows * L)
    want_added = added.download(np.int64, nn)
    g = engine.Graph(ctx)
    passes(g.stream)
    g.end()
    other = engine.CostMatrix.of_sub_indel(1, 2)
    A, B = mixed_pairs(np.random.default_rng(5), 40)
    for rep in range(3):
        # scribble over the internal rows and the counts, then replay
        scratch = np.full((rows - len(leaves)) * L, 0x5a, np.uint8)
        internal_rows = sorted(set(range(rows)) - set(int(v) for v in leaves))
        for k, r in enumerate(internal_rows):
            lib.poyg_dev_upload(ctx.h, chars.ptr + r * L, scratch[k * L:(k + 1) * L].ctypes.data, L)
        lib.poyg_dev_upload(ctx.h, chars.ptr + fin[0] * L, root.ctypes.data, L)
        added.upload(np.full(nn, 77, np.int64))
        g.launch()
        engine.run_pairs(ctx, other, engine.POYG_MODE_LINEAR_ALIGN, batch_of(A), batch_of(B))   # eager work between
        ctx.sync()
        assert np.array_equal(chars.download(np.uint8, rows * L), want_chars), rep
        assert np.array_equal(added.download(np.int64, nn), want_added), rep
    g.close()


@pytest.mark.parametrize("mode", ["affine", "linear"])
def test_plan_run_graph_replays_equal_eager(ctx, oracle, mode):
    from poyd_amd import engine, synth
    note(f"graph: {mode} plan run")
    rng = np.random.default_rng(62 + (mode == "linear"))
    A, B = mixed_pairs(rng, 400, "any" if mode == "affine" else "dna")
    w = synth.pairs_uniform(600, 50, 1500, 0.15, 0.04, 17, "graph-plan")
    A += [w.a.get(k) for k in range(600)]
    B += [w.b.get(k) for k in range(600)]
    if mode == "affine":
        cm = engine.CostMatrix.of_sub_indel(1, 1, 3)
        m = engine.POYG_MODE_AFFINE_ALIGN
        dws = None
    else:
        cm = engine.CostMatrix.of_sub_indel(1, 2)
        m = engine.POYG_MODE_LINEAR_ALIGN
        dws = engine.batch_deltaw(batch_of(A), batch_of(B))
    a, b = batch_of(A), batch_of(B)
    seq = np.concatenate([a.seq, b.seq])
    plan = engine.PairPlan(ctx, cm, m, a.off, a.ln, b.off + len(a.seq), b.ln, dws)
    d_seq = engine.DeviceBuffer.from_array(ctx, seq)
    d_cost = engine.DeviceBuffer(ctx, 4 * plan.n)
    d_out = engine.DeviceBuffer(ctx, plan.out_bytes)
    d_len = engine.DeviceBuffer(ctx, 4 * plan.n * plan.slots)
    plan.run(d_seq.ptr, d_cost.ptr, d_out.ptr, d_len.ptr)
    ctx.sync()
    want = (d_cost.download(np.int32, plan.n), d_out.download(np.uint8, plan.out_bytes),
            d_len.download(np.int32, plan.n * plan.slots))
    assert plan.status() == 0
    # a sample against the oracle, so that the eager run is itself right
    ref = oracle.CM2.sub_indel(1, 1, 3) if mode == "affine" else oracle.CM2.sub_indel(1, 2)
    for k in range(0, plan.n, 97):
        got = (oracle.align_affine_3(A[k], B[k], ref)[3] if mode == "affine"
               else oracle.cost_2(A[k], B[k], ref))
        assert want[0][k] == got, k
    g = engine.Graph(ctx)
    plan.run(d_seq.ptr, d_cost.ptr, d_out.ptr, d_len.ptr, g.stream)
    g.end()
    other = engine.CostMatrix.of_sub_indel(2, 1, 1) if mode == "affine" else engine.CostMatrix.of_sub_indel(1, 1)
    C_, D_ = mixed_pairs(np.random.default_rng(9), 60)
    for rep in range(3):
        d_cost.upload(np.full(plan.n, -7, np.int32))
        d_out.upload(np.full(plan.out_bytes, 0xee, np.uint8))
        d_len.upload(np.full(plan.n * plan.slots, -1, np.int32))
        g.launch()
        engine.run_pairs(ctx, other, engine.POYG_MODE_AFFINE_ALIGN if mode == "affine" else engine.POYG_MODE_LINEAR_ALIGN,
                         batch_of(C_), batch_of(D_))
        ctx.sync()
        assert plan.status() == 0, rep
        got = (d_cost.download(np.int32, plan.n), d_out.download(np.uint8, plan.out_bytes),
               d_len.download(np.int32, plan.n * plan.slots))
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[2], want[2]), rep
        # bytes left of each right-aligned output are unspecified: compare the outputs
        res = engine.PairResult(got[0], got[1], got[2], plan.out_off, plan.cap, plan.slots)
        ref_res = engine.PairResult(want[0], want[1], want[2], plan.out_off, plan.cap, plan.slots)
        for k in range(plan.n):
            for s in range(plan.slots):
                assert np.array_equal(res.seq(k, s), ref_res.seq(k, s)), (rep, k, s)
    g.close()
    plan.close()
