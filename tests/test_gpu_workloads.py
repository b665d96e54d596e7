"""The three-way and tree workloads at their own sizes on the GPU.

* three-way-4096x100-500: 16 seeded triples plus the three with the most
  cube cells bit-exact against the oracle's align_3 / median_3; over all
  4,096 triples, the aligned sequences strip back to the inputs and the
  median is the three-way median of the last column, once per column
  (algn_get_median_3d, src/algn.c:4160-4173).
* tree-128taxa: the Wagner build on the GPU, then the oracle's diagnosis of
  the built tree (Adjusted cost, Fitch length, root) equals the GPU's, and
  at 8 seeded steps of the build every candidate join cost equals the one
  the oracle computes on the same partial tree.
"""
import copy

import numpy as np
import pytest

from poyd_amd import synth

pytestmark = pytest.mark.gpu
GAP = 16


@pytest.fixture(scope="module")
def ctx():
    from poyd_amd import engine
    try:
        return engine.Context(0)
    except engine.PoygError as e:
        pytest.skip(f"no GPU: {e}")


def note(msg):
    print(f"[case] {msg}", flush=True)


def strip(s):
    return s[s != GAP]


@pytest.fixture(scope="module")
def cube_run(ctx):
    from poyd_amd import engine
    note("three-way-4096x100-500: whole workload on the GPU")
    w = synth.workload_three_way_4096()
    cm2 = engine.CostMatrix.of_sub_indel(1, 2)
    seq = np.concatenate([b.seq for b in w.batches])
    base = np.cumsum([0] + [len(b.seq) for b in w.batches])
    plan = engine.TriplePlan(ctx, engine.CostMatrix3(cm2),
                             *[x for q, b in enumerate(w.batches) for x in (b.off + base[q], b.ln)])
    d_seq = engine.DeviceBuffer.from_array(ctx, seq)
    d_cost = engine.DeviceBuffer(ctx, 4 * plan.n)
    d_out = engine.DeviceBuffer(ctx, plan.out_bytes)
    d_len = engine.DeviceBuffer(ctx, 4 * 4 * plan.n)
    plan.run(d_seq.ptr, d_cost.ptr, d_out.ptr, d_len.ptr)
    ctx.sync()
    assert plan.status() == 0
    res = engine.PairResult(d_cost.download(np.int32, plan.n), d_out.download(np.uint8, plan.out_bytes),
                            d_len.download(np.int32, 4 * plan.n), plan.out_off, plan.cap, 4)
    plan.close()
    return w, res


def test_three_way_4096_sample_matches_oracle(cube_run, oracle):
    w, res = cube_run
    n = len(w.batches[0].ln)
    cells = w.batches[0].ln.astype(np.int64) * w.batches[1].ln * w.batches[2].ln
    pick = list(np.argsort(-cells, kind="stable")[:3]) + \
        list(np.random.default_rng(2026).choice(n, size=16, replace=False))
    ref = oracle.CM3(oracle.CM2.sub_indel(1, 2))
    for k in pick:
        k = int(k)
        note(f"three-way triple {k} ({[len(w.get(q, k)) for q in range(3)]})")
        want = oracle.align_3(*[w.get(q, k) for q in range(3)], ref)
        assert res.cost[k] == want[0], k
        for q in range(3):
            assert np.array_equal(res.seq(k, q), want[1 + q]), (k, q)
        assert np.array_equal(res.seq(k, 3), oracle.median_3(*want[1:], ref)), k


def test_three_way_4096_all_strip_and_median(cube_run):
    from poyd_amd import engine
    w, res = cube_run
    cm3 = engine.CostMatrix3(engine.CostMatrix.of_sub_indel(1, 2))
    med_of = {}
    for k in range(len(w.batches[0].ln)):
        al = [res.seq(k, q) for q in range(3)]
        assert len(al[0]) == len(al[1]) == len(al[2]), k
        for q in range(3):
            assert np.array_equal(strip(al[q]), strip(w.get(q, k))), (k, q)
        key = (int(al[0][-1]), int(al[1][-1]), int(al[2][-1]))
        if key not in med_of:
            med_of[key] = cm3.median(*key)
        med = res.seq(k, 3)
        assert len(med) == len(al[0]) and (med == med_of[key]).all(), k


@pytest.fixture(scope="module")
def tree_build(ctx):
    from poyd_amd import engine, treebuild, treepass
    note("tree-128taxa: Wagner build on the GPU")
    w = synth.workload_tree_128()
    ds = treebuild.dataset_of(w)
    be = treepass.EngineBackend(ctx, engine.CostMatrix.of_sub_indel(1, 2))
    sample = set(int(x) for x in np.random.default_rng(128).choice(np.arange(4, ds.n_taxa + 1), 8, replace=False))
    seen = {}

    def on_step(x, t, edges, jc):
        if x in sample:
            seen[x] = (copy.deepcopy(t), list(edges), np.array(jc))

    r = treebuild.wagner(ds, be, on_step=on_step)
    d = treepass.TreeEval(ds, r.tree, be).diagnose()
    return ds, r, d, seen


def test_tree_128_diagnosis_matches_oracle(tree_build):
    from poyd_amd import treepass
    from oracle import pyoracle as po
    from oracle_backend import OracleBackend
    ds, r, d, _ = tree_build
    note("tree-128taxa: oracle diagnosis of the built tree")
    o = treepass.TreeEval(ds, r.tree, OracleBackend(po.CM2.sub_indel(1, 2), False)).diagnose()
    assert (d.cost, d.nonadd, d.root, d.downpass) == (o.cost, o.nonadd, o.root, o.downpass)


def test_tree_128_wagner_steps_match_oracle(tree_build):
    from poyd_amd import treebuild, treepass
    from oracle import pyoracle as po
    from oracle_backend import OracleBackend
    ds, r, _, seen = tree_build
    assert len(seen) == 8
    ob = OracleBackend(po.CM2.sub_indel(1, 2), False)
    for x in sorted(seen):
        t, edges, jc = seen[x]
        note(f"tree-128taxa: join costs of taxon {x} on {len(edges)} edges")
        ev = treepass.TreeEval(ds, t, ob)
        dd = ev.directional()
        emed = ev._medians([(dd[(b, a)], dd[(a, b)]) for (a, b) in edges])
        want = treebuild.join_costs(ev, x, edges, emed)
        assert np.array_equal(np.asarray(want), jc), x
