"""Tree evaluation with every alignment on the GPU (poyd_amd.treepass with
the libpoyg backend): the reference's 52 tree costs, and every vertex's
single sequence bit-identical to the oracle-backed pass."""
import numpy as np
import pytest

from poyd_amd import treepass
import tree_fixtures as tf
from oracle_backend import OracleBackend, MODELS, model_cm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from poyd_amd import engine
    return engine.Context(0)


def engine_cm(script):
    from poyd_amd import engine
    sub, indel, go = MODELS[script]
    return engine.CostMatrix.of_sub_indel(sub, indel, go)


@pytest.mark.parametrize("script,name,want", tf.cases(), ids=lambda v: str(v))
def test_gpu_tree_cost(ctx, script, name, want):
    print(f"tree {script} {name}", flush=True)
    d = treepass.TreeEval(tf.dataset(name), tf.tree(script, name),
                          treepass.EngineBackend(ctx, engine_cm(script))).diagnose()
    assert d.cost == want


@pytest.mark.parametrize("script,name", [("cc", "1.fas"), ("cc1", "9.fas"), ("cc2", "13.fas"), ("cc3", "1.fas")])
def test_gpu_tree_singles_match_oracle(ctx, script, name):
    print(f"singles {script} {name}", flush=True)
    ds, t = tf.dataset(name), tf.tree(script, name)
    g = treepass.TreeEval(ds, t, treepass.EngineBackend(ctx, engine_cm(script))).diagnose()
    o = treepass.TreeEval(ds, t, OracleBackend(model_cm(script), MODELS[script][2] is not None)).diagnose()
    assert g.root == o.root and g.downpass == o.downpass and g.cost == o.cost
    for v in o.singles:
        for c in range(len(ds.chars)):
            assert np.array_equal(g.singles[v][c], o.singles[v][c]), (v, c)


@pytest.mark.parametrize("n_taxa,n_loci,lmin,lmax,n_nonadd,seed", [
    (10, 2, 40, 120, 64, 21),
    (24, 3, 150, 400, 1000, 22),
])
def test_gpu_wagner_build_matches_oracle(ctx, n_taxa, n_loci, lmin, lmax, n_nonadd, seed):
    from poyd_amd import engine, synth, treebuild
    from oracle import pyoracle as po
    print(f"wagner {n_taxa} taxa {n_loci} loci {n_nonadd} nonadditive", flush=True)
    w = synth.tree_workload(n_taxa=n_taxa, n_loci=n_loci, lmin=lmin, lmax=lmax, n_nonadd=n_nonadd, seed=seed)
    ds = treebuild.dataset_of(w)
    gb = treepass.EngineBackend(ctx, engine.CostMatrix.of_sub_indel(1, 2))
    ob = OracleBackend(po.CM2.sub_indel(1, 2), False)
    g = treebuild.wagner(ds, gb)
    o = treebuild.wagner(ds, ob)
    assert g.join_costs == o.join_costs
    assert g.tree.nbrs == o.tree.nbrs
    gd = treepass.TreeEval(ds, g.tree, gb).diagnose()
    od = treepass.TreeEval(ds, o.tree, ob).diagnose()
    assert (gd.cost, gd.nonadd, gd.root) == (od.cost, od.nonadd, od.root)


@pytest.mark.parametrize("n_taxa,seed,affine", [(12, 4, False), (24, 6, False), (12, 8, True)])
def test_gpu_approx_d_matches_oracle(ctx, n_taxa, seed, affine):
    """Iterative ApproxD with every alignment on the GPU equals the same
    passes on the oracle: adjusted sequences, tree cost, passes kept."""
    from poyd_amd import engine, synth, treebuild
    from oracle import pyoracle as po
    print(f"approx_d {n_taxa} taxa affine={affine}", flush=True)
    w = synth.tree_workload(n_taxa=n_taxa, n_loci=2, lmin=60, lmax=300, n_nonadd=0, seed=seed)
    ds = treebuild.dataset_of(w)
    if affine:
        gb = treepass.EngineBackend(ctx, engine.CostMatrix.of_sub_indel(1, 1, 3))
        ob = OracleBackend(po.CM2.sub_indel(1, 1, 3), True)
    else:
        gb = treepass.EngineBackend(ctx, engine.CostMatrix.of_sub_indel(1, 2))
        ob = OracleBackend(po.CM2.sub_indel(1, 2), False)
    tree = treebuild.wagner(ds, ob).tree
    g_adj, g_cost, g_kept = treepass.TreeEval(ds, tree, gb).approx_d()
    o_adj, o_cost, o_kept = treepass.TreeEval(ds, tree, ob).approx_d()
    assert (g_cost, g_kept) == (o_cost, o_kept)
    for v in o_adj:
        for c in range(len(ds.chars)):
            assert np.array_equal(g_adj[v][c], o_adj[v][c]), (v, c)


@pytest.mark.parametrize("n_taxa,seed,affine", [(12, 4, False), (10, 8, True)])
def test_gpu_three_d_matches_oracle(ctx, n_taxa, seed, affine):
    """Iterative ThreeD (readjust_3d: Powell's search on the GPU, the
    Three_D median of its columns) equals the same passes on the oracle's
    restatement of the search: adjusted sequences, tree cost, passes kept."""
    from poyd_amd import engine, synth, treebuild
    from oracle import pyoracle as po
    print(f"three_d {n_taxa} taxa affine={affine}", flush=True)
    w = synth.tree_workload(n_taxa=n_taxa, n_loci=2, lmin=60, lmax=140, n_nonadd=0, seed=seed)
    ds = treebuild.dataset_of(w)
    if affine:
        gb = treepass.EngineBackend(ctx, engine.CostMatrix.of_sub_indel(1, 1, 3))
        ob = OracleBackend(po.CM2.sub_indel(1, 1, 3), True)
    else:
        gb = treepass.EngineBackend(ctx, engine.CostMatrix.of_sub_indel(1, 2))
        ob = OracleBackend(po.CM2.sub_indel(1, 2), False)
    tree = treebuild.wagner(ds, ob).tree
    g_adj, g_cost, g_kept = treepass.TreeEval(ds, tree, gb).approx_d(three_d=True)
    o_adj, o_cost, o_kept = treepass.TreeEval(ds, tree, ob).approx_d(three_d=True)
    assert (g_cost, g_kept) == (o_cost, o_kept)
    for v in o_adj:
        for c in range(len(ds.chars)):
            assert np.array_equal(g_adj[v][c], o_adj[v][c]), (v, c)
