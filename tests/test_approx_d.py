"""Iterative ApproxD (AllDirChar.adjust_tree with Sequence.readjust,
src/allDirChar.ml:488-670, src/sequence.ml:1166-1199) on the CPU oracle
backend: the layered schedule gives exactly the one-vertex-at-a-time
result, passes never raise the cost, zero passes leave the single
assignment, and the readjustment's centre rule on hand-checked triples.
The pass itself is a restatement (no reference fixture holds ApproxD
results)."""
import numpy as np
import pytest

from poyd_amd import synth, treebuild, treepass
from oracle import pyoracle as po
from oracle_backend import OracleBackend


def case(n_taxa, seed, lmax=160):
    w = synth.tree_workload(n_taxa=n_taxa, n_loci=2, lmin=60, lmax=lmax, n_nonadd=0, seed=seed)
    ds = treebuild.dataset_of(w)
    be = OracleBackend(po.CM2.sub_indel(1, 2), False)
    tree = treebuild.wagner(ds, be).tree
    return ds, tree, be


@pytest.mark.parametrize("n_taxa,seed", [(8, 3), (12, 4), (16, 5)])
def test_layered_passes_equal_sequential_passes(n_taxa, seed):
    ds, tree, be = case(n_taxa, seed)
    ev = treepass.TreeEval(ds, tree, be)
    d = ev.diagnose()
    adj_l, cost_l, kept_l = ev.approx_d(d)
    adj_s, cost_s, kept_s = treepass.TreeEval(ds, tree, be).approx_d(d, sequential=True)
    assert cost_l == cost_s and kept_l == kept_s
    for v in adj_l:
        for c in range(len(ds.chars)):
            assert np.array_equal(adj_l[v][c], adj_s[v][c]), (v, c)
    # passes are kept only when they lower the cost, and the reported cost
    # is the tree cost of the adjusted sequences
    assert cost_l <= d.cost - d.nonadd
    assert cost_l == treepass.TreeEval(ds, tree, be)._edge_cost(adj_l)
    if kept_l:
        assert cost_l < d.cost - d.nonadd


def test_zero_passes_keep_the_single_assignment():
    ds, tree, be = case(10, 7)
    ev = treepass.TreeEval(ds, tree, be)
    d = ev.diagnose()
    adj, cost, kept = ev.approx_d(d, iterations=0)
    assert kept == 0 and cost == d.cost - d.nonadd
    for v in adj:
        for c in range(len(ds.chars)):
            assert np.array_equal(adj[v][c], d.singles[v][c])


def test_readjust_centre_rule():
    """Three copies of one sequence: every pairwise median is the sequence,
    every total is 0, the centre is ab|c and closest returns the sequence;
    a vertex whose three neighbours agree takes their sequence."""
    ds, tree, be = case(8, 3)
    ev = treepass.TreeEval(ds, tree, be)
    s = np.array([16, 1, 2, 4, 8, 1, 1, 2], np.uint8)
    other = np.array([16, 8, 8, 8], np.uint8)
    (seqs, changed), = ev._readjust([([s, s], [s, s], [s, s], [other, s], None)])
    assert np.array_equal(seqs[0], s) and np.array_equal(seqs[1], s)
    assert changed == {0}


@pytest.mark.parametrize("n_taxa,seed", [(8, 3), (10, 4)])
def test_three_d_layered_passes_equal_sequential_passes(n_taxa, seed):
    """Iterative ThreeD (readjust_3d: Powell's search of the two children
    and the parent, the Three_D median of its columns): the layered schedule
    gives the one-vertex-at-a-time result, and passes never raise the cost."""
    ds, tree, be = case(n_taxa, seed, lmax=100)
    ev = treepass.TreeEval(ds, tree, be)
    d = ev.diagnose()
    adj_l, cost_l, kept_l = ev.approx_d(d, three_d=True)
    adj_s, cost_s, kept_s = treepass.TreeEval(ds, tree, be).approx_d(d, sequential=True, three_d=True)
    assert cost_l == cost_s and kept_l == kept_s
    for v in adj_l:
        for c in range(len(ds.chars)):
            assert np.array_equal(adj_l[v][c], adj_s[v][c]), (v, c)
    assert cost_l <= d.cost - d.nonadd
    assert cost_l == treepass.TreeEval(ds, tree, be)._edge_cost(adj_l)


def test_readjust_3d_rules():
    """readjust_3d keeps the vertex's sequence when the two children, the
    vertex and the parent have one length; otherwise the new sequence is the
    Three_D median of the search's columns (three copies of a sequence: the
    sequence itself, gaps dropped behind the leading gap)."""
    ds, tree, be = case(8, 3)
    ev = treepass.TreeEval(ds, tree, be)
    s = np.array([16, 1, 2, 4, 8, 1, 1, 2], np.uint8)
    t = np.array([16, 8, 8, 8, 2, 2, 2, 1], np.uint8)
    (seqs, changed), = ev._readjust_3d([([s, s], [s, t], [s, s], [t, s], None)])
    # character 0: s, s, s around t of the same length: unchanged
    assert np.array_equal(seqs[0], t) and 0 not in changed
    u = np.array([16, 1, 2, 4, 8], np.uint8)
    (seqs, changed), = ev._readjust_3d([([s], [s], [s], [u], {0})])
    assert np.array_equal(seqs[0], s) and changed == {0}
    med, cost = be.powell_median([(s, s, s)])[0]
    assert np.array_equal(med, s) and cost == 1   # the search's loop starts at cost 1
