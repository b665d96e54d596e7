"""Wagner build (poyd_amd.treebuild) with nonadditive characters, on the CPU
oracle backend: invariants that pin the join rule, and the simulator."""
import numpy as np
import pytest

from poyd_amd import synth, treebuild, treepass
from poyd_amd.treedata import Dataset
from oracle import pyoracle as po
from oracle_backend import OracleBackend


def fitch_length_brute(tree, X):
    """Textbook Fitch length of an unrooted tree (rooted at a leaf's edge)
    from explicit state sets, one pass per character."""
    leaf = min(tree.leaves)
    root_nb = tree.nbrs[leaf][0]
    total = 0
    for c in range(X.shape[1]):
        def down(v, p):
            if v in tree.leaves:
                return int(X[v - 1, c]), 0
            kids = [w for w in tree.nbrs[v] if w != p]
            (s1, c1), (s2, c2) = down(kids[0], v), down(kids[1], v)
            i = s1 & s2
            return (i, c1 + c2) if i else (s1 | s2, c1 + c2 + 1)
        s, k = down(root_nb, leaf)
        total += k + (0 if s & int(X[leaf - 1, c]) else 1)
    return total


def test_nonadd_joins_add_up_to_the_fitch_length():
    # for Fitch characters the join cost to an edge's lazy median is the exact
    # length increase, so the start length plus the joins is the final length
    w = synth.tree_workload(n_taxa=14, n_loci=0, n_nonadd=96, seed=11)
    ds = Dataset(list(w.names), [], w.nonadd)
    be = OracleBackend(po.CM2.sub_indel(1, 2), False)
    r = treebuild.wagner(ds, be)
    start = treepass.TreeEval(ds, treebuild.start_tree(ds.n_taxa), be).diagnose()
    final = treepass.TreeEval(ds, r.tree, be).diagnose()
    assert final.nonadd == start.nonadd + sum(r.join_costs)
    assert final.nonadd == fitch_length_brute(r.tree, w.nonadd)
    assert final.cost == final.nonadd


def test_build_is_a_binary_tree_over_all_taxa():
    w = synth.tree_workload(n_taxa=12, n_loci=2, lmin=30, lmax=70, n_nonadd=48, seed=12)
    ds = treebuild.dataset_of(w)
    r = treebuild.wagner(ds, OracleBackend(po.CM2.sub_indel(1, 2), False))
    t = r.tree
    n = ds.n_taxa
    assert t.leaves == set(range(1, n + 1))
    assert sorted(t.nbrs) == list(range(1, 2 * n - 1))
    assert all(len(t.nbrs[v]) == 1 for v in t.leaves)
    assert all(len(t.nbrs[v]) == 3 for v in t.nbrs if v > n)
    assert len(t.edges()) == 2 * n - 3
    assert len(r.join_costs) == n - 3 and min(r.join_costs) >= 0


def test_join_picks_the_first_least_edge():
    # a taxon identical to taxon 1 joins next to it at cost 0
    w = synth.tree_workload(n_taxa=6, n_loci=1, lmin=40, lmax=60, n_nonadd=32, seed=13)
    w.loci[0][5] = w.loci[0][0].copy()
    w.nonadd[5] = w.nonadd[0]
    ds = treebuild.dataset_of(w)
    r = treebuild.wagner(ds, OracleBackend(po.CM2.sub_indel(1, 2), False))
    assert r.join_costs[-1] == 0
    assert r.tree.nbrs[6][0] in r.tree.nbrs[1]


def test_simulator_shapes():
    w = synth.tree_workload(n_taxa=20, n_loci=3, lmin=100, lmax=200, n_nonadd=500, seed=14)
    assert len(w.loci) == 3 and all(len(l) == 20 for l in w.loci)
    assert all(s[0] == 16 and ((s[1:] & 15) != 0).all() and (s[1:] < 16).all() for l in w.loci for s in l)
    assert w.nonadd.shape == (20, 500)
    # one state per taxon and character, within the character's states
    assert (np.bitwise_and(w.nonadd, w.nonadd - 1) == 0).all()
    assert (w.nonadd < (1 << w.states)[None, :]).all()


def test_incremental_directional_medians_equal_a_full_pass():
    # after every join only the directions whose side holds the new vertex
    # are recomputed; the result must equal a full directional pass
    w = synth.tree_workload(n_taxa=13, n_loci=2, lmin=40, lmax=110, n_nonadd=64, seed=15)
    ds = treebuild.dataset_of(w)
    be = OracleBackend(po.CM2.sub_indel(1, 2), False)
    r = treebuild.wagner(ds, be)
    full = treepass.TreeEval(ds, r.tree, be).directional()
    assert set(full) == set(r.directional)
    for k, x in full.items():
        y = r.directional[k]
        assert (x.cost, x.mcc, x.na_cost) == (y.cost, y.mcc, y.na_cost), k
        assert np.array_equal(x.na, y.na), k
        for c in range(len(ds.chars)):
            assert np.array_equal(x.seqs[c], y.seqs[c]), (k, c)
