"""The reference's 52 tree-cost checks (test/cost_tests lines 11-14: scripts
cc, cc1, cc2, cc3 over test/1.fas .. 13.fas) replayed by the tree pass on the
CPU oracle.  These pin, end to end, the cost-matrix fills, the linear band
(deltaw, three-phase band, tie masks, backtrace order), the affine band and
backtrace, ancestor/median rules, closest and the distance band."""
import numpy as np
import pytest

from poyd_amd import treedata, treepass
import tree_fixtures as tf
from oracle_backend import OracleBackend, MODELS, model_cm


@pytest.mark.parametrize("script,name,want", tf.cases(), ids=lambda v: str(v))
def test_tree_cost_matches_reference(script, name, want):
    ds = tf.dataset(name)
    t = tf.tree(script, name)
    be = OracleBackend(model_cm(script), MODELS[script][2] is not None)
    d = treepass.TreeEval(ds, t, be).diagnose()
    assert d.cost == want


def test_fasta_reader_rules():
    text = ">a\nAC-GT#NN\nry\n\nb\nACGT#\n>c\nA#C\n"
    ds = treedata.read_fasta(text)
    assert ds.names == ["a", "b", "c"]
    assert len(ds.chars) == 2
    assert ds.chars[0][0].tolist() == [16, 1, 2, 4, 8]           # '-' dropped
    assert ds.chars[1][0].tolist() == [16, 15, 15, 5, 10]       # N, r, y
    assert ds.chars[1][1].tolist() == [16]                      # empty fragment
    assert treepass.is_empty(ds.chars[1][1])
    with pytest.raises(ValueError):
        treedata.read_fasta(">a\nAC#G\n>b\nACG\n")


def test_tree_codes_follow_reference_assignment():
    ds = treedata.read_fasta(">x\nA\n>y\nC\n>z\nG\n>w\nT\n")
    nw = treedata.parse_newick("(x ((y z) w))[t9.]")
    assert nw == [("x", (("y", "z"), "w"))]
    t = treedata.build_tree(nw[0], ds)
    # root takes 5 and dissolves; (y z) w gets 6, (y z) gets 7
    assert t.handle == 1
    assert t.nbrs[1] == [6]
    assert t.nbrs[6] == [1, 7, 4]
    assert t.nbrs[7] == [6, 2, 3]
    assert sorted(t.pre_order_edges()) == [(1, 6), (6, 4), (6, 7), (7, 2), (7, 3)]
    # multifurcations resolve left to right: (a b c) -> ((a b) c)
    t2 = treedata.build_tree(treedata.parse_newick("(x y z w)")[0], ds)
    assert sorted(len(v) for v in t2.nbrs.values()) == [1, 1, 1, 1, 3, 3]


def test_closest_table_rules():
    cm = model_cm("cc")
    cost = np.array([[cm.cost(a, b) for b in range(32)] for a in range(32)])
    ct = treepass.closest_table(cost)
    assert ct[1, 3] == 1           # A vs A|C -> A
    assert ct[2, 3] == 2
    assert ct[4, 3] == 1           # tie between A and C -> lowest bit
    assert ct[16, 17] == 16        # gap against A|gap keeps b, picks gap (cost 0)
    assert ct[17, 18] == 16        # both carry the gap bit -> gap
    assert ct[1, 17] == 1          # gap bit stripped
