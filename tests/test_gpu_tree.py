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
