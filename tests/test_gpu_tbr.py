"""TBR rounds on the GPU (poyd_amd.tbr over poyd_amd.seqstore.DeviceStore)
against the same rounds on the CPU oracle (OracleStore): every join cost,
every broken-tree cost and the best candidate are identical; the guided
round's exact rescoring of the candidates the model picks equals the
exhaustive round's costs of those candidates."""
import numpy as np
import pytest

from poyd_amd import synth, tbr, treebuild
from poyd_amd.treedata import Dataset
from oracle_backend import OracleBackend, OracleStore

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from poyd_amd import engine
    try:
        return engine.Context(0)
    except engine.PoygError as e:
        pytest.skip(f"no GPU: {e}")


def note(msg):
    print(f"[case] {msg}", flush=True)


def setup_round(store, ds, t):
    L = len(ds.chars)
    leaf_ids = {v: store.add([ds.chars[c][v - 1] for c in range(L)]) for v in t.leaves}
    dirs, index = tbr.directional(store, t, leaf_ids)
    sp = tbr.neighbourhood(t, index)
    cand = tbr.candidates(sp, dirs, len(dirs))
    return dirs, index, sp, cand


@pytest.mark.parametrize("n_taxa,loci,lmin,lmax,seed", [(14, 2, 40, 120, 31), (16, 3, 100, 250, 32)])
def test_gpu_tbr_round_matches_oracle(ctx, oracle, n_taxa, loci, lmin, lmax, seed):
    from poyd_amd import engine, seqstore, surrogate
    note(f"tbr {n_taxa} taxa {loci} loci {lmin}-{lmax}")
    w = synth.tree_workload(n_taxa=n_taxa, n_loci=loci, lmin=lmin, lmax=lmax, n_nonadd=0, seed=seed)
    ds = Dataset(list(w.names), [list(l) for l in w.loci], None)
    cm = oracle.CM2.sub_indel(1, 2)
    t = treebuild.wagner(ds, OracleBackend(cm, False)).tree
    ostore = OracleStore(cm, False)
    o = setup_round(ostore, ds, t)
    ores = tbr.exhaustive_round(ostore, o[2], o[0], o[3])
    gstore = seqstore.DeviceStore(ctx, engine.CostMatrix.of_sub_indel(1, 2), 1 << 28)
    g = setup_round(gstore, ds, t)
    assert np.array_equal(g[0].cost, o[0].cost)
    for k in range(len(o[0])):
        for c in range(loci):
            assert np.array_equal(gstore.get(g[0].ids[k, c]), ostore.get(o[0].ids[k, c])), (k, c)
    gres = tbr.exhaustive_round(gstore, g[2], g[0], g[3])
    assert np.array_equal(gres.dist, ores.dist)
    assert np.array_equal(gres.estimates, ores.estimates)
    assert gres.best == ores.best and gres.best_estimate == ores.best_estimate
    # a guided round with an untrained model: its exact rescoring of the 8%
    # it picks equals the exhaustive round on those candidates
    sc = surrogate.Scorer(ctx, batch_max=1024, seed=3)
    r, info = surrogate.guided_round(gstore, sc, g[2], g[0], g[1], g[3], 0.08)
    assert len(r.scored) == int(np.ceil(0.08 * len(g[3].brk)))
    assert np.array_equal(r.dist, gres.dist[r.scored])
    assert r.best_estimate == int(gres.estimates[r.scored].min())
    sc.close()
