"""The batched union kernel (poyg_union_merge) against the oracle's
union_merge restatement (src/union.c:102-164): every item and offset of
every union, for merges of random caterpillar trees with linear matrices
(column medians taken from the matrix on the device, and given) and affine
ones (align_affine_3's median with gaps given), and a merge whose children
do not match its alignment (reported, not written)."""
import numpy as np
import pytest

from test_union import random_tree_merges

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from poyd_amd import engine
    try:
        return engine.Context(0)
    except engine.PoygError as e:
        pytest.skip(f"no GPU: {e}")


class Packer:
    """Packs host arrays into one device buffer and hands out pointers."""

    def __init__(self):
        self.parts, self.size = [], 0

    def put(self, arr):
        b = np.ascontiguousarray(arr).view(np.uint8).ravel()
        at = self.size
        self.parts.append((at, b))
        self.size = (at + len(b) + 255) & ~255
        return at

    def upload(self, ctx):
        from poyd_amd import engine
        blob = np.zeros(max(self.size, 256), np.uint8)
        for at, b in self.parts:
            blob[at:at + len(b)] = b
        return engine.DeviceBuffer.from_array(ctx, blob)


def run_batch(ctx, cm, merges, give_median):
    from poyd_amd import engine
    P = Packer()
    spec = []
    for a, b, med, ua, ub, _want in merges:
        cap = len(ua[0]) + len(ub[0]) + 2
        spec.append(dict(a=P.put(a), b=P.put(b), m=P.put(med) if give_median else None, uas=P.put(ua[0]),
                         ube=P.put(ub[1].astype(np.int32)), ubs=P.put(ub[0]), uae=P.put(ua[1].astype(np.int32)),
                         os=P.put(np.zeros(cap, np.uint8)), oe=P.put(np.zeros(cap, np.int32)),
                         oa=P.put(np.zeros(cap, np.int32)), ob=P.put(np.zeros(cap, np.int32)), len=len(a),
                         na=len(ua[0]), nb=len(ub[0]), cap=cap))
    d = P.upload(ctx)
    jobs = [engine.UnionJob(d.ptr + s["a"], d.ptr + s["b"], None if s["m"] is None else d.ptr + s["m"],
                            d.ptr + s["uas"], d.ptr + s["ubs"], d.ptr + s["uae"], d.ptr + s["ube"], d.ptr + s["os"],
                            d.ptr + s["oe"], d.ptr + s["oa"], d.ptr + s["ob"], s["len"], s["na"], s["nb"], s["cap"])
            for s in spec]
    n = engine.union_merge(ctx, cm, jobs)
    blob = d.download(np.uint8, max(P.size, 256))
    d.free()
    outs = []
    for s, k in zip(spec, n):
        cap = s["cap"]
        if k < 0:
            outs.append(None)
            continue
        lo = cap - int(k)
        seq = blob[s["os"]:s["os"] + cap][lo:]
        arrs = [blob[s[x]:s[x] + 4 * cap].view(np.int32)[lo:] for x in ("oe", "oa", "ob")]
        outs.append((seq, *arrs))
    return n, outs


@pytest.mark.parametrize("affine,give_median", [(False, False), (False, True), (True, True)])
def test_batched_unions_equal_oracle(ctx, affine, give_median):
    from poyd_amd import engine
    print(f"[case] union batch affine={affine} median given={give_median}", flush=True)
    cm = engine.CostMatrix.of_sub_indel(1, 1, 3) if affine else engine.CostMatrix.of_sub_indel(1, 2)
    rng = np.random.default_rng(71 + 2 * affine + give_median)
    merges = []
    for t in range(24):
        merges += random_tree_merges(rng, affine, nleaves=6 + t % 5, length=int(rng.integers(20, 400)))
    n, outs = run_batch(ctx, cm, merges, give_median)
    assert (n >= 0).all()
    for (a, b, med, ua, ub, want), got in zip(merges, outs):
        for x, y in zip(got, want):
            assert np.array_equal(x, y)


def test_mismatched_children_reported(ctx):
    from oracle import pyoracle as po
    from poyd_amd import engine
    print("[case] union batch with a mismatched merge", flush=True)
    cm = engine.CostMatrix.of_sub_indel(1, 2)
    good = random_tree_merges(np.random.default_rng(5), False, nleaves=3, length=50)
    a = np.array([16, 1, 2, 4], np.uint8)
    b = np.array([16, 1, 16, 4], np.uint8)
    bad = (a, b, a | b, po.union_leaf([16, 1]), po.union_leaf([16, 1, 4]), None)
    n, outs = run_batch(ctx, cm, good + [bad], True)
    assert (n[:-1] >= 0).all() and n[-1] < 0
    for (_a, _b, _m, _ua, _ub, want), got in zip(good, outs):
        for x, y in zip(got, want):
            assert np.array_equal(x, y)


def random_rooted_tree(rng, n_leaves):
    """children dict of a random rooted binary tree over leaves 1..n."""
    nodes = list(range(1, n_leaves + 1))
    children, nxt = {}, n_leaves + 1
    while len(nodes) > 1:
        i, j = sorted(rng.choice(len(nodes), 2, replace=False))
        x, y = nodes[i], nodes[j]
        children[nxt] = (x, y)
        nodes = [v for k, v in enumerate(nodes) if k not in (i, j)] + [nxt]
        nxt += 1
    return children, nodes[0]


def oracle_unions(po, pcm, affine, children, leaves, mcc, root):
    med, uni = {}, {}

    def go(v):
        if v not in children:
            med[v], uni[v] = leaves[v], po.union_leaf(leaves[v])
            return
        x, y = children[v]
        go(x)
        go(y)
        if mcc[y] < mcc[x]:
            x, y = y, x
        if affine:
            m, ra, rb, _c, mwg = po.align_affine_3(med[x], med[y], pcm)
        else:
            _c, ra, rb = po.align_2(med[x], med[y], pcm)
            mwg = po.median_2_with_gaps(ra, rb, pcm)
            m = po.ancestor_2(ra, rb, pcm)
        med[v], uni[v] = m, po.union_make(ra, rb, mwg, 16, uni[x], uni[y])

    go(root)
    return med, uni


@pytest.mark.parametrize("affine", [False, True])
def test_tree_unions_equal_oracle(ctx, affine):
    from oracle import pyoracle as po
    from poyd_amd import engine, seqstore, unions
    from test_union import mutate, rand_seq
    print(f"[case] tree unions affine={affine}", flush=True)
    rng = np.random.default_rng(91 + affine)
    cm = engine.CostMatrix.of_sub_indel(1, 1, 3) if affine else engine.CostMatrix.of_sub_indel(1, 2)
    pcm = po.CM2.sub_indel(1, 1, 3) if affine else po.CM2.sub_indel(1, 2)
    for trial in range(3):
        n = 10 + 7 * trial
        root_seq = rand_seq(rng, int(rng.integers(60, 300)))
        leaves = {v: mutate(rng, mutate(rng, root_seq)) for v in range(1, n + 1)}
        children, root = random_rooted_tree(rng, n)
        mcc = {}

        def low(v):
            if v not in mcc:
                mcc[v] = v if v not in children else min(low(c) for c in children[v])
            return mcc[v]

        low(root)
        store = seqstore.DeviceStore(ctx, cm, 1 << 22)
        us = unions.UnionSet(store, 1 << 22)
        ids = store.add([leaves[v] for v in range(1, n + 1)])
        med, uni, _cost = unions.rooted_unions(store, us, children, {v: int(ids[v - 1]) for v in leaves}, mcc, root)
        wmed, wuni = oracle_unions(po, pcm, affine, children, leaves, mcc, root)
        for v in wuni:
            assert np.array_equal(store.get(med[v]), wmed[v]), v
            for x, y in zip(us.get(uni[v]), wuni[v]):
                assert np.array_equal(x, y), v
        # SeqCS.distance_union of the root's union with every leaf's
        vs = sorted(leaves)
        d = unions.distance_union(us, [uni[root]] * len(vs), [uni[v] for v in vs])
        sub = 0.8 if affine else 1.0
        for k, v in enumerate(vs):
            a, b = wuni[root][0], wuni[v][0]
            if affine:
                want = po.cost_affine_3(a, b, pcm)
            else:
                want = po.cost_2(a, b, pcm, max(8, abs(len(a) - len(b))))
            assert d[k] == sub * want, (v, d[k], want)
