"""Sequence unions (union_merge, src/union.c:102-164; Sequence.Unions,
src/sequence.ml:1810-1895).

The oracle's restatement is pinned by
* two known-answer merges worked by hand from src/union.c (a child's run of
  non-homologous items, a column whose median is the gap, and both
  children's runs in one column);
* closed forms: the union of a sequence with itself over the identity
  alignment is the sequence (OR its column medians), with no
  non-homologous item and ca = cb = the positions;
* structural invariants that follow from the structure's definition, not
  from the walk: the output length (every child item but the leading one
  is used once, the two items of a column where both children have a
  character merge into one, column 0 adds the closing gap); one homologous
  item per column whose median is not the gap, plus the closing one; runs
  of non-homologous items numbered 1..k; and every homologous item equal
  to its column's median OR the child items that the column's characters
  stand for, with ca / cb pointing at those items (the k-th character of a
  child's aligned sequence is the child union's k-th homologous item).
The C-ABI host entry point (poyg_union_make) is then compared with the
oracle, item for item and offset for offset, on random trees with linear
and affine matrices."""
import numpy as np
import pytest

GAP = 16


def rand_seq(rng, n):
    return np.concatenate([[GAP], rng.choice([1, 2, 4, 8], n)]).astype(np.uint8)


def mutate(rng, s, p=0.12):
    out = [GAP]
    for x in s[1:]:
        r = rng.random()
        if r < p:
            continue
        if r < 2 * p:
            out.append(int(rng.choice([1, 2, 4, 8])))
        out.append(int(x) if rng.random() > p else int(rng.choice([1, 2, 4, 8])))
    return np.array(out, np.uint8)


def homologous_positions(u):
    return np.nonzero(u[1] == 0)[0]


def check_invariants(a, b, med, ua, ub, out):
    seq, end, ca, cb = out
    cols = np.arange(1, len(a))
    pairs = int(np.sum((a[1:] != GAP) & (b[1:] != GAP)))
    assert len(seq) == (len(ua[0]) - 1) + (len(ub[0]) - 1) - pairs + 1
    assert len(end) == len(ca) == len(cb) == len(seq)
    hom = np.nonzero(end == 0)[0]
    assert len(hom) == 1 + int(np.sum(med[1:] != GAP))
    # runs after each homologous item are numbered 1..k
    for k in range(1, len(end)):
        assert end[k] == 0 or end[k] == end[k - 1] + 1, (k, end[max(0, k - 3):k + 3])
    assert seq[0] == GAP and end[0] == 0 and ca[0] == 0 and cb[0] == 0
    # homologous items, column by column
    ha, hb = homologous_positions(ua), homologous_positions(ub)
    ka, kb = 0, 0   # characters of a / b met so far (left to right)
    h = 1           # next homologous output item (item 0 closes the union)
    for i in cols:
        ia = ib = None
        if a[i] != GAP:
            ka += 1
            ia = int(ha[ka])
        if b[i] != GAP:
            kb += 1
            ib = int(hb[kb])
        if med[i] == GAP:
            continue
        want = int(med[i]) | (int(ua[0][ia]) if ia is not None else 0) | (int(ub[0][ib]) if ib is not None else 0)
        k = int(hom[h])
        assert seq[k] == want, (i, k)
        assert ca[k] == (ia if ia is not None else 0) and cb[k] == (ib if ib is not None else 0), (i, k)
        h += 1
    assert h == len(hom)
    assert ka == len(ha) - 1 and kb == len(hb) - 1


def test_known_answer_run_of_first_child(oracle):
    # child a: items [gap, A, A|gap, C|gap] where the last two are a run of
    # non-homologous items (end 1, 2); child b a leaf [gap, G]; one column
    ua = (np.array([16, 1, 17, 18], np.uint8), np.array([0, 0, 1, 2], np.int32))
    ub = (np.array([16, 4], np.uint8), np.array([0, 0], np.int32))
    a = np.array([16, 1], np.uint8)
    b = np.array([16, 4], np.uint8)
    med = np.array([16, 5], np.uint8)
    seq, end, ca, cb = oracle.union_make(a, b, med, GAP, ua, ub)
    # column 1 first copies a's run (gap OR each item, order kept), then
    # prepends med | A | G; column 0 closes with the gap
    assert seq.tolist() == [16, 5, 17, 18]
    assert end.tolist() == [0, 0, 1, 2]
    assert ca.tolist() == [0, 1, 0, 0]
    assert cb.tolist() == [0, 1, 0, 0]


def test_known_answer_gap_median_and_both_runs(oracle):
    # child a: [gap, A, C, T|gap(end 1)]; child b: [gap, G, C|gap(end 1)]
    ua = (np.array([16, 1, 2, 24], np.uint8), np.array([0, 0, 0, 1], np.int32))
    ub = (np.array([16, 4, 18], np.uint8), np.array([0, 0, 1], np.int32))
    # columns: 0 (gap, gap), 1 (A, G) median A|G, 2 (C, gap) median gap
    a = np.array([16, 1, 2], np.uint8)
    b = np.array([16, 4, 16], np.uint8)
    med = np.array([16, 5, 16], np.uint8)
    seq, end, ca, cb = oracle.union_make(a, b, med, GAP, ua, ub)
    # i = 2: a's run [T|gap] and b's run [C|gap] are prepended (a's first,
    # so b's lands left of it); the column's median is the gap and b has
    # none: a's C becomes a non-homologous gap|C left of them (pending 3).
    # i = 1: no run is current; A|G|med -> homologous, fills 1, 2, 3.
    # i = 0: closing gap.
    assert seq.tolist() == [16, 5, 18, 18, 24]
    assert end.tolist() == [0, 0, 1, 2, 3]
    assert ca.tolist() == [0, 1, 0, 0, 0]
    assert cb.tolist() == [0, 1, 0, 0, 0]


def test_identity_alignment_closed_form(oracle):
    rng = np.random.default_rng(5)
    cm = oracle.CM2.sub_indel(1, 2)
    for n in (1, 7, 60):
        s = rand_seq(rng, n)
        med = oracle.median_2_with_gaps(s, s, cm)
        seq, end, ca, cb = oracle.union_make(s, s, med, GAP, oracle.union_leaf(s), oracle.union_leaf(s))
        assert np.array_equal(seq, s | med)
        assert np.array_equal(seq, s)   # median(x, x) = x for single bases
        assert not end.any()
        assert np.array_equal(ca, np.arange(len(s))) and np.array_equal(cb, np.arange(len(s)))


def random_tree_merges(rng, affine, nleaves=6, length=40):
    """Merges of a caterpillar tree: (a, b, med, ua, ub) per internal node,
    the running union on the first side, a new leaf on the second."""
    from oracle import pyoracle as po
    cm = po.CM2.sub_indel(1, 1, 3) if affine else po.CM2.sub_indel(1, 2)
    root = rand_seq(rng, length)
    leaves = [mutate(rng, mutate(rng, root)) for _ in range(nleaves)]
    seq, u = leaves[0], po.union_leaf(leaves[0])
    out = []
    for leaf in leaves[1:]:
        if affine:
            med_ng, ra, rb, _cost, mwg = po.align_affine_3(seq, leaf, cm)
        else:
            _c, ra, rb = po.align_2(seq, leaf, cm)
            mwg = po.median_2_with_gaps(ra, rb, cm)
            med_ng = po.ancestor_2(ra, rb, cm)
        ul = po.union_leaf(leaf)
        new = po.union_make(ra, rb, mwg, GAP, u, ul)
        out.append((ra, rb, mwg, u, ul, new))
        seq, u = med_ng, new
    return out


@pytest.mark.parametrize("affine", [False, True])
def test_random_tree_invariants(oracle, affine):
    rng = np.random.default_rng(21 + affine)
    for _ in range(8):
        for a, b, med, ua, ub, new in random_tree_merges(rng, affine):
            check_invariants(a, b, med, ua, ub, new)


@pytest.mark.parametrize("affine", [False, True])
def test_host_entry_equals_oracle(oracle, affine):
    from poyd_amd import engine
    cm = engine.CostMatrix.of_sub_indel(1, 1, 3) if affine else engine.CostMatrix.of_sub_indel(1, 2)
    rng = np.random.default_rng(31 + affine)
    checked = 0
    for _ in range(6):
        for a, b, med, ua, ub, want in random_tree_merges(rng, affine, nleaves=8, length=70):
            got = engine.union_make(cm, a, b, med, ua, ub)
            for x, y in zip(got, want):
                assert np.array_equal(x, y)
            if not affine:   # the matrix's column medians are median_2_with_gaps
                got = engine.union_make(cm, a, b, None, ua, ub)
                for x, y in zip(got, want):
                    assert np.array_equal(x, y)
            checked += 1
    assert checked >= 40


def test_host_entry_rejects_mismatched_children():
    from poyd_amd import engine
    cm = engine.CostMatrix.of_sub_indel(1, 2)
    a = np.array([16, 1, 2, 4], np.uint8)
    b = np.array([16, 1, 16, 4], np.uint8)
    # child a's union holds one character where the alignment shows three
    with pytest.raises(engine.PoygError):
        engine.union_make(cm, a, b, None, engine.union_leaf([16, 1]), engine.union_leaf([16, 1, 4]))


def test_union_pruned_walk_with_reference_distance_visits_every_edge():
    """union_dfs_wagner (src/tabus.ml:530-551) with the reference's
    Node.union_distance (constant 0., src/node.ml:2719) scores every join
    edge; positive union distances prune subtrees."""
    from poyd_amd import unions
    rng = np.random.default_rng(3)
    for trial in range(20):
        n = int(rng.integers(3, 40))
        # the handle edge (1, 0) with random binary subtrees at both ends;
        # the walk starts on 1's side and reaches 0's side through the
        # handle edge (initialize_children_edges b a of the handle edge)
        children, nxt, frontier = {}, 2, [0, 1]
        while nxt < 2 * n:
            v = frontier.pop(int(rng.integers(len(frontier))))
            children[v] = (nxt, nxt + 1)
            frontier += [nxt, nxt + 1]
            nxt += 2
        edges = [(1, 0)] + [(a, b) for a, cs in children.items() for b in cs]
        delta = {e: float(rng.integers(0, 50)) for e in edges}

        def below(e):
            return [(e[1], c) for c in children.get(e[1], ())]

        start = [(e, 1) for e in below((0, 1))] + [((1, 0), 0)]
        seen = unions.pruned_walk(below, start, delta.__getitem__, lambda e: 0.0)
        assert sorted(seen) == sorted(edges) and len(seen) == len(edges)
        # a depth bound keeps the edges within it
        first = [(1, 0)] + below((0, 1)) + below((1, 0))
        shallow = unions.pruned_walk(below, start, delta.__getitem__, lambda e: 0.0, max_depth=1)
        assert sorted(shallow) == sorted(first)
        # union distances above every delta stop the walk after the handle
        # edge, which is expanded before any delta is known (current_delta
        # starts at max_float)
        none = unions.pruned_walk(below, start, delta.__getitem__, lambda e: 1e9)
        assert sorted(none) == sorted(first)
