"""Pins of the CPU oracle against what the mathematics fixes.

The reference's own tests hold no kernel-level vectors (its hot path is only
pinned end to end by tree-cost fixtures, see test_tree_fixtures.py), so the
oracle's kernel functions are checked here against independent formulations:
closed forms of the cost-matrix fills, textbook Needleman-Wunsch and Gotoh
dynamic programming where the reference's band covers the whole matrix,
recosting of the produced alignments, and brute-force parsimony for Fitch.
"""
import itertools

import numpy as np
import pytest

GAP = 16
A, Cb, G, T = 1, 2, 4, 8
INF = 1 << 40


def rand_dna(rng, n):
    return np.concatenate([[GAP], rng.choice([1, 2, 4, 8], size=n)]).astype(np.uint8)


def mutate(rng, s, ps=0.15, pi=0.05):
    out = [GAP]
    for x in s[1:]:
        r = rng.random()
        if r < pi / 2:
            continue
        if r < pi:
            out.append(int(rng.choice([1, 2, 4, 8])))
        out.append(int(rng.choice([1, 2, 4, 8])) if rng.random() < ps else int(x))
    return np.array(out, np.uint8)


def strip(s):
    return np.array([x for x in s if x != GAP], np.uint8)


# ---------------------------------------------------------------- R1 -------
def members(m):
    return [1 << b for b in range(5) if m & (1 << b)]


def test_tcm_linear_closed_form(oracle):
    """Metric fill: for a metric base matrix the cost of two masks is the
    minimum base-to-base cost (triangle inequality), singleton medians are
    the union of the pair, and prepend/tail are the gap rows."""
    cm = oracle.CM2.sub_indel(1, 2)
    d = {(a, b): (0 if a == b else (2 if GAP in (a, b) else 1)) for a in members(31) for b in members(31)}
    for i in range(1, 32):
        for j in range(1, 32):
            assert cm.cost(i, j) == min(d[a, b] for a in members(i) for b in members(j)), (i, j)
            if len(members(i)) == 1 and len(members(j)) == 1:
                assert cm.median(i, j) == i | j
    # hand-derived medians: argmin state sets of cost(a,v)+cost(v,b)
    assert cm.median(A, Cb | G) == A | Cb | G
    assert cm.median(A | Cb, A) == A
    assert cm.median(A | GAP, Cb) == A | Cb
    assert cm.median(A | Cb, G | T) == A | Cb | G | T
    assert cm.median(GAP | A, GAP | Cb) == GAP


def test_tcm_affine_bitwise_closed_form(oracle):
    """Bitwise fill after set_affine: shared bits cost 0 with the
    intersection as median, else the minimum member cost with the union of
    the optimal member pairs as median."""
    cm = oracle.CM2.sub_indel(1, 1, gap_open=3)
    d = {(a, b): (0 if a == b else 1) for a in members(31) for b in members(31)}
    for i in range(1, 32):
        for j in range(1, 32):
            if len(members(i)) == 1 and len(members(j)) == 1:
                assert cm.cost(i, j) == d[i, j]
                assert cm.median(i, j) == i | j
            elif i & j:
                assert cm.cost(i, j) == 0 and cm.median(i, j) == i & j
            else:
                best = min(d[a, b] for a in members(i) for b in members(j))
                med = 0
                for a in members(i):
                    for b in members(j):
                        if d[a, b] == best:
                            med |= a | b
                assert cm.cost(i, j) == best and cm.median(i, j) == med, (i, j)


def test_tcm_three_d_closed_form(oracle):
    """Three_D.of_two_dim over a metric: cost(i,j,k) = min over single states
    l of cost(l,i)+cost(l,j)+cost(l,k); the median is the lowest optimal l."""
    cm = oracle.CM2.sub_indel(1, 2)
    cm3 = oracle.CM3(cm)
    rng = np.random.default_rng(1)
    for _ in range(400):
        i, j, k = (int(x) for x in rng.integers(1, 32, size=3))
        vals = [(cm.cost(1 << l, i) + cm.cost(1 << l, j) + cm.cost(1 << l, k), 1 << l) for l in range(5)]
        best = min(v for v, _ in vals)
        assert cm3.cost(i, j, k) == best
        assert cm3.median(i, j, k) == min(s for v, s in vals if v == best)


# ---------------------------------------------------------------- R4-R7 ----
def textbook_nw(s1, s2, cm):
    m, n = len(s1), len(s2)
    D = np.zeros((m, n), np.int64)
    for j in range(1, n):
        D[0, j] = D[0, j - 1] + cm.cost(GAP, int(s2[j]))
    for i in range(1, m):
        D[i, 0] = D[i - 1, 0] + cm.cost(int(s1[i]), GAP)
        for j in range(1, n):
            D[i, j] = min(D[i - 1, j - 1] + cm.cost(int(s1[i]), int(s2[j])),
                          D[i - 1, j] + cm.cost(int(s1[i]), GAP),
                          D[i, j - 1] + cm.cost(GAP, int(s2[j])))
    return int(D[m - 1, n - 1])


@pytest.mark.parametrize("sub,indel", [(1, 2), (1, 1), (2, 1), (3, 1)])
def test_linear_full_plane_is_textbook(oracle, sub, indel):
    """When algn_fill_plane_2 fills the whole matrix the cost is textbook NW."""
    cm = oracle.CM2.sub_indel(sub, indel)
    rng = np.random.default_rng(7 + sub * 10 + indel)
    for t in range(60):
        n = int(rng.integers(1, 30))
        s2 = rand_dna(rng, n)
        s1 = rand_dna(rng, int(rng.integers(n, 60)))
        dw = oracle.deltaw(s1, s2, cm)
        assert oracle.algn_nw_cells(len(s1), len(s2), dw) == len(s1) * len(s2)
        assert oracle.cost_2(s1, s2, cm) == textbook_nw(s1, s2, cm)


def test_linear_band_invariants(oracle):
    """Banded runs: cost >= textbook optimum, the alignment recosts to the
    cost, stripping gaps gives the inputs back, medians have the lengths the
    median rules fix."""
    cm = oracle.CM2.sub_indel(1, 2)
    rng = np.random.default_rng(11)
    banded = 0
    for t in range(40):
        s1 = rand_dna(rng, int(rng.integers(150, 320)))
        s2 = mutate(rng, s1, 0.2, 0.1 if t % 2 else 0.02)
        cost, a1, a2 = oracle.align_2(s1, s2, cm)
        l1, l2 = max(len(s1), len(s2)), min(len(s1), len(s2))
        dw = oracle.deltaw(s1, s2, cm)
        if oracle.algn_nw_cells(l1, l2, dw) < l1 * l2:
            banded += 1
        assert cost == oracle.cost_2(s1, s2, cm)
        assert cost >= textbook_nw(s1, s2, cm) if t < 6 else True
        assert len(a1) == len(a2)
        assert oracle.verify_2(a1, a2, cm) == cost
        assert np.array_equal(strip(a1), strip(s1)) and np.array_equal(strip(a2), strip(s2))
        assert a1[0] == GAP and a2[0] == GAP
        mwg = oracle.median_2_with_gaps(a1, a2, cm)
        assert len(mwg) == len(a1)
        med = oracle.median_2(a1, a2, cm)
        assert len(med) == 1 + sum(1 for x in mwg if x != GAP)
        anc = oracle.ancestor_2(a1, a2, cm)
        assert np.array_equal(anc, med)
    assert banded >= 20


def test_linear_backtrace_priorities(oracle):
    """A single substitution-or-indel tie: with s1 the longer sequence the
    backtrace prefers INSERT over DELETE (swaped), else DELETE first."""
    cm = oracle.CM2.sub_indel(1, 1)
    s1 = np.array([GAP, A, Cb], np.uint8)
    s2 = np.array([GAP, G], np.uint8)
    c, dm, lo, hi = oracle.algn_nw(s1, s2, cm, 10, with_dm=True)
    assert c == 2
    r1, r2 = oracle.backtrack_2d(dm, lo, hi, s1, s2, cm, True)
    r1b, r2b = oracle.backtrack_2d(dm, lo, hi, s1, s2, cm, False)
    for x, y in ((r1, r2), (r1b, r2b)):
        assert oracle.verify_2(x, y, cm) == 2


# ---------------------------------------------------------------- R8/R9 ----
def textbook_gotoh(a, b, cm, go):
    """Three-state Gotoh with all state transitions; a gap of length L in a
    block costs go + sum of per-base gap costs."""
    m, n = len(a), len(b)
    M = np.full((m, n), INF, np.int64)
    X = np.full((m, n), INF, np.int64)   # consumes a (gap in b)
    Y = np.full((m, n), INF, np.int64)   # consumes b (gap in a)
    M[0, 0] = 0
    for i in range(1, m):
        X[i, 0] = (go if i == 1 else X[i - 1, 0]) + cm.cost(int(a[i]), GAP)
    for j in range(1, n):
        Y[0, j] = (go if j == 1 else Y[0, j - 1]) + cm.cost(GAP, int(b[j]))
    for i in range(1, m):
        for j in range(1, n):
            best = min(M[i - 1, j - 1], X[i - 1, j - 1], Y[i - 1, j - 1])
            M[i, j] = best + cm.cost(int(a[i]), int(b[j]))
            e = cm.cost(int(a[i]), GAP)
            X[i, j] = min(M[i - 1, j] + go, X[i - 1, j], Y[i - 1, j] + go) + e
            e = cm.cost(GAP, int(b[j]))
            Y[i, j] = min(M[i, j - 1] + go, Y[i, j - 1], X[i, j - 1] + go) + e
    return int(min(M[m - 1, n - 1], X[m - 1, n - 1], Y[m - 1, n - 1]))


@pytest.mark.parametrize("sub,indel,go", [(1, 1, 3), (2, 1, 1), (3, 1, 5)])
def test_affine_full_band_is_gotoh(oracle, sub, indel, go):
    """With the longer sequence at most 40 long the affine band covers the
    whole matrix; for plain DNA the cost is then textbook Gotoh."""
    cm = oracle.CM2.sub_indel(sub, indel, gap_open=go)
    rng = np.random.default_rng(3 + go)
    for t in range(50):
        si = rand_dna(rng, int(rng.integers(0, 25)))
        sj = mutate(rng, si, 0.3, 0.2) if t % 3 else rand_dna(rng, int(rng.integers(0, 38)))
        if len(sj) > 41:
            sj = sj[:41]
        ref = textbook_gotoh(si, sj, cm, go)
        med, ri, rj, cost, mwg = oracle.align_affine_3(si, sj, cm)
        assert cost == ref, (t, si, sj)
        assert oracle.cost_affine_3(si, sj, cm) == ref
        assert oracle.verify_2(ri, rj, cm) == cost


def test_affine_band_invariants(oracle):
    cm = oracle.CM2.sub_indel(1, 1, gap_open=3)
    rng = np.random.default_rng(5)
    for t in range(30):
        si = rand_dna(rng, int(rng.integers(60, 400)))
        sj = mutate(rng, si, 0.2, 0.05)
        med, ri, rj, cost, mwg = oracle.align_affine_3(si, sj, cm)
        assert cost == oracle.cost_affine_3(si, sj, cm)   # plain DNA: both fills agree
        assert cost >= (textbook_gotoh(si, sj, cm, 3) if t < 4 else 0)
        assert oracle.verify_2(ri, rj, cm) == cost
        assert np.array_equal(strip(ri), strip(si)) and np.array_equal(strip(rj), strip(sj))
        assert len(mwg) == len(ri) == len(rj)
        assert med[0] == GAP and mwg[0] == GAP
        # aligned columns give the median symbol; gap columns give x|gap
        assert len(med) == 1 + sum(1 for x in mwg[1:] if x != GAP)


# ---------------------------------------------------------------- R10/R11 --
def test_cube_single_plane_is_textbook(oracle):
    """With s1 = [gap] the cube is one plane: a textbook 2-D DP over s2, s3
    with 3-D costs of (gap, x, y)."""
    cm3 = oracle.CM3(oracle.CM2.sub_indel(1, 2))
    rng = np.random.default_rng(9)
    for _ in range(30):
        s2 = rand_dna(rng, int(rng.integers(0, 15)))
        s3 = rand_dna(rng, int(rng.integers(0, 15)))
        m, n = len(s2), len(s3)
        D = np.zeros((m, n), np.int64)
        c = lambda x, y: cm3.cost(GAP, x, y)
        for k in range(1, n):
            D[0, k] = D[0, k - 1] + c(GAP, int(s3[k]))
        for r in range(1, m):
            D[r, 0] = D[r - 1, 0] + c(int(s2[r]), GAP)
            for k in range(1, n):
                D[r, k] = min(D[r - 1, k] + c(int(s2[r]), GAP), D[r - 1, k - 1] + c(int(s2[r]), int(s3[k])),
                              D[r, k - 1] + c(GAP, int(s3[k])))
        assert oracle.cost_3(np.array([GAP], np.uint8), s2, s3, cm3) == D[m - 1, n - 1]


def test_cube_alignment_is_valid(oracle):
    cm3 = oracle.CM3(oracle.CM2.sub_indel(1, 2))
    rng = np.random.default_rng(13)
    for _ in range(20):
        anc = rand_dna(rng, int(rng.integers(5, 30)))
        s = [mutate(rng, anc) for _ in range(3)]
        cost, a1, a2, a3 = oracle.align_3(*s, cm3)
        assert len(a1) == len(a2) == len(a3)
        for x, y in zip((a1, a2, a3), s):
            assert np.array_equal(strip(x), strip(y))
        med = oracle.median_3(a1, a2, a3, cm3)
        assert len(med) == len(a1) and len(set(med.tolist())) <= 1


def _plane0(s2, s3, c):
    """Plane 0 of algn_fill_cube (src/algn.c:2940-2985): the textbook 2-D DP
    over s2, s3 with the costs of aligning against the gap of s1."""
    m, n = len(s2), len(s3)
    P = np.zeros((m, n), np.int64)
    for k in range(1, n):
        P[0, k] = P[0, k - 1] + c(GAP, GAP, int(s3[k]))
    for r in range(1, m):
        b = int(s2[r])
        P[r, 0] = P[r - 1, 0] + c(GAP, b, GAP)
        for k in range(1, n):
            P[r, k] = min(P[r - 1, k] + c(GAP, b, GAP), P[r - 1, k - 1] + c(GAP, b, int(s3[k])),
                          P[r, k - 1] + c(GAP, GAP, int(s3[k])))
    return P


def _line0(up, a, s3, c):
    """Line 0 of a plane p >= 1 (src/algn.c:2997-3016): moves from the row
    the "diagonal" pointer names (`up`) and along the line itself."""
    n = len(s3)
    X = np.zeros(n, np.int64)
    X[0] = up[0] + c(a, GAP, GAP)
    for k in range(1, n):
        x3 = int(s3[k])
        X[k] = min(up[k] + c(a, GAP, GAP), up[k - 1] + c(a, GAP, x3), X[k - 1] + c(GAP, GAP, x3))
    return X


def test_cube_one_line_planes_read_plane_zero(oracle):
    """With s2 = [gap] the line loop of algn_fill_cube (src/algn.c:3023-3053)
    never runs, so the pointer that line 0 of every plane reads (diag_m,
    src/algn.c:2991, 3002) never leaves plane 0: the cost depends on s1's
    last symbol only.  A textbook three-way recurrence fails both checks."""
    cm3 = oracle.CM3(oracle.CM2.sub_indel(1, 2))
    c = cm3.cost
    rng = np.random.default_rng(41)
    s2 = np.array([GAP], np.uint8)
    differs = 0
    for t in range(40):
        s1 = rand_dna(rng, int(rng.integers(1, 12)))
        s3 = rand_dna(rng, int(rng.integers(0, 12)))
        got = oracle.cost_3(s1, s2, s3, cm3)
        want = _line0(_plane0(s2, s3, c)[0], int(s1[-1]), s3, c)[-1]
        assert got == want, (t, s1, s3)
        assert got == oracle.cost_3(np.array([GAP, s1[-1]], np.uint8), s2, s3, cm3)
        # the textbook cost (every plane reads its predecessor) differs once
        # s1 has two or more symbols and s3 is long enough
        X = _plane0(s2, s3, c)[0]
        for a in s1[1:]:
            X = _line0(X, int(a), s3, c)
        differs += int(X[-1] != got)
    assert differs > 0


def test_cube_two_planes_lagged_previous_line(oracle):
    """Two planes (s1 = [gap, a]): in plane 1, lines j >= 1 read plane 0's
    lines j - 1 ("diagonal") and j ("upper") and, as their "previous" row,
    the flat row j - 2 rows after plane 0's end: plane 0's last line for
    j = 1, plane 1's line j - 2 after that (pointer walk of
    src/algn.c:2987-3053: diag_m, upper_m and prev_m advance n2 - 1 rows per
    plane while mm advances n2).  Moves in the reference's order: P3, P1,
    P2, S3, S1, S2, then the self pass."""
    cm3 = oracle.CM3(oracle.CM2.sub_indel(1, 2))
    c = cm3.cost
    rng = np.random.default_rng(43)
    for t in range(40):
        a = int(rng.choice([1, 2, 4, 8]))
        s1 = np.array([GAP, a], np.uint8)
        s2 = rand_dna(rng, int(rng.integers(1, 10)))
        s3 = rand_dna(rng, int(rng.integers(0, 10)))
        P0 = _plane0(s2, s3, c)
        n2, n3 = len(s2), len(s3)
        X = [_line0(P0[0], a, s3, c)]
        for j in range(1, n2):
            b = int(s2[j])
            prev = P0[n2 - 1] if j == 1 else X[j - 2]
            up, dg = P0[j], P0[j - 1]
            row = np.zeros(n3, np.int64)
            for k in range(n3):
                v = min(up[k] + c(a, GAP, GAP), prev[k] + c(GAP, b, GAP), dg[k] + c(a, b, GAP))
                if k:
                    x3 = int(s3[k])
                    v = min(v, up[k - 1] + c(a, GAP, x3), prev[k - 1] + c(GAP, b, x3), dg[k - 1] + c(a, b, x3),
                            row[k - 1] + c(GAP, GAP, x3))
                row[k] = v
            X.append(row)
        assert oracle.cost_3(s1, s2, s3, cm3) == X[-1][-1], (t, s1, s2, s3)


def test_median_3_is_the_last_columns_median(oracle):
    """algn_get_median_3d (src/algn.c:4160-4173) reads *end1, *end2, *end3
    without moving them: the median has one symbol per aligned column (the
    loop runs len(s1) times), each the three-way median of the last
    column."""
    cm3 = oracle.CM3(oracle.CM2.sub_indel(1, 2))
    rng = np.random.default_rng(47)
    for _ in range(30):
        anc = rand_dna(rng, int(rng.integers(3, 25)))
        s = [mutate(rng, anc) for _ in range(3)]
        cost, a1, a2, a3 = oracle.align_3(*s, cm3)
        med = oracle.median_3(a1, a2, a3, cm3)
        want = cm3.median(int(a1[-1]), int(a2[-1]), int(a3[-1]))
        assert len(med) == len(a1) and all(int(x) == want for x in med)
    # columns other than the last do not matter
    x = np.array([GAP, A, Cb, G], np.uint8)
    y = np.array([GAP, T, T, G], np.uint8)
    z = np.array([GAP, GAP, A, G], np.uint8)
    assert oracle.median_3(x, y, z, cm3).tolist() == [cm3.median(G, G, G)] * 4


# ---------------------------------------------------------------- R12/R13 --
def random_tree(rng, nleaves):
    """Random rooted binary tree: list of (node, left, right), root last."""
    nodes = list(range(nleaves))
    internal = []
    nxt = nleaves
    while len(nodes) > 1:
        i, j = rng.choice(len(nodes), size=2, replace=False)
        a, b = nodes[i], nodes[j]
        nodes = [x for k, x in enumerate(nodes) if k not in (i, j)] + [nxt]
        internal.append((nxt, a, b))
        nxt += 1
    return internal


@pytest.mark.parametrize("width,shift", [(1, 0), (1, 5), (2, 11), (4, 27)])
def test_fitch_matches_brute_force_parsimony(oracle, width, shift):
    """Down pass: the summed added costs equal the minimum number of changes
    over all internal state assignments.  Up pass: the final sets equal the
    states a node takes in at least one most parsimonious reconstruction.
    Characters of 1, 2 and 4 bytes (NonaddCS8/16/32), the states on bits
    shift .. shift + k - 1 so that the wider characters use their high bytes."""
    rng = np.random.default_rng(17 + shift)
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[width]
    down = (lambda x, y: oracle.nonadd_median(x, y)) if width == 1 and shift == 0 else \
        (lambda x, y: oracle.nonadd_median_w(x, y, width))
    up = (lambda *v: oracle.nonadd_median_3(*v)) if width == 1 and shift == 0 else \
        (lambda *v: oracle.nonadd_median_3_w(*v, width))
    for t in range(12):
        nleaves = int(rng.integers(3, 7))
        k = int(rng.integers(2, 4))
        nchar = 40
        tree = random_tree(rng, nleaves)
        leaf = {v: (rng.integers(1, 1 << k, size=nchar).astype(np.uint64) << shift).astype(dt)
                for v in range(nleaves)}
        prelim = dict(leaf)
        total = 0
        for v, a, b in tree:
            prelim[v], added = down(prelim[a], prelim[b])
            total += added
        root = tree[-1][0]
        final = {root: prelim[root]}
        parent = {}
        for v, a, b in tree:
            parent[a] = v
            parent[b] = v
        for v, a, b in reversed(tree[:-1]):
            p = parent[v]
            final[v] = up(final[p], prelim[v], prelim[a], prelim[b])
        internal = [v for v, _, _ in tree]
        bf_total = 0
        for ch in range(nchar):
            best = None
            mpr_states = {v: set() for v in internal}
            for assign in itertools.product(range(k), repeat=len(internal)):
                st = dict(zip(internal, assign))
                cost = 0
                for v, a, b in tree:
                    for c in (a, b):
                        if c in st:
                            cost += st[c] != st[v]
                        else:
                            cost += 0 if (int(leaf[c][ch]) >> (shift + st[v])) & 1 else 1
                if best is None or cost < best:
                    best = cost
                    mpr_states = {v: {st[v]} for v in internal}
                elif cost == best:
                    for v in internal:
                        mpr_states[v].add(st[v])
            bf_total += best
            for v in internal:
                got = {s for s in range(k) if (int(final[v][ch]) >> (shift + s)) & 1}
                assert got == mpr_states[v], (t, ch, v)
                assert int(final[v][ch]) >> (shift + k) == 0 and int(final[v][ch]) & ((1 << shift) - 1) == 0
        assert total == bf_total


def test_fitch_distance_helpers(oracle):
    rng = np.random.default_rng(19)
    a = rng.integers(1, 256, size=1000).astype(np.uint8)
    b = rng.integers(1, 256, size=1000).astype(np.uint8)
    c = rng.integers(1, 256, size=1000).astype(np.uint8)
    assert oracle.nonadd_distance(a, b) == int(((a & b) == 0).sum())
    assert oracle.nonadd_dist_2(a, b, c) == int((((a & b) == 0) & ((a & c) == 0)).sum())


# ---- R15: Powell's three-way affine model -------------------------------
def powell_states():
    """The model's states in the reference's enumeration order
    (src/ukkCommon.c:252-330): per sequence match 0 / del 1 / ins 2."""
    out = []
    for s in range(27):
        st = (s % 3, (s // 3) % 3, (s // 9) % 3)
        if st.count(0) == 0 or st.count(2) > 1:
            continue
        adv = tuple(int(t == 0) for t in st) if st.count(2) == 0 else tuple(int(t == 2) for t in st)
        out.append((st, adv))
    return out


def powell_column_cost(prev, st, chars, mm, go, ge):
    """alignmentCost's cost of one column (src/ukkCommon.c:380-445)."""
    cost = sum(go for q in range(3) if st[q] != 0 and st[q] != prev[q])
    if st.count(2):
        return cost + ge
    cost += ge * st.count(1)
    ch = [chars[q] for q in range(3) if st[q] == 0]
    for a in range(len(ch) - 1, 0, -1):
        if ch[a - 1] != ch[a]:
            cost += mm
    if len(ch) == 3 and ch[0] == ch[2] and ch[0] != ch[1]:
        cost -= mm
    return cost


def powell_brute_force(x, mm, go, ge):
    """Least cost over every sequence of states that consumes the three
    strings (branch and bound over all paths, no shared subproblems)."""
    states = powell_states()
    n = [len(s) for s in x]
    best = [None]

    def walk(pos, prev, cost):
        if best[0] is not None and cost >= best[0]:
            return
        if pos == n:
            best[0] = cost
            return
        for st, adv in states:
            nxt = [pos[q] + adv[q] for q in range(3)]
            if any(nxt[q] > n[q] for q in range(3)):
                continue
            chars = [x[q][nxt[q] - 1] if adv[q] else None for q in range(3)]
            walk(nxt, st, cost + powell_column_cost(prev, st, chars, mm, go, ge))

    walk([0, 0, 0], (0, 0, 0), 0)
    return best[0]


def test_powell_model_has_sixteen_states():
    # 7 without an insert (M/D with a match) and 3 x 3 with one
    assert len(powell_states()) == 16


@pytest.mark.parametrize("mm,go,ge", [(1, 2, 1), (2, 0, 1), (1, 3, 2), (3, 1, 1)])
def test_powell_cost_equals_brute_force(oracle, mm, go, ge):
    rng = np.random.default_rng(40 + mm + 10 * go + 100 * ge)
    bases = "ACGT"
    code = {"A": 1, "C": 2, "G": 4, "T": 8}
    for _ in range(12):
        lens = [int(v) for v in rng.integers(0, 4, size=3)]
        if sum(lens) > 6:
            lens = [min(v, 2) for v in lens]
        x = ["".join(bases[int(v)] for v in rng.integers(0, 3, size=n)) for n in lens]
        enc = [np.array([16] + [code[ch] for ch in s], np.uint8) for s in x]
        got = oracle.powell_3d(*enc, mm, go, ge)
        assert got is not None
        assert got[0] == powell_brute_force(x, mm, go, ge), (x, mm, go, ge)


def test_powell_alignment_recosts_to_its_cost(oracle):
    # the returned alignment strips back to the inputs and its columns can be
    # labelled with states whose alignmentCost total is the returned cost
    rng = np.random.default_rng(44)
    for _ in range(20):
        anc = rng.choice([1, 2, 4, 8], size=int(rng.integers(3, 12)))
        x = [np.concatenate([[16], anc[rng.random(len(anc)) > 0.2]]).astype(np.uint8) for _ in range(3)]
        cost, a, b, c = oracle.powell_3d(*x, 1, 2, 1)
        for out, src in zip((a, b, c), x):
            assert out[0] == 16 and np.array_equal(out[1:][out[1:] != 16], src[1:])
        states = powell_states()
        inv = {1: "A", 2: "C", 4: "G", 8: "T"}
        # least labelling of the fixed columns (insert columns may be I or DD)
        best = {(0, 0, 0): 0}
        for col in range(1, len(a)):
            chars = [a[col], b[col], c[col]]
            has = tuple(int(v != 16) for v in chars)
            nxt = {}
            for prev, cst in best.items():
                for st, adv in states:
                    if adv != has:
                        continue
                    v = cst + powell_column_cost(prev, st, [inv.get(int(ch)) for ch in chars], 1, 2, 1)
                    nxt[st] = min(nxt.get(st, v), v)
            best = nxt
        assert min(best.values()) == cost


def test_powell_search_cost_is_the_least_cost(oracle):
    # the restated check-point search (po_powell_ukk) ends at the model's
    # least cost (brute force on tiny inputs, the exhaustive DP beyond), but
    # at cost 1 for three equal strings: its loop starts at cost 1
    rng = np.random.default_rng(45)
    bases = "ACGT"
    code = {"A": 1, "C": 2, "G": 4, "T": 8}
    for mm, go, ge in [(1, 2, 1), (2, 0, 1), (1, 3, 2), (3, 1, 1)]:
        for _ in range(8):
            lens = [int(v) for v in rng.integers(0, 3, size=3)]
            x = ["".join(bases[int(v)] for v in rng.integers(0, 3, size=n)) for n in lens]
            enc = [np.array([16] + [code[ch] for ch in s], np.uint8) for s in x]
            got = oracle.powell_ukk(*enc, mm, go, ge)
            want = 1 if x[0] == x[1] == x[2] else powell_brute_force(x, mm, go, ge)
            assert got[0] == want, (x, mm, go, ge)
        for _ in range(6):
            anc = rng.choice([1, 2, 4, 8], size=int(rng.integers(5, 40)))
            enc = [np.concatenate([[16], anc[rng.random(len(anc)) > 0.25]]).astype(np.uint8) for _ in range(3)]
            if any(len(e) == 1 for e in enc) or all(np.array_equal(e, enc[0]) for e in enc):
                continue
            assert oracle.powell_ukk(*enc, mm, go, ge)[0] == oracle.powell_3d(*enc, mm, go, ge)[0]


def test_powell_search_alignment_recosts_and_strips(oracle):
    # the search's alignment strips back to its inputs and recosts
    # (alignmentCost over the least labelling of its columns) to its cost
    rng = np.random.default_rng(46)
    states = powell_states()
    inv = {1: "A", 2: "C", 4: "G", 8: "T"}
    for k in range(16):
        mm, go, ge = [(1, 2, 1), (2, 0, 1), (1, 3, 2), (3, 1, 1)][k % 4]
        anc = rng.choice([1, 2, 4, 8], size=int(rng.integers(4, 30)))
        x = []
        for _ in range(3):
            body = anc[rng.random(len(anc)) > 0.2].copy()
            flip = rng.random(len(body)) < 0.2
            body[flip] = rng.choice([1, 2, 4, 8], size=int(flip.sum()))
            x.append(np.concatenate([[16], body, rng.choice([1, 2, 4, 8], size=int(rng.integers(0, 3)))]).astype(np.uint8))
        cost, a, b, c = oracle.powell_ukk(*x, mm, go, ge)
        for out, src in zip((a, b, c), x):
            assert out[0] == 16 and np.array_equal(out[1:][out[1:] != 16], src[1:])
        best = {(0, 0, 0): 0}
        for col in range(1, len(a)):
            chars = [a[col], b[col], c[col]]
            has = tuple(int(v != 16) for v in chars)
            nxt = {}
            for prev, cst in best.items():
                for st, adv in states:
                    if adv != has:
                        continue
                    v = cst + powell_column_cost(prev, st, [inv.get(int(ch)) for ch in chars], mm, go, ge)
                    nxt[st] = min(nxt.get(st, v), v)
            best = nxt
        same = all(len(v) == len(x[0]) and np.array_equal(v, x[0]) for v in x)
        assert min(best.values()) == (0 if same else cost), k


def test_powell_search_plane_reuse_changes_nothing(oracle):
    # the reference keeps 2 maxSingleStep cost planes and reuses plane
    # d mod that count; keeping every plane gives the same costs and the same
    # alignments (no reused plane is read stale on these inputs)
    rng = np.random.default_rng(47)
    for k in range(24):
        mm, go, ge = [(1, 2, 1), (2, 0, 1), (1, 3, 2), (3, 1, 1)][k % 4]
        anc = rng.choice([1, 2, 4, 8], size=int(rng.integers(10, 60)))
        x = []
        for _ in range(3):
            body = anc[rng.random(len(anc)) > 0.2].copy()
            flip = rng.random(len(body)) < 0.25
            body[flip] = rng.choice([1, 2, 4, 8], size=int(flip.sum()))
            x.append(np.concatenate([[16], body]).astype(np.uint8))
        reuse = oracle.powell_ukk(*x, mm, go, ge, 0)
        keep = oracle.powell_ukk(*x, mm, go, ge, 1)
        assert reuse[0] == keep[0]
        for p, q in zip(reuse[1:], keep[1:]):
            assert np.array_equal(p, q), k


def test_powell_rejects_gap_only_characters(oracle):
    x = np.array([16, 1, 16, 2], np.uint8)
    assert oracle.powell_3d(x, x, x, 1, 2, 1) is None
