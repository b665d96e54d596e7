"""One TBR neighbourhood round over dynamic characters, scored exactly on
the GPU, and the surrogate-guided variant that scores exactly only the
candidates a learned model ranks best.

Reference behaviour restated (lealanko/poyd):

* Breaks (Ptree.do_search, src/ptree.ml:1232-1256): every edge (u, v) of
  the tree, oriented away from the handle; u's side is the tree side, v's
  the clade.  AllDirChar.break_fn (src/allDirChar.ml:1008-1120) removes the
  edge, dissolves a removed interior vertex into an edge between its two
  other neighbours (Tree.break's `Edge (rem, l1, l2)` delta) and rebuilds
  the lazy nodes and edges of each component (clear_up_over_edge,
  refresh_all_edges): a component's directional medians are those of its
  own topology, so the ones whose subtree held the removed clade change and
  every other one keeps its value.
* Joins (single_spr_round / tbr_join, src/ptree.ml:1120-1229): the clade
  is rerooted on each of its edges (reroot_on_edge: the clade's node is then
  its root median on that edge) and, for each root, joined to every edge of
  the tree side; the candidate equal to the unbroken tree (the clade's
  original edge joined to the tree side's dissolved edge, `simplify` ->
  `Same`) is skipped.  A leaf clade is joined as itself.
* Join cost (AllDirChar.cost_fn, src/allDirChar.ml:1279-1332, `Normal`):
  Node.distance of the clade's root median and the join edge's lazy median
  (create_lazy_edge: the median of the edge's two directional medians),
  i.e. per character DOS.distance (src/seqCS.ml:818-864).
* A candidate's tree cost estimate is the broken tree's cost plus the join
  cost (search#process, src/queues.ml:367-398: real_cost =
  get_cost `Adjusted pt +. cc); a component's cost here is its downpass cost
  on its dissolved edge (the median of the removed vertex, already in the
  unbroken tree's directional medians).

The round's best candidate is the first least estimate in enumeration
order: breaks in pre-order from the handle, then clade roots (the dissolved
edge first, then the clade's edges outward from it), then join edges (the
same order on the tree side).  This order and the component root are this
engine's choice: no reference fixture pins them.  The chosen join is
applied and the new tree is diagnosed exactly (poyd_amd.treepass), so the
reported tree cost is always the exact Adjusted cost.

Everything runs on device-resident sequences (poyd_amd.seqstore): all
breaks are processed together, one batch of medians per distance from the
broken spot, then all edge medians, then the join distances.
"""
from __future__ import annotations

import copy
import time
from dataclasses import dataclass, field

import numpy as np

from .treedata import Tree


# ------------------------------------------------------------------ tables
@dataclass
class Table:
    """Medians of several characters: ids[k, c] (store ids), the subtree's
    downpass cost and its smallest taxon code (the median operand order of
    Node.cs_median, src/node.ml:341-347)."""

    ids: np.ndarray
    cost: np.ndarray
    mcc: np.ndarray

    def __len__(self):
        return len(self.cost)

    def take(self, idx):
        return Table(self.ids[idx], self.cost[idx], self.mcc[idx])


def concat(tables) -> Table:
    return Table(np.concatenate([t.ids for t in tables]), np.concatenate([t.cost for t in tables]),
                 np.concatenate([t.mcc for t in tables]))


def median_rows(store, x: Table, y: Table) -> Table:
    """Row-wise median of two tables: per character DOS.median with the side
    holding the smaller taxon code first; costs add up."""
    n, L = x.ids.shape
    if n == 0:
        return Table(np.zeros((0, L), np.int64), np.zeros(0, np.int64), np.zeros(0, np.int64))
    swap = x.mcc > y.mcc
    first = np.where(swap[:, None], y.ids, x.ids)
    second = np.where(swap[:, None], x.ids, y.ids)
    ids, c = store.median(first.ravel(), second.ravel())
    return Table(ids.reshape(n, L), x.cost + y.cost + c.reshape(n, L).sum(1), np.minimum(x.mcc, y.mcc))


def directional(store, t: Tree, leaf_ids: dict) -> tuple[Table, dict]:
    """Every directional median of t (AllDirChar.internal_downpass,
    src/allDirChar.ml:722-785): index[(u, v)] = row of v's side away from
    u.  leaf_ids[v] holds the store ids of leaf v's characters.  One batch
    per height."""
    keys = [(u, v) for v in t.nbrs for u in t.nbrs[v]]
    height = {}
    for key in keys:
        stack = [key]
        while stack:
            u, v = stack[-1]
            if (u, v) in height:
                stack.pop()
                continue
            if v in t.leaves:
                height[(u, v)] = 0
                stack.pop()
                continue
            kids = [(v, w) for w in t.others(v, u)]
            todo = [k for k in kids if k not in height]
            if todo:
                stack.extend(todo)
                continue
            height[(u, v)] = 1 + max(height[k] for k in kids)
            stack.pop()
    order = sorted(keys, key=lambda k: (height[k], k))
    index = {k: i for i, k in enumerate(order)}
    L = len(next(iter(leaf_ids.values())))
    K = len(order)
    tab = Table(np.zeros((K, L), np.int64), np.zeros(K, np.int64), np.zeros(K, np.int64))
    hs = np.array([height[k] for k in order])
    for i, (u, v) in enumerate(order):
        if hs[i] == 0:
            tab.ids[i] = leaf_ids[v]
            tab.mcc[i] = v
    for h in range(1, int(hs.max()) + 1 if K else 0):
        rows = np.nonzero(hs == h)[0]
        a = np.array([index[(order[i][1], t.others(order[i][1], order[i][0])[0])] for i in rows])
        b = np.array([index[(order[i][1], t.others(order[i][1], order[i][0])[1])] for i in rows])
        m = median_rows(store, tab.take(a), tab.take(b))
        tab.ids[rows], tab.cost[rows], tab.mcc[rows] = m.ids, m.cost, m.mcc
    return tab, index


# ------------------------------------------------------- the neighbourhood
@dataclass
class Side:
    """One component of a break: the removed vertex r, the dissolved
    ("spot") edge between its other two neighbours, and every other edge as
    (c, z) with z toward the spot.  For edge e: sub[e] = row of c's subtree
    (unchanged), sib[e] = row of the subtree of z's other child (unchanged),
    up_t[e] = row of the far spot end's side when z is a spot end (else -1),
    up_e[e] = the edge (z, parent of z) otherwise (else -1)."""

    r: int
    spot: int                  # row of r's side in the unbroken tree = the spot edge's median
    spot_edge: tuple
    edges: list = field(default_factory=list)
    sub: list = field(default_factory=list)
    sib: list = field(default_factory=list)
    up_t: list = field(default_factory=list)
    up_e: list = field(default_factory=list)
    depth: list = field(default_factory=list)


def make_side(t: Tree, index: dict, r: int, away: int) -> Side:
    x, y = t.others(r, away)
    s = Side(r=r, spot=index[(away, r)], spot_edge=(x, y))
    frontier = [(x, r, index[(r, y)], -1), (y, r, index[(r, x)], -1)]   # (vertex, toward spot, up row, up edge)
    depth = 1
    while frontier:
        nxt = []
        for z, tz, upt, upe in frontier:
            if z in t.leaves:
                continue
            kids = [w for w in t.nbrs[z] if w != tz]
            for c, c2 in ((kids[0], kids[1]), (kids[1], kids[0])):
                e = len(s.edges)
                s.edges.append((c, z))
                s.sub.append(index[(z, c)])
                s.sib.append(index[(z, c2)])
                s.up_t.append(upt)
                s.up_e.append(upe)
                s.depth.append(depth)
                nxt.append((c, z, -1, e))
        frontier = nxt
        depth += 1
    return s


@dataclass
class Break:
    edge: tuple                # (u, v): u on the tree side
    tree: Side                 # the tree side (u removed)
    clade: Side | None         # the clade (v removed), None for a leaf clade
    leaf_row: int              # row of the leaf clade's data (unused otherwise)


@dataclass
class Space:
    """The TBR neighbourhood of a tree, flattened: every side's edges as one
    global list (gsub, gsib, gup_t, gup_e in global numbering, gdepth)."""

    t: Tree
    breaks: list
    gsub: np.ndarray
    gsib: np.ndarray
    gup_t: np.ndarray
    gup_e: np.ndarray
    gdepth: np.ndarray
    gedge: list                # (c, z) per global edge
    side_start: dict           # id(side) -> first global edge


def neighbourhood(t: Tree, index: dict) -> Space:
    breaks = []
    for (u, v) in t.pre_order_edges():
        if u in t.leaves:
            continue   # the tree side is one vertex: nothing to join to
        tree_side = make_side(t, index, u, v)
        clade = make_side(t, index, v, u) if v not in t.leaves else None
        breaks.append(Break((u, v), tree_side, clade, index[(u, v)]))
    sub, sib, upt, upe, dep, ged = [], [], [], [], [], []
    start = {}
    for b in breaks:
        for s in (b.tree, b.clade):
            if s is None:
                continue
            g0 = len(sub)
            start[id(s)] = g0
            sub += s.sub
            sib += s.sib
            upt += s.up_t
            upe += [g0 + e if e >= 0 else -1 for e in s.up_e]
            dep += s.depth
            ged += s.edges
    arr = lambda x: np.array(x, np.int64)
    return Space(t, breaks, arr(sub), arr(sib), arr(upt), arr(upe), arr(dep), ged, start)


def side_refs(sp: Space, s: Side | None, leaf_row: int, K: int) -> np.ndarray:
    """Candidate rows of one side in the joint numbering (directional rows
    < K, edge medians K + global edge): the spot edge first, then the
    side's edges.  A leaf clade is its own data."""
    if s is None:
        return np.array([leaf_row], np.int64)
    g0 = sp.side_start[id(s)]
    return np.concatenate([[s.spot], K + g0 + np.arange(len(s.edges))]).astype(np.int64)


@dataclass
class Candidates:
    brk: np.ndarray            # break index
    clade: np.ndarray          # joint row of the clade root's median
    edge: np.ndarray           # joint row of the join edge's median
    base: np.ndarray           # broken tree cost (both components)
    ci: np.ndarray             # clade root index within its side (0 = dissolved edge / leaf)
    ei: np.ndarray             # join edge index within the tree side (0 = dissolved edge)


def candidates(sp: Space, dirs: Table, K: int) -> Candidates:
    out = {k: [] for k in ("brk", "clade", "edge", "base", "ci", "ei")}
    for bi, b in enumerate(sp.breaks):
        cr = side_refs(sp, b.clade, b.leaf_row, K)
        er = side_refs(sp, b.tree, -1, K)
        nB, nA = len(cr), len(er)
        cl = np.repeat(cr, nA)[1:]     # drops (dissolved clade edge, dissolved tree edge): the unbroken tree
        ed = np.tile(er, nB)[1:]
        ci = np.repeat(np.arange(nB), nA)[1:]
        ei = np.tile(np.arange(nA), nB)[1:]
        base = dirs.cost[b.tree.spot] + (dirs.cost[b.clade.spot] if b.clade is not None else 0)
        out["brk"].append(np.full(len(cl), bi, np.int64))
        out["clade"].append(cl)
        out["edge"].append(ed)
        out["base"].append(np.full(len(cl), base, np.int64))
        out["ci"].append(ci)
        out["ei"].append(ei)
    return Candidates(**{k: np.concatenate(v) if v else np.zeros(0, np.int64) for k, v in out.items()})


def changed_medians(store, sp: Space, dirs: Table, need: np.ndarray | None = None) -> tuple[Table, np.ndarray]:
    """The component directional medians that differ from the unbroken
    tree's (one per non-dissolved edge, the side toward the dissolved edge),
    for every global edge or only those in `need` (closed upward here).
    Returns the table (rows by global edge; unneeded rows are -1) and the
    mask of computed rows."""
    G = len(sp.gsub)
    L = dirs.ids.shape[1]
    mask = np.ones(G, bool) if need is None else np.zeros(G, bool)
    if need is not None:
        cur = np.unique(need)
        while len(cur):
            mask[cur] = True
            cur = sp.gup_e[cur]
            cur = np.unique(cur[(cur >= 0)])
            cur = cur[~mask[cur]]
    ch = Table(np.full((G, L), -1, np.int64), np.zeros(G, np.int64), np.zeros(G, np.int64))
    dmax = int(sp.gdepth[mask].max()) if mask.any() else 0
    for d in range(1, dmax + 1):
        rows = np.nonzero(mask & (sp.gdepth == d))[0]
        if not len(rows):
            continue
        x = dirs.take(sp.gsib[rows])
        fromt = sp.gup_t[rows] >= 0
        y = Table(np.where(fromt[:, None], dirs.ids[np.maximum(sp.gup_t[rows], 0)], ch.ids[np.maximum(sp.gup_e[rows], 0)]),
                  np.where(fromt, dirs.cost[np.maximum(sp.gup_t[rows], 0)], ch.cost[np.maximum(sp.gup_e[rows], 0)]),
                  np.where(fromt, dirs.mcc[np.maximum(sp.gup_t[rows], 0)], ch.mcc[np.maximum(sp.gup_e[rows], 0)]))
        m = median_rows(store, x, y)
        ch.ids[rows], ch.cost[rows], ch.mcc[rows] = m.ids, m.cost, m.mcc
    return ch, mask


def edge_medians(store, sp: Space, dirs: Table, ch: Table, rows: np.ndarray) -> Table:
    """Lazy edge medians (create_lazy_edge, src/allDirChar.ml:58-73) of the
    given global edges in their broken component."""
    return median_rows(store, dirs.take(sp.gsub[rows]), ch.take(rows))


def join_distances(store, ids_clade: np.ndarray, ids_edge: np.ndarray) -> np.ndarray:
    """Per character DOS.distance of candidate rows: (n, L) costs."""
    n, L = ids_clade.shape
    return store.distance(ids_clade.ravel(), ids_edge.ravel()).reshape(n, L)


@dataclass
class RoundResult:
    best: int                  # candidate index of the least estimate
    best_estimate: int
    estimates: np.ndarray      # per scored candidate (exhaustive: every candidate)
    scored: np.ndarray         # candidate indices scored exactly
    dist: np.ndarray           # (scored, L) exact per-character join costs
    seconds: float
    alignments: int
    cells: int


def exhaustive_round(store, sp: Space, dirs: Table, cand: Candidates) -> RoundResult:
    """Every candidate scored by the exact DP."""
    t0 = time.perf_counter()
    a0, c0 = store.alignments, store.cells
    mark = store.mark()
    K = len(dirs)
    ch, _ = changed_medians(store, sp, dirs)
    G = len(sp.gsub)
    em = edge_medians(store, sp, dirs, ch, np.arange(G))
    joint = concat([dirs, em])
    dist = join_distances(store, joint.ids[cand.clade], joint.ids[cand.edge])
    est = cand.base + dist.sum(1)
    best = int(np.argmin(est))
    res = RoundResult(best, int(est[best]), est, np.arange(len(est)), dist, time.perf_counter() - t0,
                      store.alignments - a0, store.cells - c0)
    store.release(mark)
    return res


def rescore_round(store, sp: Space, dirs: Table, cand: Candidates, pick: np.ndarray) -> RoundResult:
    """Exact scores of the candidates in `pick` only: the changed medians
    and edge medians they need, then their distances."""
    t0 = time.perf_counter()
    a0, c0 = store.alignments, store.cells
    mark = store.mark()
    K = len(dirs)
    pick = np.sort(np.asarray(pick, np.int64))
    refs = np.concatenate([cand.clade[pick], cand.edge[pick]])
    gneed = np.unique(refs[refs >= K] - K)
    ch, _ = changed_medians(store, sp, dirs, gneed)
    em = edge_medians(store, sp, dirs, ch, gneed)
    L = dirs.ids.shape[1]
    joint_ids = np.full((K + len(sp.gsub), L), -1, np.int64)
    joint_ids[:K] = dirs.ids
    joint_ids[K + gneed] = em.ids
    dist = join_distances(store, joint_ids[cand.clade[pick]], joint_ids[cand.edge[pick]])
    est = cand.base[pick] + dist.sum(1)
    k = int(np.argmin(est))     # pick is sorted: the first least in enumeration order
    res = RoundResult(int(pick[k]), int(est[k]), est, pick, dist, time.perf_counter() - t0,
                      store.alignments - a0, store.cells - c0)
    store.release(mark)
    return res


# ------------------------------------------------------------ tree edits
def _replace(t: Tree, a: int, old: int, new: int) -> None:
    t.nbrs[a][t.nbrs[a].index(old)] = new


def apply_join(sp: Space, cand: Candidates, k: int) -> Tree:
    """The tree of candidate k: break, dissolve the removed vertices, then
    join the clade's root edge (or leaf) to the tree-side edge through the
    removed vertices (Tree.break / Tree.join, src/tree.ml)."""
    b = sp.breaks[int(cand.brk[k])]
    u, v = b.edge
    t = copy.deepcopy(sp.t)
    ea = b.tree.spot_edge if cand.ei[k] == 0 else b.tree.edges[int(cand.ei[k]) - 1]
    x, y = b.tree.spot_edge
    _replace(t, x, u, y)
    _replace(t, y, u, x)
    del t.nbrs[u]
    if b.clade is not None:
        eb = b.clade.spot_edge if cand.ci[k] == 0 else b.clade.edges[int(cand.ci[k]) - 1]
        p, q = b.clade.spot_edge
        _replace(t, p, v, q)
        _replace(t, q, v, p)
        del t.nbrs[v]
        b1, b2 = eb
        _replace(t, b1, b2, v)
        _replace(t, b2, b1, v)
        t.nbrs[v] = [b1, b2, u]
    else:
        t.nbrs[v] = [u]
    a1, a2 = ea
    _replace(t, a1, a2, u)
    _replace(t, a2, a1, u)
    t.nbrs[u] = [a1, a2, v]
    return t


def splits(t: Tree) -> frozenset:
    """The tree's non-trivial leaf bipartitions (each as the side without
    the smallest leaf)."""
    leaves = sorted(t.leaves)
    lo = leaves[0]
    out = set()
    memo = {}

    def side(u, v):   # leaves on v's side away from u
        key = (u, v)
        if key in memo:
            return memo[key]
        stack, seen, acc = [v], {u, v}, []
        while stack:
            w = stack.pop()
            if w in t.leaves:
                acc.append(w)
            for z in t.nbrs[w]:
                if z not in seen:
                    seen.add(z)
                    stack.append(z)
        memo[key] = frozenset(acc)
        return memo[key]

    for a, b in t.edges():
        s = side(a, b)
        if 1 < len(s) < len(leaves) - 1:
            out.add(s if lo not in s else frozenset(leaves) - s)
    return frozenset(out)


def check_binary(t: Tree, n_taxa: int) -> None:
    assert set(t.leaves) == set(range(1, n_taxa + 1))
    for v, ns in t.nbrs.items():
        assert len(ns) == (1 if v in t.leaves else 3), (v, ns)
        for w in ns:
            assert v in t.nbrs[w], (v, w)
    assert len(t.nbrs) == 2 * n_taxa - 2
    seen, stack = {t.handle}, [t.handle]
    while stack:
        w = stack.pop()
        for z in t.nbrs[w]:
            if z not in seen:
                seen.add(z)
                stack.append(z)
    assert len(seen) == len(t.nbrs)
