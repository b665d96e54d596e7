"""Wagner build over dynamic and nonadditive characters: taxa join the tree
in code order, each at the edge whose join cost is least.

Reference behaviour followed:

* the first two taxa form one edge, the third joins it, and every later
  taxon is tried on every edge of the current tree, keeping the best tree
  (Ptree.make_wagner_tree and its add_node_everywhere / seq_add,
  src/ptree.ml:948-1060, with the ordered sequence of taxa);
* a candidate is scored by the join cost (AllDirChar.cost_fn,
  src/allDirChar.ml:1279-1332): the distance between the taxon and the
  edge's lazy median -- the median of the two directional medians across
  the edge (create_lazy_edge, src/allDirChar.ml:58-73) -- per dynamic
  character DOS.distance (src/seqCS.ml:818-864) and per nonadditive character
  the count of empty intersections (nonadd_median's changes,
  src/nonaddCSc.c:422-462);
* the least total wins; ties go to the first edge in pre-order from the
  handle (the search manager keeps one best tree).

The order among tied joins (the search manager's edge order and its
tie-keeping) is not pinned by any reference fixture; the build is
deterministic and identical between the GPU backend and the CPU oracle
backend.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .treedata import Dataset, Tree
from .treepass import TreeEval, is_empty


@dataclass
class BuildResult:
    tree: Tree
    join_costs: list          # least join cost at each addition
    alignments: int           # pairwise alignments issued during the build
    directional: dict         # the built tree's directional medians


def start_tree(n_taxa: int) -> Tree:
    """Taxa 1, 2, 3 around the first interior vertex n_taxa + 1."""
    c = n_taxa + 1
    t = Tree()
    t.nbrs = {1: [c], 2: [c], 3: [c], c: [1, 2, 3]}
    t.leaves = {1, 2, 3}
    t.handle = 1
    return t


def join(t: Tree, edge: tuple[int, int], x: int, v: int) -> None:
    """Join leaf x to edge (a, b) (oriented away from the handle) through the
    new interior vertex v."""
    a, b = edge
    t.nbrs[a][t.nbrs[a].index(b)] = v
    t.nbrs[b][t.nbrs[b].index(a)] = v
    t.nbrs[v] = [a, b, x]
    t.nbrs[x] = [v]
    t.leaves.add(x)


def join_costs(ev: TreeEval, x: int, edges, emed) -> np.ndarray:
    """Distance of leaf x to every edge median (dynamic plus nonadditive)."""
    ds, be = ev.ds, ev.be
    total = np.zeros(len(edges), np.int64)
    pairs, where = [], []
    leaf = [ds.chars[c][x - 1] for c in range(len(ds.chars))]
    live = [c for c in range(len(ds.chars)) if not is_empty(leaf[c])]
    for k, e in enumerate(edges):
        for c in live:
            if emed[k].is_empty(c):
                continue
            pairs.append((leaf[c], emed[k].seqs[c]))
            where.append(k)
    if pairs:
        d = be.distance(pairs)
        ev.n_align += len(pairs)
        np.add.at(total, np.array(where), np.array(d, np.int64))
    if ds.nonadd is not None and edges:
        leaf = np.broadcast_to(ds.nonadd[x - 1], (len(edges), ds.nonadd.shape[1]))
        _, changes = be.fitch(np.ascontiguousarray(leaf), np.stack([m.na for m in emed]))
        total += np.asarray(changes, np.int64)
    return total


def update_directional(ev: TreeEval, d: dict, v: int) -> None:
    """Bring the directional medians d up to date after vertex v joined the
    tree.  d[(u, w)] (w's side away from u) changes exactly when that side
    holds v, i.e. for the neighbours u of w other than the one toward v; the
    others describe unchanged subtrees and are kept.  The changed medians are
    computed in order of distance from v, one batch per distance: each needs
    the one toward v, computed one batch earlier."""
    t = ev.t
    toward = {v: None}
    frontier = [v]
    while frontier:
        batch, nxt = [], []
        for w in frontier:
            if w in t.leaves:
                continue
            for u in t.nbrs[w]:
                if u != toward[w]:
                    batch.append((u, w))
        if batch:
            jobs = []
            for (u, w) in batch:
                y1, y2 = t.others(w, u)
                jobs.append((d[(w, y1)], d[(w, y2)]))
            for key, r in zip(batch, ev._medians(jobs)):
                d[key] = r
        for w in frontier:
            for u in t.nbrs[w]:
                if u not in toward:
                    toward[u] = w
                    nxt.append(u)
        frontier = nxt


def wagner(ds: Dataset, backend, progress=None, on_step=None) -> BuildResult:
    """Wagner build in taxon order.  on_step(x, tree, edges, costs), if
    given, sees each addition before it is made: the tree so far (mutated
    afterwards, so a caller that keeps it copies it), its pre-order edges
    and taxon x's join cost to each."""
    n = ds.n_taxa
    if n < 3:
        raise ValueError("a Wagner build needs at least three taxa")
    t = start_tree(n)
    nxt = n + 2
    costs, n_align = [], 0
    ev = TreeEval(ds, t, backend)
    d = ev.directional()
    for x in range(4, n + 1):
        ev.n_align = 0
        edges = t.pre_order_edges()
        emed = ev._medians([(d[(b, a)], d[(a, b)]) for (a, b) in edges])
        jc = join_costs(ev, x, edges, emed)
        if on_step is not None:
            on_step(x, t, edges, jc)
        k = int(np.argmin(jc))           # first least
        a, b = edges[k]
        v = nxt
        join(t, edges[k], x, v)
        nxt += 1
        # the joined edge's two directions now end at v; x is a leaf
        d[(v, a)] = d.pop((b, a))
        d[(v, b)] = d.pop((a, b))
        d[(v, x)] = ev.leaf_data(x)
        update_directional(ev, d, v)
        costs.append(int(jc[k]))
        n_align += ev.n_align
        if progress is not None:
            progress(x)
    return BuildResult(t, costs, n_align, d)


def dataset_of(w) -> Dataset:
    """A Dataset of a simulated tree workload (dynamic loci + nonadditive)."""
    return Dataset(list(w.names), [list(l) for l in w.loci], w.nonadd)
