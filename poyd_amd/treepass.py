"""Tree evaluation of dynamic (DO) characters: the composition POY's
AllDirChar performs around the alignment kernels, batched so that every
alignment of one tree level runs as one GPU plan.

What is computed (reference behaviour, restated):

* Directional medians (AllDirChar.internal_downpass,
  src/allDirChar.ml:722-785): every interior vertex v holds, for each
  neighbour u, the median of the two subtrees on its other sides
  ("v not_with u"). A median of two vertices orders them by the smallest
  taxon code in their subtrees (Node.cs_median, src/node.ml:341-347), and
  per character is DOS.median (src/seqCS.ml:747-776): an empty side passes
  the other through at cost 0; otherwise linear models align (align_2) and
  take ancestor_2, affine models take align_affine_3's median.  A subtree's
  cost is the sum of its medians' costs (Node.median, src/node.ml:638-678).
* Edge medians (create_lazy_edge, src/allDirChar.ml:58-73) and the root
  (create_root / pick_best_root, src/allDirChar.ml:97-125, 836-861): the
  first root is the handle's edge; then the edges, oriented away from the
  handle and sorted by decreasing (a, b), replace it whenever their total
  downpass cost is strictly smaller.
* Single assignment (assign_single, src/allDirChar.ml:288-409): the root
  single is closest(root median, handle side); every vertex then gets
  closest(parent single, its directional median away from the parent)
  (DOS.to_single, src/seqCS.ml:730-745; Sequence.Align.closest,
  src/sequence.ml:985-1033; Cost_matrix.Two_D.get_closest,
  src/cost_matrix.ml:681-700).
* Tree cost (check_cost, src/allDirChar.ml:185-213): the sum over edges of
  DOS.distance between the two singles (src/seqCS.ml:818-864: band
  deltaw = max(8, length difference) for linear models, cost_2_affine for
  affine ones; 0 when a side is empty).
* Nonadditive characters (NonaddCS8): every directional median also takes
  the Fitch median of its two sides (nonadd_median, src/nonaddCSc.c:422-462,
  through Node.cs_median, src/node.ml:264-309), and its cost adds the
  changes; the tree's nonadditive cost is the Fitch length, the cost of any
  edge median (the same on every edge).

The alignments go to a backend with three batched operations (median,
align, distance).  EngineBackend runs them on the GPU through libpoyg.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .treedata import GAP, Dataset, Tree


def is_empty(s: np.ndarray) -> bool:
    """Sequence.is_empty (src/sequence.ml:442-450): every element is a gap."""
    if len(s) > 1 and s[1] != GAP:   # the common case: a base after the leading gap
        return False
    return bool((s == GAP).all())


def closest_table(cost: np.ndarray) -> np.ndarray:
    """Cost_matrix.Two_D.get_closest for all (a, b) of a combination
    alphabet with gap bit 16: the first bit of b (lowest first) whose cost
    against a is strictly below all earlier ones, after b is reduced to the
    gap (both carry it) or stripped of it.  Entries that leave no bit are 0
    (the reference fails there)."""
    out = np.zeros((32, 32), np.uint8)
    for a in range(32):
        for b in range(32):
            if a == GAP or b == GAP:
                bb = b
            elif (a & GAP) and (b & GAP):
                bb = GAP
            else:
                bb = b & ~GAP
            best, best_cost = a, None
            for bit in range(5):
                x = 1 << bit
                if bb & x:
                    c = int(cost[a, x])
                    if best_cost is None or c < best_cost:
                        best, best_cost = x, c
            out[a, b] = best if best_cost is not None else 0
    return out


def remove_gaps(s: np.ndarray) -> np.ndarray:
    """closest's remove_gaps: drop gap elements, then prepend one gap."""
    body = s[s != GAP]
    return np.concatenate([np.array([GAP], np.uint8), body])


@dataclass
class DirData:
    """A vertex seen from one neighbour: one sequence per character, the
    subtree's downpass cost and the smallest taxon code inside it."""

    seqs: list
    cost: int
    mcc: int
    na: np.ndarray | None = None   # nonadditive state sets (Fitch median)
    na_cost: int = 0
    empty: list | None = None      # is_empty of each sequence, once known

    def is_empty(self, c: int) -> bool:
        if self.empty is None:
            self.empty = [is_empty(x) for x in self.seqs]
        return self.empty[c]


@dataclass
class Diagnosis:
    cost: int                 # Adjusted tree cost (dynamic characters) plus the nonadditive cost
    downpass: int             # total downpass cost at the chosen root
    root: tuple[int, int]
    singles: dict
    alignments: int           # pairwise alignments issued
    nonadd: int = 0           # Fitch length of the nonadditive characters


class TreeEval:
    """All-directions diagnosis of one tree over a Dataset of DO characters."""

    def __init__(self, ds: Dataset, tree: Tree, backend):
        self.ds, self.t, self.be = ds, tree, backend
        self.closest = closest_table(backend.cost_table())
        self.n_align = 0

    def leaf_data(self, v: int) -> DirData:
        na = self.ds.nonadd
        return DirData([ch[v - 1] for ch in self.ds.chars], 0, v, None if na is None else na[v - 1])

    # ---------------------------------------------------------- medians
    def _medians(self, jobs):
        """jobs: list of (x: DirData, y: DirData).  Returns DirData list."""
        pairs, where = [], []
        out = []
        L = len(self.ds.chars)
        for j, (x, y) in enumerate(jobs):
            first, second = (x, y) if x.mcc < y.mcc else (y, x)
            seqs, empty = [None] * L, [False] * L
            for c in range(L):
                if first.is_empty(c):
                    seqs[c], empty[c] = second.seqs[c], second.is_empty(c)
                elif second.is_empty(c):
                    seqs[c] = first.seqs[c]
                else:
                    pairs.append((first.seqs[c], second.seqs[c]))
                    where.append((j, c))
            out.append(DirData(seqs, x.cost + y.cost, min(x.mcc, y.mcc), empty=empty))
        if pairs:
            res = self.be.median(pairs)
            self.n_align += len(pairs)
            for (j, c), (m, cost) in zip(where, res):
                out[j].seqs[c] = m
                out[j].empty[c] = is_empty(m)
                out[j].cost += int(cost)
        if self.ds.nonadd is not None and jobs:
            med, added = self.be.fitch(np.stack([x.na for x, _ in jobs]), np.stack([y.na for _, y in jobs]))
            for j, (x, y) in enumerate(jobs):
                out[j].na = med[j]
                out[j].na_cost = x.na_cost + y.na_cost + int(added[j])
        return out

    def directional(self):
        """dir[(u, v)] = v not_with u, for every ordered adjacent pair."""
        t = self.t
        leaf = {v: self.leaf_data(v) for v in t.leaves}
        height = {}

        def h(u, v):
            key = (u, v)
            if key in height:
                return height[key]
            stack = [key]
            while stack:
                uu, vv = stack[-1]
                if vv in t.leaves:
                    height[(uu, vv)] = 0
                    stack.pop()
                    continue
                kids = [(vv, w) for w in t.others(vv, uu)]
                todo = [k for k in kids if k not in height]
                if todo:
                    stack.extend(todo)
                    continue
                height[(uu, vv)] = 1 + max(height[k] for k in kids)
                stack.pop()
            return height[key]

        keys = [(u, v) for v in t.nbrs for u in t.nbrs[v]]
        for k in keys:
            h(*k)
        d = {}
        for (u, v), hv in height.items():
            if hv == 0:
                d[(u, v)] = leaf[v]
        for level in range(1, max(height.values()) + 1):
            batch = [k for k, hv in height.items() if hv == level]
            batch.sort()
            jobs = []
            for (u, v) in batch:
                x, y = t.others(v, u)
                jobs.append((d[(v, x)], d[(v, y)]))
            for k, r in zip(batch, self._medians(jobs)):
                d[k] = r
        return d

    # ---------------------------------------------------------- singles
    def _to_single(self, jobs):
        """jobs: list of (parent seqs, mine seqs) -> list of single seqs
        (DOS.to_single per character)."""
        out = [list(m) for _, m in jobs]
        pairs, where = [], []
        for j, (p, m) in enumerate(jobs):
            for c in range(len(m)):
                mine = m[c]
                if is_empty(mine):
                    continue
                par = mine if is_empty(p[c]) else p[c]
                if par.shape == mine.shape and np.array_equal(par, mine):
                    # both equal: drop the gap bit after position 0
                    s1 = par.copy()
                    s1[1:] &= ~np.uint8(GAP)
                    out[j][c] = remove_gaps(self.closest[s1, s1])
                else:
                    pairs.append((par, mine))
                    where.append((j, c))
        if pairs:
            res = self.be.align(pairs)
            self.n_align += len(pairs)
            for (j, c), (a1, a2) in zip(where, res):
                out[j][c] = remove_gaps(self.closest[a1, a2])
        return out

    def diagnose(self) -> Diagnosis:
        t = self.t
        d = self.directional()
        # edge medians
        pre = t.pre_order_edges()
        jobs = [(d[(b, a)], d[(a, b)]) for (a, b) in pre]
        emed = dict(zip(pre, self._medians(jobs)))
        # root: the handle's edge first, then decreasing (a, b)
        h = t.handle
        root = (h, t.nbrs[h][0])
        emed_any = {frozenset(k): v for k, v in emed.items()}
        best = emed_any[frozenset(root)].cost
        for e in sorted(pre, reverse=True):
            c = emed[e].cost
            if abs(best) > abs(c):
                best, root = c, e
        a, b = root
        rmed = emed_any[frozenset(root)]
        rsingle = self._to_single([(rmed.seqs, d[(b, a)].seqs)])[0]
        singles = {}
        frontier = [(rsingle, b, a), (rsingle, a, b)]   # (parent single, parent, current)
        while frontier:
            res = self._to_single([(ps, d[(p, v)].seqs) for ps, p, v in frontier])
            nxt = []
            for (ps, p, v), s in zip(frontier, res):
                singles[v] = s
                for w in t.others(v, p):
                    nxt.append((s, v, w))
            frontier = nxt
        # adjusted cost: distances over every edge
        pairs = []
        for (p, v) in pre:
            for c in range(len(self.ds.chars)):
                x, y = singles[p][c], singles[v][c]
                if is_empty(x) or is_empty(y):
                    continue
                pairs.append((x, y))
        total = 0
        if pairs:
            total = int(sum(int(x) for x in self.be.distance(pairs)))
            self.n_align += len(pairs)
        return Diagnosis(total + rmed.na_cost, int(best), root, singles, self.n_align, rmed.na_cost)


    # ---------------------------------------------------------- ApproxD
    def post_order(self) -> list[int]:
        """Tree.post_order_node_visit from the handle (src/tree.ml:1549-1625):
        every vertex after the subtrees of its two other neighbours, taken
        in the order of its neighbour record; an interior handle visits its
        first neighbour's side, then the other two, then itself."""
        t = self.t
        out = []

        def traverse(pred, v):
            stack = [(pred, v, False)]
            while stack:
                p, x, done = stack.pop()
                if done or x in t.leaves:
                    out.append(x)
                    continue
                c1, c2 = t.others(x, p)
                stack += [(p, x, True), (x, c2, False), (x, c1, False)]

        h = t.handle
        if h in t.leaves:
            traverse(h, t.nbrs[h][0])
        else:
            for u in t.nbrs[h]:
                traverse(h, u)
        out.append(h)
        return out

    def _readjust(self, jobs):
        """Sequence.readjust (src/sequence.ml:1166-1199) through
        SeqCS.readjust's `ApproxD branch (src/seqCS.ml:680-728), per
        character, for a batch of vertices.  jobs: (ch1, ch2, parent, mine,
        chars) with per-character sequence lists and the characters to
        check (None: all).  For each character: the three pairwise medians
        ab, bc, ac (align_2 + median_2, or align_affine_3's median), each
        aligned with the third sequence; the least total gives the centre
        (make_center's order on ties: ab|c before the others, ac|b before
        bc|a), and the new
        sequence is Align.closest of the third sequence and that pairwise
        median.  Returns per job (new sequences, set of changed
        characters)."""
        L = len(self.ds.chars)
        todo = []
        for j, (a, b, p, m, chars) in enumerate(jobs):
            for c in (range(L) if chars is None else sorted(chars)):
                if is_empty(a[c]) or is_empty(b[c]) or is_empty(p[c]):
                    raise NotImplementedError("ApproxD readjustment with an empty sequence among the three")
                todo.append((j, c))
        out = [(list(m), set()) for (_a, _b, _p, m, _c) in jobs]
        if not todo:
            return out
        r1 = self.be.median([x for (j, c) in todo
                             for x in ((jobs[j][0][c], jobs[j][1][c]), (jobs[j][1][c], jobs[j][2][c]),
                                       (jobs[j][0][c], jobs[j][2][c]))])
        pw = [(r1[3 * k], r1[3 * k + 1], r1[3 * k + 2]) for k in range(len(todo))]   # ab, bc, ac
        r2 = self.be.median([x for k, (j, c) in enumerate(todo)
                             for x in ((pw[k][0][0], jobs[j][2][c]), (pw[k][1][0], jobs[j][0][c]),
                                       (pw[k][2][0], jobs[j][1][c]))])
        self.n_align += 6 * len(todo)
        picks = []
        for k, (j, c) in enumerate(todo):
            a, b, p = jobs[j][0][c], jobs[j][1][c], jobs[j][2][c]
            (ab, cab), (bc, cbc), (ac, cac) = pw[k]
            cabc = int(r2[3 * k][1]) + int(cab)
            cbca = int(r2[3 * k + 1][1]) + int(cbc)
            cacb = int(r2[3 * k + 2][1]) + int(cac)
            if cabc <= cbca:
                picks.append((p, ab) if cabc <= cacb else (b, ac))
            elif cbca < cacb:
                picks.append((a, bc))
            else:
                picks.append((b, ac))
        closest = self._to_single([([third], [med]) for third, med in picks])
        for k, (j, c) in enumerate(todo):
            s = closest[k][0]
            seqs, changed = out[j]
            old = jobs[j][3][c]
            if not (len(s) == len(old) and np.array_equal(s, old)):
                changed.add(c)
            seqs[c] = s
        return out

    def _readjust_3d(self, jobs):
        """Sequence.Align.readjust_3d (src/sequence.ml:1117-1139) through
        SeqCS.readjust's `ThreeD branch (src/seqCS.ml:680-728), per
        character, for a batch of vertices (jobs as _readjust): when the two
        children, the vertex and the parent all have the same length the
        vertex keeps its sequence; otherwise align_3_powell_inter
        (:1089-1115): Powell's three-way search of (ch1, ch2, parent) with
        mismatch cost(A, C), gap opening (affine matrices) and extension
        cost(A, gap), and the new sequence is the Three_D median of every
        column, gaps dropped, behind one leading gap."""
        L = len(self.ds.chars)
        out = [(list(m), set()) for (_a, _b, _p, m, _c) in jobs]
        todo = []
        for j, (a, b, p, m, chars) in enumerate(jobs):
            for c in (range(L) if chars is None else sorted(chars)):
                if is_empty(a[c]) or is_empty(b[c]) or is_empty(p[c]):
                    raise NotImplementedError("ThreeD readjustment with an empty sequence among the three")
                if len(a[c]) == len(b[c]) and len(m[c]) == len(p[c]) and len(a[c]) == len(m[c]):
                    continue
                todo.append((j, c))
        if not todo:
            return out
        res = self.be.powell_median([(jobs[j][0][c], jobs[j][1][c], jobs[j][2][c]) for (j, c) in todo])
        self.n_align += len(todo)
        for (j, c), (med, _cost) in zip(todo, res):
            seqs, changed = out[j]
            old = jobs[j][3][c]
            if not (len(med) == len(old) and np.array_equal(med, old)):
                changed.add(c)
            seqs[c] = med
        return out

    def _edge_cost(self, seqs: dict) -> int:
        """check_cost (src/allDirChar.ml:185-213) over the given per-vertex
        sequences: DOS.distance summed over every edge."""
        pairs = []
        for (p, v) in self.t.pre_order_edges():
            for c in range(len(self.ds.chars)):
                x, y = seqs[p][c], seqs[v][c]
                if is_empty(x) or is_empty(y):
                    continue
                pairs.append((x, y))
        if not pairs:
            return 0
        self.n_align += len(pairs)
        return int(sum(int(x) for x in self.be.distance(pairs)))

    def approx_d(self, diag: Diagnosis | None = None, iterations: int | None = None, sequential: bool = False,
                 three_d: bool = False):
        """Iterative `ApproxD (AllDirChar.downpass, src/allDirChar.ml:879-883:
        internal_downpass, pick_best_root, assign_single, then adjust_tree,
        :488-670).  Starting from the single assignment, every interior
        vertex in post order (vertices not touched by the previous pass'
        changes are skipped, and only the characters that changed next to
        them are checked) is readjusted against its three neighbours
        (Node.readjust with ch1, ch2, parent = the first, second and third
        neighbour of its record, Tree.Interior); a pass is kept only when it
        changed something and lowered the tree cost, and at most
        `iterations` passes run (None: until a pass changes nothing or does
        not lower the cost).  A vertex reads the current sequences of its
        neighbours, so the passes run in layers: a vertex waits only for its
        neighbours that come earlier in the post order (sequential=True runs
        one vertex at a time; the results are the same).  three_d: Iterative
        `ThreeD, the same passes with Sequence.Align.readjust_3d (Powell's
        three-way search) as the readjustment.  Returns (adjusted sequences
        per vertex, their tree cost, passes kept)."""
        d = diag if diag is not None else self.diagnose()
        t = self.t
        adj = {v: list(s) for v, s in d.singles.items()}
        order = self.post_order()
        prev = d.cost - d.nonadd
        affected = {v: None for v in order}
        count = (1 << 30) if iterations is None else int(iterations)
        kept = 0
        while count > 0:
            new = dict(adj)
            nxt: dict = {}
            changed = False
            verts = [v for v in order if v in affected and v not in t.leaves]
            if sequential:
                layers = [[v] for v in verts]
            else:
                lvl: dict = {}
                for v in verts:
                    lvl[v] = 1 + max([lvl[u] for u in t.nbrs[v] if u in lvl], default=-1)
                layers = [[v for v in verts if lvl[v] == k] for k in range(max(lvl.values(), default=-1) + 1)]
            for batch in layers:
                jobs = [(new[t.nbrs[v][0]], new[t.nbrs[v][1]], new[t.nbrs[v][2]], new[v], affected[v]) for v in batch]
                readjust = self._readjust_3d if three_d else self._readjust
                for v, (seqs, modified) in zip(batch, readjust(jobs)):
                    if not modified:
                        continue
                    new[v] = seqs
                    changed = True
                    for x in (*t.nbrs[v], v):
                        nxt[x] = (nxt.get(x) or set()) | modified
            if not changed:
                break
            cost = self._edge_cost(new)
            if cost >= prev:
                break
            adj, prev, affected = new, cost, nxt
            kept += 1
            count -= 1
        return adj, prev, kept


# ---------------------------------------------------------------- backends

class EngineBackend:
    """Batched alignments on the GPU through libpoyg plans."""

    def __init__(self, ctx, cm):
        from . import engine
        self.engine = engine
        self.ctx, self.cm = ctx, cm
        self.affine = cm.affine
        self.cells = 0          # DP cells of every alignment run so far
        self.fitch_chars = 0    # nonadditive character medians computed so far

    def cost_table(self) -> np.ndarray:
        return np.array([[self.cm.cost(a, b) for b in range(32)] for a in range(32)], np.int64)

    @staticmethod
    def _batch(seqs):
        from .synth import SeqBatch
        ln = np.array([len(s) for s in seqs], np.int32)
        off = np.zeros(len(seqs), np.int64)
        if len(seqs) > 1:
            off[1:] = np.cumsum(ln[:-1], dtype=np.int64)
        return SeqBatch(np.concatenate(seqs).astype(np.uint8), off, ln)

    def _run(self, mode, pairs, dws=None):
        A = self._batch([p[0] for p in pairs])
        B = self._batch([p[1] for p in pairs])
        r = self.engine.run_pairs(self.ctx, self.cm, mode, A, B, dws)
        self.cells += r.cells
        return r

    def median(self, pairs):
        # (the medians are read-only views into the run's output block)
        e = self.engine
        slot = 0 if self.affine else 2
        r = e.run_pair_list(self.ctx, self.cm, e.POYG_MODE_AFFINE_ALIGN if self.affine else e.POYG_MODE_LINEAR_ALIGN,
                            pairs)
        self.cells += r.cells
        r.out.flags.writeable = False
        return list(zip(r.seqs(slot), r.cost.tolist()))

    def align(self, pairs):
        e = self.engine
        if self.affine:
            r = self._run(e.POYG_MODE_AFFINE_ALIGN, pairs)
            return [(r.seq(k, 2).copy(), r.seq(k, 3).copy()) for k in range(len(pairs))]
        r = self._run(e.POYG_MODE_LINEAR_ALIGN, pairs)
        return [(r.seq(k, 0).copy(), r.seq(k, 1).copy()) for k in range(len(pairs))]

    def fitch(self, a, b):
        self.fitch_chars += int(np.size(a))
        return self.engine.fitch_pairs(self.ctx, a, b)

    POWELL_WORKERS = 8   # searches in flight at once (each on a context of its own)

    def powell_median(self, triples):
        """align_3_powell_inter (src/sequence.ml:1089-1115) per triple on
        the GPU: poyg_powell_3d's alignment, then the Three_D median of
        every column (host lookups), gaps dropped, one leading gap.  A
        search is a chain of small launches, so several run at once: worker
        threads, each with a context (stream) of its own, share the batch
        (results in the batch's order)."""
        e = self.engine
        if getattr(self, "_med3", None) is None:
            cm3 = e.CostMatrix3(self.cm)
            self._med3 = np.array([[[cm3.median(a, b, c) for c in range(32)] for b in range(32)] for a in range(32)],
                                  np.uint8)
        mm, ge = self.cm.cost(1, 2), self.cm.cost(1, 16)
        go = self.cm.gap_opening if self.affine else 0

        def one(ctx, t):
            cost, (x, y, z) = e.powell_3d(ctx, t[0], t[1], t[2], mm, go, ge)
            med = self._med3[x, y, z]
            return np.concatenate([[GAP], med[med != GAP]]).astype(np.uint8), cost

        k = min(self.POWELL_WORKERS, len(triples))
        if k <= 1:
            return [one(self.ctx, t) for t in triples]
        import threading
        if len(getattr(self, "_powell_ctx", [])) < k - 1:
            self._powell_ctx = [e.Context(getattr(self.ctx, "device", 0)) for _ in range(k - 1)]
        ctxs = [self.ctx] + self._powell_ctx[:k - 1]
        out = [None] * len(triples)
        errors = []

        def work(w):
            try:
                for i in range(w, len(triples), k):
                    out[i] = one(ctxs[w], triples[i])
            except Exception as exc:   # re-raised by the caller's thread
                errors.append(exc)

        threads = [threading.Thread(target=work, args=(w,)) for w in range(k)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        if errors:
            raise errors[0]
        return out

    def distance(self, pairs):
        e = self.engine
        if self.affine:
            return self._run(e.POYG_MODE_AFFINE_COST, pairs).cost.tolist()
        A = self._batch([p[0] for p in pairs])
        B = self._batch([p[1] for p in pairs])
        dws = e.batch_deltaw(A, B, np.maximum(8, np.abs(A.ln.astype(np.int64) - B.ln)))
        r = e.run_pairs(self.ctx, self.cm, e.POYG_MODE_LINEAR_COST, A, B, dws)
        self.cells += r.cells
        return r.cost.tolist()
