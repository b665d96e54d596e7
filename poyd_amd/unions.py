"""Sequence unions of device-resident sequences, and the join pruning of
the UnionBased search that reads them.

* Unions (Sequence.Unions, src/sequence.ml:1810-1895; SeqCS.union,
  src/seqCS.ml:1517-1566; Node.Union.union, src/node.ml:2355-2408): a
  leaf's union is its sequence; a vertex's union merges its two children's
  unions over the alignment behind its median (union_merge,
  src/union.c:102-164), the child with the smaller taxon code first, and an
  empty union passes the other one through.  The items of a union live in
  the DeviceStore arena (so they are sequences that plans align directly);
  the three int32 arrays live in a second device arena.
* SeqCS.distance_union (src/seqCS.ml:1569-1610): per character,
  sub_factor (0.8 for affine matrices, else 1) times cost_2 of the two
  unions' sequences with ?deltaw = max(8, length difference), 0 when one
  is empty -- DeviceStore.distance on the union sequences.
* UnionBased join pruning (union_dfs_wagner's should_continue_this_path,
  src/tabus.ml:530-551): the depth-first walk over the join edges keeps
  descending below an edge (a, b), a the parent of b and b interior, only
  while the least join delta found so far is at least
  Node.union_distance(clade, a).  Every node module of the reference
  defines union_distance as the constant 0. (src/node.ml:2719, marked for
  removal; src/allDirNode.ml:189, 572), so the test always passes and the
  search visits every join edge within the depth bound: pruned_walk below
  restates the walk and shows it, and an exhaustive round scores exactly
  that set of candidates.  The walk also accepts real union distances
  (distance_union), which the reference does not use.
"""
from __future__ import annotations

import numpy as np

from . import engine
from .seqstore import Aligned, DeviceStore


class UnionSet:
    """Unions built on the device with poyg_union_merge."""

    def __init__(self, store: DeviceStore, capacity_items: int):
        self.store, self.ctx = store, store.ctx
        self.cm = store.cm
        self.arr = engine.DeviceBuffer(self.ctx, 4 * int(capacity_items))
        self.capacity = int(capacity_items)
        self.top = 0
        self.sid = np.zeros(0, np.int64)    # store id of the union's items
        self.e_off = np.zeros(0, np.int64)  # int32 index of end[0], ca[0], cb[0]
        self.a_off = np.zeros(0, np.int64)
        self.b_off = np.zeros(0, np.int64)

    def __len__(self) -> int:
        return len(self.sid)

    def _reserve(self, count: int) -> int:
        at = (self.top + 63) & ~63
        if at + count > self.capacity:
            raise engine.PoygError(f"union arena full ({self.capacity} entries)")
        self.top = at + int(count)
        return at

    def _append(self, sid, e, a, b) -> np.ndarray:
        first = len(self.sid)
        self.sid = np.concatenate([self.sid, sid])
        self.e_off = np.concatenate([self.e_off, e])
        self.a_off = np.concatenate([self.a_off, a])
        self.b_off = np.concatenate([self.b_off, b])
        return np.arange(first, len(self.sid), dtype=np.int64)

    def n(self, u) -> np.ndarray:
        return self.store.ln[self.sid[np.asarray(u, np.int64)]]

    def leaves(self, ids) -> np.ndarray:
        """Unions.leaf of store sequences (the items are the sequences
        themselves; end = 0, ca = cb = positions)."""
        ids = np.asarray(ids, np.int64)
        ln = self.store.ln[ids].astype(np.int64)
        host, e, a, b = [], [], [], []
        base = self._reserve(int(3 * ln.sum() + 64 * len(ids)))
        at = base
        for n in ln:
            pos = np.arange(n, dtype=np.int32)
            host.append(np.concatenate([np.zeros(n, np.int32), pos, pos]))
            e.append(at)
            a.append(at + n)
            b.append(at + 2 * n)
            at += 3 * n
        blob = np.concatenate(host) if host else np.zeros(0, np.int32)
        if len(blob):
            engine._check(engine.load().poyg_dev_upload(self.ctx.h, self.arr.ptr + 4 * base, blob.ctypes.data,
                                                        blob.nbytes), "poyg_dev_upload")
        return self._append(ids, np.array(e, np.int64), np.array(a, np.int64), np.array(b, np.int64))

    def merge(self, ua, ub, al: Aligned) -> np.ndarray:
        """SeqCS.union of pairs (ua[k] the smaller taxon code side) over
        the alignments of their vertices' medians (DeviceStore.median with
        aligned=True, the same pair order).  An empty union passes the
        other one through."""
        ua = np.asarray(ua, np.int64)
        ub = np.asarray(ub, np.int64)
        st = self.store
        ea, eb = st.empty[self.sid[ua]], st.empty[self.sid[ub]]
        out = np.where(ea, ub, ua).astype(np.int64)
        work = np.nonzero(~ea & ~eb)[0]
        if len(work) == 0:
            return out
        if (al.length[work] < 0).any():
            raise engine.PoygError("a union needs the alignment behind its vertex's median "
                                   "(an empty median with non-empty unions is not supported)")
        na = self.n(ua[work]).astype(np.int64)
        nb = self.n(ub[work]).astype(np.int64)
        cap = na + nb + 2
        jobs, s_at, i_at = [], [], []
        arena, arr = st.arena.ptr, self.arr.ptr
        for k, w in enumerate(work):
            c = int(cap[k])
            so = st._reserve(c)
            io = self._reserve(3 * c)
            s_at.append(so)
            i_at.append(io)
            x, y = ua[w], ub[w]
            jobs.append(engine.UnionJob(
                arena + int(al.a_off[w]), arena + int(al.b_off[w]),
                None if al.mwg_off[w] < 0 else arena + int(al.mwg_off[w]),
                arena + int(st.off[self.sid[x]]), arena + int(st.off[self.sid[y]]),
                arr + 4 * int(self.e_off[x]), arr + 4 * int(self.e_off[y]),
                arena + so, arr + 4 * io, arr + 4 * (io + c), arr + 4 * (io + 2 * c),
                int(al.length[w]), int(na[k]), int(nb[k]), c))
        n = engine.union_merge(self.ctx, self.cm, jobs)
        if (n < 0).any():
            raise engine.PoygError(f"{int((n < 0).sum())} unions failed: the children's unions do not match "
                                   "their alignment")
        s_at = np.array(s_at, np.int64)
        i_at = np.array(i_at, np.int64)
        lo = cap - n
        soff = s_at + lo
        gaps, empty = st._stats(soff, n.astype(np.int32))
        sid = st._append(soff, n.astype(np.int32), gaps, empty)
        out[work] = self._append(sid, i_at + lo, i_at + cap + lo, i_at + 2 * cap + lo)
        return out

    def get(self, u: int):
        """(items, end, ca, cb) of union u, on the host."""
        u = int(u)
        n = int(self.n(u))
        seq = self.store.get(int(self.sid[u]))
        out = [seq]
        for off in (self.e_off[u], self.a_off[u], self.b_off[u]):
            x = np.empty(n, np.int32)
            if n:
                engine._check(engine.load().poyg_dev_download(self.ctx.h, x.ctypes.data,
                                                              self.arr.ptr + 4 * int(off), 4 * n),
                              "poyg_dev_download")
            out.append(x)
        return tuple(out)


def distance_union(us: UnionSet, ua, ub) -> np.ndarray:
    """SeqCS.distance_union of one character's union pairs (float)."""
    sub = 0.8 if us.store.affine else 1.0
    return sub * us.store.distance(us.sid[np.asarray(ua, np.int64)], us.sid[np.asarray(ub, np.int64)]).astype(
        np.float64)


def rooted_unions(store: DeviceStore, us: UnionSet, children: dict, leaf_ids: dict, mcc: dict, root: int):
    """Medians and unions of every vertex of a rooted binary tree, one
    level (height) per batch: children[v] = (x, y) for interior vertices,
    leaf_ids[v] the store id of leaf v's sequence, mcc[v] its subtree's
    smallest taxon code (the operand order of Node.cs_median and of
    Node.Union.union).  Returns (median ids, union ids, downpass cost) per
    vertex, as dicts."""
    height, order = {}, []

    def visit(v):
        stack = [(v, False)]
        while stack:
            x, done = stack.pop()
            if x not in children:
                height[x] = 0
                continue
            if done:
                height[x] = 1 + max(height[c] for c in children[x])
                order.append(x)
            else:
                stack.append((x, True))
                stack.extend((c, False) for c in children[x])

    visit(root)
    med, uni, cost = {}, {}, {}
    leaves = [v for v in height if height[v] == 0]
    lu = us.leaves([leaf_ids[v] for v in leaves])
    for v, u in zip(leaves, lu):
        med[v], uni[v], cost[v] = leaf_ids[v], int(u), 0
    for h in range(1, max(height.values()) + 1):
        level = [v for v in order if height[v] == h]
        pa, pb = [], []
        for v in level:
            x, y = children[v]
            if mcc[y] < mcc[x]:
                x, y = y, x
            pa.append(x)
            pb.append(y)
        m, c, al = store.median([med[x] for x in pa], [med[y] for y in pb], aligned=True)
        u = us.merge([uni[x] for x in pa], [uni[y] for y in pb], al)
        for k, v in enumerate(level):
            med[v], uni[v] = int(m[k]), int(u[k])
            cost[v] = int(c[k]) + cost[pa[k]] + cost[pb[k]]
    return med, uni, cost


def pruned_walk(edges_below, start_edges, delta, union_delta, max_depth=None):
    """The join-edge walk of wagner_dfs_distance_based with
    union_dfs_wagner's test (src/tabus.ml:400-520, 530-551), for one clade.

    edges_below(e) -> the two edges one step further from the handle below
    join edge e = (a, b), i.e. b's other edges (empty if b is a leaf);
    start_edges: (edge, depth) in push order -- the handle node's two other
    edges at depth 1, then the handle edge (s1, s2) at depth 0, whose
    expansion reaches s2's side (update_join, src/tabus.ml:420-436);
    delta(e): the join delta of e; union_delta(e): Node.union_distance of
    the clade and e's near vertex.  Returns the edges scored, in order.

    The walk pops edges from the to_calculate stack; an edge at depth <=
    max_depth is scored, and its delta lowers current_delta; it is expanded
    (its children edges pushed, to_compare -> to_do_later sorted by
    decreasing delta, then popped) only if current_delta >=
    union_delta(e) when it was popped."""
    to_calc = list(start_edges)   # a stack: the end of the list is its top
    to_compare, to_do_later = [], []
    current = float("inf")
    scored = []
    while True:
        if to_calc:
            e, depth = to_calc.pop()
            if max_depth is not None and depth > max_depth:
                continue
            keep = current >= union_delta(e)
            d = float(delta(e))
            scored.append(e)
            current = min(current, d)
            if keep:
                to_compare.append((e, depth, d))
            continue
        if to_compare:
            # sort_and_choose_join_edge: decreasing delta pushed onto the
            # stack, so the least delta is expanded first
            for item in sorted(to_compare, key=lambda t: -t[2]):
                to_do_later.append(item[:2])
            to_compare = []
        if not to_do_later:
            return scored
        e, depth = to_do_later.pop()
        for c in edges_below(e):
            to_calc.append((c, depth + 1))
