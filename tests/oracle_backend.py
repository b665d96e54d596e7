"""Tree-pass backend that runs every alignment on the CPU oracle (test
infrastructure: GPU tree passes are compared with it)."""
from __future__ import annotations

import numpy as np

from oracle import pyoracle as po


class OracleBackend:
    def __init__(self, cm: po.CM2, affine: bool):
        self.cm, self.affine = cm, affine

    def cost_table(self) -> np.ndarray:
        return np.array([[self.cm.cost(a, b) for b in range(32)] for a in range(32)], np.int64)

    def median(self, pairs):
        out = []
        for a, b in pairs:
            if self.affine:
                med, ri, rj, cost, mwg = po.align_affine_3(a, b, self.cm)
                out.append((med, cost))
            else:
                cost, oa, ob = po.align_2(a, b, self.cm)
                out.append((po.ancestor_2(oa, ob, self.cm), cost))
        return out

    def align(self, pairs):
        out = []
        for a, b in pairs:
            if self.affine:
                med, ri, rj, cost, mwg = po.align_affine_3(a, b, self.cm)
                out.append((ri, rj))
            else:
                cost, oa, ob = po.align_2(a, b, self.cm)
                out.append((oa, ob))
        return out

    def fitch(self, a, b):
        out, added = [], []
        for x, y in zip(a, b):
            m, c = po.nonadd_median(x, y)
            out.append(m)
            added.append(c)
        return np.array(out, np.uint8).reshape(len(a), -1), np.array(added, np.int64)

    def powell_median(self, triples):
        """align_3_powell_inter (src/sequence.ml:1089-1115) on the oracle:
        the restated search's alignment, the Three_D median of every
        column, gaps dropped, one leading gap."""
        cm3 = po.CM3(self.cm)
        mm, ge = self.cm.cost(1, 2), self.cm.cost(1, 16)
        go = self.cm.gap_open if self.affine else 0
        out = []
        for a, b, p in triples:
            cost, x, y, z = po.powell_ukk(a, b, p, mm, go, ge)
            med = [cm3.median(int(u), int(v), int(w)) for u, v, w in zip(x, y, z)]
            out.append((np.array([16] + [m for m in med if m != 16], np.uint8), cost))
        return out

    def distance(self, pairs):
        if self.affine:
            return [po.cost_affine_3(a, b, self.cm) for a, b in pairs]
        return [po.cost_2(a, b, self.cm, max(8, abs(len(a) - len(b)))) for a, b in pairs]


# the four scripts of the reference's tree-cost checks: cc (default DNA
# matrix), cc1 tcm:(1,1), cc2 tcm:(2,1) gap_opening:1, cc3 tcm:(3,1) gap_opening:5
MODELS = {
    "cc": (1, 2, None),
    "cc1": (1, 1, None),
    "cc2": (2, 1, 1),
    "cc3": (3, 1, 5),
}
TREE_SUFFIX = {"cc": ".tree", "cc1": ".1.tree", "cc2": ".2.tree", "cc3": ".3.tree"}


def model_cm(name: str) -> po.CM2:
    sub, indel, go = MODELS[name]
    return po.CM2.sub_indel(sub, indel, go)


def is_empty(s) -> bool:
    return bool((np.asarray(s) == 16).all())


class OracleStore:
    """The sequence-store interface of poyd_amd.seqstore on the CPU oracle:
    sequences in a host list, DOS.median / DOS.distance per pair."""

    def __init__(self, cm: po.CM2, affine: bool):
        self.cm, self.affine = cm, affine
        self.seqs = []
        self.alignments = 0
        self.cells = 0

    def __len__(self):
        return len(self.seqs)

    def add(self, seqs):
        first = len(self.seqs)
        self.seqs.extend(np.asarray(s, np.uint8) for s in seqs)
        return np.arange(first, len(self.seqs), dtype=np.int64)

    def get(self, i):
        return self.seqs[int(i)]

    def mark(self):
        return len(self.seqs)

    def release(self, m):
        del self.seqs[m:]

    def median(self, a, b):
        ids, costs = [], []
        for x, y in zip(np.asarray(a), np.asarray(b)):
            sx, sy = self.seqs[int(x)], self.seqs[int(y)]
            if is_empty(sx):
                ids.append(int(y))
                costs.append(0)
                continue
            if is_empty(sy):
                ids.append(int(x))
                costs.append(0)
                continue
            if self.affine:
                med, ri, rj, cost, mwg = po.align_affine_3(sx, sy, self.cm)
            else:
                cost, oa, ob = po.align_2(sx, sy, self.cm)
                med = po.ancestor_2(oa, ob, self.cm)
            self.seqs.append(med)
            ids.append(len(self.seqs) - 1)
            costs.append(cost)
            self.alignments += 1
        return np.array(ids, np.int64), np.array(costs, np.int64)

    def distance(self, a, b):
        out = []
        for x, y in zip(np.asarray(a), np.asarray(b)):
            sx, sy = self.seqs[int(x)], self.seqs[int(y)]
            if is_empty(sx) or is_empty(sy):
                out.append(0)
            elif self.affine:
                out.append(po.cost_affine_3(sx, sy, self.cm))
            else:
                out.append(po.cost_2(sx, sy, self.cm, max(8, abs(len(sx) - len(sy)))))
            self.alignments += 1
        return np.array(out, np.int64)
