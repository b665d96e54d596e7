"""Seeded synthetic DNA workloads for the direct-optimization engine.

Sequences use POY's 5-bit masks (A=1, C=2, G=4, T=8, gap=16) and carry the
leading gap POY prepends to every sequence (src/parser.ml process_sequence).

Mutation model (applied to a root of length L, vectorised over a batch):
  * every root site is substituted with probability p_s; the new base is
    uniform over the three other bases;
  * every root site starts an indel with probability p_i; insertion and
    deletion are equally likely; the length is Geometric(0.5) on {1, 2, ...};
  * a deletion removes the sites [pos, pos+len) of the root (clipped to the
    sequence); an insertion puts len uniform bases before site pos.
RNG: numpy PCG64 seeded with 20261202 + workload index.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

BASES = np.array([1, 2, 4, 8], dtype=np.uint8)
GAP = 16
SEED0 = 20261202


@dataclass
class SeqBatch:
    """Concatenated sequences: seq[off[k] : off[k] + ln[k]] is sequence k."""

    seq: np.ndarray  # uint8
    off: np.ndarray  # int64
    ln: np.ndarray   # int32

    def get(self, k: int) -> np.ndarray:
        o = int(self.off[k])
        return self.seq[o:o + int(self.ln[k])]

    def __len__(self) -> int:
        return len(self.ln)


def _pack(parts_seq: np.ndarray, seg_len: np.ndarray) -> SeqBatch:
    """Prepend the leading gap to each segment of a concatenated array."""
    n = len(seg_len)
    ln = seg_len.astype(np.int32) + 1
    off = np.zeros(n, np.int64)
    if n > 1:
        off[1:] = np.cumsum(ln[:-1], dtype=np.int64)
    out = np.empty(int(ln.sum()), np.uint8)
    out[off] = GAP
    body_pos = np.ones(len(out), bool)
    body_pos[off] = False
    out[body_pos] = parts_seq
    return SeqBatch(out, off, ln)


def random_roots(rng: np.random.Generator, lengths: np.ndarray):
    """Uniform A/C/G/T bodies (no leading gap) as one concatenated index array."""
    total = int(lengths.sum())
    return rng.integers(0, 4, size=total, dtype=np.int8)


def mutate(rng: np.random.Generator, root_idx: np.ndarray, lengths: np.ndarray,
           p_sub: np.ndarray, p_indel: np.ndarray):
    """Mutate every segment of a concatenated base-index array.

    Returns (mutated_index_array, mutated_lengths)."""
    n = len(lengths)
    total = len(root_idx)
    seg = np.repeat(np.arange(n), lengths)
    seg_start = np.zeros(n, np.int64)
    if n > 1:
        seg_start[1:] = np.cumsum(lengths[:-1], dtype=np.int64)
    seg_end = seg_start + lengths
    ps = p_sub[seg]
    pi = p_indel[seg]
    out = root_idx.copy()
    sub = rng.random(total) < ps
    out[sub] = (out[sub] + rng.integers(1, 4, size=int(sub.sum()), dtype=np.int8)) % 4
    starts = np.nonzero(rng.random(total) < pi)[0]
    is_ins = rng.random(len(starts)) < 0.5
    glen = rng.geometric(0.5, size=len(starts)).astype(np.int64)
    # deletions: difference array over the global index space
    d_st = starts[~is_ins]
    d_en = np.minimum(d_st + glen[~is_ins], seg_end[seg[d_st]])
    diff = np.zeros(total + 1, np.int64)
    np.add.at(diff, d_st, 1)
    np.add.at(diff, d_en, -1)
    keep = np.cumsum(diff[:-1]) == 0
    # insertions before the chosen root sites
    i_pos = starts[is_ins]
    i_len = glen[is_ins]
    ins_at = np.repeat(i_pos, i_len)
    ins_val = rng.integers(0, 4, size=len(ins_at), dtype=np.int8)
    # assemble: kept root sites plus inserted bases, ordered by root position
    # (inserted bases sort before the site they precede)
    kept_pos = np.nonzero(keep)[0]
    key = np.concatenate([kept_pos * 2 + 1, ins_at * 2])
    val = np.concatenate([out[kept_pos], ins_val])
    order = np.argsort(key, kind="stable")
    mutated = val[order]
    seg_of = np.concatenate([seg[kept_pos], seg[ins_at]])[order]
    new_len = np.bincount(seg_of, minlength=n).astype(np.int64)
    return mutated, new_len


@dataclass
class PairWorkload:
    a: SeqBatch
    b: SeqBatch
    name: str


def pairs_uniform(n: int, lmin: int, lmax: int, p_sub: float, p_indel: float, seed: int,
                  name: str = "pairs") -> PairWorkload:
    """Root length uniform on {lmin..lmax}, partner by the mutation model."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lengths = rng.integers(lmin, lmax + 1, size=n).astype(np.int64)
    root = random_roots(rng, lengths)
    mut, mlen = mutate(rng, root, lengths, np.full(n, p_sub), np.full(n, p_indel))
    return PairWorkload(_pack(BASES[root], lengths), _pack(BASES[mut], mlen), name)


def pairs_loguniform(n: int, lmin: int, lmax: int, dmin: float, dmax: float, seed: int,
                     name: str = "pairs") -> PairWorkload:
    """Root length round(exp(U[ln lmin, ln lmax])), divergence d ~ U[dmin, dmax],
    p_sub = d, p_indel = d / 5."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lengths = np.rint(np.exp(rng.uniform(np.log(lmin), np.log(lmax), size=n))).astype(np.int64)
    d = rng.uniform(dmin, dmax, size=n)
    root = random_roots(rng, lengths)
    mut, mlen = mutate(rng, root, lengths, d, d / 5.0)
    return PairWorkload(_pack(BASES[root], lengths), _pack(BASES[mut], mlen), name)


@dataclass
class TripleWorkload:
    batches: list          # three SeqBatch, one per member of the triples
    name: str

    @property
    def ln(self):
        return [b.ln for b in self.batches]

    def get(self, q: int, k: int) -> np.ndarray:
        return self.batches[q].get(k)


def triples_uniform(n: int, lmin: int, lmax: int, p_sub: float, p_indel: float, seed: int,
                    name: str = "triples") -> TripleWorkload:
    """Ancestor length uniform on {lmin..lmax}; three independent descendants."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lengths = rng.integers(lmin, lmax + 1, size=n).astype(np.int64)
    root = random_roots(rng, lengths)
    out = []
    for _ in range(3):
        mut, mlen = mutate(rng, root, lengths, np.full(n, p_sub), np.full(n, p_indel))
        out.append(_pack(BASES[mut], mlen))
    return TripleWorkload(out, name)


# The workloads measured by bench.py and replayed by the parity tests.
def workload_linear_1000(n: int = 1000) -> PairWorkload:
    """1,000 pairs, L ~ U{200..400}, 10% substitutions, 2% indels."""
    return pairs_uniform(n, 200, 400, 0.10, 0.02, SEED0 + 0, "pairwise-linear-1000x200-400")


def workload_affine_65536(n: int = 65536) -> PairWorkload:
    """65,536 pairs, L log-uniform on [50, 5000], divergence U[5%, 30%]."""
    return pairs_loguniform(n, 50, 5000, 0.05, 0.30, SEED0 + 1, "pairwise-affine-65536x50-5000")


def workload_three_way_4096(n: int = 4096):
    """4,096 triples, L ~ U{100..500}, 15% substitutions, 3% indels."""
    return triples_uniform(n, 100, 500, 0.15, 0.03, SEED0 + 2, "three-way-4096x100-500")


# ---------------------------------------------------------------- trees
@dataclass
class TreeWorkload:
    """A simulated dataset: taxon codes 1..N, `loci[c][t-1]` the DNA
    sequence of taxon t for locus c (leading gap included), `nonadd[t-1]` its
    nonadditive characters as one-bit state sets (u8), and the simulated
    tree (for reference only: the build starts from the data)."""

    names: list
    loci: list
    nonadd: np.ndarray
    states: np.ndarray     # number of states of each nonadditive character
    name: str


def yule_tree(rng: np.random.Generator, n_taxa: int, mean_bl: float):
    """A Yule (pure-birth) rooted binary tree: start from two lineages below
    the root, split a uniformly chosen extant lineage until there are n_taxa.
    Returns (parent, bl, leaf_nodes): node 0 is the root, parent[v] < v,
    bl[v] the length of the edge above v (Exp with mean mean_bl)."""
    parent = [-1, 0, 0]
    extant = [1, 2]
    while len(extant) < n_taxa:
        k = int(rng.integers(len(extant)))
        v = extant[k]
        a, b = len(parent), len(parent) + 1
        parent += [v, v]
        extant[k] = a
        extant.append(b)
    parent = np.array(parent, np.int64)
    bl = rng.exponential(mean_bl, size=len(parent))
    bl[0] = 0.0
    return parent, bl, np.array(extant, np.int64)


def evolve_tree_seqs(rng, parent, bl, root_len):
    """Evolve one locus down a rooted tree (nodes in parent-before-child
    order) with the substitution/indel model, p_sub = bl, p_indel = bl / 5."""
    n = len(parent)
    seqs = [None] * n
    seqs[0] = random_roots(rng, np.array([root_len]))
    for v in range(1, n):
        par = seqs[parent[v]]
        mut, ml = mutate(rng, par, np.array([len(par)]), np.array([bl[v]]), np.array([bl[v] / 5.0]))
        seqs[v] = mut
    return seqs


def evolve_tree_mk(rng, parent, bl, states):
    """Mk evolution of characters with k states (k >= 2) down the tree: on an
    edge of length t a character changes with probability
    (k-1)/k * (1 - exp(-k t / (k-1))), to a uniform other state."""
    n, m = len(parent), len(states)
    out = np.zeros((n, m), np.int64)
    out[0] = rng.integers(0, states)
    k = states.astype(np.float64)
    for v in range(1, n):
        pc = (k - 1) / k * (1.0 - np.exp(-k * bl[v] / (k - 1)))
        ch = rng.random(m) < pc
        shift = rng.integers(1, states)
        out[v] = np.where(ch, (out[parent[v]] + shift) % states, out[parent[v]])
    return out


def tree_workload(n_taxa=128, n_loci=5, lmin=300, lmax=1500, n_nonadd=10000, kmin=2, kmax=8,
                  mean_bl=0.05, seed=SEED0 + 3, name=None) -> TreeWorkload:
    rng = np.random.Generator(np.random.PCG64(seed))
    parent, bl, leaves = yule_tree(rng, n_taxa, mean_bl)
    # taxon codes: a seeded permutation of the leaves (the addition order of the build)
    leaves = leaves[rng.permutation(len(leaves))]
    loci = []
    for _ in range(n_loci):
        L = int(rng.integers(lmin, lmax + 1))
        seqs = evolve_tree_seqs(rng, parent, bl, L)
        loci.append([np.concatenate([[GAP], BASES[seqs[v]]]).astype(np.uint8) for v in leaves])
    states = rng.integers(kmin, kmax + 1, size=n_nonadd)
    mk = evolve_tree_mk(rng, parent, bl, states)
    nonadd = (np.left_shift(1, mk[leaves])).astype(np.uint8)
    names = [f"t{k + 1}" for k in range(n_taxa)]
    return TreeWorkload(names, loci, nonadd, states,
                        name or f"tree-{n_taxa}taxa-{n_loci}loci-{lmin}-{lmax}-nonadd{n_nonadd}")


def workload_tree_128() -> TreeWorkload:
    """128-taxon Yule tree (branch lengths Exp, mean 0.05), 5 DNA loci of
    300-1,500 bp, 10,000 nonadditive characters of 2-8 states."""
    return tree_workload()
