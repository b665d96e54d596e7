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
def workload_c1(n: int = 1000) -> PairWorkload:
    """1,000 pairs, L ~ U{200..400}, 10% substitutions, 2% indels."""
    return pairs_uniform(n, 200, 400, 0.10, 0.02, SEED0 + 0, "pairwise-linear-1000x200-400")


def workload_c2(n: int = 65536) -> PairWorkload:
    """65,536 pairs, L log-uniform on [50, 5000], divergence U[5%, 30%]."""
    return pairs_loguniform(n, 50, 5000, 0.05, 0.30, SEED0 + 1, "pairwise-affine-65536x50-5000")


def workload_c3(n: int = 4096):
    """4,096 triples, L ~ U{100..500}, 15% substitutions, 3% indels."""
    return triples_uniform(n, 100, 500, 0.15, 0.03, SEED0 + 2, "three-way-4096x100-500")
