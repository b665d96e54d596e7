"""Input side of a tree evaluation: FASTA characters, newick trees and the
vertex codes POY gives them.

* FASTA (Parser.Fasta, src/parser.ml:219-372): a record starts at a line that
  begins with '>' or follows an empty line; '@', '|' and '#' split a record
  into chromosomes, loci and fragments, and every piece becomes its own
  dynamic character (Data.process_parsed_sequences flattens them,
  src/data.ml:1047-1101).  Whitespace is skipped, '-' is dropped (the file is
  not prealigned) and the leading gap is prepended (process_sequence,
  src/parser.ml:278-292).
* Taxon codes are 1..N in file order (Data.process_taxon_code,
  src/data.ml:860-891).
* Trees (Tree.convert_to / add_tree_to, src/tree.ml:722-866): interior codes
  come from N+1..2N in pre-order, the root takes the first code and is then
  dissolved into an edge between its two children; the first child is the
  component's handle.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

GAP = 16

# Alphabet.nucleotides (src/alphabet.ml:139-186): 5-bit masks, A C G T gap.
IUPAC = {
    "A": 1, "C": 2, "G": 4, "T": 8, "U": 8,
    "M": 3, "R": 5, "W": 9, "S": 6, "Y": 10, "K": 12,
    "V": 7, "H": 11, "D": 13, "B": 14, "N": 15, "X": 15,
    "-": 16,
    "1": 17, "2": 18, "3": 19, "4": 20, "5": 21, "6": 22, "7": 23, "8": 24, "9": 25,
    "0": 26, "!": 27, "^": 28, "$": 29, "*": 31, "?": 31,
}
_LUT = np.full(256, 255, np.uint8)
for _k, _v in IUPAC.items():
    _LUT[ord(_k)] = _v
    _LUT[ord(_k.lower())] = _v


def encode(text: str) -> np.ndarray:
    """One fragment's text -> u8 masks with POY's leading gap; gaps in the
    text are removed (the file is read as unaligned)."""
    raw = np.frombuffer(text.encode("latin-1"), np.uint8)
    raw = raw[raw > 32]
    codes = _LUT[raw]
    if (codes == 255).any():
        bad = chr(int(raw[np.argmax(codes == 255)]))
        raise ValueError(f"illegal character {bad!r} in a molecular sequence")
    codes = codes[codes != GAP]
    return np.concatenate([np.array([GAP], np.uint8), codes])


@dataclass
class Dataset:
    """Taxa (codes 1..N) and their dynamic characters: chars[c][t-1] is the
    sequence of taxon code t for character c."""

    names: list[str]
    chars: list[list[np.ndarray]]
    # nonadditive characters (NonaddCS8): nonadd[t-1] holds taxon t's state
    # sets, one bit per state, one byte per character
    nonadd: np.ndarray | None = None

    @property
    def n_taxa(self) -> int:
        return len(self.names)

    def code(self, name: str) -> int:
        return self.names.index(name) + 1


def _split(lines: list[str], sep: str) -> list[list[str]]:
    pieces: list[list[str]] = [[]]
    for line in lines:
        parts = line.split(sep)
        pieces[-1].append(parts[0])
        for p in parts[1:]:
            pieces.append([p])
    return pieces


def read_fasta(text: str) -> Dataset:
    records: list[tuple[str, list[str]]] = []
    taxon, seq, last_empty = "", [], True
    for line in text.splitlines():
        if line.strip() == "":
            last_empty = True
            continue
        if line.startswith(">") or last_empty:
            if taxon:
                records.append((taxon, seq))
            taxon = (line[1:] if line.startswith(">") else line).strip()
            seq, last_empty = [], False
        else:
            seq.append(line)
    if taxon:
        records.append((taxon, seq))
    per_taxon = []
    for name, lines in records:
        frags = []
        for chrom in _split(lines, "@"):
            for locus in _split(chrom, "|"):
                for frag in _split(locus, "#"):
                    frags.append(encode("".join(frag)))
        per_taxon.append((name, frags))
    nchar = len(per_taxon[0][1]) if per_taxon else 0
    for name, frags in per_taxon:
        if len(frags) != nchar:
            raise ValueError(f"sequence {name} has an inconsistent number of fragments "
                             f"({len(frags)}, expected {nchar})")
    names = [n for n, _ in per_taxon]
    if len(set(names)) != len(names):
        raise ValueError("duplicate taxon names")
    chars = [[frags[c] for _, frags in per_taxon] for c in range(nchar)]
    return Dataset(names, chars)


# ---------------------------------------------------------------- newick

def parse_newick(text: str):
    """POY tree files: '(a (b c))' with blanks or commas between members,
    an optional '[...]' annotation and ';' or '*' between trees.  Returns a
    list of nested tuples/str."""
    trees, stack, tok, i = [], [], "", 0

    def flush():
        nonlocal tok
        if tok:
            if not stack:
                trees.append(tok)
            else:
                stack[-1].append(tok)
            tok = ""

    while i < len(text):
        ch = text[i]
        if ch == "[":
            flush()
            j = text.find("]", i)
            i = len(text) if j < 0 else j + 1
            continue
        if ch == "(":
            flush()
            stack.append([])
        elif ch == ")":
            flush()
            node = tuple(stack.pop())
            if stack:
                stack[-1].append(node)
            else:
                trees.append(node)
        elif ch in " \t\r\n,;*":
            flush()
        elif ch == ":":
            flush()
            while i + 1 < len(text) and text[i + 1] not in " \t\r\n,;()[]":
                i += 1
        else:
            tok += ch
        i += 1
    flush()
    return trees


@dataclass
class Tree:
    """An unrooted binary tree with POY's vertex codes.  nbrs[v] lists the
    neighbours of v (one for a leaf, three for an interior vertex, in the
    order (parent, child a, child b) of Tree.Interior)."""

    nbrs: dict[int, list[int]] = field(default_factory=dict)
    handle: int = 0
    leaves: set[int] = field(default_factory=set)

    def edges(self) -> list[tuple[int, int]]:
        out = []
        for v, ns in self.nbrs.items():
            for u in ns:
                if v < u:
                    out.append((v, u))
        return out

    def pre_order_edges(self) -> list[tuple[int, int]]:
        """Edge(pred, nd) for every edge, oriented away from the handle
        (Tree.pre_order_edge_visit, src/tree.ml:1347-1420)."""
        out, stack = [], [(self.handle, u) for u in self.nbrs[self.handle]]
        while stack:
            p, v = stack.pop()
            out.append((p, v))
            stack.extend((v, u) for u in self.nbrs[v] if u != p)
        return out

    def others(self, v: int, u: int) -> list[int]:
        return [w for w in self.nbrs[v] if w != u]


def build_tree(newick, ds: Dataset) -> Tree:
    """Tree.add_tree_to on a fresh tree (src/tree.ml:722-836)."""
    n = ds.n_taxa
    avail = iter(range(n + 1, 2 * n + 1))
    t = Tree()

    def resolve(children):
        children = list(children)
        if len(children) < 2:
            raise ValueError("a tree vertex with fewer than two members")
        while len(children) > 2:
            children = [tuple(children[:2])] + children[2:]
        return children

    def assign(node, parent):
        if isinstance(node, str):
            code = ds.code(node)
            if code in t.leaves:
                raise ValueError(f"taxon {node} occurs twice in a tree")
            t.nbrs[code] = [parent]
            t.leaves.add(code)
            return code
        sc = next(avail)
        a, b = resolve(node)
        ca = assign(a, sc)
        cb = assign(b, sc)
        t.nbrs[sc] = [parent, ca, cb]
        return sc

    if isinstance(newick, str) or len(newick) == 1:
        raise ValueError("trees need more than two taxa")
    sc = next(avail)
    a, b = resolve(newick)
    ca = assign(a, sc)
    cb = assign(b, sc)
    t.nbrs[ca][0] = cb
    t.nbrs[cb][0] = ca
    t.handle = ca
    return t
