"""Loader for the reference's tree-cost fixtures kept in tests/golden/tree_costs."""
from __future__ import annotations

import gzip
import json
import os

from poyd_amd import treedata

DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tree_costs")
SUFFIX = {"cc": ".tree", "cc1": ".1.tree", "cc2": ".2.tree", "cc3": ".3.tree"}


def index():
    with open(os.path.join(DIR, "costs.json")) as f:
        return json.load(f)


def cases():
    idx = index()
    return [(script, name, idx["costs"][script][k])
            for script in ("cc", "cc1", "cc2", "cc3")
            for k, name in enumerate(idx["files"])]


_DS = {}


def dataset(name: str) -> treedata.Dataset:
    if name not in _DS:
        with gzip.open(os.path.join(DIR, name + ".gz"), "rt") as f:
            _DS[name] = treedata.read_fasta(f.read())
    return _DS[name]


def tree(script: str, name: str) -> treedata.Tree:
    with open(os.path.join(DIR, name + SUFFIX[script])) as f:
        trees = treedata.parse_newick(f.read())
    return treedata.build_tree(trees[0], dataset(name))
