"""Copies the reference's tree-cost fixtures (data files of its test suite)
into tests/golden/tree_costs/: the FASTA inputs test/1.fas .. 13.fas
(gzip-compressed), the trees test/N.fas{,.1,.2,.3}.tree and the expected
costs test/cc{,1,2,3}.costs.  The reference's checks (test/cost_tests,
lines 11-14) run test/cc*.poy over test/first_fasta: read the FASTA, apply
the script's transform, read the tree, and compare the Adjusted tree cost.

Run in the container that holds /root/reference:
    python tests/golden/make_tree_fixtures.py
"""
import gzip
import json
import os
import shutil

SRC = "/root/reference/test"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tree_costs")


def main():
    os.makedirs(DST, exist_ok=True)
    files = open(os.path.join(SRC, "first_fasta")).read().split()
    costs = {}
    for script in ("cc", "cc1", "cc2", "cc3"):
        costs[script] = [int(x) for x in open(os.path.join(SRC, script + ".costs")).read().split()]
    for name in files:
        with open(os.path.join(SRC, name), "rb") as f, gzip.open(os.path.join(DST, name + ".gz"), "wb", 9) as g:
            g.write(f.read())
        for suffix in (".tree", ".1.tree", ".2.tree", ".3.tree"):
            shutil.copyfile(os.path.join(SRC, name + suffix), os.path.join(DST, name + suffix))
    with open(os.path.join(DST, "costs.json"), "w") as f:
        json.dump({"files": files, "costs": costs,
                   "scripts": {"cc": "read; default DNA matrix (sub 1, indel 2)",
                               "cc1": "transform (tcm:(1,1))",
                               "cc2": "transform (tcm:(2,1), gap_opening:1)",
                               "cc3": "transform (tcm:(3,1), gap_opening:5)"}}, f, indent=1)


if __name__ == "__main__":
    main()
