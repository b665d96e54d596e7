"""The C-ABI library builds, loads without a GPU and exports every entry
point include/poyg.h declares; the host-side cost matrices equal the
oracle's (no device work is done here)."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "poyg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(poyg_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_symbols_exported():
    from poyd_amd import engine
    lib = engine.load()
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from poyd_amd import engine
    names = set(declared_functions())
    assert set(engine.SIGNATURES) <= names | {"poyg_plan_status"}


@pytest.mark.parametrize("sub,indel,go", [(1, 2, None), (1, 1, None), (2, 1, 1), (3, 1, 5), (1, 1, 3)])
def test_engine_tcm_equals_oracle(oracle, sub, indel, go):
    from poyd_amd import engine
    cm = engine.CostMatrix.of_sub_indel(sub, indel, go)
    ref = oracle.CM2.sub_indel(sub, indel, go)
    for a in range(1, 32):
        assert cm.prepend(a) == ref.cost(16, a) and cm.tail(a) == ref.cost(a, 16)
        for b in range(1, 32):
            assert cm.cost(a, b) == ref.cost(a, b)
            assert cm.median(a, b) == ref.median(a, b)
            assert cm.worst(a, b) == ref.worst(a, b)


def test_engine_tcm3_equals_oracle(oracle):
    from poyd_amd import engine
    cm3 = engine.CostMatrix3(engine.CostMatrix.of_sub_indel(1, 2))
    ref = oracle.CM3(oracle.CM2.sub_indel(1, 2))
    rng = np.random.default_rng(0)
    for _ in range(2000):
        a, b, c = (int(x) for x in rng.integers(1, 32, 3))
        assert cm3.cost(a, b, c) == ref.cost(a, b, c)
        assert cm3.median(a, b, c) == ref.median(a, b, c)


def test_engine_deltaw_equals_oracle(oracle):
    from poyd_amd import engine
    cm = oracle.CM2.sub_indel(1, 2)
    rng = np.random.default_rng(1)
    for _ in range(200):
        a = np.concatenate([[16], rng.integers(1, 32, int(rng.integers(0, 400)))]).astype(np.uint8)
        b = np.concatenate([[16], rng.integers(1, 32, int(rng.integers(0, 400)))]).astype(np.uint8)
        assert engine.deltaw(a, b) == oracle.deltaw(a, b, cm)
        assert engine.deltaw(a, b, 7) == oracle.deltaw(a, b, cm, 7)


def test_batch_deltaw_equals_oracle(oracle):
    """The vectorised band parameter of a batch, with and without ?deltaw."""
    from poyd_amd import engine, synth
    cm = oracle.CM2.sub_indel(1, 2)
    rng = np.random.default_rng(2)

    def batch(seqs):
        ln = np.array([len(s) for s in seqs], np.int32)
        off = np.concatenate([[0], np.cumsum(ln[:-1])]).astype(np.int64)
        return synth.SeqBatch(np.concatenate(seqs).astype(np.uint8), off, ln)

    A, B = [], []
    for _ in range(150):
        A.append(np.concatenate([[16], rng.integers(1, 32, int(rng.integers(0, 300)))]).astype(np.uint8))
        B.append(np.concatenate([[16], rng.integers(1, 32, int(rng.integers(0, 300)))]).astype(np.uint8))
    dws = rng.integers(0, 40, len(A))
    got = engine.batch_deltaw(batch(A), batch(B))
    got_dw = engine.batch_deltaw(batch(A), batch(B), dws)
    for k, (a, b) in enumerate(zip(A, B)):
        assert got[k] == oracle.deltaw(a, b, cm), k
        assert got_dw[k] == oracle.deltaw(a, b, cm, int(dws[k])), k


def test_host_column_rules_equal_oracle(oracle):
    """median_2 / with_gaps / ancestor_2 / verify_2 / worst_2 are host code
    in the engine; they run without a GPU."""
    import ctypes as C
    from poyd_amd import engine
    L = engine.load()
    rng = np.random.default_rng(2)
    for go in (None, 3):
        cm = engine.CostMatrix.of_sub_indel(1, 1, go)
        ref = oracle.CM2.sub_indel(1, 1, go)
        for _ in range(50):
            n = int(rng.integers(1, 60))
            a = rng.integers(1, 32, n).astype(np.uint8)
            b = rng.integers(1, 32, n).astype(np.uint8)
            a[0] = b[0] = 16
            out = np.zeros(n + 1, np.uint8)
            ln = C.c_int32(0)
            for fn, rfn in ((L.poyg_algn_median_2_no_gaps, oracle.median_2),
                            (L.poyg_algn_median_2_with_gaps, oracle.median_2_with_gaps),
                            (L.poyg_algn_ancestor_2, oracle.ancestor_2)):
                assert fn(cm.h, engine._p8(a), engine._p8(b), n, engine._p8(out), n + 1, C.byref(ln)) == 0
                assert np.array_equal(out[n + 1 - ln.value:], rfn(a, b, ref))
            v = C.c_int32(0)
            L.poyg_algn_verify_2(cm.h, engine._p8(a), engine._p8(b), n, C.byref(v))
            assert v.value == oracle.verify_2(a, b, ref)
            L.poyg_algn_worst_2(cm.h, engine._p8(a), engine._p8(b), n, C.byref(v))
            assert v.value == oracle.worst_2(a, b, ref)


def be32(b, off):
    return int.from_bytes(b[off:off + 4], "big", signed=True)


@pytest.mark.parametrize("sub,indel,go", [(1, 2, None), (1, 1, 3), (3, 1, 5)])
def test_cm_wire_format(sub, indel, go):
    # the byte layout of cm_CAML_serialize (src/cm.c:909-932): eight 4-byte
    # big-endian fields, then cost, median (bytes), worst, prepend, tail over
    # 2 * 32 * 32 entries indexed (a << 5) + b
    from poyd_amd import engine
    cm = engine.CostMatrix.of_sub_indel(sub, indel, go)
    b = cm.serialize()
    L = 2 * 32 * 32
    assert len(b) == 32 + L * 17
    lib = engine.load()
    assert [be32(b, 4 * k) for k in range(8)][:3] == [31, 5, 16]
    assert be32(b, 12) == lib.poyg_cm_get_affine(cm.h)
    assert be32(b, 20) == lib.poyg_cm_get_gap_opening(cm.h)
    for a in range(1, 32):
        for c in range(1, 32):
            k = (a << 5) + c
            assert be32(b, 32 + 4 * k) == cm.cost(a, c)
            assert b[32 + 4 * L + k] == cm.median(a, c)
            assert be32(b, 32 + 5 * L + 4 * k) == cm.worst(a, c)
        assert be32(b, 32 + 9 * L + 4 * a) == cm.prepend(a)
        assert be32(b, 32 + 13 * L + 4 * a) == cm.tail(a)
    back = engine.CostMatrix.deserialize(b)
    assert back.compare(cm) == 0 and cm.compare(back) == 0
    assert back.serialize() == b


def test_cm_compare_order_and_errors():
    from poyd_amd import engine
    a = engine.CostMatrix.of_sub_indel(1, 1, 3)
    b = engine.CostMatrix.of_sub_indel(1, 1, 5)
    c = engine.CostMatrix.of_sub_indel(1, 2)
    # fields first (gap_open, then the model), then the tables
    assert a.compare(b) < 0 < b.compare(a)
    assert a.compare(c) != 0 and (a.compare(c) > 0) == (c.compare(a) < 0)
    d = engine.CostMatrix.of_sub_indel(1, 3)
    assert (c.compare(d) < 0) == (c.cost(1, 16) < d.cost(1, 16))
    blob = a.serialize()
    with pytest.raises(engine.PoygError):
        engine.CostMatrix.deserialize(blob[:100])
    bad = bytearray(blob)
    bad[4:8] = (7).to_bytes(4, "big")
    with pytest.raises(engine.PoygError):
        engine.CostMatrix.deserialize(bytes(bad))


def test_cm3_wire_format():
    # cm_CAML_serialize_3d (src/cm.c:934-952): seven fields, then cost and
    # median over 64 * 64 * 64 entries indexed ((a << 5) + b) << 5 + c
    from poyd_amd import engine
    cm3 = engine.CostMatrix3(engine.CostMatrix.of_sub_indel(1, 2))
    b = cm3.serialize()
    L = 64 ** 3
    assert len(b) == 28 + 5 * L
    assert [be32(b, 4 * k) for k in range(3)] == [31, 5, 16]
    for (x, y, z) in [(1, 2, 4), (16, 1, 1), (3, 5, 16), (31, 31, 31), (8, 8, 8)]:
        k = (((x << 5) + y) << 5) + z
        assert be32(b, 28 + 4 * k) == cm3.cost(x, y, z)
        assert b[28 + 4 * L + k] == cm3.median(x, y, z)
    back = engine.CostMatrix3.deserialize(b)
    assert back.serialize() == b
