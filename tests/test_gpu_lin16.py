"""The packed 16-bit linear fill (lin_cost16_kernel), which takes every
linear pair whose shorter sequence has 2 .. 2,048 positions and whose costs
stay below 0xffff, against the oracle's cost_2 / align_2
(Sequence.Align.cost_2 and align_2, src/sequence.ml:690-723, 1040-1060)
and against the 32-bit systolic fill (tuning knob lin_cost_kernel = 1);
with POYG_MODE_LINEAR_ALIGN it also writes the direction band that the
backtrace reads (at its own row phase), so the aligned pairs and medians
are compared too:

* every band case of algn_fill_plane_2 (full plane, the two banded cases
  with their three row sections), lengths 1 .. 2,048 on the short side,
  default and explicit ?deltaw, DNA, ambiguity codes and gap-bearing
  characters, three matrices;
* strips: row counts around multiples of 128 (the rows of one wave pass),
  so the last strip holds 1, 2, 127 or 128 rows;
* a batch of 3,000 pairs of 500-2,000 positions (the TBR round's
  distances), every cost equal to the systolic fill's;
* pairs whose costs could reach 0xffff go to the systolic fill, and a batch
  mixing both kinds still equals the oracle."""
import numpy as np
import pytest

from test_gpu_parity import GAP, batch_of, mutate, note, rand_seq

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from poyd_amd import engine
    try:
        return engine.Context(0)
    except engine.PoygError as e:
        pytest.skip(f"no GPU: {e}")


def with_kernels(ctx, run):
    """run() under the packed fill's choice (default), one pair per wave (3),
    four waves per pair (4) and the systolic fill only (1)."""
    out = {}
    for knob in (-1, 3, 4, 1):
        ctx.set_tuning("lin_cost_kernel", knob)
        try:
            out[knob] = run()
        finally:
            ctx.set_tuning("lin_cost_kernel", -1)
    return out


def both_kernels(ctx, cm, A, B, dws=None):
    """Costs of a LINEAR_COST batch with the packed fill (default; the one
    pair per wave and four waves per pair forms give the same costs) and with
    the systolic fill only."""
    from poyd_amd import engine
    r = with_kernels(ctx, lambda: engine.run_pairs(ctx, cm, engine.POYG_MODE_LINEAR_COST, batch_of(A), batch_of(B),
                                                   dws).cost.copy())
    assert np.array_equal(r[3], r[-1]) and np.array_equal(r[4], r[-1]), \
        np.nonzero((r[3] != r[-1]) | (r[4] != r[-1]))[0][:10]
    return r[-1], r[1]


def shape_pairs(rng, alphabet):
    A, B = [], []
    shapes = [(1, 1), (2, 1), (1, 2), (2, 2), (3, 2), (5, 5), (40, 39), (60, 8), (8, 60), (129, 128),
              (130, 129), (256, 200), (257, 257), (383, 380), (384, 384), (385, 100), (900, 300), (300, 900),
              (1000, 999), (2047, 2047), (2048, 2047), (2600, 2048), (2048, 1400), (4000, 2048), (3000, 700)]
    for la, lb in shapes:
        a = rand_seq(rng, la - 1, alphabet)
        b = mutate(rng, a, 0.15, 0.06)[:lb] if lb <= la and alphabet == "dna" else rand_seq(rng, lb - 1, alphabet)
        A.append(a)
        B.append(b)
    for k in range(160):
        la = int(rng.integers(2, 900))
        a = rand_seq(rng, la - 1, alphabet)
        if k % 3 == 0:
            b = rand_seq(rng, int(rng.integers(1, 900)), alphabet)
        else:
            b = mutate(rng, a, 0.2, 0.1)
        A.append(a)
        B.append(b)
    return A, B


@pytest.mark.parametrize("sub,indel", [(1, 2), (1, 1), (3, 2)])
@pytest.mark.parametrize("alphabet", ["dna", "ambig", "any"])
def test_packed_fill_every_band_case(ctx, oracle, sub, indel, alphabet):
    from poyd_amd import engine
    note(f"lin16 shapes sub={sub} indel={indel} alphabet={alphabet}")
    rng = np.random.default_rng(1000 + 10 * sub + indel + {"dna": 0, "ambig": 100, "any": 200}[alphabet])
    A, B = shape_pairs(rng, alphabet)
    cm = engine.CostMatrix.of_sub_indel(sub, indel)
    ref = oracle.CM2.sub_indel(sub, indel)
    packed, systolic = both_kernels(ctx, cm, A, B)
    for k, (a, b) in enumerate(zip(A, B)):
        want = oracle.cost_2(a, b, ref)
        assert packed[k] == want, (k, len(a), len(b), packed[k], want)
        assert systolic[k] == want, (k, len(a), len(b))


def test_packed_fill_explicit_deltaw(ctx, oracle):
    from poyd_amd import engine
    note("lin16 explicit deltaw")
    rng = np.random.default_rng(17)
    A, B = shape_pairs(rng, "dna")
    cm = engine.CostMatrix.of_sub_indel(1, 2)
    ref = oracle.CM2.sub_indel(1, 2)
    dws = [int(max(8, abs(len(a) - len(b)))) for a, b in zip(A, B)]
    bands = np.array([engine.deltaw(a, b, d) for a, b, d in zip(A, B, dws)], np.int32)
    packed, systolic = both_kernels(ctx, cm, A, B, bands)
    for k, (a, b) in enumerate(zip(A, B)):
        want = oracle.cost_2(a, b, ref, dws[k])
        assert packed[k] == want and systolic[k] == want, (k, len(a), len(b))


def test_packed_fill_strip_boundaries(ctx, oracle):
    """Row counts around multiples of the 128 rows of a wave's strip."""
    from poyd_amd import engine
    note("lin16 strip boundaries")
    rng = np.random.default_rng(23)
    A, B = [], []
    for rows in (128, 129, 130, 255, 256, 257, 258, 383, 384, 385, 511, 512, 513, 1025):
        for cols in (2, 64, rows - 30, rows):
            if cols < 2:
                continue
            a = rand_seq(rng, rows - 1)
            b = mutate(rng, a, 0.1, 0.05)[:cols] if cols < rows else rand_seq(rng, cols - 1)
            A.append(a)
            B.append(b)
    cm = engine.CostMatrix.of_sub_indel(1, 2)
    ref = oracle.CM2.sub_indel(1, 2)
    packed, systolic = both_kernels(ctx, cm, A, B)
    for k, (a, b) in enumerate(zip(A, B)):
        want = oracle.cost_2(a, b, ref)
        assert packed[k] == want and systolic[k] == want, (k, len(a), len(b))


def test_packed_fill_tbr_batch(ctx, oracle):
    """3,000 pairs of 500-2,000 positions with DOS.distance's band, as the
    TBR round's join distances: identical to the systolic fill, and to the
    oracle on a sample."""
    from poyd_amd import engine, synth
    note("lin16 TBR-size batch")
    w = synth.pairs_uniform(3000, 500, 2000, 0.12, 0.04, 31, "lin16-tbr")
    A = [w.a.get(k) for k in range(3000)]
    B = [w.b.get(k) for k in range(3000)]
    dws = [int(max(8, abs(len(a) - len(b)))) for a, b in zip(A, B)]
    bands = np.array([engine.deltaw(a, b, d) for a, b, d in zip(A, B, dws)], np.int32)
    cm = engine.CostMatrix.of_sub_indel(1, 2)
    ref = oracle.CM2.sub_indel(1, 2)
    packed, systolic = both_kernels(ctx, cm, A, B, bands)
    assert np.array_equal(packed, systolic)
    for k in np.random.default_rng(3).choice(3000, 30, replace=False):
        assert packed[k] == oracle.cost_2(A[k], B[k], ref, dws[k]), k


def test_costs_beyond_16_bits_use_the_systolic_fill(ctx, oracle):
    from poyd_amd import engine
    note("lin16 cost range")
    rng = np.random.default_rng(29)
    A, B = [], []
    for la, lb in [(1500, 1400), (300, 200), (2000, 2000), (50, 40)]:
        a = rand_seq(rng, la - 1)
        A.append(a)
        B.append(rand_seq(rng, lb - 1))
    cm = engine.CostMatrix.of_sub_indel(120, 100)
    ref = oracle.CM2.sub_indel(120, 100)
    packed, systolic = both_kernels(ctx, cm, A, B)
    for k, (a, b) in enumerate(zip(A, B)):
        want = oracle.cost_2(a, b, ref)
        assert packed[k] == want and systolic[k] == want, (k, len(a), len(b), want)
    assert max(packed) > 0xffff


def both_align(ctx, cm, A, B):
    """LINEAR_ALIGN results of the packed fill (default; its two forms give
    the same bytes) and of the systolic fill only."""
    from poyd_amd import engine
    r = with_kernels(ctx, lambda: engine.run_pairs(ctx, cm, engine.POYG_MODE_LINEAR_ALIGN, batch_of(A), batch_of(B)))
    for knob in (3, 4):
        assert np.array_equal(r[knob].cost, r[-1].cost), knob
        for k in range(len(A)):
            for slot in range(3):
                assert np.array_equal(r[knob].seq(k, slot), r[-1].seq(k, slot)), (knob, k, slot)
    return r[-1], r[1]


@pytest.mark.parametrize("sub,indel,alphabet", [(1, 2, "dna"), (1, 1, "ambig"), (3, 2, "any")])
def test_packed_align_equals_oracle(ctx, oracle, sub, indel, alphabet):
    """Aligned pairs and medians from the packed fill's direction band."""
    from poyd_amd import engine
    note(f"lin16 align sub={sub} indel={indel} alphabet={alphabet}")
    rng = np.random.default_rng(3000 + 10 * sub + indel)
    A, B = shape_pairs(rng, alphabet)
    for rows in (128, 129, 256, 257, 385, 513):
        a = rand_seq(rng, rows - 1, alphabet)
        A.append(a)
        B.append(mutate(rng, a, 0.1, 0.05) if alphabet == "dna" else rand_seq(rng, rows - 40, alphabet))
    cm = engine.CostMatrix.of_sub_indel(sub, indel)
    ref = oracle.CM2.sub_indel(sub, indel)
    packed, systolic = both_align(ctx, cm, A, B)
    for k, (a, b) in enumerate(zip(A, B)):
        cost, oa, ob = oracle.align_2(a, b, ref)
        for res in (packed, systolic):
            assert res.cost[k] == cost, (k, len(a), len(b))
            assert np.array_equal(res.seq(k, 0), oa), (k, len(a), len(b))
            assert np.array_equal(res.seq(k, 1), ob), (k, len(a), len(b))
            assert np.array_equal(res.seq(k, 2), oracle.median_2(oa, ob, ref)), k


def test_packed_align_large_batch(ctx, oracle):
    """2,000 pairs of 200-1,500 positions: every output byte of the packed
    fill's backtrace equals the systolic fill's, and the oracle's on a
    sample."""
    from poyd_amd import engine, synth
    note("lin16 align batch")
    w = synth.pairs_uniform(2000, 200, 1500, 0.12, 0.04, 37, "lin16-align")
    A = [w.a.get(k) for k in range(2000)]
    B = [w.b.get(k) for k in range(2000)]
    cm = engine.CostMatrix.of_sub_indel(1, 2)
    ref = oracle.CM2.sub_indel(1, 2)
    packed, systolic = both_align(ctx, cm, A, B)
    assert np.array_equal(packed.cost, systolic.cost)
    for k in range(2000):
        for slot in range(3):
            assert np.array_equal(packed.seq(k, slot), systolic.seq(k, slot)), (k, slot)
    for k in np.random.default_rng(5).choice(2000, 20, replace=False):
        cost, oa, ob = oracle.align_2(A[k], B[k], ref)
        assert packed.cost[k] == cost and np.array_equal(packed.seq(k, 0), oa), k


@pytest.mark.parametrize("kernel", [-1, 3, 1])
def test_direction_matrix_equals_oracle(ctx, oracle, kernel):
    """poyg_algn_direction_matrix (the Matrix.default readers' view) equals
    the oracle's algn_nw direction matrix cell for cell, for the packed and
    the systolic fill and for the row kernels (s2 longer than 2,048)."""
    from poyd_amd import engine
    note(f"direction matrix kernel={kernel}")
    rng = np.random.default_rng(41)
    cm = engine.CostMatrix.of_sub_indel(1, 2)
    ref = oracle.CM2.sub_indel(1, 2)
    al = engine.Align(ctx)
    ctx.set_tuning("lin_cost_kernel", kernel)
    try:
        for la, lb in [(2, 2), (60, 40), (130, 129), (300, 120), (700, 650), (1500, 1490), (2600, 2300)]:
            a = rand_seq(rng, la - 1)
            b = mutate(rng, a, 0.1, 0.05)[:lb] if lb < la else rand_seq(rng, lb - 1)
            if len(b) > len(a):
                a, b = b, a
            dw = engine.deltaw(a, b)
            cost, dm = al.direction_matrix(a, b, cm, dw)
            wcost, wdm, lo, hi = oracle.algn_nw(a, b, ref, dw, with_dm=True)
            assert cost == wcost, (la, lb)
            assert np.array_equal(dm, wdm), (la, lb, np.argwhere(dm != wdm)[:5])
    finally:
        ctx.set_tuning("lin_cost_kernel", -1)
