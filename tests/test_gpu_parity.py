"""GPU parity: every engine path against the CPU oracle, bit for bit.

All calls go through the C ABI (libpoyg.so).  Inputs are seeded; each case
prints its name before running so that a failing run names it.
"""
import sys

import numpy as np
import pytest

from poyd_amd import synth

pytestmark = pytest.mark.gpu
GAP = 16


@pytest.fixture(scope="module")
def ctx():
    from poyd_amd import engine
    try:
        return engine.Context(0)
    except engine.PoygError as e:
        pytest.skip(f"no GPU: {e}")


def note(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def batch_of(seqs):
    ln = np.array([len(s) for s in seqs], np.int32)
    off = np.zeros(len(seqs), np.int64)
    if len(seqs) > 1:
        off[1:] = np.cumsum(ln[:-1])
    return synth.SeqBatch(np.concatenate(seqs).astype(np.uint8) if seqs else np.zeros(0, np.uint8), off, ln)


def rand_seq(rng, n, alphabet="dna"):
    if alphabet == "dna":
        body = rng.choice([1, 2, 4, 8], size=n)
    elif alphabet == "ambig":
        body = rng.integers(1, 16, size=n)
    else:   # any 5-bit mask, gap-bearing symbols included
        body = rng.integers(1, 32, size=n)
    return np.concatenate([[GAP], body]).astype(np.uint8)


def mutate(rng, s, ps, pi):
    out = [GAP]
    for x in s[1:]:
        r = rng.random()
        if r < pi / 2:
            continue
        if r < pi:
            out.append(int(rng.choice([1, 2, 4, 8])))
        out.append(int(rng.choice([1, 2, 4, 8])) if rng.random() < ps else int(x))
    return np.array(out, np.uint8)


def mixed_pairs(rng, n, alphabet="dna"):
    A, B = [], []
    edge = [(0, 0), (0, 5), (7, 0), (1, 1), (39, 40), (40, 40), (41, 41), (45, 120), (300, 12), (10, 700),
            (700, 10), (250, 250), (1, 300)]
    for la, lb in edge:
        A.append(rand_seq(rng, la, alphabet))
        B.append(rand_seq(rng, lb, alphabet))
    for k in range(n):
        L = int(rng.integers(20, 600))
        a = rand_seq(rng, L, alphabet)
        b = mutate(rng, a, 0.2, 0.08) if k % 4 else rand_seq(rng, int(rng.integers(1, 600)), alphabet)
        if alphabet != "dna" and k % 2:
            b = rand_seq(rng, len(b) - 1, alphabet)
        A.append(a)
        B.append(b)
    return A, B


AFFINE_MATRICES = [(1, 1, 3), (2, 1, 1), (3, 1, 5)]


def test_concurrent_copies_agree(ctx, oracle):
    """Thousands of copies of a few wide-band pairs in one batch: every copy
    must give the oracle's result (many waves resident per workgroup)."""
    from poyd_amd import engine
    note("concurrent copies")
    rng = np.random.default_rng(31)
    base = [(rand_seq(rng, 388), rand_seq(rng, 542)), (rand_seq(rng, 300), rand_seq(rng, 420))]
    a0 = rand_seq(rng, 900)
    base.append((a0, mutate(rng, a0, 0.2, 0.06)))
    A = [base[k % 3][0] for k in range(3000)]
    B = [base[k % 3][1] for k in range(3000)]
    cm = engine.CostMatrix.of_sub_indel(1, 1, 3)
    ref = oracle.CM2.sub_indel(1, 1, 3)
    res = engine.run_pairs(ctx, cm, engine.POYG_MODE_AFFINE_ALIGN, batch_of(A), batch_of(B))
    want = [oracle.align_affine_3(a, b, ref) for a, b in base]
    bad = [k for k in range(3000) if res.cost[k] != want[k % 3][3]]
    assert not bad, (len(bad), bad[:10])
    for k in range(0, 3000, 7):
        med, ri, rj, cost, mwg = want[k % 3]
        for slot, w in enumerate((med, mwg, ri, rj)):
            assert np.array_equal(res.seq(k, slot), w), (k, slot)
    lcm = engine.CostMatrix.of_sub_indel(1, 2)
    lref = oracle.CM2.sub_indel(1, 2)
    res = engine.run_pairs(ctx, lcm, engine.POYG_MODE_LINEAR_ALIGN, batch_of(A), batch_of(B))
    lw = [oracle.align_2(a, b, lref) for a, b in base]
    bad = [k for k in range(3000) if res.cost[k] != lw[k % 3][0]]
    assert not bad, (len(bad), bad[:10])


@pytest.mark.parametrize("sub,indel,go", AFFINE_MATRICES)
@pytest.mark.parametrize("alphabet", ["dna", "any"])
def test_affine_batch_matches_oracle(ctx, oracle, sub, indel, go, alphabet):
    from poyd_amd import engine
    note(f"affine batch sub={sub} indel={indel} go={go} alphabet={alphabet}")
    rng = np.random.default_rng(100 + sub * 10 + go)
    A, B = mixed_pairs(rng, 120, alphabet)
    cm = engine.CostMatrix.of_sub_indel(sub, indel, go)
    ref = oracle.CM2.sub_indel(sub, indel, go)
    res = engine.run_pairs(ctx, cm, engine.POYG_MODE_AFFINE_ALIGN, batch_of(A), batch_of(B))
    cres = engine.run_pairs(ctx, cm, engine.POYG_MODE_AFFINE_COST, batch_of(A), batch_of(B))
    for k, (a, b) in enumerate(zip(A, B)):
        med, ri, rj, cost, mwg = oracle.align_affine_3(a, b, ref)
        assert res.cost[k] == cost, (k, len(a), len(b))
        assert cres.cost[k] == oracle.cost_affine_3(a, b, ref), (k, len(a), len(b))
        for slot, want in enumerate((med, mwg, ri, rj)):
            got = res.seq(k, slot)
            assert np.array_equal(got, want), (k, slot, len(a), len(b))


def block_indel_pairs(rng):
    """Pairs whose alignment path crosses many band lanes: long sequences
    with block insertions and deletions (band widths of 40 to 250 diagonal
    pairs), long runs of one state, several 1,024-step output batches, and
    sequences where one side ends early (tail gaps)."""
    A, B = [], []
    for k, (L, ins, dele) in enumerate([(1500, 150, 0), (1500, 0, 140), (2200, 300, 0), (900, 60, 60),
                                         (3000, 420, 30), (1200, 5, 5), (2500, 0, 0), (400, 200, 0),
                                         (1800, 90, 90), (700, 0, 380)]):
        a = rand_seq(rng, L)
        body = list(a[1:])
        if dele:
            at = int(rng.integers(0, L - dele))
            del body[at:at + dele]
        if ins:
            at = int(rng.integers(0, len(body)))
            body[at:at] = list(rng.choice([1, 2, 4, 8], size=ins))
        b = mutate(rng, np.array([GAP] + body, np.uint8), 0.1, 0.02)
        A.append(a)
        B.append(b)
        A.append(b)   # the same pair with the sides swapped
        B.append(a)
    return A, B


@pytest.mark.parametrize("sub,indel,go", AFFINE_MATRICES)
def test_affine_block_indels(ctx, oracle, sub, indel, go):
    from poyd_amd import engine
    note(f"affine block indels sub={sub} indel={indel} go={go}")
    rng = np.random.default_rng(700 + sub * 10 + go)
    A, B = block_indel_pairs(rng)
    cm = engine.CostMatrix.of_sub_indel(sub, indel, go)
    ref = oracle.CM2.sub_indel(sub, indel, go)
    res = engine.run_pairs(ctx, cm, engine.POYG_MODE_AFFINE_ALIGN, batch_of(A), batch_of(B))
    for k, (a, b) in enumerate(zip(A, B)):
        med, ri, rj, cost, mwg = oracle.align_affine_3(a, b, ref)
        assert res.cost[k] == cost, (k, len(a), len(b))
        for slot, want in enumerate((med, mwg, ri, rj)):
            assert np.array_equal(res.seq(k, slot), want), (k, slot, len(a), len(b))


@pytest.mark.parametrize("alphabet", ["dna", "any"])
def test_affine_packed_fill_saturation(ctx, oracle, alphabet):
    """The packed 16-bit narrow-band fill holds values as 4 v saturated at
    0xffff: pairs whose cost reaches 0x3fff are queued for the 32-bit fill.
    Costs of 40 / 30 / 50 put short DNA pairs below that bound and longer
    ones above it, in one batch; every output equals the oracle's and the
    32-bit fill's."""
    from poyd_amd import engine
    note(f"affine packed fill saturation alphabet={alphabet}")
    rng = np.random.default_rng(911 + (alphabet == "any"))
    A, B = [], []
    for k in range(96):
        L = int(rng.integers(20, 900))
        a = rand_seq(rng, L, alphabet)
        b = mutate(rng, a, 0.3, 0.05) if k % 3 else rand_seq(rng, L + int(rng.integers(-15, 15)), alphabet)
        A.append(a)
        B.append(b)
    cm = engine.CostMatrix.of_sub_indel(40, 30, 50)
    ref = oracle.CM2.sub_indel(40, 30, 50)
    res = engine.run_pairs(ctx, cm, engine.POYG_MODE_AFFINE_ALIGN, batch_of(A), batch_of(B))
    cres = engine.run_pairs(ctx, cm, engine.POYG_MODE_AFFINE_COST, batch_of(A), batch_of(B))
    ctx.set_tuning("aff_narrow_kernel", 1)
    try:
        res32 = engine.run_pairs(ctx, cm, engine.POYG_MODE_AFFINE_ALIGN, batch_of(A), batch_of(B))
    finally:
        ctx.set_tuning("aff_narrow_kernel", -1)
    costs = []
    for k, (a, b) in enumerate(zip(A, B)):
        med, ri, rj, cost, mwg = oracle.align_affine_3(a, b, ref)
        costs.append(cost)
        assert res.cost[k] == cost and res32.cost[k] == cost, (k, len(a), len(b))
        assert cres.cost[k] == oracle.cost_affine_3(a, b, ref), (k, len(a), len(b))
        for slot, want in enumerate((med, mwg, ri, rj)):
            assert np.array_equal(res.seq(k, slot), want), (k, slot, len(a), len(b))
            assert np.array_equal(res32.seq(k, slot), want), (k, slot, len(a), len(b))
    costs = np.array(costs)
    if alphabet == "dna":   # gap-bearing characters make the other alphabet's alignments cheap
        assert (costs < 0x3fff).sum() >= 10 and (costs >= 0x3fff).sum() >= 5, costs


@pytest.mark.parametrize("sub,indel", [(1, 2), (1, 1), (2, 1)])
@pytest.mark.parametrize("alphabet", ["dna", "ambig"])
def test_linear_batch_matches_oracle(ctx, oracle, sub, indel, alphabet):
    from poyd_amd import engine
    note(f"linear batch sub={sub} indel={indel} alphabet={alphabet}")
    rng = np.random.default_rng(200 + sub * 10 + indel)
    A, B = mixed_pairs(rng, 120, alphabet)
    cm = engine.CostMatrix.of_sub_indel(sub, indel)
    ref = oracle.CM2.sub_indel(sub, indel)
    res = engine.run_pairs(ctx, cm, engine.POYG_MODE_LINEAR_ALIGN, batch_of(A), batch_of(B))
    cres = engine.run_pairs(ctx, cm, engine.POYG_MODE_LINEAR_COST, batch_of(A), batch_of(B))
    for k, (a, b) in enumerate(zip(A, B)):
        cost, oa, ob = oracle.align_2(a, b, ref)
        assert res.cost[k] == cost, (k, len(a), len(b))
        assert cres.cost[k] == oracle.cost_2(a, b, ref), (k, len(a), len(b))
        assert np.array_equal(res.seq(k, 0), oa), (k, len(a), len(b))
        assert np.array_equal(res.seq(k, 1), ob), (k, len(a), len(b))
        assert np.array_equal(res.seq(k, 2), oracle.median_2(oa, ob, ref)), k


def test_linear_explicit_deltaw(ctx, oracle):
    """DOS.distance passes ?deltaw (src/seqCS.ml:856-866); the batch takes the
    resulting band parameter per pair."""
    from poyd_amd import engine
    note("linear explicit deltaw")
    rng = np.random.default_rng(7)
    A, B = mixed_pairs(rng, 40)
    cm = engine.CostMatrix.of_sub_indel(1, 2)
    ref = oracle.CM2.sub_indel(1, 2)
    dws = [max(8, abs(len(a) - len(b))) for a, b in zip(A, B)]
    bands = np.array([engine.deltaw(a, b, d) for a, b, d in zip(A, B, dws)], np.int32)
    res = engine.run_pairs(ctx, cm, engine.POYG_MODE_LINEAR_COST, batch_of(A), batch_of(B), bands)
    for k, (a, b) in enumerate(zip(A, B)):
        assert res.cost[k] == oracle.cost_2(a, b, ref, dws[k])


def test_linear_long_columns(ctx, oracle):
    """Column sequences around 2,048: the systolic kernel takes up to 2,048
    columns (a full-plane band of that width included), the row kernels
    longer ones (narrow bands in LDS rings, a wide full-plane band)."""
    from poyd_amd import engine
    note("linear long columns")
    rng = np.random.default_rng(11)
    shapes = [(2047, 2047), (2048, 2048), (2049, 2049), (2100, 2200), (3072, 2048), (3100, 2049), (3600, 2300)]
    A, B = [], []
    for la, lb in shapes:
        a = rand_seq(rng, la)
        b = mutate(rng, a, 0.15, 0.05)[:lb] if lb <= la else rand_seq(rng, lb)
        A.append(a)
        B.append(b)
    cm = engine.CostMatrix.of_sub_indel(1, 2)
    ref = oracle.CM2.sub_indel(1, 2)
    res = engine.run_pairs(ctx, cm, engine.POYG_MODE_LINEAR_ALIGN, batch_of(A), batch_of(B))
    cres = engine.run_pairs(ctx, cm, engine.POYG_MODE_LINEAR_COST, batch_of(A), batch_of(B))
    for k, (a, b) in enumerate(zip(A, B)):
        cost, oa, ob = oracle.align_2(a, b, ref)
        assert res.cost[k] == cost, (k, len(a), len(b))
        assert cres.cost[k] == cost, (k, len(a), len(b))
        assert np.array_equal(res.seq(k, 0), oa), (k, len(a), len(b))
        assert np.array_equal(res.seq(k, 1), ob), (k, len(a), len(b))


def test_per_call_entry_points(ctx, oracle):
    from poyd_amd import engine
    note("per-call entry points")
    rng = np.random.default_rng(9)
    al = engine.Align(ctx)
    cm = engine.CostMatrix.of_sub_indel(1, 2)
    ref = oracle.CM2.sub_indel(1, 2)
    cma = engine.CostMatrix.of_sub_indel(1, 1, 3)
    refa = oracle.CM2.sub_indel(1, 1, 3)
    # short pairs, then the lengths of the per-call benchmark and longer
    sizes = [int(rng.integers(1, 200)) for _ in range(12)] + [300, 300, 1000, 1000, 2500]
    for k, n in enumerate(sizes):
        a = rand_seq(rng, n)
        b = mutate(rng, a, 0.1, 0.05) if k % 2 else rand_seq(rng, n if n >= 300 else int(rng.integers(1, 200)))
        assert al.cost_2(a, b, cm) == oracle.cost_2(a, b, ref)
        x, y, c = al.align_2(a, b, cm)
        c2, ox, oy = oracle.align_2(a, b, ref)
        assert c == c2 and np.array_equal(x, ox) and np.array_equal(y, oy)
        got = al.align_affine_3(a, b, cma)
        want = oracle.align_affine_3(a, b, refa)
        assert got[3] == want[3]
        for g, w in zip(got, want):
            if isinstance(g, np.ndarray):
                assert np.array_equal(g, w)
        assert al.cost_2(a, b, cma) == oracle.cost_affine_3(a, b, refa)


def test_three_d_matches_oracle(ctx, oracle):
    import ctypes as C
    from poyd_amd import engine
    note("three-dimensional alignment")
    rng = np.random.default_rng(11)
    cm3 = engine.CostMatrix3(engine.CostMatrix.of_sub_indel(1, 2))
    ref = oracle.CM3(oracle.CM2.sub_indel(1, 2))
    L = engine.load()
    for k in range(16):
        anc = rand_seq(rng, int(rng.integers(0, 40)))
        s = [mutate(rng, anc, 0.15, 0.1) for _ in range(3)]
        if k == 0:
            s = [rand_seq(rng, 0), rand_seq(rng, 0), rand_seq(rng, 0)]
        if k == 1:
            s = [rand_seq(rng, 0), rand_seq(rng, 5), rand_seq(rng, 9)]
        cap = sum(len(x) for x in s) + 1
        outs = [np.zeros(cap, np.uint8) for _ in range(3)]
        ln = C.c_int32(0)
        cost = C.c_int32(0)
        rc = L.poyg_algn_align_3d(ctx.h, cm3.h, *[v for x in s for v in (engine._p8(x), len(x))],
                                  *[engine._p8(o) for o in outs], cap, C.byref(ln), C.byref(cost))
        assert rc == 0, engine.load().poyg_last_error()
        want = oracle.align_3(*s, ref)
        assert cost.value == want[0], k
        for o, w in zip(outs, want[1:]):
            assert np.array_equal(o[cap - ln.value:], w), k


def run_triples(ctx, cm2, T):
    """One triple plan over T (lists of three sequences): costs, aligned
    triples and medians as numpy arrays."""
    from poyd_amd import engine
    cm3 = engine.CostMatrix3(cm2)
    seqs = [x for t in T for x in t]
    bt = batch_of(seqs)
    off = bt.off.reshape(-1, 3)
    ln = bt.ln.reshape(-1, 3)
    plan = engine.TriplePlan(ctx, cm3, off[:, 0], ln[:, 0], off[:, 1], ln[:, 1], off[:, 2], ln[:, 2])
    d_seq = engine.DeviceBuffer.from_array(ctx, bt.seq)
    d_cost = engine.DeviceBuffer(ctx, 4 * len(T))
    d_out = engine.DeviceBuffer(ctx, plan.out_bytes)
    d_len = engine.DeviceBuffer(ctx, 16 * len(T))
    plan.run(d_seq.ptr, d_cost.ptr, d_out.ptr, d_len.ptr)
    ctx.sync()
    import ctypes as C
    bad = C.c_int32(0)
    engine.load().poyg_plan_status(plan.h, C.byref(bad))
    assert bad.value == 0
    cost = d_cost.download(np.int32, len(T))
    out = d_out.download(np.uint8, plan.out_bytes)
    lens = d_len.download(np.int32, 4 * len(T))
    res = []
    for k, t in enumerate(T):
        cap = ((sum(len(x) for x in t) + 1 + 15) // 16) * 16
        got = []
        for q in range(4):
            n = lens[4 * k + q]
            base = plan.out_off[k] + q * cap
            got.append(out[base + cap - n: base + cap].copy())
        res.append((int(cost[k]), got))
    plan.close()
    return res


def check_triples(ctx, oracle, sub, indel, T):
    from poyd_amd import engine
    cm2 = engine.CostMatrix.of_sub_indel(sub, indel)
    ref = oracle.CM3(oracle.CM2.sub_indel(sub, indel))
    got = run_triples(ctx, cm2, T)
    for k, t in enumerate(T):
        want = oracle.align_3(*t, ref)
        assert got[k][0] == want[0], (k, [len(x) for x in t])
        for q in range(3):
            assert np.array_equal(got[k][1][q], want[1 + q]), (k, q, [len(x) for x in t])
        assert np.array_equal(got[k][1][3], oracle.median_3(*want[1:], ref)), k


def test_three_d_batch_matches_oracle(ctx, oracle):
    note("three-dimensional batch")
    rng = np.random.default_rng(12)
    T = []
    for k in range(40):
        anc = rand_seq(rng, int(rng.integers(5, 60)))
        T.append([mutate(rng, anc, 0.15, 0.06) for _ in range(3)])
    check_triples(ctx, oracle, 1, 2, T)


def shaped_triples(rng, shapes):
    T = []
    for n1, n2, n3 in shapes:
        anc = rand_seq(rng, max(n1, n2, n3) - 1)
        t = []
        for n in (n1, n2, n3):
            x = mutate(rng, anc, 0.15, 0.04)[:n]
            if len(x) < n:
                x = np.concatenate([x, rand_seq(rng, n - len(x))[1:]]).astype(np.uint8)
            t.append(x)
        T.append(t)
    return T


# s3 lengths across every row width of the row kernel (64 * 4, 8, 12, 16
# cells) and past it (the chunked kernel), with degenerate planes and lines
CUBE_SHAPES = [(1, 1, 1), (1, 1, 9), (1, 7, 1), (6, 1, 1), (1, 5, 300), (4, 1, 520), (2, 2, 2),
               (9, 11, 255), (9, 11, 256), (9, 11, 257), (12, 8, 300), (7, 13, 511), (7, 13, 512),
               (7, 13, 513), (10, 9, 700), (5, 6, 768), (5, 6, 769), (6, 5, 1000), (4, 4, 1024),
               (3, 5, 1025), (4, 3, 1300), (40, 35, 90), (30, 45, 130)]


def test_three_d_row_classes(ctx, oracle):
    note("three-dimensional row classes, int16 rows")
    rng = np.random.default_rng(31)
    check_triples(ctx, oracle, 1, 2, shaped_triples(rng, CUBE_SHAPES))


def test_three_d_wide_costs(ctx, oracle):
    note("three-dimensional row classes, int32 rows (costs past 2^15)")
    rng = np.random.default_rng(32)
    check_triples(ctx, oracle, 90, 200, shaped_triples(rng, CUBE_SHAPES[:18]))


def test_fitch_matches_oracle(ctx, oracle):
    from poyd_amd import engine
    note("fitch down/up")
    rng = np.random.default_rng(13)
    for length in (1, 17, 10000, 4096):
        a, b, c, d = (rng.integers(1, 256, length).astype(np.uint8) for _ in range(4))
        res = np.zeros(length, np.uint8)
        import ctypes as C
        added = C.c_int64(0)
        L = engine.load()
        assert L.poyg_nonadd_median(ctx.h, engine._p8(a), engine._p8(b), engine._p8(res), length, C.byref(added)) == 0
        want, wadd = oracle.nonadd_median(a, b)
        assert np.array_equal(res, want) and added.value == wadd
        assert L.poyg_nonadd_median_3(ctx.h, engine._p8(a), engine._p8(b), engine._p8(c), engine._p8(d),
                                      engine._p8(res), length) == 0
        assert np.array_equal(res, oracle.nonadd_median_3(a, b, c, d))


def test_fitch_host_batches_match_oracle(ctx, oracle):
    """poyg_fitch_pairs_host (engine.fitch_pairs): batches of row pairs over
    host arrays, every median row and change count equal to the oracle's
    nonadd_median, for odd lengths, one row and many, and an empty batch."""
    from poyd_amd import engine
    note("fitch host batches")
    rng = np.random.default_rng(14)
    for n, length in ((1, 1), (3, 17), (40, 1001), (700, 64), (5, 10000)):
        a = rng.integers(1, 256, (n, length)).astype(np.uint8)
        b = rng.integers(1, 256, (n, length)).astype(np.uint8)
        b[:, ::3] = a[:, ::3]   # some characters agree
        res, added = engine.fitch_pairs(ctx, a, b)
        for k in range(n):
            want, wadd = oracle.nonadd_median(a[k], b[k])
            assert np.array_equal(res[k], want) and int(added[k]) == wadd, (n, length, k)
    res, added = engine.fitch_pairs(ctx, np.zeros((0, 8), np.uint8), np.zeros((0, 8), np.uint8))
    assert res.shape == (0, 8) and added.shape == (0,)


def test_workload_pairwise_linear_1000(ctx, oracle):
    """The 1,000-pair linear workload (L 200-400, 10% subs, 2% indels),
    every pair bit-exact: cost, both aligned sequences and the median."""
    from poyd_amd import engine
    note("workload pairwise-linear-1000")
    w = synth.workload_linear_1000()
    cm = engine.CostMatrix.of_sub_indel(1, 2)
    ref = oracle.CM2.sub_indel(1, 2)
    res = engine.run_pairs(ctx, cm, engine.POYG_MODE_LINEAR_ALIGN, w.a, w.b)
    for k in range(len(w.a)):
        a, b = w.a.get(k), w.b.get(k)
        cost, oa, ob = oracle.align_2(a, b, ref)
        assert res.cost[k] == cost
        assert np.array_equal(res.seq(k, 0), oa) and np.array_equal(res.seq(k, 1), ob)
        assert np.array_equal(res.seq(k, 2), oracle.median_2(oa, ob, ref))


def test_workload_pairwise_affine_sample(ctx, oracle):
    """The 65,536-pair affine workload: a seeded sample of 1,500 pairs
    bit-exact (all four outputs and the cost) and, for all pairs, the
    size-independent properties: recost of the alignment equals the cost and
    stripping gaps returns the inputs."""
    from poyd_amd import engine
    note("workload pairwise-affine-65536")
    w = synth.workload_affine_65536()
    cm = engine.CostMatrix.of_sub_indel(1, 1, 3)
    ref = oracle.CM2.sub_indel(1, 1, 3)
    res = engine.run_pairs(ctx, cm, engine.POYG_MODE_AFFINE_ALIGN, w.a, w.b)
    rng = np.random.default_rng(5)
    sample = rng.choice(len(w.a), size=1500, replace=False)
    for k in sample:
        a, b = w.a.get(k), w.b.get(k)
        med, ri, rj, cost, mwg = oracle.align_affine_3(a, b, ref)
        assert res.cost[k] == cost, k
        for slot, want in enumerate((med, mwg, ri, rj)):
            assert np.array_equal(res.seq(k, slot), want), (k, slot)
    bad = 0
    for k in range(len(w.a)):
        ri, rj = res.seq(k, 2), res.seq(k, 3)
        if oracle.verify_2(ri, rj, ref) != res.cost[k]:
            bad += 1
        elif not (np.array_equal(ri[ri != GAP], w.a.get(k)[1:]) and np.array_equal(rj[rj != GAP], w.b.get(k)[1:])):
            bad += 1
    assert bad == 0


@pytest.mark.parametrize("mm,go,ge", [(1, 2, 1), (2, 0, 1), (1, 3, 2)])
def test_powell_matches_oracle(ctx, oracle, mm, go, ge):
    from poyd_amd import engine
    note(f"powell_3d mm={mm} go={go} ge={ge}")
    rng = np.random.default_rng(50 + mm + 7 * go + 13 * ge)
    shapes = [(0, 0, 0), (0, 3, 5), (4, 0, 0), (1, 1, 1), (7, 7, 7)] + \
             [tuple(int(v) for v in rng.integers(1, 60, size=3)) for _ in range(12)] + [(150, 140, 160)]
    for n1, n2, n3 in shapes:
        anc = rand_seq(rng, max(n1, n2, n3, 1))
        t = []
        for n in (n1, n2, n3):
            x = mutate(rng, anc, 0.15, 0.05)[:n + 1]
            if len(x) < n + 1:
                x = np.concatenate([x, rand_seq(rng, n + 1 - len(x))[1:]]).astype(np.uint8)
            t.append(x)
        cost, al = engine.powell_3d(ctx, *t, mm, go, ge)
        # the reference's own search picks the alignment (its cost is the
        # model's least cost, except that three equal strings cost 1: the
        # search's loop starts at cost 1)
        want = oracle.powell_ukk(*t, mm, go, ge)
        least = oracle.powell_3d(*t, mm, go, ge)[0]
        same = all(len(x) == len(t[0]) and np.array_equal(x, t[0]) for x in t)
        assert cost == want[0] == (1 if same else least), (n1, n2, n3)
        for q in range(3):
            assert np.array_equal(al[q], want[1 + q]), (n1, n2, n3, q)


def test_powell_search_choices(ctx, oracle):
    # divergent triples, where least-cost alignments are many and the
    # check-point recursion runs several levels deep: the engine's alignment
    # is the reference search's, column for column
    from poyd_amd import engine
    note("powell_3d search choices")
    rng = np.random.default_rng(58)
    for k in range(8):
        mm, go, ge = [(1, 2, 1), (2, 0, 1), (1, 3, 2), (3, 1, 1), (1, 3, 1)][k % 5]
        anc = rand_seq(rng, int(rng.integers(40, 100)))
        t = [mutate(rng, anc, 0.3, 0.1) for _ in range(3)]
        cost, al = engine.powell_3d(ctx, *t, mm, go, ge)
        want = oracle.powell_ukk(*t, mm, go, ge)
        assert cost == want[0] == oracle.powell_3d(*t, mm, go, ge)[0], k
        for q in range(3):
            assert np.array_equal(al[q], want[1 + q]), (k, q)


def test_powell_rejects_gap_only_characters(ctx):
    from poyd_amd import engine
    x = np.array([16, 1, 16, 2], np.uint8)
    with pytest.raises(engine.PoygError):
        engine.powell_3d(ctx, x, x, x, 1, 2, 1)


def test_device_cache_reuse(ctx):
    """Blocks released to the per-device cache come back only after a
    synchronization, and a reused block holds what its new owner wrote."""
    from poyd_amd import engine
    note("device cache reuse")
    rng = np.random.default_rng(5)
    live = []
    for cycle in range(40):
        size = int(rng.choice([100, 256, 4000, 70000, 1 << 20]))
        data = rng.integers(0, 256, size, dtype=np.uint8)
        b = engine.DeviceBuffer.from_array(ctx, data)
        live.append((b, data))
        if cycle % 3 == 2:
            # release one buffer without a sync: a block allocated now must not be it
            old, _ = live.pop(0)
            ptr = old.ptr
            old.free()
            fresh = engine.DeviceBuffer(ctx, 16)
            assert fresh.ptr != ptr
            fresh.free()
        if cycle % 7 == 6:
            ctx.sync()
        for b2, d2 in live:
            assert np.array_equal(b2.download(np.uint8, len(d2)), d2), cycle
    for b2, _ in live:
        b2.free()
    ctx.sync()


def test_linear_fill_wait_gives_up_reported(ctx, oracle, monkeypatch):
    """The systolic fill's wait for the previous strip of a pair is a bounded
    poll.  With the poll limit forced to zero from the host, the waits give
    up at once: the kernel still ends, the plan reports the pairs as failed
    and the engine raises; the next run with the default limit is exact."""
    from poyd_amd import engine
    note("linear fill wait gives up")
    rng = np.random.default_rng(17)
    A = [rand_seq(rng, 1500) for _ in range(4)]
    B = [mutate(rng, a, 0.1, 0.03) for a in A]
    cm = engine.CostMatrix.of_sub_indel(1, 2)
    ref = oracle.CM2.sub_indel(1, 2)
    monkeypatch.setenv("POYG_SYS_SPIN_LIMIT", "0")
    # these pairs would go to the packed fill, which has no waits: the
    # systolic fill takes them here
    ctx.set_tuning("lin_cost_kernel", 1)
    try:
        # each failed pair counts once, however many of its waves gave up
        with pytest.raises(engine.PoygError, match=r"\b[1-4] pairs failed"):
            engine.run_pairs(ctx, cm, engine.POYG_MODE_LINEAR_ALIGN, batch_of(A), batch_of(B))
    finally:
        ctx.set_tuning("lin_cost_kernel", -1)
    monkeypatch.delenv("POYG_SYS_SPIN_LIMIT")
    res = engine.run_pairs(ctx, cm, engine.POYG_MODE_LINEAR_ALIGN, batch_of(A), batch_of(B))
    for k, (a, b) in enumerate(zip(A, B)):
        cost, oa, ob = oracle.align_2(a, b, ref)
        assert res.cost[k] == cost, k
        assert np.array_equal(res.seq(k, 0), oa), k


def test_plan_rejects_what_the_kernels_cannot_hold(ctx):
    """Planning refuses, before any allocation or launch: a linear matrix
    with a negative entry (POYG_EINVAL: the kernels' out-of-band marker must
    stay above every cell value), and a linear alignment whose direction band
    reaches the kernels' 32-bit offsets (POYG_EUNSUPPORTED)."""
    import ctypes as C
    from poyd_amd import engine
    note("plan guards")
    L = engine.load()
    cm = engine.CostMatrix.of_sub_indel(1, 2)
    neg = engine.CostMatrix.of_list([[0, -1, 1, 1, 2], [-1, 0, 1, 1, 2], [1, 1, 0, 1, 2], [1, 1, 1, 0, 2],
                                     [2, 2, 2, 2, 0]])

    def plan(m, mode, la, lb, dw):
        h = C.c_void_p()
        rc = L.poyg_plan_pairs(ctx.h, m.h, mode, 1, (C.c_int64 * 1)(0), (C.c_int32 * 1)(la),
                               (C.c_int64 * 1)(0), (C.c_int32 * 1)(lb), (C.c_int32 * 1)(dw), C.byref(h))
        if rc == 0:
            L.poyg_plan_destroy(h)
        return rc

    assert plan(cm, engine.POYG_MODE_LINEAR_ALIGN, 300, 280, 40) == 0
    assert plan(neg, engine.POYG_MODE_LINEAR_COST, 300, 280, 40) == -1
    assert plan(cm, engine.POYG_MODE_LINEAR_ALIGN, 60000, 40000, 100) == -4
    assert plan(cm, engine.POYG_MODE_LINEAR_COST, 60000, 40000, 100) == 0   # no direction band


@pytest.mark.parametrize("width", [1, 2, 4])
def test_fitch_widths_match_oracle(ctx, oracle, width):
    """NonaddCS8/16/32: the down and up passes at 1, 2 and 4 bytes per
    character, with state sets over every bit of the width, row lengths
    that are and are not multiples of 16 bytes."""
    import ctypes as C
    from poyd_amd import engine
    note(f"fitch width {width}")
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[width]
    rng = np.random.default_rng(40 + width)
    L = engine.load()
    for length in (1, 7, 1000, 4096, 10003):
        a, b, c, d = (rng.integers(1, 1 << (8 * width), size=length, dtype=np.uint64).astype(dt) for _ in range(4))
        # sparse sets too, so that empty intersections occur at every width
        a[::3] = (np.uint64(1) << rng.integers(0, 8 * width, size=len(a[::3])).astype(np.uint64)).astype(dt)
        b[::3] = (np.uint64(1) << rng.integers(0, 8 * width, size=len(b[::3])).astype(np.uint64)).astype(dt)
        res = np.zeros(length, dt)
        added = C.c_int64(0)
        assert L.poyg_nonadd_median_w(ctx.h, width, a.ctypes.data, b.ctypes.data, res.ctypes.data, length,
                                      C.byref(added)) == 0
        want, wadd = oracle.nonadd_median_w(a, b, width)
        assert np.array_equal(res, want) and added.value == wadd, length
        assert L.poyg_nonadd_median_3_w(ctx.h, width, a.ctypes.data, b.ctypes.data, c.ctypes.data, d.ctypes.data,
                                        res.ctypes.data, length) == 0
        assert np.array_equal(res, oracle.nonadd_median_3_w(a, b, c, d, width)), length
