"""Benchmark of the direct-optimization engine on one MI355X.

Default workload: 65,536 DNA pairs, lengths log-uniform on [50, 5000],
divergence U[5%, 30%] (substitutions d, indels d/5), affine gaps (open 3,
extend 1, substitution 1).  One step = one batched align_affine_3 over all
pairs: cost, both aligned sequences, median and median-with-gaps, with the
sequences resident in HBM before timing starts.  The metric is band-GCUPS:
DP cells the reference algorithm evaluates (its band) per second.

Prints one JSON line: the metric, its value and unit, the run shape, the
workload, the fill stage's roofline and the single-core CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DP GCUPS (batched NW); TBR joins/s surrogate+exact vs exhaustive, same cost"

# int32 VALU operations per DP cell counted from the reference recurrences
# (src/algn.c:1740-1977 affine, :87-453 linear, :2812-3055 cube).
OPS_PER_CELL = {"affine": 30, "linear": 8, "cube": 20}
# MI355X: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz int32 VALU ops; HBM3E 8 TB/s.
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
HBM_PEAK_GBS = 8000.0


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def workload(name: str):
    from poyd_amd import synth
    if name == "affine":
        return synth.workload_affine_65536(), "affine"
    if name == "linear":
        return synth.workload_linear_1000(), "linear"
    if name == "cube":
        return synth.workload_three_way_4096(), "cube"
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(w, kind: str, budget_s: float):
    """Single-core oracle (the C restatement of the reference algorithm,
    built -O2 without -march like the reference) over a seeded sample of
    the same pairs, timed until budget_s of CPU work has been spent.
    Returns the baseline record and the oracle cost of every sampled item
    (item index -> cost), which the caller compares with the GPU's."""
    from oracle import pyoracle as po
    cm = po.CM2.sub_indel(1, 1, 3) if kind == "affine" else po.CM2.sub_indel(1, 2)
    if kind == "cube":
        cm3 = po.CM3(cm)
        n = len(w.ln[0])
    else:
        n = len(w.a)
    rng = np.random.default_rng(7)
    order = rng.permutation(n)
    try:
        old_aff = os.sched_getaffinity(0)
        core = min(old_aff)
        os.sched_setaffinity(0, {core})
    except (AttributeError, OSError):
        old_aff = None
    cells = 0
    done = 0
    want = {}
    t0 = time.perf_counter()
    for k in order:
        if kind == "cube":
            x = [w.get(q, int(k)) for q in range(3)]
            want[int(k)] = po.align_3(*x, cm3)[0]
            cells += len(x[0]) * len(x[1]) * len(x[2])
        else:
            a, b = w.a.get(int(k)), w.b.get(int(k))
            if kind == "affine":
                want[int(k)] = po.align_affine_3(a, b, cm)[3]
                cells += po.affine_cells(len(a), len(b))
            else:
                want[int(k)] = po.align_2(a, b, cm)[0]
                l1, l2 = max(len(a), len(b)), min(len(a), len(b))
                cells += po.algn_nw_cells(l1, l2, po.deltaw(a, b, cm))
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    if old_aff is not None:
        os.sched_setaffinity(0, old_aff)
    return {
        "value": cells / dt / 1e9,
        "unit": "GCUPS",
        "cores": 1,
        "kind": "port",
        "sample": f"{done} of {n} {'triples' if kind == 'cube' else 'pairs'} (seeded random order), "
                  f"{cells} {'cube' if kind == 'cube' else 'band'} cells in {dt:.1f} s, "
                  f"{'align_3 + median_3 per triple' if kind == 'cube' else 'align + median per pair'}, "
                  f"1 pinned core of {cpu_model()}",
    }, want


def oracle_sample_costs(w, kind: str, count: int) -> dict:
    """Oracle costs of `count` seeded random items (item index -> cost)."""
    from oracle import pyoracle as po
    cm = po.CM2.sub_indel(1, 1, 3) if kind == "affine" else po.CM2.sub_indel(1, 2)
    cm3 = po.CM3(cm) if kind == "cube" else None
    n = len(w.ln[0]) if kind == "cube" else len(w.a)
    want = {}
    for k in np.random.default_rng(11).choice(n, size=min(count, n), replace=False):
        k = int(k)
        if kind == "cube":
            want[k] = po.align_3(*[w.get(q, k) for q in range(3)], cm3)[0]
        elif kind == "affine":
            want[k] = po.align_affine_3(w.a.get(k), w.b.get(k), cm)[3]
        else:
            want[k] = po.align_2(w.a.get(k), w.b.get(k), cm)[0]
    return want


def make_plan(ctx, kind: str, w):
    """The plan of a workload on ctx (its knobs as currently set), the
    concatenated sequences and the per-sequence length arrays."""
    from poyd_amd import engine
    if kind == "cube":
        cm = engine.CostMatrix.of_sub_indel(1, 2)
        cm3 = engine.CostMatrix3(cm)
        seq = np.concatenate([b.seq for b in w.batches])
        base = np.cumsum([0] + [len(b.seq) for b in w.batches])
        plan = engine.TriplePlan(ctx, cm3, *[x for q, b in enumerate(w.batches) for x in (b.off + base[q], b.ln)])
        return plan, seq, [b.ln for b in w.batches]
    if kind == "affine":
        cm = engine.CostMatrix.of_sub_indel(1, 1, 3)
        mode, dws = engine.POYG_MODE_AFFINE_ALIGN, None
    else:
        cm = engine.CostMatrix.of_sub_indel(1, 2)
        mode, dws = engine.POYG_MODE_LINEAR_ALIGN, engine.batch_deltaw(w.a, w.b)
    seq = np.concatenate([w.a.seq, w.b.seq])
    plan = engine.PairPlan(ctx, cm, mode, w.a.off, w.a.ln, w.b.off + len(w.a.seq), w.b.ln, dws)
    return plan, seq, [w.a.ln, w.b.ln]


def tune_workload(ctx, kind: str, w):
    """Times launch-shape trials of the workload's own plan (poyd_amd.autotune)
    and leaves the best knobs set on ctx."""
    from poyd_amd import autotune, engine
    lengths = np.concatenate([np.asarray(x) for x in (w.ln if kind == "cube" else (w.a.ln, w.b.ln))])
    seq = None
    d = {}

    def trial():
        nonlocal seq
        plan, s, _ = make_plan(ctx, kind, w)
        if seq is None:
            seq = s
            d["seq"] = engine.DeviceBuffer.from_array(ctx, seq)
        d_cost = engine.DeviceBuffer(ctx, 4 * plan.n)
        d_out = engine.DeviceBuffer(ctx, plan.out_bytes)
        d_len = engine.DeviceBuffer(ctx, 4 * plan.n * plan.slots)
        plan.run(d["seq"].ptr, d_cost.ptr, d_out.ptr, d_len.ptr)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(2):
            plan.run(d["seq"].ptr, d_cost.ptr, d_out.ptr, d_len.ptr)
        ctx.sync()
        ms = (time.perf_counter() - t0) * 1e3 / 2
        plan.close()
        for b in (d_cost, d_out, d_len):
            b.free()
        ctx.sync()
        return ms

    res = autotune.tune(ctx, kind, lengths, trial, repeats=2, log=log)
    d["seq"].free()
    return res


def run_workload(kind: str, steps: int, warmup: int, tune: str = "table"):
    """Plans, runs and times one pair / triple workload; returns the
    workload, the plan's numbers and a verification of the run.  tune:
    "table" applies the tuned launch shapes for this device and batch,
    "default" the built-in ones, "trials" tunes them on this batch first."""
    from poyd_amd import autotune, engine
    w, kind = workload(kind)
    ctx = engine.Context(0)
    lengths = np.concatenate([np.asarray(x) for x in (w.ln if kind == "cube" else (w.a.ln, w.b.ln))])
    tuning = None
    if tune == "trials":
        tuning = tune_workload(ctx, kind, w)
    elif tune == "table":
        autotune.apply(ctx, kind, lengths)
    else:
        autotune.reset(ctx, kind)
    knobs = {k: ctx.tuning(k) for k, _ in autotune.KNOBS[kind]}
    t_plan = time.perf_counter()
    plan, seq, seq_lens = make_plan(ctx, kind, w)
    plan_ms = (time.perf_counter() - t_plan) * 1e3
    d_seq = engine.DeviceBuffer.from_array(ctx, seq)
    d_cost = engine.DeviceBuffer(ctx, 4 * plan.n)
    d_out = engine.DeviceBuffer(ctx, plan.out_bytes)
    d_len = engine.DeviceBuffer(ctx, 4 * plan.n * plan.slots)
    stream = load_stream(ctx)

    def step():
        plan.run(d_seq.ptr, d_cost.ptr, d_out.ptr, d_len.ptr, stream)

    for _ in range(warmup):
        step()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    ctx.sync()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / steps
    costs = d_cost.download(np.int32, plan.n)
    failed = plan.status()

    # kernel timings from HIP events recorded on the run stream
    plan.set_timing(True)
    fills, traces = [], []
    for _ in range(max(3, min(steps, 10))):
        step()
        f, t = plan.last_times()
        fills.append(f)
        traces.append(t)
    plan.set_timing(False)
    failed += plan.status()
    # the timing runs repeat the step: the same costs again
    deterministic = bool(np.array_equal(d_cost.download(np.int32, plan.n), costs))
    seq_bytes = int(sum(int(x.sum()) for x in seq_lens))
    out_bytes = int(d_len.download(np.int32, plan.n * plan.slots).astype(np.int64).sum())
    r = dict(w=w, kind=kind, n=plan.n, cells=int(plan.cells), ms=ms, plan_ms=plan_ms,
             fill_ms=float(np.mean(fills)), trace_ms=float(np.mean(traces)), costs=costs, failed=int(failed),
             deterministic=deterministic, seq_bytes=seq_bytes, out_bytes=out_bytes, knobs=knobs, tuning=tuning)
    plan.close()
    return r


def verify(r, want: dict) -> dict:
    """Plan status (no failed pair in any timed or timing run) and the GPU
    cost of every oracle-sampled item."""
    bad = [k for k, c in want.items() if int(r["costs"][k]) != int(c)]
    return {"plan_status_failed": r["failed"], "oracle_checked": len(want), "oracle_mismatches": len(bad),
            "verified": r["failed"] == 0 and not bad and r["deterministic"]}


def sub_result(kind: str, steps: int, warmup: int, n_check: int) -> dict:
    """A compact line for a second workload run in the same process."""
    r = run_workload(kind, steps, warmup)
    want = oracle_sample_costs(r["w"], r["kind"], n_check)
    ops = OPS_PER_CELL[r["kind"]]
    valu = r["cells"] * ops / (r["fill_ms"] / 1e3) / 1e12
    return {
        "workload": r["w"].name,
        "value": round(r["cells"] / (r["ms"] / 1e3) / 1e9, 3),
        "unit": "GCUPS",
        "ms_per_step": round(r["ms"], 3),
        "steps": steps,
        ("cube_cells" if r["kind"] == "cube" else "band_cells"): r["cells"],
        "plan_ms": round(r["plan_ms"], 1),
        "fill_ms": round(r["fill_ms"], 3),
        "valu_frac": round(valu / VALU_PEAK_TOPS, 4),
        **verify(r, want),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 50 for affine, 10 for the other pair workloads; tree and tbr: 1)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default 2; tree: 0)")
    ap.add_argument("--workload", default="affine",
                    choices=["affine", "linear", "cube", "tree", "fitch", "tbr", "percall", "powell"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sub", action="store_true", help="skip the linear and three-way sub-results")
    ap.add_argument("--tune", choices=["table", "default", "trials"], default="table",
                    help="launch shapes: the tuned table (default), the built-in ones, or fresh trials")
    ap.add_argument("--seeds", type=int, default=None, help="tbr: seeded datasets for the identical-cost rate")
    ap.add_argument("--taxa", type=int, default=None, help="tbr: taxa per dataset (default 256)")
    ap.add_argument("--epochs", type=int, default=8, help="tbr: training epochs of the join scorer")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device once (the reference dispatch that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(pairs are independent: N GPUs would each take 1/N of the batch)", file=sys.stderr)
        return 2
    tree = args.workload == "tree"
    if args.steps is None:
        # the default affine run times 50 steps (about 0.5 s of GPU work, so
        # that a utilisation sampler beside the run sees it)
        args.steps = 1 if tree or args.workload == "tbr" else (50 if args.workload == "affine" else 10)
    if args.warmup is None:
        args.warmup = 0 if tree or args.workload == "tbr" else 2
    if tree:
        return bench_tree(args)
    if args.workload == "fitch":
        return bench_fitch(args)
    if args.workload == "tbr":
        return bench_tbr(args)
    if args.workload == "percall":
        return bench_percall(args)
    if args.workload == "powell":
        return bench_powell(args)

    r = run_workload(args.workload, args.steps, args.warmup, args.tune)
    w, kind, ms, cells = r["w"], r["kind"], r["ms"], r["cells"]
    fill_ms, trace_ms = r["fill_ms"], r["trace_ms"]
    gcups = cells / (ms / 1e3) / 1e9
    ops = OPS_PER_CELL[kind]
    valu_tops = cells * ops / (fill_ms / 1e3) / 1e12
    # algorithmic HBM bytes of the fill: one direction byte per band cell,
    # both sequences read once, one int32 cost written per pair; the affine
    # fills also run the backtrace of their pairs, which reads both
    # sequences again and writes the aligned outputs
    alg_bytes = cells + r["seq_bytes"] + 4 * r["n"]
    if kind == "affine":
        alg_bytes += r["seq_bytes"] + r["out_bytes"]
    fill_gbs = alg_bytes / (fill_ms / 1e3) / 1e9
    if kind == "cube":
        matrix_cells = cells
    else:
        matrix_cells = int(((w.a.ln.astype(np.int64) - 1) * (w.b.ln.astype(np.int64) - 1)).sum())
    kernel = ("fill stage: aff_pk_kernel (eight pairs per wave in 16-bit halves; main and side-stream "
              "launches), the one-set and lane-set fills (each wave tracing its pairs after filling them) "
              "and the 32-bit fill of the pairs whose cost saturates 16 bits; one span of HIP events on "
              "the main stream (the packed fill's backtraces, aff_trace_list_kernel, are backtrace_ms)") \
        if kind == "affine" \
        else "cube_rows_kernel (every row class)" if kind == "cube" \
        else "lin_cost16_kernel (packed 16-bit halves, one pair per wave, 128-row strips)"
    out = {
        "metric": METRIC,
        "value": round(gcups, 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": w.name,
            ("triples" if kind == "cube" else "pairs"): r["n"],
            ("cube_cells" if kind == "cube" else "band_cells"): int(cells),
            "matrix_equivalent_gcups": round(matrix_cells / (ms / 1e3) / 1e9, 3),
            "cost_model": "affine open 3 extend 1 sub 1" if kind == "affine" else "linear sub 1 indel 2",
            "outputs": ("cost, three aligned sequences, median" if kind == "cube" else
                        "cost, aligned pair, median" + (", median with gaps" if kind == "affine" else "")),
            "plan_ms": round(r["plan_ms"], 1),
            "launch_shapes": r["knobs"],
            "tuning_trials": r["tuning"],
            "fill_ms": round(fill_ms, 3),
            "backtrace_ms": round(trace_ms, 3),
            "deterministic": r["deterministic"],
        },
        "roofline": {
            "bound": "valu",
            "achieved": round(valu_tops, 3),
            "peak": round(VALU_PEAK_TOPS, 2),
            "unit": "TOP/s (int32)",
            "frac": round(valu_tops / VALU_PEAK_TOPS, 4),
            "traffic": measured_traffic(kind),
            "ops_per_cell": ops,
            "kernel": kernel,
        },
        "hbm_roofline": {
            "bound": "hbm",
            "achieved": round(fill_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(fill_gbs / HBM_PEAK_GBS, 5),
            "algorithmic_bytes": int(alg_bytes),
        },
    }
    want = {}
    if not args.no_cpu:
        cb, want = cpu_baseline(w, kind, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["config"]["gpu_vs_cpu_1core"] = round(gcups / cb["value"], 1)
    else:
        want = oracle_sample_costs(w, kind, 16)
    out["config"].update(verify(r, want))
    if not args.no_sub and args.workload == "affine":
        out["sub_results"] = {"linear_1000": sub_result("linear", 20, 3, 64),
                              "three_way_4096": sub_result("cube", 1, 1, 3)}
    print(json.dumps(out), flush=True)
    return 0 if out["config"]["verified"] else 1


def bench_tree(args) -> int:
    """Tree workload: 128-taxon Yule tree data (5 DNA loci of 300-1,500 bp,
    10,000 nonadditive characters); one step = a Wagner build in taxon order
    (every taxon tried on every edge: directional medians, edge medians and
    join distances, each tree level one batch on the GPU) plus the final
    tree's diagnosis (medians, single assignment, Adjusted cost, Fitch
    length).  value = band-GCUPS over every alignment of the step."""
    from poyd_amd import engine, synth, treebuild, treepass
    w = synth.workload_tree_128()
    ds = treebuild.dataset_of(w)
    ctx = engine.Context(0)
    be = treepass.EngineBackend(ctx, engine.CostMatrix.of_sub_indel(1, 2))

    def step():
        t0 = time.perf_counter()
        def progress(x):
            if x % 32 == 0:
                print(f"[tree] {x} of {ds.n_taxa} taxa joined, {time.perf_counter() - t0:.1f} s",
                      file=sys.stderr, flush=True)

        r = treebuild.wagner(ds, be, progress)
        t1 = time.perf_counter()
        d = treepass.TreeEval(ds, r.tree, be).diagnose()
        return r, d, t1 - t0, time.perf_counter() - t1

    for _ in range(args.warmup):
        step()
    ctx.sync()
    c0, f0 = be.cells, be.fitch_chars
    costs, b_s, d_s = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r, d, tb, td = step()
        costs.append((d.cost, d.nonadd))
        b_s.append(tb)
        d_s.append(td)
    ctx.sync()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    cells = (be.cells - c0) / args.steps
    fitch = (be.fitch_chars - f0) / args.steps
    gcups = cells / (ms / 1e3) / 1e9
    ops = OPS_PER_CELL["linear"]
    valu = cells * ops / (ms / 1e3) / 1e12
    out = {
        "metric": METRIC,
        "value": round(gcups, 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": w.name,
            "taxa": ds.n_taxa,
            "loci": len(ds.chars),
            "nonadditive_chars": int(ds.nonadd.shape[1]),
            "cost_model": "linear sub 1 indel 2; nonadditive Fitch, weight 1",
            "step": "Wagner build in taxon order + diagnosis of the built tree",
            "tree_cost": int(costs[-1][0]),
            "nonadditive_cost": int(costs[-1][1]),
            "band_cells_per_step": int(cells),
            "fitch_char_medians_per_step": int(fitch),
            "build_ms": round(1e3 * float(np.mean(b_s)), 1),
            "diagnosis_ms": round(1e3 * float(np.mean(d_s)), 1),
            "deterministic": len(set(costs)) == 1,
        },
        "roofline": {
            "bound": "valu",
            "achieved": round(valu, 4),
            "peak": round(VALU_PEAK_TOPS, 2),
            "unit": "TOP/s (int32)",
            "frac": round(valu / VALU_PEAK_TOPS, 6),
            "traffic": None,
            "kernel": "whole step (one plan per tree level: a few pairs per launch, each launch as long as "
                      "its longest pair takes on one workgroup)",
        },
    }
    # Iterative ApproxD on the built tree (outside the timed steps): the
    # readjustment passes of AllDirChar.adjust_tree from its diagnosis
    ta = time.perf_counter()
    ev = treepass.TreeEval(ds, r.tree, be)
    _adj, a_cost, kept = ev.approx_d(d, iterations=2)
    ctx.sync()
    out["config"]["approx_d"] = {
        "passes": 2, "passes_kept": int(kept), "ms": round(1e3 * (time.perf_counter() - ta), 1),
        "tree_cost_before": int(d.cost), "tree_cost_after": int(a_cost + d.nonadd),
        "alignments": int(ev.n_align),
    }
    # one Iterative ThreeD pass (Powell's search per vertex and locus)
    ta = time.perf_counter()
    ev = treepass.TreeEval(ds, r.tree, be)
    _adj, t_cost, kept = ev.approx_d(d, iterations=1, three_d=True)
    ctx.sync()
    out["config"]["three_d"] = {
        "passes": 1, "passes_kept": int(kept), "ms": round(1e3 * (time.perf_counter() - ta), 1),
        "tree_cost_after": int(t_cost + d.nonadd), "powell_searches": int(ev.n_align),
    }
    if not args.no_cpu:
        cb = cpu_baseline_tree(ds, r.tree, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["config"]["gpu_vs_cpu_1core"] = round(gcups / cb["value"], 1)
    print(json.dumps(out), flush=True)
    return 0


def bench_fitch(args) -> int:
    """Nonadditive characters: Fitch down pass then up pass over a 1,024-taxon
    Yule tree with 1,048,576 characters of 2-8 states (one byte each), every
    node's sets resident in HBM.  One launch per tree level; value = character
    medians (down and up) per second.  Algorithmic bytes: 3 per character per
    down-pass node (two children read, one median written), 5 per up-pass
    node (parent final, own preliminary, two children read, final written)."""
    import ctypes as C
    from poyd_amd import engine, synth
    n_taxa, L = 1024, 1 << 20
    rng = np.random.default_rng(synth.SEED0 + 5)
    parent, _, leaves = synth.yule_tree(rng, n_taxa, 0.05)
    nn = len(parent)
    kids = [[] for _ in range(nn)]
    for v in range(1, nn):
        kids[parent[v]].append(v)
    internal = [v for v in range(nn) if kids[v]]
    height = np.zeros(nn, np.int64)
    for v in range(nn - 1, -1, -1):
        if kids[v]:
            height[v] = 1 + max(height[c] for c in kids[v])
    depth = np.zeros(nn, np.int64)
    for v in range(1, nn):
        depth[v] = depth[parent[v]] + 1
    # rows: preliminary sets of every node (0..nn-1), final sets of internal nodes (nn + index)
    fin = {v: nn + q for q, v in enumerate(internal)}
    rows = nn + len(internal)
    ctx = engine.Context(0)
    d_chars = engine.DeviceBuffer(ctx, rows * L)
    states = rng.integers(2, 9, size=L).astype(np.uint8)
    lib = engine.load()
    sample = []
    for q, v in enumerate(leaves):
        row = np.left_shift(np.uint8(1), rng.integers(0, 8, size=L, dtype=np.uint8) % states).astype(np.uint8)
        lib.poyg_dev_upload(ctx.h, d_chars.ptr + int(v) * L, row.ctypes.data, L)
        if q < 2:
            sample.append(row)
    down = []
    for h in range(1, int(height.max()) + 1):
        vs = [v for v in internal if height[v] == h]
        down.append(np.array([[kids[v][0], kids[v][1], v] for v in vs], np.int32))
    up = []
    for dpt in range(1, int(depth.max()) + 1):
        vs = [v for v in internal if depth[v] == dpt]
        if vs:
            up.append(np.array([[fin[parent[v]], v, kids[v][0], kids[v][1], fin[v]] for v in vs], np.int32))
    dev_lv = [[engine.DeviceBuffer.from_array(ctx, np.ascontiguousarray(a[:, c])) for c in range(3)] for a in down]
    dev_up = [[engine.DeviceBuffer.from_array(ctx, np.ascontiguousarray(a[:, c])) for c in range(5)] for a in up]
    d_added = engine.DeviceBuffer(ctx, 8 * nn)
    stream = load_stream(ctx)
    root = 0

    def step(st=None):
        st = stream if st is None else st
        engine.dev_zero(ctx, d_added.ptr, 8 * nn, st)
        for a, (ia, ib, ir) in zip(down, dev_lv):
            engine._check(lib.poyg_fitch_down(ctx.h, d_chars.ptr, L, len(a), ia.ptr, ib.ptr, ir.ptr,
                                              d_added.ptr + 8 * 0, st), "poyg_fitch_down")
        for a, (iA, iN, i1, i2, iF) in zip(up, dev_up):
            engine._check(lib.poyg_fitch_up(ctx.h, d_chars.ptr, L, len(a), iA.ptr, iN.ptr, i1.ptr, i2.ptr,
                                            iF.ptr, st), "poyg_fitch_up")

    # the root's final sets are its preliminary ones: copied once after a
    # first down pass (the leaves never change)
    for _ in range(max(1, args.warmup)):
        step()
    ctx.sync()
    host_root = np.empty(L, np.uint8)
    lib.poyg_dev_download(ctx.h, host_root.ctypes.data, d_chars.ptr + root * L, L)
    lib.poyg_dev_upload(ctx.h, d_chars.ptr + fin[root] * L, host_root.ctypes.data, L)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.sync()
    eager_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # the same step recorded once as a graph (one launch per step)
    g = engine.Graph(ctx)
    step(g.stream)
    g.end()
    for _ in range(max(1, args.warmup)):
        g.launch(stream)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.launch(stream)
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    g.close()
    n_down = sum(len(a) for a in down)
    n_up = sum(len(a) for a in up)
    medians = (n_down + n_up) * L
    alg_bytes = (3 * n_down + 5 * n_up) * L
    gbs = alg_bytes / (ms / 1e3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(medians / (ms / 1e3) / 1e9, 3),
        "unit": "G character medians/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": f"fitch-{n_taxa}taxa-{L}chars",
            "down_pass_nodes": n_down,
            "up_pass_nodes": n_up,
            "levels": len(down) + len(up),
            "launch": "the step's level launches recorded once as a HIP graph (poyg_graph_*), one graph "
                      "launch per step",
            "eager_ms_per_step": round(eager_ms, 3),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "fitch_down_kernel + fitch_up_kernel, one kernel per tree level, replayed as one graph",
        },
    }
    if not args.no_cpu:
        from oracle import pyoracle as po
        try:
            old_aff = os.sched_getaffinity(0)
            os.sched_setaffinity(0, {min(old_aff)})
        except (AttributeError, OSError):
            old_aff = None
        a, b = sample[0], sample[1]
        done, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < min(args.cpu_seconds, 10.0):
            po.nonadd_median(a, b)
            done += 1
        dt = time.perf_counter() - t0
        if old_aff is not None:
            os.sched_setaffinity(0, old_aff)
        out["cpu_baseline"] = {"value": done * L / dt / 1e9, "unit": "G character medians/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{done} down-pass medians of {L} characters in {dt:.1f} s, 1 pinned core "
                                         f"of {cpu_model()}"}
    print(json.dumps(out), flush=True)
    return 0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def tbr_dataset(n_taxa: int, seed: int):
    """256 taxa (or n_taxa) on a seeded Yule tree, 3 DNA loci of 500-2,000
    bp, no nonadditive characters; the same simulator as the tree workload."""
    from poyd_amd import synth
    from poyd_amd.treedata import Dataset
    w = synth.tree_workload(n_taxa=n_taxa, n_loci=3, lmin=500, lmax=2000, n_nonadd=0, seed=seed,
                            name=f"tbr-{n_taxa}taxa-3loci-500-2000-seed{seed}")
    return Dataset(list(w.names), [list(l) for l in w.loci], None), w.name


class TbrCase:
    """One dataset: its Wagner tree (built on the GPU), the unbroken tree's
    directional medians in a device store, and the TBR candidates."""

    def __init__(self, ctx, cm, n_taxa, seed, arena_gb):
        from poyd_amd import seqstore, tbr, treebuild, treepass
        self.ds, self.name = tbr_dataset(n_taxa, seed)
        t0 = time.perf_counter()
        self.tree = treebuild.wagner(self.ds, treepass.EngineBackend(ctx, cm)).tree
        self.build_s = time.perf_counter() - t0
        self.store = seqstore.DeviceStore(ctx, cm, int(arena_gb * (1 << 30)))
        L = len(self.ds.chars)
        leaf_ids = {v: self.store.add([self.ds.chars[c][v - 1] for c in range(L)]) for v in self.tree.leaves}
        t0 = time.perf_counter()
        self.dirs, self.index = tbr.directional(self.store, self.tree, leaf_ids)
        self.dir_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        self.sp = tbr.neighbourhood(self.tree, self.index)
        self.cand = tbr.candidates(self.sp, self.dirs, len(self.dirs))
        self.enum_s = time.perf_counter() - t0

    def close(self):
        self.store.arena.free()


def tbr_training_rows(case, res, n_rows, rng):
    """(clade id, edge id, exact distance) rows of the model's training
    set: the model reads the unbroken tree's medians, the label is the
    exact per-character join cost of the exhaustive round."""
    from poyd_amd import surrogate
    te, g_te = surrogate.unbroken_edge_medians(case.store, case.sp, case.dirs, case.index)
    K = len(case.dirs)
    ca = surrogate.approx_ids(case.cand.clade, case.dirs, te, g_te, K)
    ea = surrogate.approx_ids(case.cand.edge, case.dirs, te, g_te, K)
    flat = rng.choice(ca.size, size=min(n_rows, ca.size), replace=False)
    return ca.ravel()[flat], ea.ravel()[flat], res.dist.ravel()[flat], te


def bench_tbr(args) -> int:
    """Surrogate-guided TBR against exhaustive exact TBR.  One step = one
    TBR neighbourhood round of a Wagner tree: every break x clade root x
    join edge candidate.  Exhaustive: every candidate scored by the exact
    DP (the changed component medians, every edge median, every join
    distance).  Guided: the learned model scores every candidate from 4-mer
    profiles of the unbroken tree's medians; the best 8% are scored exactly.
    value = candidate joins per second of the guided round; the exhaustive
    round's rate, the best-join recall and the identical-best-cost rate over
    the seeds are in config.  The model is trained first, on the exact
    costs of the exhaustive round of a separate training dataset."""
    from poyd_amd import engine, surrogate, tbr, treepass
    n_taxa = args.taxa or 256
    seeds = args.seeds or 1
    ctx = engine.Context(0)
    cm = engine.CostMatrix.of_sub_indel(1, 2)
    arena_gb = 64 if n_taxa >= 200 else 16
    rng = np.random.default_rng(2026)
    # ---- training: the exhaustive round of a training dataset labels 1M pairs
    t0 = time.perf_counter()
    tc = TbrCase(ctx, cm, n_taxa, synth_seed(100), arena_gb)
    log(f"training dataset {tc.name}: Wagner build {tc.build_s:.1f} s, {len(tc.cand.brk)} candidates")
    tres = tbr.exhaustive_round(tc.store, tc.sp, tc.dirs, tc.cand)
    ia, ib, lab, _ = tbr_training_rows(tc, tres, 1 << 20, rng)
    prof = surrogate.Profiles(tc.store, np.concatenate([ia, ib]))
    x = prof.features(ctx, ia, ib)
    tgt = engine.DeviceBuffer.from_array(ctx, (lab / surrogate.Y_SCALE).astype(np.float32))
    n_train = len(lab)
    scorer = surrogate.Scorer(ctx, batch_max=4096, seed=7)
    t1 = time.perf_counter()
    losses = scorer.train(x, tgt, n_train, epochs=args.epochs, lr=1e-3, seed=3, log=log)
    ctx.sync()
    train_s = time.perf_counter() - t1
    data_s = t1 - t0
    x.free()
    tgt.free()
    prof.counts.free()
    tc.close()
    del tc
    ctx.sync()
    # ---- evaluation seeds
    rows = []
    for s in range(seeds):
        case = TbrCase(ctx, cm, n_taxa, synth_seed(4) + s, arena_gb)
        J = len(case.cand.brk)
        log(f"seed {s} {case.name}: Wagner build {case.build_s:.1f} s, {J} candidates, "
            f"{len(case.sp.breaks)} breaks")
        ctx.sync()
        ex = tbr.exhaustive_round(case.store, case.sp, case.dirs, case.cand)
        ctx.sync()
        gr, info = surrogate.guided_round(case.store, scorer, case.sp, case.dirs, case.index, case.cand, 0.08)
        ctx.sync()
        # the same budget without the model: a seeded random 8%, and the plain
        # L1 distance of the 4-mer profiles in place of the network
        n_pick = len(gr.scored)
        rnd = np.sort(np.random.default_rng(s).choice(J, size=n_pick, replace=False))
        l1_score = profile_l1_score(case)
        l1 = np.sort(np.argpartition(l1_score, n_pick - 1)[:n_pick])
        model_score = info["predicted"].sum(1) + case.cand.base
        # where the exhaustive round's best candidate ranks under each ordering
        # (0 = first), and so at which budgets each ordering would keep it
        rank_model = int((model_score < model_score[ex.best]).sum())
        rank_l1 = int((l1_score < l1_score[ex.best]).sum())
        be = treepass.EngineBackend(ctx, cm)

        def exact_cost(k):
            return treepass.TreeEval(case.ds, tbr.apply_join(case.sp, case.cand, int(k)), be).diagnose().cost

        start_cost = treepass.TreeEval(case.ds, case.tree, be).diagnose().cost
        ex_cost = exact_cost(ex.best)
        gr_cost = ex_cost if gr.best == ex.best else exact_cost(gr.best)
        best_set = np.nonzero(ex.estimates == ex.best_estimate)[0]
        row = {
            "seed": s, "candidates": J, "start_cost": int(start_cost),
            "exhaustive_s": ex.seconds + case.enum_s, "guided_s": info["total_s"] + case.enum_s,
            "guided_model_s": info["model_s"], "guided_exact_s": info["exact_s"],
            "guided_edge_medians_s": info["edge_medians_s"], "enum_s": case.enum_s,
            "exhaustive_alignments": ex.alignments, "guided_alignments": gr.alignments,
            "exhaustive_cells": ex.cells, "guided_cells": gr.cells,
            "best_estimate": ex.best_estimate, "guided_best_estimate": gr.best_estimate,
            "best_cost": int(ex_cost), "guided_cost": int(gr_cost),
            "recall": bool(np.isin(ex.best, gr.scored)),
            "recall_any_best": bool(np.isin(best_set, gr.scored).any()),
            "random_recall": bool(np.isin(ex.best, rnd)), "l1_recall": bool(np.isin(ex.best, l1)),
            "random_best_estimate": int(ex.estimates[rnd].min()), "l1_best_estimate": int(ex.estimates[l1].min()),
            "model_rank_of_best": rank_model, "l1_rank_of_best": rank_l1,
            "budget_recall": {f"{f:g}": {"model": bool(rank_model < int(np.ceil(f * J))),
                                         "l1": bool(rank_l1 < int(np.ceil(f * J)))} for f in BUDGETS},
        }
        rows.append(row)
        log(json.dumps(row))
        case.close()
        del case
        ctx.sync()
    r0 = rows[0]
    g_rate = r0["candidates"] / r0["guided_s"]
    e_rate = r0["candidates"] / r0["exhaustive_s"]
    same = [r["guided_cost"] == r["best_cost"] for r in rows]
    out = {
        "metric": METRIC,
        "value": round(g_rate, 1),
        "unit": "joins/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(1e3 * r0["guided_s"], 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 DP; bf16 model (f32 accumulate)",
        "data": "synthetic",
        "config": {
            "workload": f"tbr-{n_taxa}taxa-3loci-500-2000",
            "taxa": n_taxa,
            "loci": 3,
            "cost_model": "linear sub 1 indel 2",
            "start_tree": "Wagner build in taxon order (GPU)",
            "round": "one TBR neighbourhood: every break x clade root x join edge",
            "candidates": r0["candidates"],
            "exhaustive_joins_per_s": round(e_rate, 1),
            "guided_joins_per_s": round(g_rate, 1),
            "speedup": round(g_rate / e_rate, 2),
            "median_over_seeds": {
                "exhaustive_joins_per_s": round(float(np.median([r["candidates"] / r["exhaustive_s"] for r in rows])), 1),
                "guided_joins_per_s": round(float(np.median([r["candidates"] / r["guided_s"] for r in rows])), 1),
                "speedup": round(float(np.median([r["exhaustive_s"] / r["guided_s"] for r in rows])), 2)},
            "rescored_fraction": 0.08,
            "seeds": seeds,
            "best_join_recall": round(float(np.mean([r["recall"] for r in rows])), 4),
            "identical_best_cost_rate": round(float(np.mean(same)), 4),
            "random_8pct_recall": round(float(np.mean([r["random_recall"] for r in rows])), 4),
            "profile_l1_8pct_recall": round(float(np.mean([r["l1_recall"] for r in rows])), 4),
            "recall_by_budget": {f"{f:g}": {"model": round(float(np.mean([r["budget_recall"][f"{f:g}"]["model"]
                                                                         for r in rows])), 4),
                                            "profile_l1": round(float(np.mean([r["budget_recall"][f"{f:g}"]["l1"]
                                                                              for r in rows])), 4),
                                            "random": f} for f in BUDGETS},
            "median_rank_of_best": {"model": float(np.median([r["model_rank_of_best"] for r in rows])),
                                    "profile_l1": float(np.median([r["l1_rank_of_best"] for r in rows]))},
            "model": "MLP 256-256-256-1, ReLU, bf16 weights, f32 accumulate, Adam, MSE",
            "training_pairs": int(n_train),
            "training_loss": [round(float(l), 5) for l in losses],
            "training_s": round(train_s, 2),
            "training_data_s": round(data_s, 1),
            "per_seed": rows,
        },
    }
    ex_cells, ex_s = r0["exhaustive_cells"], r0["exhaustive_s"]
    valu = ex_cells * OPS_PER_CELL["linear"] / ex_s / 1e12
    out["roofline"] = {"bound": "valu", "achieved": round(valu, 3), "peak": round(VALU_PEAK_TOPS, 2),
                       "unit": "TOP/s (int32)", "frac": round(valu / VALU_PEAK_TOPS, 4), "traffic": None,
                       "kernel": "exhaustive round: every linear DP of the round (lin_cost16_kernel / "
                                 "lin_c16mw_kernel fills, "
                                 "lin_trace_wave_kernel), 8 ops per band cell over the round's wall time"}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_tbr(n_taxa, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


def synth_seed(k: int) -> int:
    from poyd_amd import synth
    return synth.SEED0 + k


BUDGETS = (0.005, 0.01, 0.02, 0.04, 0.08)


def profile_l1_score(case):
    """The model's inputs without the model: broken cost plus half the
    summed L1 distance of the 4-mer profiles (the unbroken tree's medians;
    about one edit per two changed 4-mer counts)."""
    from poyd_amd import surrogate
    te, g_te = surrogate.unbroken_edge_medians(case.store, case.sp, case.dirs, case.index)
    K = len(case.dirs)
    ca = surrogate.approx_ids(case.cand.clade, case.dirs, te, g_te, K)
    ea = surrogate.approx_ids(case.cand.edge, case.dirs, te, g_te, K)
    prof = surrogate.Profiles(case.store, np.concatenate([ca.ravel(), ea.ravel()]))
    counts = prof.counts.download(np.float32, 256 * len(prof.ids)).reshape(len(prof.ids), 256)
    prof.counts.free()
    ra, rb = prof.rows(ca.ravel()), prof.rows(ea.ravel())
    l1 = np.empty(len(ra), np.float32)
    for c0 in range(0, len(ra), 1 << 18):
        l1[c0:c0 + (1 << 18)] = np.abs(counts[ra[c0:c0 + (1 << 18)]] - counts[rb[c0:c0 + (1 << 18)]]).sum(1)
    return case.cand.base + l1.reshape(ca.shape).sum(1) * 0.5


def cpu_baseline_tbr(n_taxa: int, budget_s: float):
    """The oracle scoring whole breaks of a smaller round exactly (the same
    simulator, 32 taxa: the oracle cannot finish a 256-taxon round in the
    bench's time) on one pinned core, breaks in a seeded random order until
    the budget is spent: candidate joins per second, the component medians
    of each break included."""
    from oracle import pyoracle as po
    from poyd_amd import tbr, treebuild
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_backend import OracleBackend, OracleStore
    ds, name = tbr_dataset(32, synth_seed(4))
    cm = po.CM2.sub_indel(1, 2)
    tree = treebuild.wagner(ds, OracleBackend(cm, False)).tree
    store = OracleStore(cm, False)
    leaf_ids = {v: store.add([ds.chars[c][v - 1] for c in range(3)]) for v in tree.leaves}
    dirs, index = tbr.directional(store, tree, leaf_ids)
    sp = tbr.neighbourhood(tree, index)
    cand = tbr.candidates(sp, dirs, len(dirs))
    try:
        old_aff = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {min(old_aff)})
    except (AttributeError, OSError):
        old_aff = None
    done, nb, t0 = 0, 0, time.perf_counter()
    for b in np.random.default_rng(5).permutation(len(sp.breaks)):
        if time.perf_counter() - t0 >= budget_s:
            break
        pick = np.nonzero(cand.brk == b)[0]
        tbr.rescore_round(store, sp, dirs, cand, pick)
        done += len(pick)
        nb += 1
    dt = time.perf_counter() - t0
    if old_aff is not None:
        os.sched_setaffinity(0, old_aff)
    return {"value": done / dt, "unit": "joins/s", "cores": 1, "kind": "port",
            "sample": f"{nb} of {len(sp.breaks)} breaks ({done} of {len(cand.brk)} candidate joins) of the TBR "
                      f"round of {name}, each break scored exactly with its component medians, in {dt:.1f} s "
                      f"on 1 pinned core of {cpu_model()}"}


def bench_percall(args) -> int:
    """Latency of the per-call entry points the OCaml stubs use (one pair
    per call, host buffers, synchronous): algn_CAML_simple_2 ->
    poyg_algn_simple_2 and algn_CAML_align_affine_3 ->
    poyg_algn_align_affine_3 on 300 bp and 1 kb pairs (10% substitutions,
    2% indels), against the oracle doing the same call on one core."""
    import ctypes as C
    from oracle import pyoracle as po
    from poyd_amd import engine, synth
    ctx = engine.Context(0)
    L = engine.load()
    lin, aff = engine.CostMatrix.of_sub_indel(1, 2), engine.CostMatrix.of_sub_indel(1, 1, 3)
    plin, paff = po.CM2.sub_indel(1, 2), po.CM2.sub_indel(1, 1, 3)
    rows = []
    for length in (300, 1000):
        w = synth.pairs_uniform(64, length, length, 0.10, 0.02, 77 + length, f"percall-{length}")
        pairs = [(w.a.get(k), w.b.get(k)) for k in range(len(w.a))]
        for name in ("simple_2", "align_affine_3"):
            def gpu_call(a, b):
                if name == "simple_2":
                    x, y = (a, b) if len(a) >= len(b) else (b, a)
                    out = C.c_int32(0)
                    engine._check(L.poyg_algn_simple_2(ctx.h, lin.h, engine._p8(x), len(x), engine._p8(y), len(y),
                                                       engine.deltaw(x, y), C.byref(out)), "simple_2")
                    return out.value
                return engine.Align(ctx).align_affine_3(a, b, aff)[3]

            def cpu_call(a, b):
                if name == "simple_2":
                    return po.cost_2(a, b, plin)
                return po.align_affine_3(a, b, paff)[3]

            for a, b in pairs[:4]:
                gpu_call(a, b)
            t0 = time.perf_counter()
            got = [gpu_call(a, b) for a, b in pairs]
            g_us = (time.perf_counter() - t0) / len(pairs) * 1e6
            t0 = time.perf_counter()
            want = [cpu_call(a, b) for a, b in pairs]
            c_us = (time.perf_counter() - t0) / len(pairs) * 1e6
            rows.append({"call": name, "length": length, "gpu_us": round(g_us, 1), "cpu_oracle_us": round(c_us, 1),
                         "same_costs": got == want})
            log(json.dumps(rows[-1]))
    print(json.dumps({"metric": "per-call latency of the stub entry points", "unit": "us per call",
                      "rows": rows, "cpu": cpu_model()}), flush=True)
    return 0


def bench_powell(args) -> int:
    """Powell's three-way search (powell_3D_align -> poyg_powell_3d), one
    triple per call: three descendants of a random ancestor (substitutions
    and indels at the given rates), mismatch 1, gap opening 2, extension 1;
    against the oracle's restatement of the same search on one core where
    that finishes in seconds (the shorter, closer triples), with the
    alignments compared column for column."""
    from oracle import pyoracle as po
    from poyd_amd import engine
    ctx = engine.Context(0)
    rng = np.random.default_rng(5)

    def descendant(anc, ps, pi):
        out = [16]
        for x in anc:
            r = rng.random()
            if r < pi / 2:
                continue
            if r < pi:
                out.append(int(rng.choice([1, 2, 4, 8])))
            out.append(int(rng.choice([1, 2, 4, 8])) if rng.random() < ps else int(x))
        return np.array(out, np.uint8)

    rows = []
    for n, ps, pi in [(100, 0.1, 0.04), (300, 0.1, 0.04), (300, 0.2, 0.08), (1000, 0.05, 0.02), (1000, 0.1, 0.04)]:
        anc = rng.choice([1, 2, 4, 8], size=n)
        t = [descendant(anc, ps, pi) for _ in range(3)]
        engine.powell_3d(ctx, *t, 1, 2, 1)
        t0 = time.perf_counter()
        for _ in range(3):
            cost, al = engine.powell_3d(ctx, *t, 1, 2, 1)
        g_ms = (time.perf_counter() - t0) / 3 * 1e3
        row = {"length": n, "substitutions": ps, "indels": pi, "cost": cost, "gpu_ms": round(g_ms, 1)}
        if n <= 300 and ps <= 0.1 and not args.no_cpu:
            t0 = time.perf_counter()
            want = po.powell_ukk(*t, 1, 2, 1)
            row["cpu_oracle_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
            row["same_alignment"] = bool(want[0] == cost and all(np.array_equal(p, q) for p, q in zip(al, want[1:])))
        rows.append(row)
        log(json.dumps(row))
    print(json.dumps({"metric": "Powell three-way search latency", "unit": "ms per triple", "rows": rows,
                      "cpu": cpu_model()}), flush=True)
    return 0


def cpu_baseline_tree(ds, tree, budget_s: float):
    """The oracle on one pinned core: the alignments of the built tree's
    diagnosis (directional and edge medians, single assignment, distances),
    in the same order, until budget_s is spent; band cells per second."""
    from oracle import pyoracle as po
    from poyd_amd import treepass
    cm = po.CM2.sub_indel(1, 2)

    class Stop(Exception):
        pass

    class CpuBackend:
        def __init__(self):
            self.cells, self.t0 = 0, time.perf_counter()

        def _tick(self):
            if time.perf_counter() - self.t0 > budget_s:
                raise Stop

        def cost_table(self):
            return np.array([[cm.cost(a, b) for b in range(32)] for a in range(32)], np.int64)

        def _cells(self, a, b, dw=None):
            l1, l2 = max(len(a), len(b)), min(len(a), len(b))
            return po.algn_nw_cells(l1, l2, po.deltaw(a, b, cm, dw))

        def median(self, pairs):
            out = []
            for a, b in pairs:
                self._tick()
                cost, oa, ob = po.align_2(a, b, cm)
                self.cells += self._cells(a, b)
                out.append((po.ancestor_2(oa, ob, cm), cost))
            return out

        def align(self, pairs):
            out = []
            for a, b in pairs:
                self._tick()
                cost, oa, ob = po.align_2(a, b, cm)
                self.cells += self._cells(a, b)
                out.append((oa, ob))
            return out

        def distance(self, pairs):
            out = []
            for a, b in pairs:
                self._tick()
                dw = max(8, abs(len(a) - len(b)))
                out.append(po.cost_2(a, b, cm, dw))
                self.cells += self._cells(a, b, dw)
            return out

        def fitch(self, a, b):
            m, c = zip(*[po.nonadd_median(x, y) for x, y in zip(a, b)])
            return np.array(m, np.uint8), np.array(c, np.int64)

    try:
        old_aff = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {min(old_aff)})
    except (AttributeError, OSError):
        old_aff = None
    be = CpuBackend()
    done = "whole diagnosis"
    try:
        treepass.TreeEval(ds, tree, be).diagnose()
    except Stop:
        done = "part of the diagnosis"
    dt = time.perf_counter() - be.t0
    if old_aff is not None:
        os.sched_setaffinity(0, old_aff)
    return {
        "value": be.cells / dt / 1e9,
        "unit": "GCUPS",
        "cores": 1,
        "kind": "port",
        "sample": f"{done} of the built {ds.n_taxa}-taxon tree on the oracle: {be.cells} band cells in "
                  f"{dt:.1f} s, 1 pinned core of {cpu_model()}",
    }


def measured_traffic(kind: str):
    """HBM bytes per step of the fill stage from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled by a calibration copy), or None."""
    if kind != "affine":
        return None
    path = os.path.join(ROOT, "profiles", "affine_65536_fill_traffic_pmc.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_step:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def copy_calibration(mib: int) -> int:
    """One plain device copy of known size (reads and writes mib MiB)."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "checksum_ok": bool(dst[-1].item() == 1)}), flush=True)
    return 0


def load_stream(ctx):
    from poyd_amd import engine
    return engine.load().poyg_ctx_stream(ctx.h)


if __name__ == "__main__":
    sys.exit(main())
