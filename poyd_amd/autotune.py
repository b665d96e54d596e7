"""Launch-shape autotuner of the batched plans.

The plans' results never depend on these knobs (poyg_ctx_set_tuning: which
affine bands go to the four-pairs-per-wave fill, how many CUs the side
streams get, the packed or 32-bit fill of the narrow affine bands, one or
four waves per pair of the packed linear fill, resident workgroups per CU of
the linear fill and backtrace, waves per CU of the three-way cube); their speed does, per device and per
batch shape.  `tune` times trials of a caller's batch (coordinate descent
over the knobs of its kind, each trial the median of a few plan runs) and
stores the best settings per device, kind and length bucket in a JSON table;
`apply` sets a context's knobs from that table.  The table shipped with the
package (tuned.json, made by `tune` on an MI355X) is read first; a table in
$POYD_AMD_TUNE_CACHE (default ~/.cache/poyd_amd/tuned.json) overrides it.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SHIPPED = os.path.join(HERE, "tuned.json")

KNOBS = {
    "affine": [("aff_quad_max_lanes", [44, 46, 48]), ("aff_side_div", [3, 4, 6, 8, 16]),
               ("aff_narrow_kernel", [1])],
    "linear": [("lin_cost_kernel", [3, 4]), ("lin_sys_blocks_per_cu", [1, 2, 3, 4]),
               ("lin_trace_blocks_per_cu", [2, 4, 7])],
    "cube": [("cube_waves_per_cu", [4, 6, 8, 12, 16])],
}


def cache_path() -> str:
    return os.environ.get("POYD_AMD_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "poyd_amd",
                                                              "tuned.json"))


def bucket(lengths) -> str:
    """Length bucket of a batch: its median sequence length."""
    m = float(np.median(np.asarray(lengths))) if len(lengths) else 0.0
    return "short" if m < 500 else ("medium" if m < 2000 else "long")


def _load(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def table() -> dict:
    t = _load(SHIPPED)
    for dev, kinds in _load(cache_path()).items():
        t.setdefault(dev, {}).update(kinds)
    return t


def lookup(ctx, kind: str, lengths) -> dict:
    return table().get(ctx.device_name(), {}).get(f"{kind}/{bucket(lengths)}", {})


def apply(ctx, kind: str, lengths) -> dict:
    """Set the context's knobs for a batch of this kind and lengths from the
    tuned table (knobs it does not list keep their defaults)."""
    got = lookup(ctx, kind, lengths)
    for knob, _ in KNOBS[kind]:
        ctx.set_tuning(knob, int(got.get(knob, -1)))
    return got


def reset(ctx, kind: str) -> None:
    for knob, _ in KNOBS[kind]:
        ctx.set_tuning(knob, -1)


def tune(ctx, kind: str, lengths, trial, repeats: int = 3, save: bool = True, log=None) -> dict:
    """Coordinate descent over the knobs of `kind`: trial() builds and runs
    the caller's plan with the context's current knobs and returns its time
    in ms; each setting's time is the median of `repeats` trials.  Returns
    {"settings", "ms", "default_ms", "trials"} and stores the settings."""
    reset(ctx, kind)

    def timed():
        return float(np.median([trial() for _ in range(repeats)]))

    default_ms = timed()
    best = {}
    best_ms = default_ms
    n_trials = repeats
    for knob, values in KNOBS[kind]:
        for v in values:
            ctx.set_tuning(knob, v)
            ms = timed()
            n_trials += repeats
            if log:
                log(f"tune {kind}: {knob}={v} -> {ms:.3f} ms (best {best_ms:.3f})")
            if ms < best_ms * 0.98:   # a clear gain, not noise
                best_ms, best[knob] = ms, v
        ctx.set_tuning(knob, best.get(knob, -1))
    if save:
        path = cache_path()
        os.makedirs(os.path.dirname(path), exist_ok=True)
        t = _load(path)
        t.setdefault(ctx.device_name(), {})[f"{kind}/{bucket(lengths)}"] = dict(best)
        with open(path, "w") as f:
            json.dump(t, f, indent=1, sort_keys=True)
    return {"settings": best, "ms": best_ms, "default_ms": default_ms, "trials": n_trials}
