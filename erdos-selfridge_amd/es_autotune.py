"""Autotuner of the sieve launch shape for one wheel.

Searches strip-mine length (mi_strip_target), workgroup size (threads), LDS table layout
(number of word masks, masks served through L1) and the lead-ring swap (mi_ring_swap) by
timing a probe window on the GPU (device
time from HIP events), and returns the fastest setting.  The setting changes speed only: every
candidate setting scans the same positions with the same exact predicate.  Results can be kept
in a JSON cache keyed by (k, wheel, device).

Two search methods: "grid" times every point of the space; "coord" (coordinate descent) starts
from the default setting and sweeps one dimension at a time with the others held, keeping a
change only when it beats the current best by more than `margin`, until a full sweep changes
nothing -- 14 probes per sweep over the default space instead of 600.
"""
from __future__ import annotations

import copy
import itertools
import json
import os

import es_plan
import gpu_sieve_mi355x as G

SPACE = {
    "mi_strip_target": [64, 256, 1024, 4096, 16384],
    "threads": ["auto", 64, 128, 256, 512],
    "mi_masks": ["auto", 4, 6, 8],
    "mi_global_masks": [0, 2],
    "mi_ring_swap": ["auto", 0, 1],
}


def _key(k, wheel, device):
    return f"{k}|{json.dumps([[p, T] for p, T in wheel])}|{device}"


def probe(plan, tune: dict, t: int, lo: int, n: int, repeats: int = 2) -> float:
    """Positions per second of one setting on positions [lo, lo+n) of block t (best of
    `repeats` timed scans after a warm-up scan)."""
    p = copy.copy(plan)
    p.tun = dict(plan.tun, **tune)
    eng = G.GpuSieve(p)
    eng.scan_window(t, lo, lo + min(n, 1 << 26))        # warm the code object and caches
    eng.sync()
    eng.kernel_time()
    best = 0.0
    for _ in range(max(1, repeats)):
        eng.scan_window(t, lo, lo + n)
        ms, _, pos = eng.kernel_time()
        best = max(best, pos / (ms / 1e3))
    return best


def _allowed(plan, tune) -> bool:
    # L1-served masks exist for ring-0 radix <= 32 only
    return not (tune.get("mi_global_masks", 0) and plan.radix and plan.radix[0] > 32)


def autotune(k: int, wheel=None, space: dict | None = None, probe_log2: int = 30, t: int = 1,
             cache: str | None = None, log=None, method: str = "coord",
             margin: float = 0.02) -> tuple[dict, list]:
    """Best setting for k's wheel (the default planner's when wheel is None) and the trials."""
    space = space or SPACE
    plan = es_plan.Plan(k, None, wheel)
    wheel = [[p, T] for p, T in plan.wheel]
    dev = G.nat.device_query()["name"] if cache else ""
    if cache and os.path.exists(cache):
        db = json.load(open(cache))
        hit = db.get(_key(k, wheel, dev))
        if hit:
            return hit["best"], hit["trials"]
    n = min(1 << probe_log2, plan.R)
    lo = (plan.R - n) // 2
    trials = []
    seen = {}

    def rate_of(tune):
        key = json.dumps(tune, sort_keys=True)
        if key not in seen:
            seen[key] = probe(plan, tune, t, lo, n)
            trials.append({"tune": dict(tune), "rate": seen[key]})
            if log:
                log(f"k={k} {tune} {seen[key]:.3e} positions/s")
        return seen[key]

    names = list(space)
    if method == "grid":
        for vals in itertools.product(*(space[x] for x in names)):
            tune = dict(zip(names, vals))
            if _allowed(plan, tune):
                rate_of(tune)
        best = max(trials, key=lambda x: x["rate"])["tune"]
    elif method == "coord":
        best = {x: plan.tun.get(x, space[x][0]) for x in names}
        best_rate = rate_of(best)
        changed = True
        while changed:
            changed = False
            for x in names:
                for v in space[x]:
                    tune = dict(best, **{x: v})
                    if tune == best or not _allowed(plan, tune):
                        continue
                    r = rate_of(tune)
                    if r > best_rate * (1 + margin):
                        best, best_rate, changed = tune, r, True
    else:
        raise ValueError(f"unknown method {method!r}")
    if cache:
        db = json.load(open(cache)) if os.path.exists(cache) else {}
        db[_key(k, wheel, dev)] = {"best": best, "trials": trials}
        json.dump(db, open(cache, "w"), indent=1)
    return best, trials
