"""Autotuner search logic on the CPU (probe replaced by a synthetic rate surface) and, on the
GPU, that a tuned setting scans the same survivors as the default one."""
from __future__ import annotations

import pytest

import es_autotune as AT
import es_plan as EP


def _surface(best):
    def probe(plan, tune, t, lo, n, repeats=2):
        # separable, single peak at `best`; every coordinate matters
        r = 1.0
        for x, v in tune.items():
            opts = AT.SPACE[x]
            r *= 1.0 / (1 + abs(opts.index(v) - opts.index(best[x])))
        return r
    return probe


@pytest.mark.parametrize("method", ["coord", "grid"])
def test_search_finds_peak(monkeypatch, method):
    best = {"mi_strip_target": 4096, "threads": 128, "mi_masks": 8, "mi_global_masks": 0,
            "mi_ring_swap": 0}
    monkeypatch.setattr(AT, "probe", _surface(best))
    got, trials = AT.autotune(30, probe_log2=12, method=method)
    assert got == best
    n_grid = 1
    for opts in AT.SPACE.values():
        n_grid *= len(opts)
    if method == "coord":
        assert len(trials) < 45                  # far fewer probes than the full grid
    else:
        assert len(trials) in (n_grid, n_grid // 2)


def test_coord_keeps_default_within_margin(monkeypatch):
    monkeypatch.setattr(AT, "probe", lambda *a, **k: 1.0)   # flat surface
    got, _ = AT.autotune(30, probe_log2=12, method="coord")
    plan = EP.Plan(30)
    assert got == {x: plan.tun[x] for x in AT.SPACE}


def test_timed_rate_windows():
    """Layout timing: windows grow from 2^24 positions toward the target at a fixed device
    rate, stop once the next window would not grow by a quarter, and a layout below the
    floor stops after its second window."""
    import gpu_sieve_mi355x as G
    calls = []

    def scan_at(rate):
        def scan(lo, hi):
            calls.append(hi - lo)
            return rate
        return scan
    R = 1 << 50
    assert G.timed_rate(scan_at(1e12), R, 1 << 16, max_s=0.05) == 1e12
    assert calls[0] == 1 << 24 and calls[-1] == 1 << 34 and len(calls) == 2
    calls.clear()
    G.timed_rate(scan_at(1e9), R, 1 << 16, max_s=0.05)      # 5e7 positions fit in 50 ms
    assert calls == [1 << 24, int(1e9 * 0.05)]
    calls.clear()
    G.timed_rate(scan_at(1e9), R, 1 << 16, max_s=0.05, floor=2e9)
    assert len(calls) == 2
    calls.clear()
    G.timed_rate(scan_at(1e9), R, 1 << 16, max_s=0.1, floor=2e9)
    assert len(calls) == 2                                    # would grow, but below the floor


def test_cache_roundtrip(monkeypatch, tmp_path):
    calls = []

    def probe(*a, **k):
        calls.append(1)
        return 1.0
    monkeypatch.setattr(AT, "probe", probe)
    monkeypatch.setattr(AT.G.nat, "device_query", lambda: {"name": "cpu-test"})
    cache = str(tmp_path / "tune.json")
    a, _ = AT.autotune(30, probe_log2=12, cache=cache)
    n = len(calls)
    b, _ = AT.autotune(30, probe_log2=12, cache=cache)
    assert a == b and len(calls) == n


@pytest.mark.gpu
def test_tuned_setting_same_survivors():
    import gpu_sieve_mi355x as G
    k = 61
    space = {"mi_strip_target": [64, 1024], "threads": [64, 256], "mi_masks": [6, 8],
             "mi_global_masks": [0]}
    best, trials = AT.autotune(k, space=space, probe_log2=26)
    assert len(trials) >= 4 and all(t["rate"] > 0 for t in trials)
    a = G.GpuSieve(EP.Plan(k))
    b = G.GpuSieve(EP.Plan(k, best))
    R = a.plan.R
    lo, hi = R // 4, min(R, R // 4 + (1 << 26))
    sa, _ = a.scan_window(3, lo, hi)
    sb, _ = b.scan_window(3, lo, hi)
    assert sorted(sa) == sorted(sb)
