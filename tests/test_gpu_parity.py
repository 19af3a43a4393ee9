"""GPU parity: survivors of the HIP engine (through the C ABI) against the CPU oracle.

Bar: bit-exact survivor ID sets (compared sorted), exact minima, b-file values.  Every case
prints its name, plan and window before touching the GPU."""
from __future__ import annotations

import json
import os
import random

import numpy as np
import pytest

import es_oracle as O
import es_plan as EP
from conftest import GOLDEN, bfile

pytestmark = pytest.mark.gpu

C1_WHEEL = [[2, 5], [3, 4], [5, 3], [7, 2], [11, 2], [13, 2]]
K399_PRIMARY = [[107, 1], [3, 6], [211, 1], [139, 1], [83, 1], [59, 1], [103, 1], [137, 1],
                [41, 1], [2, 9], [37, 1], [29, 1], [101, 1], [13, 1], [31, 1], [67, 1], [5, 2]]
K399_CONFIRM = [[223, 1], [3, 6], [211, 1], [139, 1], [83, 1], [59, 1], [103, 1], [137, 1],
                [5, 4], [2, 9], [37, 1], [29, 1], [101, 1], [13, 1], [31, 1], [67, 1]]


def engine(k, wheel=None, **tune):
    import gpu_sieve_mi355x as G
    plan = EP.Plan(k, dict(tune), wheel)
    return G.GpuSieve(plan), plan


def say(*a):
    print(*a, flush=True)


def window_check(k, wheel, t, lo, hi, limit=None, **tune):
    if limit is not None:
        tune["mi_filter_limit"] = limit
    tune.setdefault("max_survivors", 1 << 22)
    eng, plan = engine(k, wheel, **tune)
    filt = plan.rest if limit is None else plan.rest[:limit]
    say(f"case k={k} t={t} window=[{lo},{hi}) filters={len(filt)} masks={eng.info.n_masks_used} "
        f"strip={eng.strip_len} r0={eng.r0}")
    ids, mn = eng.scan_window(t, lo, hi)
    got = np.sort(np.asarray(ids, dtype=np.uint64))
    n, want, want_min = O.c_scan_min(k, plan.wheel, filt, t, lo, hi, threads=16, cap=1 << 22)
    assert n == len(want)
    assert len(got) == n, f"count {len(got)} != oracle {n}"
    assert np.array_equal(got, want)
    # the device-side minimum over n > k+1 equals the oracle's
    assert mn == want_min, f"minimum {mn} != oracle {want_min}"
    return n


def test_config1_window_dense():
    assert window_check(30, C1_WHEEL, 0, 0, 1 << 24) > 100000


def test_config1_unaligned_windows():
    rng = random.Random(1)
    for _ in range(4):
        lo = rng.randrange(0, 230947200 - (1 << 20))
        hi = lo + rng.randrange(1, 1 << 20)
        window_check(30, C1_WHEEL, rng.randrange(0, 50), lo, hi)


def test_later_block_minimum_counts_every_residue():
    """In block t >= 1 every survivor counts for the minimum, s <= k+1 included: block 14 of the
    C1 wheel starts with a survivor at idx 0, n = 14 M + 30."""
    eng, plan = engine(30, C1_WHEEL)
    assert window_check(30, C1_WHEEL, 14, 0, 1 << 16) > 0
    _, mn = eng.scan_window(14, 0, 1 << 16)
    assert mn == 14 * plan.M + 30
    _, mn0 = eng.scan_window(0, 0, 1 << 16)   # block 0: s <= 31 left out
    assert mn0 is None or mn0 > 31


def test_minimum_needs_the_block():
    """Constants uploaded as tconst (block unknown): IDs still collect, the minimum is refused."""
    import ctypes as C
    import es_native as nat
    eng, plan = engine(30, C1_WHEEL)
    say("case k=30 tconst upload, minimum refused")
    tc = np.asarray([((3 % q) * (plan.M % q)) % q for q in eng.rest_q], dtype=np.uint32)
    L = nat.lib()
    nat.check(L.es_engine_set_tconst(eng._h, nat.ptr32(tc), len(tc)))
    nat.check(L.es_engine_reset(eng._h))
    nat.check(L.es_engine_scan_window(eng._h, 0, 1 << 16))
    n = C.c_uint64(0)
    ids = np.zeros(1 << 16, dtype=np.uint64)
    nat.check(L.es_engine_collect(eng._h, nat.ptr64(ids), 1 << 16, C.byref(n), None, None))
    want_n, want = O.c_scan(30, plan.wheel, plan.rest, 3, 0, 1 << 16, threads=4)
    assert n.value == want_n and np.array_equal(np.sort(ids[:n.value]), want)
    ms = np.zeros(2, dtype=np.uint64)
    has = C.c_int(0)
    rc = L.es_engine_collect(eng._h, nat.ptr64(ids), 1 << 16, C.byref(n), nat.ptr64(ms), C.byref(has))
    assert rc == nat.ES_ERR_INVALID and b"block index" in L.es_last_error()
    _, mn = eng.scan_window(3, 0, 1 << 16)     # set_block makes the block known again
    assert mn == O.c_scan_min(30, plan.wheel, plan.rest, 3, 0, 1 << 16, threads=4)[2]


def test_config1_full_block_min():
    eng, plan = engine(30, C1_WHEEL, max_survivors=1 << 24)
    say("case k=30 full block 0")
    ids, mn = eng.scan_window(0, 0, plan.R)
    n, want = O.c_scan(30, plan.wheel, plan.rest, 0, 0, plan.R, threads=16, cap=1 << 24)
    assert len(ids) == n == 12654618
    assert np.array_equal(np.sort(np.asarray(ids, dtype=np.uint64)), want)
    assert mn == 58782


@pytest.mark.parametrize("k", [100, 200, 399])
def test_filter_prefix_windows(k):
    """Only the first few filters: dense survivors at large k, random block and offset."""
    rng = random.Random(k)
    wheel = K399_PRIMARY if k == 399 else None
    plan = EP.Plan(k, None, wheel)
    span = min(1 << 22, plan.R // 2)
    for limit in (3, 5):
        t = rng.randrange(0, 1 << 20)
        lo = rng.randrange(0, plan.R - span)
        window_check(k, wheel, t, lo, lo + span, limit=limit)


def test_k399_confirmation_prefix():
    plan = EP.Plan(399, None, K399_CONFIRM)
    rng = random.Random(7)
    lo = rng.randrange(0, plan.R - (1 << 22))
    window_check(399, K399_CONFIRM, 3, lo, lo + (1 << 22), limit=4)


@pytest.mark.parametrize("ks", [(1, 100), (101, 200)])
def test_bfile_gate(ks):
    """GpuSieve.search (default planner) returns the b-file g(k) (harness/gate.py:5-17)."""
    import gpu_sieve_mi355x as G
    b = bfile()
    for k in range(ks[0], ks[1] + 1):
        plan = EP.Plan(k, {"max_survivors": 1 << 24})
        say(f"case gate k={k} M={plan.M} R={plan.R}")
        res = G.GpuSieve(plan).search(max_blocks=400)
        assert res["status"] == "FOUND", res
        assert res["g"] == b[k], (k, res["g"], b[k])


@pytest.mark.parametrize("ks", [(201, 240), (241, 280), (281, 320), (321, 360), (361, 400)])
def test_default_wheels_every_k(ks):
    """Every default-planner wheel of k = 201..400, as the engine sets it up by itself (its own
    layout timing, masks and ring order), scans a seeded window of 2^22 positions in a seeded
    block to the oracle's survivors and minimum.  The first 8 filters leave 15 to several
    hundred survivors per window (median 124, 28,983 over the 200 windows), so the masks, the
    carry forms and the exact re-test all carry weight."""
    for k in range(ks[0], ks[1] + 1):
        plan = EP.Plan(k, None)
        rng = random.Random(1000 + k)
        span = min(1 << 22, plan.R // 2)
        lo = rng.randrange(0, plan.R - span)
        window_check(k, None, rng.randrange(0, 1 << 20), lo, lo + span, limit=8)


def test_random_tunables():
    """150 seeded random cases: k in [30, 400], the default wheel under random engine tunables
    (layout, mask count, workgroup size, strip target, ring order, masks in global memory,
    filter prefix) and a random window of 1 to 2^22 positions in a random block, against the
    oracle's survivors and minimum.  The only set-up allowed to fail is a forced row-lane layout
    that the wheel cannot take, with the engine's own message."""
    import gpu_sieve_mi355x as G
    rng = random.Random(20261204)
    ran = 0
    for i in range(150):
        k = rng.randrange(30, 401)
        tune = {"max_survivors": 1 << 22}
        if rng.random() < 0.5:
            tune["mi_layout"] = rng.choice(["auto", "strips", "rows"])
        if rng.random() < 0.4:
            tune["mi_masks"] = rng.randrange(1, 9)
        if rng.random() < 0.3:
            tune["threads"] = rng.choice([256, 512, 1024])
        if rng.random() < 0.3:
            tune["mi_strip_target"] = rng.choice([256, 512, 2048, 4096])
        if rng.random() < 0.3:
            tune["mi_ring_swap"] = rng.choice(["auto", 0, 1])
        if rng.random() < 0.15:
            tune["mi_global_masks"] = rng.choice([0, 2, 4])
        limit = rng.choice([None, 3, 5, 8, 12])
        if limit is not None:
            tune["mi_filter_limit"] = limit
        plan = EP.Plan(k, dict(tune))
        span = min(rng.choice([1, 100, 4096, 1 << 16, 1 << 20, 1 << 22]), plan.R)
        lo = rng.randrange(0, plan.R - span + 1)
        t = rng.randrange(0, 1 << 20)
        say(f"case random {i} k={k} t={t} window=[{lo},{lo + span}) tune={tune}")
        try:
            eng = G.GpuSieve(plan)
        except AssertionError as e:   # the engine's set-up errors, as the reference's asserts
            assert tune.get("mi_layout") == "rows" and "row-lane layout needs" in str(e), e
            continue
        ids, mn = eng.scan_window(t, lo, lo + span)
        filt = plan.rest if limit is None else plan.rest[:limit]
        n, want, wmin = O.c_scan_min(k, plan.wheel, filt, t, lo, lo + span, threads=16,
                                     cap=1 << 22)
        got = np.sort(np.asarray(ids, dtype=np.uint64))
        assert len(got) == n and np.array_equal(got, want)
        assert mn == wmin
        ran += 1
    assert ran >= 120


def _frontier():
    return json.load(open(os.path.join(GOLDEN, "runs_frontier.json")))["terms"]


def _recorded_runs():
    """(k, role, mi_ring_swap): both forced orders for a wheel whose ring 0 could trade places
    with a later ring (the engine's timing picks one of them), the default otherwise."""
    import gpu_sieve_mi355x as G
    out = []
    for t in _frontier():
        for role in ("primary", "confirmation"):
            P = EP.Plan(t["k"], None, t[role]["wheel"])
            opts = G.GpuSieve._layout_options(P)
            swaps = sorted({s for _, s, _ in opts})
            swaps = swaps if len(swaps) > 1 else ["auto"]
            out += [(t["k"], role, s) for s in swaps]
    return out


@pytest.mark.parametrize("k,role,swap", _recorded_runs())
def test_recorded_survivor_windows(k, role, swap):
    """Every recorded run of evidence/through400/runs.json (46 wheels: both roles of k=378..400),
    in both kernel ring orders where its ring 0 could give way to a lead block: around every
    recorded survivor of every block, a 2^24-position window returns exactly the recorded
    survivors inside it and the oracle's survivors, and its minimum agrees with the oracle's
    (full filter set)."""
    term = next(x for x in _frontier() if x["k"] == k)
    run = term[role]
    # a forced order also keeps ring 0 whole (k=378's 3^6 ring could otherwise be timed as
    # digit rings, which have no lead block)
    eng, plan = engine(k, run["wheel"], max_survivors=1 << 16,
                       **({} if swap == "auto" else {"mi_ring_swap": swap, "mi_inner_radix": 128}))
    assert str(plan.M) == run["N"] and plan.R == run["residues_per_block"]
    if swap != "auto":
        assert (eng._d.lead > 0) == (swap == 1)
    windows = 0
    for audit in run["block_audit"]:
        t = audit["t"]
        listed = audit["survivors_n"] or []
        recorded = sorted(plan.n_to_idx(int(x))[0] for x in listed)
        for sv in listed:
            idx, tt = plan.n_to_idx(int(sv))
            assert tt == t
            lo = max(0, idx - (1 << 23))
            hi = min(plan.R, lo + (1 << 24))
            say(f"case k={k} {role} t={t} survivor idx={idx} window=[{lo},{hi}) "
                f"masks={eng.info.n_masks_used} r0={eng.info.r0}")
            ids, mn = eng.scan_window(t, lo, hi)
            assert sorted(ids) == [i for i in recorded if lo <= i < hi]
            cnt, want, want_min = O.c_scan_min(k, plan.wheel, plan.rest, t, lo, hi, threads=16)
            assert sorted(ids) == [int(x) for x in want] and mn == want_min
            windows += 1
    assert windows >= 1
    # the recorded answer's window returns g(k) as its minimum
    g = int(term["n"])
    gi, gt = plan.n_to_idx(g)
    assert gt == run["block"]


@pytest.mark.parametrize("k", [270, 275, 280])
def test_full_block_known_answer(k):
    """Block 0 of the default-planner wheel: recorded survivor count and g
    (evidence/logs/2026-08-22/logs/gate_v9_270_280.log)."""
    rec = json.load(open(os.path.join(GOLDEN, "gate_v9_270_280.json")))[str(k)]
    eng, plan = engine(k, rec["plan"]["wheel"])
    assert plan.R == rec["residues_checked"]
    say(f"case full block k={k} R={plan.R}")
    ids, mn = eng.scan_window(0, 0, plan.R)
    ns = sorted(plan.idx_to_n(i, 0) for i in ids)
    assert len(ns) == rec["survivors_in_block"]
    assert all(O.is_good(n, k) or n <= k + 1 for n in ns)
    good = [n for n in ns if O.is_good(n, k)]
    assert good[0] == int(rec["g"]) == mn


def test_lds_short_masks_stay_exact(monkeypatch):
    """Masks the library cannot stage (LDS budget) must fall back to the exact re-test:
    the caller's slow list omits fully covered filters, the library appends them."""
    import gpu_sieve_mi355x as G
    orig = G.choose_masks
    monkeypatch.setattr(G, "choose_masks",
                        lambda plan, n, limit=None, lds_budget=0, r0=None, q0=None, stop=0.0,
                        rows=False: orig(plan, n, limit, 1 << 30, r0, q0, stop, rows=rows))
    for k, limit in ((100, 5), (101, None)):
        plan = EP.Plan(k)
        span = min(1 << 21, plan.R)
        n = window_check(k, None, 3, 0, span, limit=limit, mi_lds_budget=8192)
        say(f"  k={k} survivors={n}")


def test_tconst_paths_agree():
    """tconst uploaded by a driver (sieve/gpu_sieve_v9.py:512-514) == es_engine_set_block."""
    import ctypes as C
    import es_native as nat
    eng, plan = engine(399, K399_PRIMARY, mi_filter_limit=4, max_survivors=1 << 22)
    lo = 123456789
    for t in (0, 1, 6, 1 << 19):
        a, _ = eng.scan_window(t, lo, lo + (1 << 22))          # Python tconst path
        nat.check(nat.lib().es_engine_set_block(eng._h, t))
        nat.check(nat.lib().es_engine_reset(eng._h))
        nat.check(nat.lib().es_engine_scan_window(eng._h, lo, lo + (1 << 22)))
        b, _ = eng.collect(t)
        assert sorted(a) == sorted(b) and len(a) > 0


def test_scan_outer_units_equal_window():
    """_launch(ostart, ocount) over outer units == scan_window over the same positions."""
    eng, plan = engine(200, None, mi_filter_limit=3, max_survivors=1 << 22)
    eng.set_block(5)
    eng.out_n.fill(0)
    o0 = plan.R // eng.r0 // 3
    eng._launch(o0, 100000)
    eng._launch(o0 + 100000, 50000)
    n = int(eng.out_n.get()[0])
    a = sorted(int(x) for x in eng.out[:n].get())
    b, _ = eng.scan_window(5, o0 * eng.r0, (o0 + 150000) * eng.r0)
    assert a == sorted(b) and n > 0


def test_overflow_is_reported():
    eng, plan = engine(30, C1_WHEEL, max_survivors=1000)
    with pytest.raises(RuntimeError, match="overflow"):
        eng.scan_window(0, 0, 1 << 20)


@pytest.mark.parametrize("k", [100, 399])
def test_mask_counts_agree(k):
    """Every mask count scans the same survivors: 0 masks (every ring-0 slot queued for the
    exact re-test), 1, 3, the automatic count and 12 (dense prefix windows, against the
    oracle)."""
    wheel = K399_PRIMARY if k == 399 else None
    plan = EP.Plan(k, None, wheel)
    rng = random.Random(k + 1)
    span = min(1 << 21, plan.R // 2)
    lo = rng.randrange(0, plan.R - span)
    for masks in (0, 1, 3, "auto", 12):
        window_check(k, wheel, 2, lo, lo + span, limit=6, mi_masks=masks)


def test_masks_not_prime_to_M_are_skipped(monkeypatch):
    """A mask of a wheel prime (its modulus shares a factor with M) has no row order in which
    a wrap past M is one row on; the library must leave it unstaged and keep its filter in the
    exact re-test, with survivors unchanged."""
    import gpu_sieve_mi355x as G
    plan = EP.Plan(30)
    wheel_p = [p for p in plan.rest if plan.wheelT.get(p, 0) and plan.wheelT[p] < plan.tp[p]]
    assert wheel_p, "k=30 default plan has a partial wheel prime"
    p = wheel_p[0]
    fi = plan.rest.index(p)
    extra = (fi, plan.wheelT[p] + 1, p ** (plan.wheelT[p] + 1), 0.5)
    orig = G.choose_masks
    monkeypatch.setattr(G, "choose_masks", lambda *a, **kw: [extra] + orig(*a, **kw))
    eng, _ = engine(30, None)
    offered = len(eng._d.masks)
    say(f"case k=30 extra mask {extra[2]} (p={p} in the wheel): offered {offered}, "
        f"staged {eng.info.n_masks_used}")
    assert eng._d.masks[0] == extra and eng.info.n_masks_used == offered - 1
    window_check(30, None, 1, 0, min(1 << 22, plan.R))


@pytest.mark.parametrize("k", [301, 399])
def test_auto_workgroup_size(k):
    """threads="auto" picks the workgroup size of 256, 512, 1024 that keeps the most waves
    resident per CU: k=301's default wheel, kept in the plan's ring order, stages 55-62 KB of
    two-word mask rows, which at 256 threads leaves two workgroups (8 waves) per CU, so a
    larger size must win there.  Every size scans the oracle's survivors."""
    wheel = K399_PRIMARY if k == 399 else None
    keep = {"mi_ring_swap": 0} if k == 301 else {}
    eng, plan = engine(k, wheel, **keep)
    say(f"case k={k} auto threads={eng.threads} lds={eng.info.lds_bytes} r0={eng.r0}")
    assert eng.threads in (256, 512, 1024)
    if k == 301:
        assert eng.threads >= 512 and 163840 // eng.info.lds_bytes == 2
    rng = random.Random(k)
    span = min(1 << 22, plan.R // 2)
    lo = rng.randrange(0, plan.R - span)
    for th in ("auto", 256, 1024):
        window_check(k, wheel, 1, lo, lo + span, limit=8, threads=th, **keep)


@pytest.mark.parametrize("k,tune", [(399, {}), (399, {"mi_global_masks": 2}),
                                    (300, {"mi_layout": "strips"}), (300, {"mi_layout": "rows"})])
def test_carry_table_forms(k, tune):
    """Both ring-0 carry forms of one-word rows scan the oracle's survivors: the byte table of
    suffix starts (k=399, q0 = 107 <= 256), word rows (the same wheel with two masks served
    from global memory, an instance without the byte form), word rows by size in the
    strip-lane layout (k=300, q0 = 2^9 > 256) and the byte table at that size in the row-lane
    layout."""
    wheel = K399_PRIMARY if k == 399 else None
    plan = EP.Plan(k, None, wheel)
    if "mi_layout" in tune:
        eng, _ = engine(k, wheel, mi_filter_limit=8, **tune)
        assert eng.layout == tune["mi_layout"] and eng.info.q0 > 256 and eng.info.words == 1
    rng = random.Random(7 * k + len(str(tune)))
    span = min(1 << 22, plan.R // 2)
    lo = rng.randrange(0, plan.R - span)
    window_check(k, wheel, 3, lo, lo + span, limit=8, **tune)


@pytest.mark.parametrize("k", [301, 309, 312, 378, 363, 386, 372, 383])
def test_swapped_lead_rings(k):
    """A wheel whose ring 0 has 33..128 residues runs with its first ring j of at most 32
    residues as the kernel's ring 0 and rings 0..j-1 as a lead block behind it (j = 1, or 2 at
    k=363, 383 and 386), one-word mask rows, while idx keeps the plan's order: unaligned
    windows (with their minimum), outer-unit launches and the unswapped engine all give the
    oracle's survivors.  Both orders are forced (mi_ring_swap 1 and 0), whatever the engine's
    timing would choose; k=372 and 383 are wheels the fill rule alone would not swap.  (k=378's
    ring 0 is 3^6 with 108 residues: the inner radix is fixed at 128 to keep it whole.)"""
    import gpu_sieve_mi355x as G
    keep = {"mi_inner_radix": 128} if k == 378 else {}
    keep["mi_ring_swap"] = 1
    eng, plan = engine(k, None, mi_filter_limit=8, max_survivors=1 << 22, **keep)
    say(f"case k={k} radix {plan.radix[:3]} kernel r0={eng._d.r0} lead={eng._d.lead}")
    j = G.swap_lead(plan.radix, force=True)
    assert j and eng._d.lead == np.prod(plan.radix[:j]) and eng._d.r0 == plan.radix[j] <= 32
    assert (eng.inner_bound, eng.ring_swap) == (keep.get("mi_inner_radix", 128), 1)
    rng = random.Random(k)
    for span in (1 << 22, 12345, 1):
        lo = rng.randrange(0, plan.R - span)
        window_check(k, None, 2, lo, lo + span, limit=8, **keep)
    eng.set_block(4)
    eng.out_n.fill(0)
    o0 = plan.R // eng.r0 // 3
    eng._launch(o0, 70001)
    eng._launch(o0 + 70001, 3)
    n = int(eng.out_n.get()[0])
    a = np.sort(np.asarray([int(x) for x in eng.out[:n].get()], dtype=np.uint64))
    lo, hi = o0 * eng.r0, (o0 + 70004) * eng.r0
    nw, want = O.c_scan(k, plan.wheel, plan.rest[:8], 4, lo, hi, threads=16, cap=1 << 22)
    assert n == nw and np.array_equal(a, want)
    b, _ = eng.scan_window(4, lo, hi)
    assert np.array_equal(np.sort(np.asarray(b, dtype=np.uint64)), want)
    keep["mi_ring_swap"] = 0
    eng2, _ = engine(k, None, mi_filter_limit=8, max_survivors=1 << 22, **keep)
    assert eng2._d.lead == 0
    c, _ = eng2.scan_window(4, lo, hi)
    assert np.array_equal(np.sort(np.asarray(c, dtype=np.uint64)), want)


@pytest.mark.parametrize("k,wheel", [(399, K399_CONFIRM), (383, None), (400, None)])
def test_timed_ring_order(k, wheel):
    """With mi_ring_swap "auto" the engine times both kernel ring orders of a wheel whose ring 0
    could give way to a lead block (k=399's confirmation wheel, 47 then 24; k=383 and 400's
    default wheels, lead blocks of two rings) and keeps the faster: its descriptor follows the
    choice, and the chosen engine scans the oracle's survivors."""
    import gpu_sieve_mi355x as G
    tune = {"mi_filter_limit": 8, "max_survivors": 1 << 22}
    plan = EP.Plan(k, dict(tune), wheel)
    assert len(G.GpuSieve._layout_options(plan)) >= 2
    eng = G.GpuSieve(plan)
    say(f"case timed ring order k={k}: swap={eng.ring_swap} lead={eng._d.lead} "
        f"layout={eng.layout} r0={eng.info.r0}")
    assert eng.ring_swap in (0, 1) and (eng._d.lead > 0) == (eng.ring_swap == 1)
    rng = random.Random(31 * k)
    span = min(1 << 22, plan.R // 2)
    lo = rng.randrange(0, plan.R - span)
    ids, mn = eng.scan_window(5, lo, lo + span)
    n, want, want_min = O.c_scan_min(k, plan.wheel, plan.rest[:8], 5, lo, lo + span, threads=16,
                                     cap=1 << 22)
    assert sorted(ids) == [int(x) for x in want] and mn == want_min


@pytest.mark.parametrize("k", [100, 309])
def test_window_edges(k):
    """Edge windows against the oracle: empty, the block's first and last positions, and a
    window crossing an outer-index boundary by one position (k=309 runs swapped lead rings,
    forced, whose windows cover whole runs of lead * r0 positions)."""
    tune = {"mi_ring_swap": 1} if k == 309 else {}
    eng, plan = engine(k, None, mi_filter_limit=8, max_survivors=1 << 22, **tune)
    assert (eng._d.lead > 0) == (k == 309)
    ids, mn = eng.scan_window(2, 1000, 1000)
    assert list(ids) == [] and mn is None
    r = plan.radix[0]
    span = min(1 << 20, plan.R // 2)
    for lo, hi in ((0, span), (plan.R - span, plan.R), (plan.R - 3, plan.R),
                   (5 * r - 1, 5 * r + 1)):
        window_check(k, None, 2, lo, hi, limit=8, **tune)


def test_block_overflow_falls_back_to_per_launch_drain():
    """scan_block of C1 block 0 (12,654,618 survivors) with a 2^20-ID buffer and 2^24-position
    launches (~0.9M survivors each): the block overflows the buffer, no launch does, so the
    per-launch drain (_scan_block_chunked) must return exactly the oracle's survivors."""
    eng, plan = engine(30, C1_WHEEL, max_survivors=1 << 20, mi_launch_log2=24)
    say("case k=30 block 0 overflow drain")
    ids = eng.scan_block(0)
    n, want = O.c_scan(30, plan.wheel, plan.rest, 0, 0, plan.R, threads=16, cap=1 << 24)
    assert n == 12654618 > eng.out_cap
    got = np.sort(np.asarray(ids, dtype=np.uint64))
    assert len(got) == n and np.array_equal(got, want)


@pytest.mark.parametrize("k,wheel", [(30, C1_WHEEL), (200, "fullpower"), (304, None), (309, None),
                                     (318, None), (320, None), (378, None), (399, K399_PRIMARY),
                                     (399, K399_CONFIRM)])
def test_row_lane_layout_equals_strip_lane(k, wheel):
    """The row-lane layout (lanes over grouped strip rows, bank-neutral row map) and the
    strip-lane layout return the same survivors and minimum on the same windows: dense
    filter-prefix windows at unaligned edges and several launches, and the oracle agrees.
    k=304 and 320 stage a mask mod 2^9 (2 is not in their wheels): the power-of-2 table keeps
    the one-row wrap map."""
    import gpu_sieve_mi355x as G
    if wheel == "fullpower":
        wheel = [list(x) for x in EP.fullpower_wheel(200)]
    rng = random.Random(k)
    plan = EP.Plan(k, None, wheel)
    for limit in (3, 6, None):
        tune = {"max_survivors": 1 << 22, "mi_launch_log2": 26}
        if limit is not None:
            tune["mi_filter_limit"] = limit
        rows = G.GpuSieve(EP.Plan(k, dict(tune, mi_layout="auto"), wheel))
        strips = G.GpuSieve(EP.Plan(k, dict(tune, mi_layout="strips"), wheel))
        assert strips.layout == "strips"
        span = min(1 << 24, plan.R // 3)
        t = rng.randrange(0, 1 << 16)
        lo = rng.randrange(0, plan.R - span)
        hi = lo + span - rng.randrange(0, 1000)
        say(f"case layouts k={k} limit={limit} t={t} window=[{lo},{hi}) layout={rows.layout} "
            f"steps={rows.info.steps} groups={rows.info.groups} strip={rows.strip_len}")
        a, amin = rows.scan_window(t, lo, hi)
        b, bmin = strips.scan_window(t, lo, hi)
        assert sorted(a) == sorted(b) and amin == bmin
        filt = plan.rest if limit is None else plan.rest[:limit]
        n, want, wmin = O.c_scan_min(k, plan.wheel, filt, t, lo, hi, threads=16, cap=1 << 22)
        assert sorted(a) == [int(x) for x in want] and amin == wmin
    if rows.info.words == 1 and rows.strip_len >= 32:
        assert rows.layout == "rows"


@pytest.mark.parametrize("k,wheel", [(392, None), (384, None), (328, None), (385, None),
                                     (399, K399_CONFIRM), (383, None)])
def test_row_lane_wider_rows(k, wheel):
    """Rings 0 of 33..64 residues kept whole (mi_inner_radix 128: no digit rings) in the plan's
    order (mi_ring_swap 0) run the row-lane layout with two words per ring-0 row (k=384: 49
    residues, k=392 and 328: 64, the k=399 confirmation wheel: 47, k=383: 56, k=385: 42), one
    table per word: the same survivors and minimum as the strip-lane layout and the oracle on
    dense filter-prefix windows at unaligned edges, over several launches, and for an
    outer-unit launch sequence.  A three-word ring 0 asked for row-lane is refused."""
    import gpu_sieve_mi355x as G
    rng = random.Random(5 * k + 1)
    plan = EP.Plan(k, None, wheel)
    tune = {"mi_inner_radix": 128}
    for limit in (3, 6, None):
        base = dict(tune, max_survivors=1 << 22, mi_launch_log2=26, mi_ring_swap=0)
        if limit is not None:
            base["mi_filter_limit"] = limit
        rows = G.GpuSieve(EP.Plan(k, dict(base, mi_layout="rows"), wheel))
        strips = G.GpuSieve(EP.Plan(k, dict(base, mi_layout="strips"), wheel))
        assert rows.layout == "rows" and strips.layout == "strips" and rows._d.lead == 0
        assert rows.info.words == (plan.radix[0] + 31) // 32 >= 2
        span = min(1 << 24, plan.R // 3)
        t = rng.randrange(0, 1 << 16)
        lo = rng.randrange(0, plan.R - span)
        hi = lo + span - rng.randrange(0, 1000)
        say(f"case wide rows k={k} W={rows.info.words} limit={limit} t={t} window=[{lo},{hi}) "
            f"masks={rows.info.n_masks_used} steps={rows.info.steps} groups={rows.info.groups}")
        a, amin = rows.scan_window(t, lo, hi)
        b, bmin = strips.scan_window(t, lo, hi)
        assert sorted(a) == sorted(b) and amin == bmin
        filt = plan.rest if limit is None else plan.rest[:limit]
        n, want, wmin = O.c_scan_min(k, plan.wheel, filt, t, lo, hi, threads=16, cap=1 << 22)
        assert sorted(a) == [int(x) for x in want] and amin == wmin
    # outer-unit launches (the drivers' _launch) over a dense prefix
    eng = G.GpuSieve(EP.Plan(k, dict(tune, mi_filter_limit=4, max_survivors=1 << 22,
                                     mi_ring_swap=0, mi_layout="rows"), wheel))
    eng.set_block(3)
    eng.out_n.fill(0)
    o0 = plan.R // eng.r0 // 5
    eng._launch(o0, 30001)
    eng._launch(o0 + 30001, 7)
    n = int(eng.out_n.get()[0])
    got = np.sort(np.asarray([int(x) for x in eng.out[:n].get()], dtype=np.uint64))
    lo, hi = o0 * eng.r0, (o0 + 30008) * eng.r0
    nw, want = O.c_scan(k, plan.wheel, plan.rest[:4], 3, lo, hi, threads=16, cap=1 << 22)
    assert n == nw and np.array_equal(got, want)
    if k == 392:
        with pytest.raises(Exception, match="two words"):
            G.GpuSieve(EP.Plan(400, {"mi_inner_radix": 128, "mi_ring_swap": 0,
                                     "mi_layout": "rows"}))


def _windows_against_oracle(eng, plan, k, filt, t, windows):
    got = eng.scan_windows(t, windows)
    assert len(got) == len(windows)
    total = 0
    for (lo, hi), (ids, mn) in zip(windows, got):
        n, want, wmin = O.c_scan_min(k, plan.wheel, filt, t, lo, hi, threads=16, cap=1 << 22)
        assert len(ids) == n, f"window [{lo},{hi}): count {len(ids)} != oracle {n}"
        assert np.array_equal(np.sort(np.asarray(ids, dtype=np.uint64)), want)
        assert mn == wmin, f"window [{lo},{hi}): minimum {mn} != oracle {wmin}"
        total += n
    return total


@pytest.mark.parametrize("k,wheel,tune", [
    (30, C1_WHEEL, {}),                                             # strip-lane, dense
    (399, K399_PRIMARY, {"mi_filter_limit": 4}),                    # row-lane, one word
    (309, None, {"mi_filter_limit": 8, "mi_ring_swap": 1}),         # lead block
    (399, K399_CONFIRM, {"mi_filter_limit": 4, "mi_ring_swap": 0,   # two-word rows
                         "mi_inner_radix": 128, "mi_layout": "rows"}),
])
def test_scan_windows_pipelined(k, wheel, tune):
    """es_engine_scan_windows (each window queued before the previous one is read back, two
    result slots) returns per window exactly the oracle's survivors and minimum: dense,
    empty, unaligned and one-position windows, an odd number of windows (both slots reused),
    block 0 (s <= k+1 left out of the minimum) and a later block."""
    eng, plan = engine(k, wheel, max_survivors=1 << 20, **tune)
    filt = plan.rest[:tune["mi_filter_limit"]] if "mi_filter_limit" in tune else plan.rest
    rng = random.Random(k)
    base = plan.R // 7
    span = 1 << 18 if k == 30 else 1 << 20
    windows = [(base, base + span), (base + span, base + span), (base + span, base + 2 * span + 17),
               (base + 5 * span + 3, base + 5 * span + 4)]
    windows += [(lo, lo + rng.randrange(1, span)) for lo in
                (rng.randrange(0, plan.R - span) for _ in range(3))]
    for t in (0, 6):
        say(f"case scan_windows k={k} t={t} layout={eng.layout} windows={len(windows)} "
            f"r0={eng.r0} W={eng.info.words}")
        assert _windows_against_oracle(eng, plan, k, filt, t, windows) > 0
    # the first positions of block 0 (the floor) and the engine's own buffer left untouched
    eng.scan_window(0, 0, span, collect=False)
    mine = eng.collect(0)
    _windows_against_oracle(eng, plan, k, filt, 0, [(0, span), (span, 2 * span)])
    assert eng.collect(0) == mine


def test_scan_windows_overflow_and_errors():
    eng, plan = engine(30, C1_WHEEL, max_survivors=1 << 16)
    say("case scan_windows overflow k=30")
    lo = 1 << 22
    with pytest.raises(RuntimeError, match="overflow"):       # one window above out_cap
        eng.scan_windows(0, [(0, 1 << 24)])
    with pytest.raises(RuntimeError, match="overflow"):       # the windows together above cap
        eng.scan_windows(0, [(lo, lo + (1 << 16)), (lo + (1 << 16), lo + (1 << 17))], cap=100)
    with pytest.raises(Exception, match="outside"):
        eng.scan_windows(0, [(0, plan.R + 1)])
    assert eng.scan_windows(0, []) == []
    # the engine recovers: the next call is exact
    _windows_against_oracle(eng, plan, 30, plan.rest, 2, [(lo, lo + (1 << 14))] * 3)
