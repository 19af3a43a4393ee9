"""k = 200 on its full-power wheel (full prime powers for p <= 23, extended to M <= 2^100):
windows of 2^36 residue positions with all 37 filter primes.

* around g(200) in block 0: the minimum over n > k+1 is g(200) (b-file), every survivor is
  good by the oracle's Kummer test, and a 2^24 sub-window around it is bit-exact;
* at a seeded block t in [0, 2^20) and seeded offset: the first 2^28 positions are bit-exact
  against the oracle scan; the full window's survivor hash is the same under two different
  launch splittings (2^30- and 2^37-position launches) and equals the hash of the union of
  the window's four quarter scans; every survivor is good and lies in [tM, (t+1)M).
"""
from __future__ import annotations

import random

import numpy as np
import pytest

import es_oracle as O
import es_plan as EP
from conftest import bfile

pytestmark = pytest.mark.gpu

K = 200
N = 1 << 36


def say(*a):
    print(*a, flush=True)


def _engine(**tune):
    import gpu_sieve_mi355x as G
    tune.setdefault("max_survivors", 1 << 20)
    plan = EP.Plan(K, tune, EP.fullpower_wheel(K))
    return G.GpuSieve(plan), plan


def _ns(plan, ids, t):
    return sorted(plan.idx_to_n(int(i), t) for i in ids)


def test_c2_window_at_g200():
    eng, plan = _engine()
    g = bfile()[K]
    gi, gt = plan.n_to_idx(g)
    assert gt == 0
    lo = max(0, gi - N // 2)
    hi = min(plan.R, lo + N)
    say(f"c2 g200 window k={K} t=0 [{lo}, {hi}) R={plan.R} filters={len(plan.rest)}")
    ids, mn = eng.scan_window(0, lo, hi)
    ns = _ns(plan, ids, 0)
    assert gi in set(int(i) for i in ids)
    assert min(n for n in ns if n > K + 1) == g
    assert all(O.is_good(n, K) or n <= K + 1 for n in ns)
    sl, sh = gi - (1 << 23), gi + (1 << 23)
    got, _ = eng.scan_window(0, sl, sh)
    cnt, want = O.c_scan(K, plan.wheel, plan.rest, 0, sl, sh, threads=16)
    assert cnt == len(want)
    assert np.array_equal(np.sort(np.asarray(got, dtype=np.uint64)), want)


def test_c2_seeded_window_hash():
    rng = random.Random(200)
    t = rng.randrange(1 << 20)
    e37, plan = _engine(mi_launch_log2=37)
    lo = rng.randrange(plan.R - N)
    hi = lo + N
    say(f"c2 seeded window k={K} t={t} [{lo}, {hi})")
    ids, _ = e37.scan_window(t, lo, hi)
    ns = _ns(plan, ids, t)
    e30, _ = _engine(mi_launch_log2=30)
    ids30, _ = e30.scan_window(t, lo, hi)
    q = N // 4
    quarters = []
    for j in range(4):
        part, _ = e37.scan_window(t, lo + j * q, lo + (j + 1) * q)
        quarters += list(part)
    h = O.survivor_hash(ns)
    assert O.survivor_hash(_ns(plan, ids30, t)) == h
    assert O.survivor_hash(_ns(plan, quarters, t)) == h
    assert all(O.is_good(n, K) for n in ns)
    assert all(t * plan.M <= n < (t + 1) * plan.M for n in ns)
    say(f"c2 seeded window survivors={len(ns)} hash={h}")
    sub = 1 << 28
    cnt, want = O.c_scan(K, plan.wheel, plan.rest, t, lo, lo + sub, threads=16)
    got = np.sort(np.asarray([i for i in ids if lo <= int(i) < lo + sub], dtype=np.uint64))
    assert cnt == len(want) and np.array_equal(got, want)


def test_c2_digit_rings_dense_parity():
    """Ring 0 (13^3, radix 1056) swept as digit rings (88 x 12) vs swept whole, with only the first 6
    filters (dense survivors): both equal the oracle on a 2^26 window, and the driver path
    (_launch in plan outer units) covers the same positions."""
    rng = random.Random(2001)
    t = rng.randrange(1 << 20)
    split, plan = _engine(mi_filter_limit=6, max_survivors=1 << 22, mi_inner_radix=128)
    whole, _ = _engine(mi_filter_limit=6, max_survivors=1 << 22, mi_inner_radix=0)
    assert split._f == 12 and whole._f == 1
    r0 = plan.radix[0]
    o_lo = rng.randrange(plan.R // r0 - (1 << 16))
    lo, hi = o_lo * r0, (o_lo + (1 << 16)) * r0
    say(f"c2 digit rings k={K} t={t} [{lo}, {hi}) filters=6")
    cnt, want = O.c_scan(K, plan.wheel, plan.rest[:6], t, lo, hi, threads=16, cap=1 << 22)
    assert cnt == len(want) and cnt > 100
    for eng in (split, whole):
        got, _ = eng.scan_window(t, lo, hi)
        assert np.array_equal(np.sort(np.asarray(got, dtype=np.uint64)), want)
        eng.set_block(t)
        eng.out_n.fill(0)
        eng._launch(o_lo, 1 << 16)
        n = int(eng.out_n.get()[0])
        assert np.array_equal(np.sort(eng.out[:n].get()), want)
