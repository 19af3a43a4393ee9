"""Batched launches (k = 300..320 shapes): every job's survivors equal the
unbatched engine's on the same block and window."""
from __future__ import annotations

import random

import pytest

import es_plan as EP
from conftest import bfile

pytestmark = pytest.mark.gpu


def _engine(k, **tune):
    import gpu_sieve_mi355x as G
    return G.GpuSieve(EP.Plan(k, dict(tune)))


def test_batch_equals_unbatched_dense():
    """Filter-prefix engines (dense survivors) of mixed shapes: one-word rows, swapped lead
    rings (k=301, 308, forced) and two-word rows (k=301 kept in the plan's ring order), several
    F."""
    import es_batch
    rng = random.Random(2026)
    b = es_batch.Batch(out_cap_per_job=1 << 20)
    jobs = []
    for k, tune in ((300, {}), (301, {"mi_ring_swap": 1}), (301, {"mi_ring_swap": 0}),
                    (308, {"mi_ring_swap": 1}), (317, {})):
        eng = _engine(k, mi_filter_limit=4, max_survivors=1 << 20, **tune)
        for _ in range(2):
            t = rng.randrange(1 << 20)
            lo = rng.randrange(eng.plan.R - (1 << 22))
            hi = lo + rng.randrange(1 << 21, 1 << 22)
            jobs.append((b.add(eng, t, lo, hi), eng, t, lo, hi))
    print(f"case batch dense jobs={len(jobs)}", flush=True)
    ms, launches = b.run()
    assert launches >= 2
    total = 0
    for jid, eng, t, lo, hi in jobs:
        want, _ = eng.scan_window(t, lo, hi)
        got = b.collect(jid)
        assert got == sorted(want) and len(got) > 0
        total += len(got)
    print(f"  survivors={total} ms={ms:.2f}", flush=True)


def test_batch_full_filters_finds_g():
    """k = 300..320, default planner, full filters: a window around g(k) in its block returns
    g(k), and seeded windows agree with the unbatched engine."""
    import es_batch
    bf = bfile()
    rng = random.Random(7)
    b = es_batch.Batch()
    jobs = []
    for k in range(300, 321):
        eng = _engine(k)
        idx, t = eng.plan.n_to_idx(bf[k])
        lo = max(0, idx - (1 << 26))
        hi = min(eng.plan.R, lo + (1 << 27))
        jobs.append((b.add(eng, t, lo, hi), eng, t, lo, hi, bf[k]))
        t2 = rng.randrange(1 << 20)
        lo2 = rng.randrange(eng.plan.R - (1 << 27))
        jobs.append((b.add(eng, t2, lo2, lo2 + (1 << 27)), eng, t2, lo2, lo2 + (1 << 27), None))
    print(f"case batch full jobs={len(jobs)}", flush=True)
    b.run()
    for jid, eng, t, lo, hi, g in jobs:
        got = b.collect(jid)
        want, mn = eng.scan_window(t, lo, hi)
        assert got == sorted(want)
        if g is not None:
            assert eng.plan.n_to_idx(g)[0] in got and mn == g


def test_graph_sweep_matches_eager_every_replay():
    """Block sweep through a captured graph: jobs of mixed shapes moved to new blocks and
    replayed several times, with the engines' own eager scans between replays; every replay's
    survivors equal the eager engine on the same block and window, and a job moved back to its
    first block returns its first survivors again."""
    import es_batch
    rng = random.Random(77)
    b = es_batch.Batch(out_cap_per_job=1 << 20)
    jobs = []
    for k in (300, 308, 317):
        eng = _engine(k, mi_filter_limit=4, max_survivors=1 << 20)
        lo = rng.randrange(eng.plan.R - (1 << 22))
        jobs.append((b.add(eng, 0, lo, lo + (1 << 21)), eng, lo, lo + (1 << 21)))
    launches = b.capture()
    print(f"case batch graph jobs={len(jobs)} launches={launches}", flush=True)
    blocks = [5, 1 << 19, 12345, 5]
    first = None
    for rep, t in enumerate(blocks):
        if rep % 2:
            b.set_blocks([t] * len(jobs))
        else:
            for jid, _, _, _ in jobs:
                b.set_block(jid, t)
        b.replay()
        got = [b.collect(jid) for jid, _, _, _ in jobs]
        for (jid, eng, lo, hi), g in zip(jobs, got):
            want, _ = eng.scan_window(t, lo, hi)          # eager work between replays
            assert g == sorted(want) and len(g) > 0, (rep, t, jid)
        if rep == 0:
            first = got
    assert got == first                                    # block 5 again
    # the eager batch run agrees with the last replay
    b.run()
    assert [b.collect(jid) for jid, _, _, _ in jobs] == got


def test_config3_full_size_autotuned():
    """Config 3 at its size: k = 300..320 with autotuned launch shapes, 4 seeded blocks x 2^34
    positions per k in one batch; every job equals the unbatched engine, and a seeded 2^28
    sub-window of every job equals the oracle (full filter set)."""
    import es_autotune
    import es_batch
    import es_oracle as O
    rng = random.Random(300320)
    N = 1 << 34
    space = dict(es_autotune.SPACE, mi_global_masks=[0])     # batches read masks from LDS
    b = es_batch.Batch()
    jobs = []
    for k in range(300, 321):
        tuned, _ = es_autotune.autotune(k, space=space, probe_log2=30, method="coord")
        eng = _engine(k, **tuned)
        print(f"case config3 k={k} tuned={tuned}", flush=True)
        for _ in range(4):
            t = rng.randrange(1 << 20)
            lo = rng.randrange(eng.plan.R - N)
            jobs.append((b.add(eng, t, lo, lo + N), eng, t, lo, lo + N))
    print(f"case config3 batch jobs={len(jobs)}", flush=True)
    ms, launches = b.run()
    print(f"  batch {ms:.1f} ms, {launches} launches", flush=True)
    for jid, eng, t, lo, hi in jobs:
        got = b.collect(jid)
        want, _ = eng.scan_window(t, lo, hi)
        assert got == sorted(want), (eng.plan.k, t, lo)
        sub = lo + rng.randrange(N - (1 << 28))
        plan = eng.plan
        print(f"  job k={plan.k} t={t} oracle window [{sub}, +2^28)", flush=True)
        n, ids = O.c_scan(plan.k, plan.wheel, plan.rest, t, sub, sub + (1 << 28), threads=16)
        assert [i for i in got if sub <= i < sub + (1 << 28)] == [int(x) for x in ids]


def test_batch_random_mixes():
    """30 seeded random batches of 1..8 engines (k in [30, 400], filter prefixes of 4..12,
    forced ring orders and layouts) with 1..3 windows each of 1 to 2^22 positions in random
    blocks: every job returns the oracle's survivors."""
    import es_batch
    import es_oracle as O
    import gpu_sieve_mi355x as G
    rng = random.Random(5)
    for r in range(30):
        b = es_batch.Batch(out_cap_per_job=1 << 20)
        jobs = []
        for _ in range(rng.randrange(1, 9)):
            k = rng.randrange(30, 401)
            tune = {"max_survivors": 1 << 20, "mi_filter_limit": rng.choice([4, 6, 8, 12])}
            if rng.random() < 0.3:
                tune["mi_ring_swap"] = rng.choice([0, 1])
            if rng.random() < 0.3:
                tune["mi_layout"] = rng.choice(["strips", "auto"])
            eng = G.GpuSieve(EP.Plan(k, dict(tune)))
            for _ in range(rng.randrange(1, 4)):
                span = min(rng.choice([1, 1000, 1 << 16, 1 << 20, 1 << 22]), eng.plan.R)
                lo = rng.randrange(0, eng.plan.R - span + 1)
                t = rng.randrange(1 << 20)
                jobs.append((b.add(eng, t, lo, lo + span), eng, k, tune, t, lo, lo + span))
        print(f"case batch random {r}: " + " ".join(f"k={k}/t={t}/[{lo},{hi})"
                                                     for _, _, k, _, t, lo, hi in jobs),
              flush=True)
        b.run()
        for jid, eng, k, tune, t, lo, hi in jobs:
            filt = eng.plan.rest[:tune["mi_filter_limit"]]
            n, want = O.c_scan(k, eng.plan.wheel, filt, t, lo, hi, threads=16, cap=1 << 20)
            assert b.collect(jid) == [int(x) for x in want], (r, k, t, lo, hi)
