"""Bounded n-window enumeration (meet in the middle): the survivors of a window
[t*M + a, t*M + b) are exactly the n there that dominate k for every p <= k, checked against
the oracle's brute-force range scan (independent of any wheel), in both walk modes; the
bound-growing search returns the b-file g(k)."""
from __future__ import annotations

import math
import random

import pytest

import es_mitm as MM
import es_oracle as O
from conftest import bfile


def test_range_oracle_matches_python_kummer():
    n, got = O.c_good_range(30, 0, 300_000, threads=4)
    want = [x for x in range(300_000) if all(O.dominates(x, 30, p) for p in O.primes_upto(30))]
    assert got == want and n == len(want) and 58782 in got
    lo = (1 << 70) + 12345
    n, got = O.c_good_range(12, lo, lo + 200_000, threads=4)
    assert got == [x for x in range(lo, lo + 200_000)
                   if all(O.dominates(x, 12, p) for p in O.primes_upto(12))]


@pytest.mark.parametrize("k,mem", [(60, 1e-3), (100, 0.01), (200, 1.0), (300, 8.0)])
def test_split_and_wheel_choice(k, mem):
    info = MM._info(k)
    log2E = -sum(math.log2(len(r) / q) for q, r in info.values())
    w = MM.choose_wheel_mitm(k, info, log2E, mem * 1e9)
    assert w
    w1, w2, n1, n2 = MM.split_wheel(w, info)
    assert sorted(w1 + w2) == sorted(w) and not set(w1) & set(w2)
    assert n1 * n2 == math.prod(len(info[p][1]) for p in w)
    assert min(n1, n2) <= mem * 1e9 / 16
    logM = sum(math.log2(info[p][0]) for p in w)
    assert logM >= log2E + 2 - 1e-6 or logM >= sum(math.log2(info[p][0]) for p in info) - 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("walk,sweep", [(1, 1), (2, 1), (0, 2), (8, 2), (16, 2), (32, 2), (64, 2)])
def test_windows_equal_range_oracle(walk, sweep):
    """walk 1 / 2: lane or wave per s1 with binary searches; sweep 2: both halves sorted, two
    run pointers per group of `walk` lanes (0: by run length) moving down the table."""
    rng = random.Random(31 + walk + 10 * sweep)
    for k, mem in ((30, 1e-4), (60, 1e-3), (100, 0.01)):
        s = MM.MitmSieve(k, mem_gb=mem, walk=walk, sweep=sweep, out_cap=1 << 20)
        for _ in range(3):
            t = rng.randrange(0, 4)
            span = min(s.M, rng.randrange(1 << 20, 1 << 24))
            a = rng.randrange(0, s.M - span + 1)
            b = a + span
            print(f"case mitm k={k} walk={walk} sweep={sweep} wheel={s.wheel} t={t} [{a}, {b})",
                  flush=True)
            got = s.scan(t, a, b)
            n, want = O.c_good_range(k, t * s.M + a, t * s.M + b, threads=16)
            assert got == want and len(got) == n
        # a window reaching the top of the block (run wraps around the table)
        a = s.M - (1 << 20) if s.M > (1 << 20) else 0
        got = s.scan(1, a, s.M)
        assert got == O.c_good_range(k, s.M + a, 2 * s.M, threads=16)[1]
        if s.M < (1 << 24):
            # long runs (half the table and all of it per s1), and a run crossing the table end
            for lo, hi in ((s.M // 3, s.M // 3 + s.M // 2), (0, s.M), (s.M // 5, s.M)):
                print(f"case mitm k={k} walk={walk} sweep={sweep} t=2 [{lo}, {hi})", flush=True)
                got = s.scan(2, lo, hi)
                assert got == O.c_good_range(k, 2 * s.M + lo, 2 * s.M + hi, threads=16)[1]


@pytest.mark.gpu
@pytest.mark.parametrize("k,mem", [(60, 1e-3), (100, 0.01), (150, 0.1), (180, 1.0)])
def test_search_finds_bfile_g(k, mem):
    r = MM.mitm_search(k, mem_gb=mem)
    print(f"case mitm search k={k}: {r['status']} bound={r.get('bound')} "
          f"candidates={r['candidates']} ms={r['device_ms']:.1f}", flush=True)
    assert r["status"] == "FOUND" and r["g"] == bfile()[k]


@pytest.mark.gpu
def test_cli_json_lines(capsys):
    """argv and JSON lines of the reference's MITM CLI (sieve/gpu_mitm.py:225-239)."""
    import json
    assert MM.main(["100", "150", "--mem-gb", "0.1"]) == 0
    lines = [json.loads(x) for x in capsys.readouterr().out.splitlines() if x.startswith("{")]
    assert [r["k"] for r in lines] == [100, 150]
    for r in lines:
        assert r["status"] == "FOUND" and r["matches_bfile"] is True
        assert int(r["g"]) == bfile()[r["k"]]
