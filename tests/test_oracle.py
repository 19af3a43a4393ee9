"""Pin the CPU oracle before trusting it: b-file values, exact binomials, brute force,
closed forms and invariants.  A plausible mistake (wrong digit order, dropped filter, wrong
ring order, off-by-one in a window) fails at least one of these."""
from __future__ import annotations

import json
import os
import random
from math import comb, prod

import numpy as np
import pytest

import es_oracle as O
from conftest import GOLDEN, bfile


def test_primes_and_digits():
    assert O.primes_upto(1) == [] and O.primes_upto(2) == [2]
    assert O.primes_upto(30) == [2, 3, 5, 7, 11, 13, 17, 19, 23, 29]
    assert len(O.primes_upto(400)) == 78
    assert O.digits(0, 7) == [] and O.digits(399, 107) == [78, 3]
    for _ in range(200):
        x, p = random.randrange(10 ** 30), random.choice([2, 3, 5, 7, 101, 397])
        assert sum(d * p ** i for i, d in enumerate(O.digits(x, p))) == x


def test_kummer_matches_exact_binomials():
    """dominates(n,k,p) <=> p does not divide C(n,k) (Kummer), checked with math.comb."""
    rng = random.Random(5)
    for _ in range(400):
        k = rng.randrange(1, 60)
        n = rng.randrange(k, 3000)
        c = comb(n, k)
        for p in O.primes_upto(k):
            assert O.dominates(n, k, p) == (c % p != 0)


def test_bfile_terms_are_good_and_exact():
    b = bfile()
    assert len(b) == 400
    for k, n in b.items():
        assert O.is_good(n, k), k
    for k in range(1, 41):  # exact binomial prime-divisor check (harness/verify_release.py:102-107)
        c = comb(b[k], k)
        assert all(c % p for p in O.primes_upto(k))


def test_bfile_minimality_small_k():
    b = bfile()
    for k in range(1, 25):
        assert O.g_bruteforce(k) == b[k]
    for k in range(1, 41):
        assert O.g_crt(k) == b[k]


def test_ring_residue_counts_closed_form():
    """|residues mod p^T| = prod_{i<T} (p - a_i)  (harness/verify_release.py:56-59)."""
    for k in (30, 97, 200, 399):
        for p in O.primes_upto(k)[:12]:
            kd = O.digits(k, p)
            for T in range(1, len(kd) + 1):
                q, res = O.low_residues(k, p, T)
                assert q == p ** T and res == sorted(res)
                assert len(res) == prod(p - d for d in kd[:T])


def test_crt_basis_properties():
    mods = [2 ** 5, 3 ** 4, 5 ** 3, 7 ** 2, 11 ** 2, 13 ** 2]
    M, basis = O.crt_basis(mods)
    assert M == 324648324000
    for j, b in enumerate(basis):
        for i, q in enumerate(mods):
            assert b % q == (1 if i == j else 0)


def test_survivor_hash_recipe():
    """Recorded hashes of runs.json (harness/run_k.py:42-44 recipe), incl. the empty list."""
    terms = json.load(open(os.path.join(GOLDEN, "runs_frontier.json")))["terms"]
    nb = 0
    for term in terms:
        for role in ("primary", "confirmation"):
            for blk in term[role]["block_audit"]:
                assert O.survivor_hash(int(x) for x in blk["survivors_n"]) == blk["sha256_survivors"]
                nb += 1
    assert nb >= 100
    assert O.survivor_hash([]) == "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"


def test_recorded_survivors_admissible_and_small_survivors():
    terms = json.load(open(os.path.join(GOLDEN, "runs_frontier.json")))["terms"]
    for term in terms:
        k = term["k"]
        small = O.small_survivors(k)
        assert k in small
        for role in ("primary", "confirmation"):
            run = term[role]
            for blk in run["block_audit"]:
                for s in blk["survivors_n"]:
                    n = int(s)
                    assert all(O.dominates(n, k, p) for p in O.primes_upto(k))
                    if n <= k + 1:
                        assert n in small


# ---------------------------------------------------------------- C restatement

def _ring_order(k, wheel):
    return sorted(((p, T) for p, T in wheel),
                  key=lambda it: (-prod(it[0] - d for d in O.digits(k, it[0])[:it[1]]), it[0]))


def test_c_geometry_matches_python():
    for k, wheel in ((30, [[2, 5], [3, 4], [5, 3], [7, 2], [11, 2], [13, 2]]),
                     (399, [[107, 1], [3, 6], [211, 1], [139, 1], [83, 1], [59, 1], [103, 1],
                            [137, 1], [41, 1], [2, 9], [37, 1], [29, 1], [101, 1], [13, 1],
                            [31, 1], [67, 1], [5, 2]])):
        ro = _ring_order(k, wheel)
        wm = O.WheelMap(k, ro)
        M, R, rad = O.c_geometry(k, ro)
        assert (M, R, rad) == (wm.M, wm.R, wm.radix)
        rng = random.Random(k)
        for _ in range(200):
            idx, t = rng.randrange(R), rng.randrange(1 << 20)
            assert O.c_idx_to_n(k, ro, idx, t) == wm.idx_to_n(idx, t)
    # k=399 wheel of runs.json: M and R as recorded
    assert M == 242767881445433682148391005401600 and R == 11908694063185920


def test_c_scan_brute_force_small_blocks():
    """Every position of small blocks, checked with the Python Kummer test on n itself; the
    block covers [tM, (t+1)M) exactly once (tests/test_enumeration.py:37-44)."""
    for k, wheel in ((12, [[2, 2], [3, 1], [5, 1]]), (20, [[2, 3], [3, 2], [5, 1], [7, 1]]),
                     (30, [[2, 2], [3, 2], [5, 2], [7, 1]])):
        ro = _ring_order(k, wheel)
        wm = O.WheelMap(k, ro)
        wheelT = dict(ro)
        ps = O.primes_upto(k)
        filt = [p for p in ps if wheelT.get(p, 0) < len(O.digits(k, p))]
        for t in (0, 1, 7):
            ns = [wm.idx_to_n(i, t) for i in range(wm.R)]
            assert len(set(ns)) == wm.R and all(t * wm.M <= n < (t + 1) * wm.M for n in ns)
            want = [i for i, n in enumerate(ns) if all(O.dominates(n, k, p) for p in filt)]
            cnt, got = O.c_scan(k, ro, filt, t, 0, wm.R, threads=4)
            assert cnt == len(want) and [int(x) for x in got] == want
            # every admissible n of the block is enumerated
            adm = [n for n in range(t * wm.M, (t + 1) * wm.M) if all(O.dominates(n, k, p) for p in ps)]
            assert sorted(wm.idx_to_n(i, t) for i in want) == adm


def test_c_scan_window_split_invariance():
    k, ro = 30, _ring_order(30, [[2, 5], [3, 4], [5, 3], [7, 2], [11, 2], [13, 2]])
    filt = [17, 19, 23, 29]
    n_all, ids = O.c_scan(k, ro, filt, 2, 1000, 1000 + 300000, threads=8)
    n_a, a = O.c_scan(k, ro, filt, 2, 1000, 77777, threads=3)
    n_b, b = O.c_scan(k, ro, filt, 2, 77777, 1000 + 300000, threads=1)
    assert n_all == n_a + n_b and np.array_equal(ids, np.concatenate([a, b]))


def test_c_scan_config1_minimum():
    """Config 1 (k=30): the least survivor n > 31 of block 0 is g(30) = 58782, and it lies at
    idx 13,258,355 < 2^24."""
    k, ro = 30, _ring_order(30, [[2, 5], [3, 4], [5, 3], [7, 2], [11, 2], [13, 2]])
    wm = O.WheelMap(k, ro)
    n, ids = O.c_scan(k, ro, [17, 19, 23, 29], 0, 0, 1 << 24, threads=8, cap=1 << 21)
    ns = [(wm.idx_to_n(int(i), 0), int(i)) for i in ids]
    best = min(x for x in ns if x[0] > k + 1)
    assert best == (58782, 13258355)


def test_c_dominates_u128():
    rng = random.Random(9)
    for _ in range(300):
        n = rng.randrange(1 << 127)
        k = rng.randrange(2, 401)
        p = rng.choice(O.primes_upto(k))
        assert O.lib().eso_dominates(n & ((1 << 64) - 1), n >> 64, k, p) == int(O.dominates(n, k, p))


@pytest.mark.parametrize("k", [6, 13, 29])
def test_word_mask_tables_definition(k):
    rng = random.Random(k)
    p = O.primes_upto(k)[-1]
    q, res = O.allowed_residues(k, p)
    terms = [rng.randrange(1 << 100) for _ in range(37)]
    tab = O.word_mask_phase_tables(q, res, terms)
    for ph in rng.sample(range(q), min(q, 20)):
        for i, tv in enumerate(terms):
            bit = (tab[ph][i // 32] >> (i % 32)) & 1
            assert bit == int((ph + tv) % q in set(res))


def test_scan_min_pinned():
    """The oracle's window minimum: g(30) in block 0 of the C1 wheel (s <= k+1 left out there),
    every later block's least survivor counted whatever its residue, and both equal to the
    Python restatement's idx -> n map over the survivor list."""
    import es_plan as EP
    plan = EP.Plan(30, None, [[2, 5], [3, 4], [5, 3], [7, 2], [11, 2], [13, 2]])
    wm = O.WheelMap(30, [tuple(x) for x in plan.wheel])
    n, ids, mn = O.c_scan_min(30, plan.wheel, plan.rest, 0, 0, 1 << 24, threads=8)
    assert mn == 58782 and n == len(ids) > 0
    # block 14 starts with a survivor at idx 0: n = 14 M + 30, counted although s = 30 <= k+1
    n14, ids14, mn14 = O.c_scan_min(30, plan.wheel, plan.rest, 14, 0, 1 << 16, threads=4)
    assert ids14[0] == 0 and mn14 == 14 * wm.M + 30
    for t, lo in ((0, 0), (3, 5_000_000), (14, 0)):
        n, ids, mn = O.c_scan_min(30, plan.wheel, plan.rest, t, lo, lo + (1 << 14), threads=2)
        ns = [wm.idx_to_n(int(i), t) for i in ids]
        valid = [x for x in ns if x > 31]
        assert mn == (min(valid) if valid else None)
