"""Candidate filter pass (k = 397, filter primes 101..397): CPU checks of the tables, the
generator specification and the oracle; GPU parity of the pass against the oracle."""
from __future__ import annotations

import math
import random

import numpy as np
import pytest

import es_filter_pass as FP
import es_oracle as O
import es_plan as EP

K = 397
PRIMES = [p for p in O.primes_upto(K) if p >= 101]


def test_filter_tables_match_oracle_residues():
    order = FP.filter_order(K, PRIMES)
    assert sorted(order) == PRIMES and len(order) == 53
    qs, offs, words = FP.filter_tables(K, order[:5])
    for p, q, off in zip(order[:5], qs, offs):
        q2, res = O.allowed_residues(K, p)
        assert q == q2 == p * p
        got = [r for r in range(q) if (int(words[off + (r >> 6)]) >> (r & 63)) & 1]
        assert got == res


def test_generator_spec_bounds_and_survival():
    mods, resl, j_span, surv = FP.generator_spec(K, FP.filter_order(K, PRIMES))
    P = math.prod(mods)
    assert P < 1 << 108 and P * j_span <= 1 << 113 and j_span >= 16
    for a in range(len(mods)):
        for b in range(a + 1, len(mods)):
            assert math.gcd(mods[a], mods[b]) == 1
    assert 5e-7 < surv < 1e-5                       # about the most values below 2^113 allow
    for m, res in zip(mods, resl):
        p = min(x for x in PRIMES if m % x == 0)
        d = round(math.log(m, p))
        assert res == O.low_residues(K, p, d)[1]


def _spec_values(spec, n, seed):
    """Python draw of values of a generator spec (the device draw uses another RNG)."""
    mods, resl, j_span, _ = spec
    P = math.prod(mods)
    rng = random.Random(seed)
    basis = [(P // m) * pow(P // m, -1, m) for m in mods]
    out = []
    for _ in range(n):
        s = sum(rng.choice(r) * b for r, b in zip(resl, basis)) % P
        out.append(s + rng.randrange(j_span) * P)
    return out


def test_conditional_order_puts_fixed_digits_late():
    spec = FP.generator_spec(K, FP.filter_order(K, PRIMES))
    order = FP.conditional_order(K, PRIMES, spec)
    fixed = {min(x for x in PRIMES if m % x == 0) for m in spec[0]}
    assert sorted(order) == PRIMES
    assert not fixed & set(order[:5])


def test_generator_spec_survival_monte_carlo():
    """Dense variant (3 filters, 24-bit budget): measured survival matches the spec."""
    prim = [101, 103, 107]
    spec = FP.generator_spec(K, prim, mod_bits=24)
    vals = _spec_values(spec, 20000, 3)
    keep = O.c_filter_values(K, prim, vals, np.arange(len(vals)))
    frac = len(keep) / len(vals)
    assert abs(frac - spec[3]) < 5 * math.sqrt(spec[3] / len(vals))


def test_oracle_filter_matches_python_kummer():
    rng = random.Random(11)
    prim = PRIMES[:6]
    vals = [rng.randrange(1 << 113) for _ in range(3000)]
    # plant admissible residues for the 6 primes on a third of the values
    spec = FP.generator_spec(K, prim, mod_bits=90)
    vals[::3] = _spec_values(spec, len(vals[::3]), 4)
    ids = np.arange(len(vals), dtype=np.uint32)[::-1].copy()
    got = O.c_filter_values(K, prim, vals, ids)
    want = [int(i) for i in ids if all(O.dominates(vals[i], K, p) for p in prim)]
    assert [int(x) for x in got] == want and len(want) > 100


# ------------------------------------------------------------------ GPU parity

@pytest.mark.gpu
def test_gpu_pass_uploaded_stream_dense():
    prim = FP.filter_order(K, PRIMES[:8])
    fp = FP.FilterPass(K, prim)
    spec = FP.generator_spec(K, fp.primes, mod_bits=100)
    vals = _spec_values(spec, 1 << 16, 5) + [random.Random(6).randrange(1 << 120) for _ in range(1 << 14)]
    ids = np.random.default_rng(7).permutation(len(vals)).astype(np.uint32)
    print(f"case fpass uploaded n={len(vals)} filters={len(prim)}", flush=True)
    fp.upload(vals, ids)
    got = fp.run()
    want = np.sort(O.c_filter_values(K, fp.primes, vals, ids))
    assert len(want) > 1000 and np.array_equal(got, want)


@pytest.mark.gpu
def test_gpu_pass_generated_stream_all_filters():
    fp = FP.FilterPass(K)
    n = 1 << 22
    print(f"case fpass generated n={n} filters=53", flush=True)
    fp.generate(n, seed=20260901)
    got = fp.run()
    vals, ids = fp.read_stream(0, n)
    assert sorted(int(i) for i in ids) == list(range(n))        # a permutation of the stream
    assert all(v < (1 << 113) for v in vals[:4096])
    want = np.sort(O.c_filter_values(K, fp.primes, vals, ids))
    assert np.array_equal(got, want)
    # same seed, same stream
    fp.generate(n, seed=20260901)
    v2, _ = fp.read_stream(0, 1024)
    assert v2 == vals[:1024]


@pytest.mark.gpu
def test_gpu_pass_conditional_order_same_survivors():
    """Filter order changes speed only: the conditional order returns the same survivors."""
    spec = FP.generator_spec(K, FP.filter_order(K, PRIMES))
    a = FP.FilterPass(K)
    b = FP.FilterPass(K, FP.conditional_order(K, PRIMES, spec), order="given")
    n = 1 << 20
    print(f"case fpass orders n={n}", flush=True)
    a.generate(n, seed=77, spec=spec)
    b.generate(n, seed=77, spec=spec)
    assert np.array_equal(a.run(), b.run())


@pytest.mark.gpu
def test_gpu_pass_overflow_reported():
    prim = [101]
    fp = FP.FilterPass(K, prim, out_cap=16)
    vals = [x for x in range(1000)]
    fp.upload(vals, np.arange(1000))
    with pytest.raises(RuntimeError, match="overflow"):
        fp.run()


@pytest.mark.gpu
def test_gpu_pass_shapes_agree_ragged():
    """Stage-1 depth and unroll change speed only: every shape returns the oracle's survivors,
    on a ragged stream (length not a multiple of a wave) where stage 1 keeps everything
    (n_first = 0 sends every candidate through the wave lists), some or few."""
    prim = FP.filter_order(K, PRIMES[:8])
    fp = FP.FilterPass(K, prim)
    spec = FP.generator_spec(K, fp.primes, mod_bits=100)
    vals = _spec_values(spec, 100_003, 9)
    ids = np.random.default_rng(10).permutation(len(vals)).astype(np.uint32)
    print(f"case fpass shapes n={len(vals)} filters={len(prim)}", flush=True)
    fp.upload(vals, ids)
    want = np.sort(O.c_filter_values(K, fp.primes, vals, ids))
    assert len(want) > 1000
    for unroll in (2, 4, 8):
        for n_first in (0, 1, 2, 5, 8):
            fp.set_shape(unroll, n_first)
            assert np.array_equal(fp.run(), want), (unroll, n_first)


@pytest.mark.gpu
def test_gpu_pass_empty_and_tiny_lists():
    """No IDs (nothing launched), one ID, and 1, 63, 65, 257 IDs into a longer stream: the
    lanes past the end of the list load a clamped ID and are masked, in every shape."""
    prim = FP.filter_order(K, PRIMES[:8])
    fp = FP.FilterPass(K, prim)
    spec = FP.generator_spec(K, fp.primes, mod_bits=100)
    vals = _spec_values(spec, 4096, 11)
    fp.upload(vals, np.zeros(0, dtype=np.uint32))
    print("case fpass empty list", flush=True)
    assert fp.run().size == 0
    rng = np.random.default_rng(12)
    for n in (1, 63, 65, 257):
        ids = rng.choice(len(vals), size=n, replace=False).astype(np.uint32)
        print(f"case fpass tiny n={n}", flush=True)
        fp.upload(vals, ids)
        want = np.sort(O.c_filter_values(K, fp.primes, vals, ids))
        for unroll in (2, 8):
            fp.set_shape(unroll, 3)
            assert np.array_equal(fp.run(), want), (n, unroll)


@pytest.mark.gpu
def test_gpu_pass_arbitrary_bitmaps():
    """The pass tests membership in whatever bitmaps it is given, in LDS or in global memory.
    Here the fourth filter's bitmap loses one admissible residue and gains one inadmissible one
    (no longer a Kummer set) and the LDS budget holds only the first table: the pass must
    return exactly the values whose residues are in every bitmap (membership checked in
    Python)."""
    prim = FP.filter_order(K, PRIMES[:8])
    qs, offs, words = FP.filter_tables(K, prim)
    words = words.copy()
    q3, o3 = qs[3], offs[3]
    bit = lambda r: int(words[o3 + (r >> 6)] >> np.uint64(r & 63)) & 1
    on = next(r for r in range(q3) if bit(r))
    off = next(r for r in range(q3) if not bit(r))
    words[o3 + (on >> 6)] ^= np.uint64(1) << np.uint64(on & 63)
    words[o3 + (off >> 6)] ^= np.uint64(1) << np.uint64(off & 63)
    fp = FP.FilterPass(K, prim, lds_budget=(qs[0] + 63) // 64 * 8, tables=(qs, offs, words))
    spec = FP.generator_spec(K, prim, mod_bits=100)
    vals = _spec_values(spec, 1 << 14, 13)
    # values that hit the changed residues of filter 3, so both edits matter
    vals += [on + q3 * i for i in range(1, 200)] + [off + q3 * i for i in range(1, 200)]
    # edges of the u128 range and of every modulus (the residue's exactness and its range)
    top = (1 << 128) - 1
    vals += [0, 1, top, top - 1, 1 << 127, (1 << 96) - 1, 1 << 64, (1 << 32) - 1]
    for q in qs:
        vals += [q - 1, q, q + 1, top - top % q, top - top % q - 1, (1 << 100) // q * q]
    ids = np.random.default_rng(14).permutation(len(vals)).astype(np.uint32)
    print(f"case fpass forms n={len(vals)} q3={q3}", flush=True)
    fp.upload(vals, ids)

    def member(v):
        return all((int(words[o + ((v % q) >> 6)]) >> ((v % q) & 63)) & 1 for q, o in zip(qs, offs))
    want = np.sort(np.asarray([i for i in ids if member(vals[i])], dtype=np.uint32))
    assert len(want) > 100
    assert np.array_equal(fp.run(), want)


@pytest.mark.gpu
def test_gpu_pass_config4_full_size():
    """Config 4 at its full size: 2^30 seeded candidates of the conditioned u128 stream, all 53
    filters in the conditional order the bench uses; every survivor ID equals the oracle's
    Kummer test of every stream value (16 host threads)."""
    n = 1 << 30
    spec = FP.generator_spec(K, FP.filter_order(K, PRIMES))
    fp = FP.FilterPass(K, FP.conditional_order(K, PRIMES, spec), order="given")
    print(f"case fpass full size n=2^30 filters={len(fp.primes)}", flush=True)
    fp.generate(n, seed=20260901, spec=spec)
    got = fp.run()
    lohi, ids = fp.read_stream_raw(0, n)
    assert ids.size == n
    want = O.c_filter_lohi(K, fp.primes, lohi, ids, threads=16)
    print(f"  survivors={len(got)} oracle={len(want)}", flush=True)
    assert len(want) > 100 and np.array_equal(got, want)
