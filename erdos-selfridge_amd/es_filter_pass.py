"""Candidate filter pass: Kummer admissibility of explicit candidates against many filters.

The late stage of the word-mask engine (sieve/es_word_masks.py:87-92): the candidates a mask
pass keeps are re-tested against the full admissibility tables and emitted with their original
IDs.  Here the candidates are u32 IDs into a stream of u128 values resident in HBM; the pass
keeps every ID whose value passes every filter.  The default workload is k = 397 with the
filter primes 101..397 (q = p^2, 53 tables of 64-bit words, ~429 KiB).

The synthetic stream is conditioned so that its values pass as many filters as u128 values
near 10^34 allow: a knapsack picks, per prime, how many low base-p digits (0, 1 or 2) are fixed
to random admissible values by CRT, maximising the survival probability while the product of
the fixed moduli stays below 2^108; a random multiple of that product lifts values to < 2^113.
That gives a survival of ~6e-7 over all 53 filters: a candidate stream of u128 values cannot be
conditioned on every filter (that needs ~400 bits), so ~1e-3 is out of reach at this size.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

import es_native as nat
from es_plan import base_digits, partial_count, partial_residues, primes_upto


def filter_order(k: int, primes: list[int]) -> list[int]:
    """Most selective first (smallest admissible fraction), ties by p."""
    def rate(p):
        nd = len(base_digits(k, p))
        return partial_count(k, p, nd) / p ** nd
    return sorted(primes, key=lambda p: (rate(p), p))


def filter_tables(k: int, primes: list[int]):
    """(q list, word offsets, u64 words) of the full admissibility tables, in the given order."""
    qs, offs, parts = [], [], []
    n = 0
    for p in primes:
        q, res = partial_residues(k, p, len(base_digits(k, p)))
        w = np.zeros((q + 63) // 64, dtype=np.uint64)
        r = np.asarray(res, dtype=np.int64)
        np.bitwise_or.at(w, r >> 6, np.uint64(1) << (r & 63).astype(np.uint64))
        qs.append(q)
        offs.append(n)
        parts.append(w)
        n += w.size
    words = np.concatenate(parts) if parts else np.zeros(1, dtype=np.uint64)
    return qs, offs, words


def generator_spec(k: int, primes: list[int], mod_bits: float = 108.0, value_bits: int = 113,
                   scale: int = 64):
    """Choose per prime the number of fixed low digits (0, 1, 2, ...) maximising the survival
    probability of a random value, with prod(fixed moduli) < 2^mod_bits (group knapsack over
    `scale` cells per bit).  Returns (mods, residue lists, j_span, survival)."""
    cap = int(mod_bits * scale)
    NEG = -1e300
    best = [NEG] * (cap + 1)
    best[0] = 0.0
    picks = []
    for p in primes:
        nd = len(base_digits(k, p))
        full = math.log2(partial_count(k, p, nd) / p ** nd)
        opts = []
        for d in range(0, nd + 1):
            w = 0 if d == 0 else int(math.ceil(d * math.log2(p) * scale))
            # conditional pass with d digits fixed admissible
            val = full - (math.log2(partial_count(k, p, d) / p ** d) if d else 0.0)
            opts.append((w, val, d))
        nxt = [NEG] * (cap + 1)
        choice = [0] * (cap + 1)
        for c in range(cap + 1):
            if best[c] == NEG:
                continue
            for w, val, d in opts:
                if c + w <= cap and best[c] + val > nxt[c + w]:
                    nxt[c + w] = best[c] + val
                    choice[c + w] = d
        best = nxt
        picks.append(choice)
    c = max(range(cap + 1), key=lambda i: best[i])
    survival = 2.0 ** best[c]
    depth = [0] * len(primes)
    for i in reversed(range(len(primes))):
        d = picks[i][c]
        depth[i] = d
        if d:
            c -= int(math.ceil(d * math.log2(primes[i]) * scale))
    mods, resl = [], []
    for p, d in zip(primes, depth):
        if d:
            m, res = partial_residues(k, p, d)
            mods.append(m)
            resl.append(res)
    P = math.prod(mods)
    j_span = max(1, (1 << value_bits) // P)
    return mods, resl, j_span, survival


def conditional_order(k: int, primes: list[int], spec) -> list[int]:
    """Filter order for a stream drawn from a generator spec: least pass rate given the digits
    the generator fixed first (a fixed-digit filter almost always passes, so testing it early
    keeps every lane of a wave alive)."""
    mods = spec[0]
    fixed = {}
    for m in mods:
        p = next(x for x in primes if m % x == 0)
        d = 0
        while m % p == 0:
            m //= p
            d += 1
        fixed[p] = d

    def rate(p):
        nd = len(base_digits(k, p))
        d = fixed.get(p, 0)
        done = partial_count(k, p, d) / p ** d if d else 1.0
        return (partial_count(k, p, nd) / p ** nd) / done
    return sorted(primes, key=lambda p: (rate(p), p))


class FilterPass:
    """One filter set on the current GPU (libes_sieve.so es_fpass_*).  Filters are tested in
    the order given (order="given") or most selective first (order="rate")."""

    def __init__(self, k: int, primes: list[int] | None = None, lds_budget: int = 16384,
                 out_cap: int = 1 << 24, order: str = "rate", tables=None):
        """tables: (moduli, word offsets, u64 words) of other admissible-residue bitmaps, one
        per prime in the order given, in place of the Kummer sets of k (the pass tests
        membership in whatever sets it is given)."""
        self.k = k
        if primes is None:
            primes = [p for p in primes_upto(k) if p >= 101]
        self.primes = list(primes) if order == "given" or tables is not None \
            else filter_order(k, primes)
        qs, offs, words = tables if tables is not None else filter_tables(k, self.primes)
        self.qs = qs
        self._keep = (np.asarray(qs, dtype=np.uint32), np.asarray(offs, dtype=np.uint32), words)
        h = C.c_void_p()
        nat.check(nat.lib().es_fpass_create(len(qs), nat.ptr32(self._keep[0]),
                                            nat.ptr32(self._keep[1]), nat.ptr64(words),
                                            int(words.size), lds_budget, out_cap, C.byref(h)))
        self._h = h
        self.out_cap = out_cap
        self.n_values = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                nat.lib().es_fpass_destroy(h)
            except Exception:
                pass
            self._h = None

    def upload(self, values: list[int], ids) -> None:
        v = np.zeros(2 * max(1, len(values)), dtype=np.uint64)
        for i, x in enumerate(values):
            v[2 * i] = x & ((1 << 64) - 1)
            v[2 * i + 1] = x >> 64
        ids = np.ascontiguousarray(np.asarray(ids, dtype=np.uint32))
        nat.check(nat.lib().es_fpass_upload(self._h, nat.ptr64(v), len(values), nat.ptr32(ids),
                                            int(ids.size)))
        self.n_values = len(values)

    def generate(self, n: int, seed: int, spec=None) -> None:
        mods, resl, j_span, _ = spec or generator_spec(self.k, self.primes)
        m = np.asarray(mods or [2], dtype=np.uint32)
        cnt = np.asarray([len(r) for r in resl] or [1], dtype=np.uint32)
        res = np.asarray([x for r in resl for x in r] or [0], dtype=np.uint32)
        nat.check(nat.lib().es_fpass_generate(self._h, n, seed, len(mods), nat.ptr32(m),
                                              nat.ptr32(cnt), nat.ptr32(res), j_span))
        self.n_values = n

    def run(self) -> np.ndarray:
        """Survivor IDs, sorted."""
        n = C.c_uint64(0)
        nat.check(nat.lib().es_fpass_run(self._h, C.byref(n)))
        ids = np.zeros(max(1, n.value), dtype=np.uint32)
        nat.check(nat.lib().es_fpass_read(self._h, nat.ptr32(ids), n.value))
        return np.sort(ids[:n.value])

    def set_shape(self, unroll: int = 2, n_first: int = 3) -> None:
        """Launch shape (speed only): candidates per lane per round, leading filters tested on
        every candidate before a wave packs its survivors for the rest."""
        nat.check(nat.lib().es_fpass_set_shape(self._h, unroll, n_first))

    def last_ms(self) -> float:
        ms = C.c_double(0)
        nat.check(nat.lib().es_fpass_last_ms(self._h, C.byref(ms)))
        return ms.value

    def read_stream_raw(self, offset: int, count: int):
        """Stream entries [offset, offset+count) as numpy arrays: values as (lo, hi) u64 pairs
        ([2 * count]) and their candidate IDs."""
        v = np.zeros(2 * max(1, count), dtype=np.uint64)
        ids = np.zeros(max(1, count), dtype=np.uint32)
        nat.check(nat.lib().es_fpass_read_stream(self._h, offset, count, nat.ptr64(v), nat.ptr32(ids)))
        return v[:2 * count], ids[:count]

    def read_stream(self, offset: int, count: int):
        """(values as python ints, ids) of stream entries [offset, offset+count)."""
        v = np.zeros(2 * max(1, count), dtype=np.uint64)
        ids = np.zeros(max(1, count), dtype=np.uint32)
        nat.check(nat.lib().es_fpass_read_stream(self._h, offset, count, nat.ptr64(v), nat.ptr32(ids)))
        vals = [(int(v[2 * i + 1]) << 64) | int(v[2 * i]) for i in range(count)]
        return vals, ids[:count]
