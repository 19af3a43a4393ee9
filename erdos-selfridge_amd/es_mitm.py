"""Bounded n-window enumeration by meet in the middle (sieve/gpu_mitm.py, harness/mitm_search.py).

The block sieve scans whole wheel periods [tM, (t+1)M), so its modulus must stay near
E = 1 / (density of good n).  Here only the n of a window [t*M + a, t*M + b) are formed: the
wheel (whole primes at their full modulus q_p) is split into two halves, half 2 becomes a
sorted device table of CRT partial sums, half 1 is decoded per index in the kernel, and each
s1 meets the contiguous (circular) run of s2 with (s1 + s2) mod M in [a, b).  M is then free to
exceed E, so the wheel can absorb more primes.  The window's survivors are every n in it that
passes the remaining primes' tables -- with the wheel primes whole, every n of the window that
dominates k for all p <= k.

`choose_wheel_mitm` and `split_wheel` follow harness/mitm_search.py:52-120 (DP over 64 cells per
bit of log2 M for the least log2 R, cost max(log2 R / 2, log2 E + log2 R - log2 M), M >= 4E,
smaller half within the memory budget; greedy balanced split); `mitm_search` follows the
bound-growing loop of sieve/gpu_mitm.py:190-222.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

import es_data

import es_native as nat
import es_plan as EP

SCALE = 64


class MitmDesc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32), ("M_lo", C.c_uint64), ("M_hi", C.c_uint64),
        ("n1", C.c_uint32), ("radix1", nat.u32p), ("term1_lo", nat.u64p), ("term1_hi", nat.u64p),
        ("n2", C.c_uint32), ("radix2", nat.u32p), ("term2_lo", nat.u64p), ("term2_hi", nat.u64p),
        ("n_filters", C.c_uint32), ("filt_q", nat.u32p), ("filt_off", nat.u32p),
        ("filt_bits", nat.u32p), ("n_filt_words", C.c_uint32), ("out_cap", C.c_uint64),
        ("walk", C.c_uint32), ("sweep", C.c_uint32),
    ]


def _info(k):
    return {p: EP.allowed_residues(k, p) for p in EP.primes_upto(k)}


def split_wheel(wheel, info):
    """Two halves with |S1| and |S2| as balanced as the greedy allows: primes by descending
    residue count, each to the currently smaller half (harness/mitm_search.py:109-118)."""
    w1, w2, n1, n2 = [], [], 1, 1
    for p in sorted(wheel, key=lambda p: -len(info[p][1])):
        if n1 <= n2:
            w1.append(p)
            n1 *= len(info[p][1])
        else:
            w2.append(p)
            n2 *= len(info[p][1])
    return sorted(w1), sorted(w2), n1, n2


def choose_wheel_mitm(k, info, log2E, mem_bytes, min_m_over_e=4.0):
    """Whole-prime wheel for a meet-in-the-middle search (harness/mitm_search.py:52-107):
    0/1 knapsack over SCALE cells per bit of log2 M for the least log2 R per cell; cells with
    log2 M >= log2 E + log2(min_m_over_e) (where the primes allow) and log2 R within twice the
    table budget are ranked by max(log2 R / 2, log2 E + log2 R - log2 M), and the first whose
    actual split keeps the smaller half within mem_bytes / 16 entries wins.  None if none fits."""
    ps = list(info)
    w = [max(1, round(math.log2(info[p][0]) * SCALE)) for p in ps]
    a = [math.log2(len(info[p][1])) for p in ps]
    cap = min(4096 * SCALE, sum(w))
    dp = np.full(cap + 1, np.inf)
    dp[0] = 0.0
    take = []
    for wi, ai in zip(w, a):
        t = np.zeros(cap + 1, dtype=bool)
        if wi <= cap:
            cand = dp[:cap + 1 - wi] + ai            # previous row: 0/1 knapsack
            win = cand < dp[wi:]
            dp[wi:][win] = cand[win]
            t[wi:] = win
        take.append(t)
    cap_logR = 2 * math.log2(mem_bytes / 16)
    total_logM = sum(w) / SCALE
    need_logM = min(log2E + math.log2(min_m_over_e), total_logM)

    def rebuild(c):
        out = []
        for i in range(len(ps) - 1, -1, -1):
            if take[i][c]:
                out.append(ps[i])
                c -= w[i]
        return sorted(out)

    cells = np.nonzero(np.isfinite(dp) & (dp <= cap_logR))[0]
    logM = cells / SCALE
    keep = logM >= need_logM
    cells, logM = cells[keep], logM[keep]
    logR = dp[cells]
    cost = np.maximum(logR / 2.0, log2E + logR - logM)
    order = np.lexsort((cells, cost))
    budget = mem_bytes / 16
    for c in cells[order]:
        wh = rebuild(int(c))
        _, _, n1, n2 = split_wheel(wh, info)
        if min(n1, n2) <= budget:
            return wh
    return None


class MitmSieve:
    """One whole-prime wheel for k on the current GPU (es_mitm_* in libes_sieve.so)."""

    def __init__(self, k: int, wheel=None, mem_gb: float = 8.0, out_cap: int = 1 << 20,
                 walk: int = 0, sweep: int = 0):
        self.k = k
        info = _info(k)
        self.info = info
        dens = 1.0
        for q, res in info.values():
            dens *= len(res) / q
        self.log2E = -math.log2(dens)
        if wheel is None:
            wheel = choose_wheel_mitm(k, info, self.log2E, mem_gb * 1e9)
            assert wheel, "no meet-in-the-middle wheel fits the memory budget"
        self.wheel = sorted(wheel)
        self.M = math.prod(info[p][0] for p in self.wheel)
        assert self.M < (1 << EP.MAX_M_BITS), "wheel modulus too large for the 128-bit path"
        basis = {}
        for p in self.wheel:
            q = info[p][0]
            Mi = self.M // q
            basis[p] = Mi * pow(Mi, -1, q)
        w1, w2, n1, n2 = split_wheel(self.wheel, info)
        if n2 > n1:                                   # the table is the smaller half
            w1, w2, n1, n2 = w2, w1, n2, n1
        self.halves = (w1, w2)
        self.sizes = (n1, n2)
        self.rest = sorted((p for p in info if p not in self.wheel),
                           key=lambda p: (len(info[p][1]) / info[p][0], p))

        def half(primes):
            if not primes:
                return [1], [0]
            radix, terms = [], []
            for p in primes:
                radix.append(len(info[p][1]))
                terms.extend((r * basis[p]) % self.M for r in info[p][1])
            return radix, terms
        r1, t1 = half(w1)
        r2, t2 = half(w2)
        M64 = (1 << 64) - 1
        keep = {}
        keep["r1"] = np.asarray(r1, dtype=np.uint32)
        keep["t1l"] = np.asarray([t & M64 for t in t1], dtype=np.uint64)
        keep["t1h"] = np.asarray([t >> 64 for t in t1], dtype=np.uint64)
        keep["r2"] = np.asarray(r2, dtype=np.uint32)
        keep["t2l"] = np.asarray([t & M64 for t in t2], dtype=np.uint64)
        keep["t2h"] = np.asarray([t >> 64 for t in t2], dtype=np.uint64)
        import gpu_sieve_mi355x as G
        qs, offs, words, nw = [], [], [], 0
        for p in self.rest:
            q = info[p][0]
            wd = G._table_words(k, p, len(EP.base_digits(k, p)))
            qs.append(q)
            offs.append(nw)
            words.append(wd)
            nw += wd.size
        keep["q"] = np.asarray(qs or [2], dtype=np.uint32)
        keep["off"] = np.asarray(offs or [0], dtype=np.uint32)
        keep["bits"] = np.concatenate(words) if words else np.zeros(1, dtype=np.uint32)
        d = MitmDesc()
        d.abi_version = nat.ES_ABI_VERSION
        d.M_lo, d.M_hi = self.M & M64, self.M >> 64
        d.n1, d.radix1 = len(r1), nat.ptr32(keep["r1"])
        d.term1_lo, d.term1_hi = nat.ptr64(keep["t1l"]), nat.ptr64(keep["t1h"])
        d.n2, d.radix2 = len(r2), nat.ptr32(keep["r2"])
        d.term2_lo, d.term2_hi = nat.ptr64(keep["t2l"]), nat.ptr64(keep["t2h"])
        d.n_filters = len(qs)
        d.filt_q, d.filt_off = nat.ptr32(keep["q"]), nat.ptr32(keep["off"])
        d.filt_bits, d.n_filt_words = nat.ptr32(keep["bits"]), int(keep["bits"].size)
        d.out_cap = out_cap
        d.walk = walk
        d.sweep = sweep
        self._keep = keep
        self.out_cap = out_cap
        h = C.c_void_p()
        nat.check(nat.lib().es_mitm_create(C.byref(d), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                nat.lib().es_mitm_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def swept(self) -> bool:
        """True when half 1 is a sorted table swept with two pointers (no binary searches)."""
        n1, n2, sw = C.c_uint64(0), C.c_uint64(0), C.c_uint32(0)
        nat.check(nat.lib().es_mitm_sizes(self._h, C.byref(n1), C.byref(n2), C.byref(sw)))
        return bool(sw.value)

    def scan(self, t: int, a: int, b: int) -> list[int]:
        """Survivors n of [t*M + a, t*M + b), 0 <= a < b <= M, ascending."""
        M64 = (1 << 64) - 1
        n = C.c_uint64(0)
        nat.check(nat.lib().es_mitm_scan(self._h, int(t), a & M64, a >> 64, b & M64, b >> 64,
                                         C.byref(n)))
        buf = np.zeros(2 * max(1, n.value), dtype=np.uint64)
        nat.check(nat.lib().es_mitm_read(self._h, nat.ptr64(buf), n.value))
        base = t * self.M
        return sorted(base + ((int(buf[2 * i + 1]) << 64) | int(buf[2 * i])) for i in range(n.value))

    def last(self):
        """(device ms, candidates formed) of the last scan."""
        ms = C.c_double(0)
        c = C.c_uint64(0)
        nat.check(nat.lib().es_mitm_last(self._h, C.byref(ms), C.byref(c)))
        return ms.value, c.value


def mitm_search(k: int, mem_gb: float = 8.0, bound_mult: float = 3.0, max_rounds: int = 8,
                sieve: MitmSieve | None = None, log=None) -> dict:
    """g(k) by windows [0, B) with B = bound_mult * E, quadrupled until a good n > k+1 is found
    or B reaches M (sieve/gpu_mitm.py:190-222)."""
    s = sieve or MitmSieve(k, mem_gb=mem_gb)
    B = max(int(bound_mult * 2.0 ** s.log2E), 1)
    total_ms, total_cand = 0.0, 0
    for _ in range(max_rounds):
        B = min(B, s.M)
        ns = s.scan(0, 0, B)
        ms, cand = s.last()
        total_ms += ms
        total_cand += cand
        good = [n for n in ns if n > k + 1 and EP.is_good(n, k)]
        if log:
            log(f"k={k} B=2^{math.log2(B):.1f} survivors={len(ns)} good={len(good)} "
                f"{ms:.1f} ms")
        if good:
            return {"k": k, "status": "FOUND", "g": min(good), "wheel": s.wheel,
                    "M_bits": s.M.bit_length(), "S1": s.sizes[0], "S2": s.sizes[1], "bound": B,
                    "device_ms": total_ms, "candidates": total_cand}
        if B >= s.M:
            return {"k": k, "status": "MITM_OUT_OF_RANGE", "M_bits": s.M.bit_length(),
                    "device_ms": total_ms, "candidates": total_cand}
        B = min(B * 4, s.M)
    return {"k": k, "status": "ROUND_LIMIT", "device_ms": total_ms, "candidates": total_cand}


def main(argv=None) -> int:
    """CLI with the argv and JSON lines of sieve/gpu_mitm.py:225-239: one JSON object per k
    (status, g, bound, wheel, sizes, and the b-file comparison when the k is in it)."""
    import argparse
    import json
    import os
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("k", type=int, nargs="+")
    ap.add_argument("--mem-gb", type=float, default=8.0)
    ap.add_argument("--bound-mult", type=float, default=3.0)
    ap.add_argument("--chunk-log2", type=int, default=24,
                    help="accepted for argv compatibility; launches are sized by the engine")
    ap.add_argument("--bfile", default=None, help="b-file (k n per line) to compare against")
    a = ap.parse_args(argv)
    bf = {}
    here = os.path.dirname(os.path.abspath(__file__))
    path = a.bfile or es_data.BFILE
    if os.path.exists(path):
        for line in open(path):
            t = line.split()
            if len(t) == 2 and t[0].isdigit():
                bf[int(t[0])] = int(t[1])
    for k in a.k:
        r = mitm_search(k, mem_gb=a.mem_gb, bound_mult=a.bound_mult)
        if r.get("status") == "FOUND":
            r["bfile"] = str(bf[k]) if k in bf else None
            r["matches_bfile"] = (bf[k] == r["g"]) if k in bf else None
            r["g"] = str(r["g"])
            r["bound"] = str(r["bound"])
        print(json.dumps(r, default=str), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
