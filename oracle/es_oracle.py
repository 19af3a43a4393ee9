"""CPU ORACLE for the CRT-wheel sieve -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module,
and only as the checker (or the timed CPU baseline).  The engine never imports it.

Pure-Python restatement of the reference's referee and enumeration, each function citing the
reference lines it follows, plus a ctypes front end to the C restatement (es_oracle.c) that
scans windows of positions.  Pinning (what checks each function, tests/test_oracle.py):
  primes_upto, digits, dominates, is_good  -- b-file g(k) for k <= 400 (exact Kummer check of
      every published term) and exact math.comb prime-divisor checks for small k
  g_bruteforce, g_crt                       -- b-file values (k <= 24 / k <= 60)
  allowed_residues / ring enumeration       -- brute force over [0, q) with math.comb-free
      digit checks and the closed form prod (p - a_i) of harness/verify_release.py:56-59
  crt_basis                                 -- basis_j = 1 mod q_j, 0 mod q_i (i != j)
  idx_to_n / C scan                         -- coverage of [0, R) exactly once, brute-force
      Kummer on every n of small blocks, recorded survivor lists (runs.json windows, k=270..280
      block counts), survivor hashes
  word_mask_phase_tables                    -- per-bit recomputation from the definition
  mlp_* (surrogate)                         -- finite differences
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
from itertools import product
from math import prod

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libes_oracle.so")


# ------------------------------------------------------------------ ref/erdos_ref.py restated

def primes_upto(k: int) -> list[int]:
    """ref/erdos_ref.py:20-28 (sieve of Eratosthenes)."""
    if k < 2:
        return []
    is_p = [True] * (k + 1)
    is_p[0] = is_p[1] = False
    for i in range(2, int(k ** 0.5) + 1):
        if is_p[i]:
            for j in range(i * i, k + 1, i):
                is_p[j] = False
    return [i for i in range(k + 1) if is_p[i]]


def digits(x: int, p: int) -> list[int]:
    """ref/erdos_ref.py:31-37: base-p digits, least significant first."""
    d = []
    while x:
        d.append(x % p)
        x //= p
    return d


def dominates(n: int, k: int, p: int) -> bool:
    """ref/erdos_ref.py:40-48: every base-p digit of k <= the matching digit of n."""
    dk, dn = digits(k, p), digits(n, p)
    dn += [0] * (len(dk) - len(dn))
    return all(a <= b for a, b in zip(dk, dn))


def is_good(n: int, k: int) -> bool:
    """ref/erdos_ref.py:51-55."""
    return n > k + 1 and all(dominates(n, k, p) for p in primes_upto(k))


def g_bruteforce(k: int, limit: int | None = None) -> int | None:
    """ref/erdos_ref.py:58-66."""
    n = k + 2
    ps = primes_upto(k)
    while limit is None or n <= limit:
        if all(dominates(n, k, p) for p in ps):
            return n
        n += 1
    return None


def low_residues(k: int, p: int, T: int) -> tuple[int, list[int]]:
    """sieve/gpu_sieve_v9.py:127-138 / ref/erdos_ref.py:69-81: residues mod p^T whose T low
    digits dominate k's, ascending -- here by brute force over [0, p^T)."""
    q = p ** T
    kd = (digits(k, p) + [0] * T)[:T]
    out = []
    for r in range(q):
        rd = (digits(r, p) + [0] * T)[:T]
        if all(a <= b for a, b in zip(kd, rd)):
            out.append(r)
    return q, out


def allowed_residues(k: int, p: int) -> tuple[int, list[int]]:
    return low_residues(k, p, len(digits(k, p)))


def crt_basis(moduli: list[int]) -> tuple[int, list[int]]:
    """ref/erdos_ref.py:84-90."""
    M = prod(moduli)
    return M, [(M // q) * pow(M // q, -1, q) for q in moduli]


def g_crt(k: int, max_wheel_modulus: int = 10 ** 9) -> int:
    """ref/erdos_ref.py:93-130: exhaustive CRT-wheel search in increasing blocks."""
    import math
    ps = primes_upto(k)
    info = {p: allowed_residues(k, p) for p in ps}
    scored = sorted(ps, key=lambda p: -math.log(info[p][0] / len(info[p][1])) / math.log(info[p][0]))
    wheel, M = [], 1
    for p in scored:
        q = info[p][0]
        if M * q <= max_wheel_modulus:
            wheel.append(p)
            M *= q
    rest = sorted((p for p in ps if p not in wheel), key=lambda p: len(info[p][1]) / info[p][0])
    M, basis = crt_basis([info[p][0] for p in wheel])
    residues = sorted(sum(a * b for a, b in zip(combo, basis)) % M
                      for combo in product(*[info[p][1] for p in wheel]))
    t = 0
    while True:
        for r in residues:
            n = t * M + r
            if n > k + 1 and all(dominates(n, k, p) for p in rest):
                return n
        t += 1


def survivor_hash(ns) -> str:
    """harness/run_k.py:42-44: sha256 of '\\n'.join(sorted decimal n), no trailing newline."""
    return hashlib.sha256("\n".join(str(int(n)) for n in sorted(ns)).encode()).hexdigest()


def small_survivors(k: int) -> list[int]:
    """harness/reconstruct_audit.py:39-42: every n <= k+1 dominating k for all p <= k."""
    ps = primes_upto(k)
    return [n for n in range(0, k + 2) if all(dominates(n, k, p) for p in ps)]


class WheelMap:
    """Mixed-radix idx <-> n map of a wheel in ring order (sieve/gpu_sieve_v9.py:385-390),
    built from the oracle's own residue lists and basis."""

    def __init__(self, k: int, wheel_ring_order: list[tuple[int, int]]):
        self.k = k
        self.rings = [low_residues(k, p, T) for p, T in wheel_ring_order]
        self.M, self.basis = crt_basis([q for q, _ in self.rings])
        self.radix = [len(r) for _, r in self.rings]
        self.R = prod(self.radix)

    def idx_to_n(self, idx: int, t: int) -> int:
        s = 0
        for (q, res), rad, b in zip(self.rings, self.radix, self.basis):
            idx, d = divmod(idx, rad)
            s += res[d] * b
        return t * self.M + s % self.M


# ------------------------------------------------------------------ word masks (config 4)

def word_mask_phase_tables(q: int, allowed: list[int], terms: list[int]) -> list[list[int]]:
    """sieve/es_word_masks.py:18-36 for one filter: table[phase][word] bit i is set iff
    (phase + terms[i]) mod q is admissible."""
    ok = set(allowed)
    width = (len(terms) + 31) // 32
    table = []
    for ph in range(q):
        row = [0] * width
        for i, tv in enumerate(terms):
            if (ph + tv % q) % q in ok:
                row[i // 32] |= 1 << (i % 32)
        table.append(row)
    return table


# ------------------------------------------------------------------ C restatement front end

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run oracle/build_oracle.py")
        L = C.CDLL(LIB_PATH)
        u32p, u64p = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)
        L.eso_scan.restype = C.c_int64
        L.eso_scan.argtypes = [C.c_uint32, C.c_uint32, u32p, u32p, C.c_uint32, u32p, C.c_uint64,
                               C.c_uint64, C.c_uint64, C.c_uint32, u64p, C.c_uint64]
        L.eso_scan_min.restype = C.c_int64
        L.eso_scan_min.argtypes = [C.c_uint32, C.c_uint32, u32p, u32p, C.c_uint32, u32p,
                                   C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, u64p,
                                   C.c_uint64, u64p]
        L.eso_geometry.restype = C.c_int
        L.eso_geometry.argtypes = [C.c_uint32, C.c_uint32, u32p, u32p, u64p, u64p, u32p]
        L.eso_idx_to_s.restype = C.c_int
        L.eso_idx_to_s.argtypes = [C.c_uint32, C.c_uint32, u32p, u32p, C.c_uint64, u64p]
        L.eso_dominates.restype = C.c_int
        L.eso_dominates.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32]
        L.eso_filter_values.restype = C.c_int64
        L.eso_filter_values.argtypes = [C.c_uint32, C.c_uint32, u32p, u64p, C.c_uint64, u32p,
                                        C.c_uint64, u32p, C.c_uint64]
        L.eso_good_range.restype = C.c_int64
        L.eso_good_range.argtypes = [C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                     C.c_uint32, u64p, C.c_uint64]
        _lib = L
    return _lib


def _u32(xs):
    arr = (C.c_uint32 * max(1, len(xs)))(*xs)
    return arr


def c_scan(k: int, wheel_ring_order, filt_primes, t: int, idx_lo: int, idx_hi: int,
           threads: int = 8, cap: int = 1 << 22):
    """Survivor IDs (ascending numpy u64) of positions [idx_lo, idx_hi) of block t.
    Returns (count, ids); count may exceed cap, then ids holds the first cap."""
    import numpy as np
    L = lib()
    rp = _u32([p for p, _ in wheel_ring_order])
    rT = _u32([T for _, T in wheel_ring_order])
    fp = _u32(list(filt_primes))
    out = np.zeros(max(1, cap), dtype=np.uint64)
    n = L.eso_scan(k, len(wheel_ring_order), rp, rT, len(filt_primes), fp, t, idx_lo, idx_hi,
                   threads, out.ctypes.data_as(C.POINTER(C.c_uint64)), cap)
    if n < 0:
        raise ValueError("oracle rejected the plan")
    return int(n), out[:min(n, cap)].copy()


def c_scan_min(k: int, wheel_ring_order, filt_primes, t: int, idx_lo: int, idx_hi: int,
               threads: int = 8, cap: int = 1 << 22):
    """As c_scan, plus the least survivor n = t*M + s > k+1 of the window (None if none):
    (count, ids, min_n)."""
    import numpy as np
    L = lib()
    rp = _u32([p for p, _ in wheel_ring_order])
    rT = _u32([T for _, T in wheel_ring_order])
    fp = _u32(list(filt_primes))
    out = np.zeros(max(1, cap), dtype=np.uint64)
    m = (C.c_uint64 * 3)()
    n = L.eso_scan_min(k, len(wheel_ring_order), rp, rT, len(filt_primes), fp, t, idx_lo,
                       idx_hi, threads, out.ctypes.data_as(C.POINTER(C.c_uint64)), cap, m)
    if n < 0:
        raise ValueError("oracle rejected the plan")
    mn = None
    if m[2]:
        M, _, _ = c_geometry(k, wheel_ring_order)
        mn = t * M + ((int(m[1]) << 64) | int(m[0]))
    return int(n), out[:min(n, cap)].copy(), mn


def c_geometry(k: int, wheel_ring_order):
    L = lib()
    w = len(wheel_ring_order)
    rp = _u32([p for p, _ in wheel_ring_order])
    rT = _u32([T for _, T in wheel_ring_order])
    M = (C.c_uint64 * 2)()
    R = C.c_uint64(0)
    rad = (C.c_uint32 * max(1, w))()
    if L.eso_geometry(k, w, rp, rT, M, C.byref(R), rad) != 0:
        raise ValueError("oracle rejected the plan")
    return (int(M[1]) << 64) | int(M[0]), R.value, [int(rad[i]) for i in range(w)]


def c_idx_to_n(k: int, wheel_ring_order, idx: int, t: int) -> int:
    L = lib()
    w = len(wheel_ring_order)
    rp = _u32([p for p, _ in wheel_ring_order])
    rT = _u32([T for _, T in wheel_ring_order])
    s = (C.c_uint64 * 2)()
    if L.eso_idx_to_s(k, w, rp, rT, idx, s) != 0:
        raise ValueError("oracle rejected the plan")
    M, _, _ = c_geometry(k, wheel_ring_order)
    return t * M + ((int(s[1]) << 64) | int(s[0]))


def c_filter_values(k: int, primes, values, ids, cap: int = 1 << 22):
    """Survivor IDs (input order) of the candidate filter pass: ids whose value dominates k in
    base p for every listed prime."""
    import numpy as np
    L = lib()
    v = np.zeros(2 * max(1, len(values)), dtype=np.uint64)
    for i, x in enumerate(values):
        v[2 * i] = x & ((1 << 64) - 1)
        v[2 * i + 1] = x >> 64
    ids = np.ascontiguousarray(np.asarray(ids, dtype=np.uint32))
    out = np.zeros(max(1, cap), dtype=np.uint32)
    u32p, u64p = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)
    n = L.eso_filter_values(k, len(primes), _u32(list(primes)), v.ctypes.data_as(u64p),
                            len(values), ids.ctypes.data_as(u32p), ids.size,
                            out.ctypes.data_as(u32p), cap)
    if n < 0:
        raise ValueError("oracle rejected the input")
    return out[:min(n, cap)].copy()


def c_filter_lohi(k: int, primes, lohi, ids, threads: int = 8, cap: int = 1 << 24):
    """c_filter_values on a stream held as (lo, hi) u64 pairs (numpy), the IDs split over
    `threads` concurrent calls (ctypes releases the GIL).  Returns the survivor IDs sorted."""
    import threading

    import numpy as np
    L = lib()
    lohi = np.ascontiguousarray(lohi, dtype=np.uint64)
    ids = np.ascontiguousarray(np.asarray(ids, dtype=np.uint32))
    nv = lohi.size // 2
    u32p, u64p = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)
    pr = _u32(list(primes))
    parts = np.array_split(ids, max(1, threads))
    outs = [None] * len(parts)
    errs = []

    def work(i):
        part = np.ascontiguousarray(parts[i])
        out = np.zeros(max(1, cap), dtype=np.uint32)
        n = L.eso_filter_values(k, len(primes), pr, lohi.ctypes.data_as(u64p), nv,
                                part.ctypes.data_as(u32p), part.size, out.ctypes.data_as(u32p),
                                cap)
        if n < 0 or n > cap:
            errs.append(n)
        outs[i] = out[:max(0, min(n, cap))]
    th = [threading.Thread(target=work, args=(i,)) for i in range(len(parts))]
    for x in th:
        x.start()
    for x in th:
        x.join()
    if errs:
        raise ValueError(f"oracle rejected the input or overflowed: {errs}")
    return np.sort(np.concatenate(outs)) if outs else np.zeros(0, dtype=np.uint32)


def c_good_range(k: int, n_lo: int, n_hi: int, threads: int = 8, cap: int = 1 << 20):
    """(count, ascending list) of every n in [n_lo, n_hi) dominating k for all p <= k
    (eso_good_range: brute force, independent of any wheel)."""
    import numpy as np
    L = lib()
    out = np.zeros(2 * max(1, cap), dtype=np.uint64)
    M64 = (1 << 64) - 1
    n = L.eso_good_range(k, n_lo & M64, n_lo >> 64, n_hi & M64, n_hi >> 64, threads,
                         out.ctypes.data_as(C.POINTER(C.c_uint64)), cap)
    if n < 0:
        raise ValueError("oracle rejected the range")
    m = min(n, cap)
    return n, [(int(out[2 * i + 1]) << 64) | int(out[2 * i]) for i in range(m)]
