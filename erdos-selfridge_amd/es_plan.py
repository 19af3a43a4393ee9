"""Wheel geometry for the CRT-wheel sieve of the Erdos-Selfridge function g(k).

g(k) is the least n > k+1 such that C(n, k) has no prime factor <= k.  By Kummer's theorem
p does not divide C(n, k) iff every base-p digit of k is <= the matching digit of n, so the
condition for one prime p only looks at n mod q_p, q_p = p^(number of base-p digits of k).

A *wheel* is a set of items (p, T), 1 <= T <= t_p: the T low base-p digits of n are fixed by
the wheel ring mod p^T.  The admissible residues of a ring are enumerated by a mixed-radix
index (ring 0 fastest), their CRT combination is one residue s mod M = prod p^T, and block t
is n in [tM, (t+1)M).  Every prime whose full q_p is not covered by the wheel is a *filter*,
checked against its full admissibility table mod q_p.

This module is the host-side plan object the engine and the drivers share.  Its surface
mirrors the reference engine module (sieve/gpu_sieve_v9.py:127-413): ``partial_residues``,
``partial_count``, ``plan_wheel`` (group-knapsack planner, truecost or knapsack objective),
``order_filters`` and ``Plan`` with ``idx_to_n`` / ``summary``.  The planner's floating-point
steps are kept operation-for-operation alike so plans match the recorded ones bit for bit.
"""
from __future__ import annotations

import functools
import math
from itertools import product
from math import prod

import numpy as np

MAX_M_BITS = 126            # CRT sums of two residues stay below 2^127
U64 = (1 << 64) - 1

# Tunables: same keys and meaning as the reference engine (sieve/gpu_sieve_v9.py:99-120,
# sieve/gpu_sieve_v8.py:98-112, sieve/gpu_sieve_v8_masks.py:15).  Keys that only make
# sense on the reference's hardware (maxrregcount, smem_budget_words) are accepted and
# ignored.  The MI355X engine adds its own launch-shape keys (prefixed "mi_").
TUNABLES = {
    "planner": "truecost",     # "truecost" | "knapsack"
    "n_mult": 1.0,             # knapsack planner: capacity = log2(n_mult * E[g])
    "cost_scale": 256,         # DP cells per bit of log2(M)
    "m_cap_bits": 120,         # hard cap on log2(M)
    "max_R_bits": 62,          # residues per block must index in uint64
    "target_g": 0,             # > 0: a KNOWN g (confirmation runs): the wheel minimises the
                               # positions scanned to cover [0, g), ceil(g / M) * R
    "threads": "auto",         # workgroup size ("auto": the engine picks the size of 256,
                               # 512, 1024 that keeps the most waves resident per CU)
    "chunk_log2": 30,          # positions per launch unit of the reference (drivers read ochunk)
    "max_survivors": 1 << 20,  # survivor buffer capacity (u64 indices)
    "fast_filters": 4,         # filters evaluated as ring-0 word masks
    "word_masks": "auto",      # "auto" | "force" | "off"  (accepted; masks are always exact)
    "mi_launch_log2": 37,      # positions per engine launch (outer units = 2^this / r0)
    "mi_strip_target": 1024,   # target outer indices per uniform strip
    "mi_masks": "auto",        # word-mask prefilters staged in LDS ("auto": by candidate rate)
    "mi_lds_budget": 65536,    # bytes of LDS tables per workgroup
    "mi_global_masks": 0,      # masks read through L1 instead of LDS (0, 2, 4; r0 <= 32)
    "mi_layout": "auto",       # "auto" | "strips" | "rows": lanes over strips, or over strip
                               # rows grouped by LDS bank ("auto": rows where they apply)
    "mi_inner_radix": "auto",  # ring 0 mod p^T wider than this is swept as digit rings (0:
                               # never; "auto": the engine times the distinct splits)
    "mi_ring_swap": "auto",    # a ring 0 of 33..128 residues trading places with the first
                               # later ring of <= 32 in the kernel (one-word mask rows): 1
                               # wherever possible, 0 never, "rule" where it fills its word
                               # clearly better, "auto": the engine times both orders
}


# --------------------------------------------------------------------------- number theory

def primes_upto(n: int) -> list[int]:
    """All primes <= n (ascending)."""
    if n < 2:
        return []
    flags = bytearray([1]) * (n + 1)
    flags[0] = flags[1] = 0
    p = 2
    while p * p <= n:
        if flags[p]:
            flags[p * p::p] = bytes(len(range(p * p, n + 1, p)))
        p += 1
    return [i for i, f in enumerate(flags) if f]


def base_digits(x: int, p: int) -> list[int]:
    """Base-p digits of x, least significant first; [] for x == 0."""
    out = []
    while x > 0:
        x, r = divmod(x, p)
        out.append(r)
    return out


def kummer_ok(n: int, k: int, p: int) -> bool:
    """True iff p does not divide C(n, k): each base-p digit of k <= that of n."""
    while k:
        if k % p > n % p:
            return False
        k //= p
        n //= p
    return True


def is_good(n: int, k: int) -> bool:
    """True iff n > k+1 and C(n, k) has no prime factor <= k."""
    return n > k + 1 and all(kummer_ok(n, k, p) for p in primes_upto(k))


@functools.lru_cache(maxsize=512)
def partial_residues(k: int, p: int, T: int) -> tuple[int, list[int]]:
    """(p^T, ascending residues r mod p^T whose T low base-p digits dominate k's).  Memoised
    (a few k values' tables): callers must not modify the returned list."""
    kd = base_digits(k, p)
    assert 1 <= T <= len(kd)
    # digit i ranges over [kd[i], p); residues = all sums of digit * p^i
    vals = np.zeros(1, dtype=np.int64)
    for i in range(T):
        digit_vals = np.arange(kd[i], p, dtype=np.int64) * p ** i
        vals = np.add.outer(digit_vals, vals).ravel()
    vals.sort()
    return p ** T, vals.tolist()


def allowed_residues(k: int, p: int) -> tuple[int, list[int]]:
    """Full-period table of prime p: (q_p, admissible residues mod q_p)."""
    return partial_residues(k, p, len(base_digits(k, p)))


def partial_count(k: int, p: int, T: int) -> int:
    """Number of admissible residues mod p^T: prod_{i<T} (p - a_i)."""
    return prod(p - d for d in base_digits(k, p)[:T])


# --------------------------------------------------------------------------- planner

def _knapsack_tables(k: int, ps: list[int], scale: int, max_M_bits: int):
    """Group knapsack over DP cells (scale cells per bit of log2 M), one exponent per prime.

    best[c] is the least log2(R) over wheels whose rounded weights sum to exactly c cells;
    pick[i][c] is the exponent chosen for ps[i] on that optimum (0 = not in the wheel).
    Same recurrence, rounding and tie rule (strict improvement only, exponents tried in
    increasing order) as sieve/gpu_sieve_v9.py:146-173."""
    ncell = int(max_M_bits * scale) + 1
    best = np.full(ncell, np.inf)
    best[0] = 0.0
    pick, cells = [], []
    for p in ps:
        nd = len(base_digits(k, p))
        nxt = best.copy()
        chosen = np.zeros(ncell, dtype=np.int8)
        wmap = {}
        for T in range(1, nd + 1):
            w = max(1, round(T * math.log2(p) * scale))
            gain = math.log2(partial_count(k, p, T))
            wmap[T] = w
            if w >= ncell:
                continue
            trial = best[:ncell - w] + gain
            win = trial < nxt[w:]
            nxt[w:][win] = trial[win]
            chosen[w:][win] = T
        best = nxt
        pick.append(chosen)
        cells.append(wmap)
    return best, pick, cells


def _walk_back(ps, pick, cells, c) -> list[tuple[int, int]]:
    items = []
    for i in reversed(range(len(ps))):
        T = int(pick[i][c])
        if T:
            items.append((ps[i], T))
            c -= cells[i][T]
    assert c == 0
    return sorted(items)


def plan_wheel(k: int, ps: list[int], log2E: float, tun: dict) -> list[tuple[int, int]]:
    """Choose the wheel items for k from candidate primes ps (sieve/gpu_sieve_v9.py:187-227).

    truecost: minimise log2(R) - log2(1 - exp(-M/E)) over the DP cells (expected checks
    including block overshoot); knapsack: maximise log2(M) - log2(R) under log2(n_mult*E);
    target_g > 0 (g already known, confirmation runs): minimise ceil(g/M) * R.
    The exact integers M < 2^126 and R < 2^62 are then enforced by dropping top digits."""
    scale = int(tun.get("cost_scale", 256))
    max_M_bits = min(int(tun.get("m_cap_bits", 120)), MAX_M_BITS)
    max_R_bits = float(tun.get("max_R_bits", 62))
    if not ps:
        return []
    best, pick, cells = _knapsack_tables(k, ps, scale, max_M_bits)
    last = len(best) - 1
    cidx = np.arange(last + 1)
    ok = np.isfinite(best) & (best <= max_R_bits)
    if tun.get("planner", "truecost") == "knapsack":
        target = min(last, int(math.floor((log2E + math.log2(float(tun.get("n_mult", 1.0)))) * scale)))
        ok &= cidx <= max(target, 0)
        score = np.where(ok, cidx / scale - best, -np.inf)
        c = int(np.argmax(score))
        if not ok[c]:
            c = 0
    elif int(tun.get("target_g", 0)) > 0:
        # g known (sieve/gpu_sieve_v11.py:199-213): covering [0, g) takes ceil(g / M) blocks
        # of R positions, so the cost of cell c is log2 R + log2 ceil(2^(log2 g - c/scale))
        over = math.log2(int(tun["target_g"])) - cidx / scale
        with np.errstate(over="ignore", invalid="ignore"):
            blocks = np.where(over <= 0.0, 1.0, np.ceil(np.exp2(np.clip(over, -200.0, 200.0))))
        with np.errstate(divide="ignore", invalid="ignore"):
            cost = np.where(ok & np.isfinite(blocks), best + np.log2(blocks), np.inf)
        c = int(np.argmin(cost))
        if not np.isfinite(cost[c]):
            c = 0
    else:
        x = np.minimum(2.0 ** np.clip(cidx / scale - log2E, -200.0, 60.0), 1e300)
        hit = -np.expm1(-x)
        with np.errstate(divide="ignore", invalid="ignore"):
            cost = np.where(ok & (hit > 0), best - np.log2(hit), np.inf)
        c = int(np.argmin(cost))
        if not np.isfinite(cost[c]):
            c = 0
    wheel = _walk_back(ps, pick, cells, c)

    def drop_weakest_digit(wh):
        # the item whose top digit filters least per bit of modulus loses that digit
        def value_per_bit(it):
            p, T = it
            hi, lo = partial_count(k, p, T), (partial_count(k, p, T - 1) if T > 1 else 1)
            return (math.log2(p) - math.log2(hi / lo)) / math.log2(p)
        it = min(wh, key=value_per_bit)
        wh.remove(it)
        if it[1] > 1:
            wh.append((it[0], it[1] - 1))
            wh.sort()

    while wheel and (prod(p ** T for p, T in wheel) >= (1 << MAX_M_BITS)
                     or prod(partial_count(k, p, T) for p, T in wheel) >= (1 << 62)):
        drop_weakest_digit(wheel)
    return wheel


def fullpower_wheel(k: int, p_full: int = 23, m_cap_bits: float = 100.0,
                    r_cap_bits: int = 62) -> list[tuple[int, int]]:
    """Wheel of full prime powers q_p for every p <= p_full, extended by the exponent choice
    (p, T) for larger primes that maximises M/R (integers covered per position) subject to
    M <= 2^m_cap_bits and R < 2^r_cap_bits.  The search is exhaustive: an item can only enter
    when its radix fits the R budget the full powers leave, which keeps the candidate set small
    (at k = 200: 29, 41, 53, 67, 101, 103 with T = 1).  Ties go to the smaller R, then to the
    lexicographically smaller item list."""
    ps = primes_upto(k)
    base = [(p, len(base_digits(k, p))) for p in ps if p <= p_full]
    logM = sum(T * math.log2(p) for p, T in base)
    R0 = prod(partial_count(k, p, T) for p, T in base)
    M0 = prod(p ** T for p, T in base)
    assert logM <= m_cap_bits and R0 < (1 << r_cap_bits), "full powers exceed the caps"
    r_room = ((1 << r_cap_bits) - 1) // R0          # R0 * extra < 2^r_cap_bits
    m_room = (1 << int(math.floor(m_cap_bits))) // M0 if m_cap_bits == int(m_cap_bits) else None
    items = []                                      # (p, T, q, count)
    for p in ps:
        if p <= p_full:
            continue
        for T in range(1, len(base_digits(k, p)) + 1):
            c = partial_count(k, p, T)
            if c <= r_room and math.log2(p) * T <= m_cap_bits - logM:
                items.append((p, T, p ** T, c))
    best = (0.0, 1, ())                             # (log2 M/R gain, R factor, items)
    budget = [1 << 20]

    def fits_m(mq):
        if m_room is not None:
            return mq <= m_room
        return math.log2(M0 * mq) <= m_cap_bits

    def dfs(i, mq, rc, gain, chosen, used):
        nonlocal best
        budget[0] -= 1
        assert budget[0] > 0, "wheel search too large"
        key = (gain, -rc, tuple(chosen))
        if (key[0] > best[0] + 1e-12 or (abs(key[0] - best[0]) <= 1e-12 and
                                         (rc < best[1] or (rc == best[1] and
                                                           tuple(chosen) < best[2])))):
            best = (gain, rc, tuple(chosen))
        for j in range(i, len(items)):
            p, T, q, c = items[j]
            if p in used or rc * c > r_room or not fits_m(mq * q):
                continue
            used.add(p)
            chosen.append((p, T))
            dfs(j + 1, mq * q, rc * c, gain + math.log2(q / c), chosen, used)
            chosen.pop()
            used.discard(p)

    dfs(0, 1, 1, 0.0, [], set())
    return sorted(base + list(best[2]))


def order_filters(k: int, filt: list[int], info: dict, wheelT: dict[int, int]) -> list[int]:
    """Filter order: smallest conditional pass rate first; a split prime's low digits were
    already fixed by its wheel ring, so its rate is conditioned on them
    (sieve/gpu_sieve_v9.py:230-237).  Stable sort: ties keep ascending p."""
    def rate(p):
        q, res = info[p]
        T = wheelT.get(p, 0)
        return (len(res) / q) / (partial_count(k, p, T) / p ** T if T else 1.0)
    return sorted(filt, key=rate)


def normalise_wheel(wheel) -> list[tuple[int, int | None]]:
    """Accept [[p, T], ...] or bare p (= full power) items (sieve/gpu_sieve_v9.py:334-341)."""
    out = []
    for it in wheel:
        if isinstance(it, (list, tuple)):
            out.append((int(it[0]), int(it[1])))
        else:
            out.append((int(it), None))
    return out


class Plan:
    """Wheel plan for one k (sieve/gpu_sieve_v9.py:344-413).

    Attributes used by engines and drivers: k, tun, info {p: (q_p, residues)}, tp, wheel (ring
    order: radix descending, then p), wheelT, rings [(p^T, residues)], moduli, M, rest (filter
    primes in kernel order), basis, radix, R, density_all/wheel/rest."""

    def __init__(self, k: int, tun: dict | None = None, wheel=None):
        self.k = k
        self.tun = dict(TUNABLES, **(tun or {}))
        ps = primes_upto(k)
        self.info = {p: allowed_residues(k, p) for p in ps}
        self.tp = {p: len(base_digits(k, p)) for p in ps}
        self.density_all = prod(len(v[1]) / v[0] for v in self.info.values()) if ps else 1.0
        log2E = -math.log2(self.density_all) if self.density_all > 0 else 0.0
        if wheel is not None:
            self.wheel = sorted((p, self.tp[p] if T is None else T) for p, T in normalise_wheel(wheel))
        else:
            self.wheel = plan_wheel(k, ps, log2E, self.tun)
        for p, T in self.wheel:
            assert p in self.tp and 1 <= T <= self.tp[p], f"bad wheel item {(p, T)}"
        assert len({p for p, _ in self.wheel}) == len(self.wheel), "prime used twice in wheel"
        self.wheelT = dict(self.wheel)
        # ring 0 (swept inside a thread) is the largest ring; the order is part of the idx map
        self.wheel.sort(key=lambda it: (-partial_count(k, it[0], it[1]), it[0]))
        self.rings = [partial_residues(k, p, T) for p, T in self.wheel]
        self.moduli = [q for q, _ in self.rings]
        self.M = prod(self.moduli)
        assert self.M < (1 << MAX_M_BITS), "wheel modulus too large for 128-bit path"
        filt = [p for p in ps if self.wheelT.get(p, 0) < self.tp[p]]
        self.rest = order_filters(k, filt, self.info, self.wheelT)
        self.basis = []
        for q in self.moduli:
            Mi = self.M // q
            self.basis.append(Mi * pow(Mi, -1, q))
        self.radix = [len(res) for _, res in self.rings]
        self.R = prod(self.radix)
        assert self.R < (1 << 62), "too many residues per block"
        self.density_wheel = self.R / self.M if self.M else 1.0
        self.density_rest = self.density_all / self.density_wheel

    # -- index maps --------------------------------------------------------------------
    def ring_terms(self) -> list[list[int]]:
        """CRT terms (a * basis_j) mod M of every ring, in ring order."""
        return [[(a * b) % self.M for a in res] for (_, res), b in zip(self.rings, self.basis)]

    def idx_to_s(self, idx: int) -> int:
        """Residue s in [0, M) of mixed-radix position idx (ring 0 fastest)."""
        r = 0
        for (q, res), rad, b in zip(self.rings, self.radix, self.basis):
            idx, d = divmod(idx, rad)
            r += res[d] * b
        return r % self.M

    def idx_to_n(self, idx: int, t: int) -> int:
        return t * self.M + self.idx_to_s(idx)

    def n_to_idx(self, n: int) -> tuple[int, int] | None:
        """Inverse map: (idx, t) of n, or None if n mod some ring is not admissible."""
        t, s = divmod(n, self.M)
        idx, place = 0, 1
        for (q, res), rad in zip(self.rings, self.radix):
            r = s % q
            lo, hi = 0, len(res)
            while lo < hi:
                mid = (lo + hi) // 2
                if res[mid] < r:
                    lo = mid + 1
                else:
                    hi = mid
            if lo == len(res) or res[lo] != r:
                return None
            idx += lo * place
            place *= rad
        return idx, t

    # -- reporting ---------------------------------------------------------------------
    def summary(self) -> dict:
        E = 1 / self.density_all
        return {"k": self.k, "wheel": [[p, T] for p, T in self.wheel],
                "wheel_primes": [p for p, _ in self.wheel],
                "split_primes": [[p, T, self.tp[p]] for p, T in self.wheel if T < self.tp[p]],
                "N_bits": self.M.bit_length(), "M_bits": self.M.bit_length(), "M": str(self.M),
                "N": str(self.M), "n_filters": len(self.rest),
                "residues_per_block": self.R, "density_wheel": self.density_wheel,
                "density_rest": self.density_rest, "expected_g": E, "E_g": E,
                "expected_blocks": E / self.M, "N_over_E": self.M / E,
                "expected_residue_checks": self.density_wheel / self.density_all,
                "expected_checks_true": self._true_cost()}

    def _true_cost(self) -> float:
        """R / (1 - e^{-M/E}): expected residue checks including block overshoot."""
        try:
            E = 1.0 / self.density_all
            x = self.M / E
            denom = -math.expm1(-x) if x < 40 else 1.0
            return self.R / denom
        except (ZeroDivisionError, OverflowError):
            return float("inf")

    def filter_pass_rates(self) -> list[float]:
        """Conditional pass rate of each filter in kernel order."""
        out = []
        for p in self.rest:
            q, res = self.info[p]
            T = self.wheelT.get(p, 0)
            out.append((len(res) / q) / (partial_count(self.k, p, T) / p ** T if T else 1.0))
        return out

    def expected_filter_evals(self) -> float:
        """E_f = sum_j prod_{i<j} pi_i: filter evaluations per position with early exit."""
        e, run = 0.0, 1.0
        for pi in self.filter_pass_rates():
            e += run
            run *= pi
        return e
