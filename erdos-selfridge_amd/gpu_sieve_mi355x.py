"""MI355X engine module: drop-in for the reference's engine modules.

Same module surface as sieve/gpu_sieve_v9.py / sieve/gpu_sieve_v8_masks.py, which drivers load
by import name (harness/run_k.py:34, harness/run_k_masks.py:36) or run as a CLI
(harness/replay_release.py:25): ``Plan``, ``plan_wheel``, ``TUNABLES``, ``GpuSieve(plan)``,
``free_vram_mib()``, ``load_known(*paths)`` and ``main()``.  ``GpuSieve`` keeps the engine
object contract those drivers read and write (sieve/gpu_sieve_v9.py:425-568): ``rest_q``,
``rest_Mmodq``, ``tconst_h`` / ``tconst.set``, ``out_n.fill(0)`` / ``out_n.get()[0]``,
``out[:n].get()``, ``out_cap``, ``_launch(ostart, ocount)``, ``residues_checked``,
``launches``, ``r0``, ``R_outer``, ``ochunk``, ``chunk``, ``mask_mode``, ``scan_block(t)`` and
``search(max_blocks, log)``.  All sieving runs in libes_sieve.so (HIP, gfx950).
"""
from __future__ import annotations

import argparse
import ctypes as C
import functools
import json
import math
import os
import signal
import sys
import time

import numpy as np

import es_data
import es_native as nat
from es_plan import (MAX_M_BITS, TUNABLES, Plan, allowed_residues, base_digits, is_good,  # noqa: F401
                     order_filters, partial_count, partial_residues, plan_wheel, primes_upto)

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
MASK_MOD_CAP = 4096          # largest prefilter modulus staged in LDS
MAX_MASKS = 12
ROWS_MAX_MASKS = 11          # row-lane kernel instances exist for 1..11 masks
AUTO_MASK_STOP = 0.01        # "auto" masks: stop below this many candidate slots per outer index


# --------------------------------------------------------------------------- table building

def pack_bits(q: int, residues) -> np.ndarray:
    """One bit per residue of [0, q), padded to whole u32 words."""
    words = np.zeros((q + 31) // 32, dtype=np.uint32)
    r = np.asarray(residues, dtype=np.int64)
    if r.size:
        np.bitwise_or.at(words, r >> 5, (np.uint32(1) << (r & 31).astype(np.uint32)))
    return words


def kernel_rings(plan: Plan, max_inner: int = 64):
    """Rings as the kernel sweeps them: the plan's rings, except that a ring 0 mod p^T (T >= 2)
    whose radix exceeds max_inner is passed as digit rings -- its lowest digits grouped while
    their radix stays <= max_inner (at least one digit), then one ring per further digit.  The
    ascending admissible residues of p^T are the mixed radix over the digits (digit 0 fastest,
    digit j over [a_j, p)), so the idx order is unchanged.  Returns (moduli, radices, terms,
    inner_mod, f) with inner_mod = p^T (0 when ring 0 is not split) and f = plan ring-0 radix
    divided by the kernel's (outer indices of the plan = f outer indices of the kernel)."""
    if not plan.rings:
        return [1], [1], [[0]], 0, 1
    mods = [q for q, _ in plan.rings]
    radix = list(plan.radix)
    terms = plan.ring_terms()
    p, T = plan.wheel[0]
    if max_inner <= 0 or radix[0] <= max_inner or T < 2:
        return mods, radix, terms, 0, 1
    a = base_digits(plan.k, p)[:T]
    groups = [[0]]
    rad = p - a[0]
    for j in range(1, T):
        if len(groups) == 1 and rad * (p - a[j]) <= max_inner:
            groups[0].append(j)
            rad *= p - a[j]
        else:
            groups.append([j])
    M, b0 = plan.M, plan.basis[0]
    g_mods, g_radix, g_terms = [], [], []
    for g in groups:
        rhos = [0]
        for j in g:                      # lowest digit of the group fastest
            rhos = [r + d * p ** j for d in range(a[j], p) for r in rhos]
        g_mods.append(p ** len(g))
        g_radix.append(len(rhos))
        g_terms.append([(r * b0) % M for r in rhos])
    f = radix[0] // g_radix[0]
    return g_mods + mods[1:], g_radix + radix[1:], g_terms + terms[1:], p ** T, f


LEAD_CAP = 8192   # positions of the lead block (the rings before the one brought forward)
LEAD_GAIN = 1.15  # least gain in mask-word use (slots per word bit) that brings a ring forward


def swap_lead(radix, inner_mod: int = 0, force: bool = False) -> int:
    """j >= 1 when the kernel sweeps the plan's ring j as its ring 0 (one word per mask row)
    and the rings 0..j-1 before it as its first strip rings (the lead block): when ring 0 has
    33..128 residues (two to four words per row) and is not split into digit rings, ring j is
    the first ring of at most 32, the lead block holds at most LEAD_CAP positions, and (unless
    `force`) ring j fills its word at least LEAD_GAIN times as well as ring 0 fills its words
    (or ring 0 needs four words).  0 otherwise.  idx keeps the plan's order; the engine maps
    it.  The fill rule is a cheap guess: GpuSieve times both orders where a swap is possible
    (on 2^36 windows of the 84 frontier wheels that can swap, the rule took the slower order
    by more than 5 % for 11: k=372/383 run 1.31-1.33x faster swapped against it, k=325 1.12x
    faster unswapped; k=378c/400c would lose 20 % swapped and the rule keeps them)."""
    if inner_mod or len(radix) < 2 or not 32 < radix[0] <= 128:
        return 0
    L = 1
    for j in range(1, len(radix)):
        L *= radix[j - 1]
        if L > LEAD_CAP:
            return 0
        if radix[j] <= 32:
            if force:
                return j
            fill0 = radix[0] / (32 * -(-radix[0] // 32))
            return j if radix[0] > 96 or radix[j] / 32 >= LEAD_GAIN * fill0 else 0
    return 0


def swap_mode(tun: dict):
    """The mi_ring_swap tunable as 0, 1, "rule" or "auto" (ints may come as strings)."""
    v = tun.get("mi_ring_swap", "auto")
    if v in ("auto", "rule"):
        return v
    v = int(v)
    if v not in (0, 1):
        raise ValueError(f"mi_ring_swap must be 'auto', 'rule', 0 or 1, not {v!r}")
    return v


def lead_order(n: int, j: int) -> list[int]:
    """Kernel ring order for a lead block of j rings: ring j, then rings 0..j-1, then the
    rest (identity for j = 0)."""
    return list(range(n)) if not j else [j] + list(range(j)) + list(range(j + 1, n))


@functools.lru_cache(maxsize=512)
def _table_words(k: int, p: int, T: int) -> np.ndarray:
    """pack_bits of the admissible residues mod p^T (memoised; read-only)."""
    q, res = partial_residues(k, p, T)
    w = pack_bits(q, res)
    w.setflags(write=False)
    return w


def engine_geometry(radix, lead_rings: int = 0, strip_target: int = 1024):
    """Strip length S, strip ring count m and steps per unit V the engine derives from the
    kernel's ring radices (es_engine_create): strip rings 1..m are the leading product nearest
    the target in log scale (a lead block is always whole), the sub rings after them the
    product up to 1024."""
    W = (radix[0] + 31) // 32 if radix else 1
    S, m = 1, 0
    while 1 + m < len(radix):
        nxt = S * radix[1 + m]
        if nxt > (1 << 22) or nxt * W >= (1 << 26):
            break
        if m > 0 and m >= lead_rings and nxt > strip_target:
            if nxt * S < strip_target * strip_target:
                S, m = nxt, m + 1
            break
        S, m = nxt, m + 1
    V, j = 1, 1 + m
    while j < len(radix) and V * radix[j] <= 1024:
        V *= radix[j]
        j += 1
    return S, m, V


def rows_layout(radix, lead_rings: int, tun: dict) -> bool:
    """Whether the engine sweeps this ring order in the row-lane layout (every mask in LDS, a
    strip of >= 32 rows and >= 16 steps; one-word rows under "auto", two-word rows too under
    "rows")."""
    mode = str(tun.get("mi_layout", "auto"))
    if mode == "strips" or int(tun.get("mi_global_masks", 0)):
        return False
    if not radix or radix[0] > (64 if mode == "rows" else 32):
        return False
    S, _, V = engine_geometry(radix, lead_rings, int(tun.get("mi_strip_target", 1024)))
    return S >= 32 and V >= 16


def choose_masks(plan: Plan, n_masks: int, limit: int | None = None,
                 lds_budget: int = 65536, r0: int | None = None, q0: int | None = None,
                 stop: float = 0.0, rows: bool = False):
    """Word-mask prefilters: low-digit conditions mod p^D (D <= t_p, p^D <= MASK_MOD_CAP) of the
    filter primes outside the wheel -- each a necessary condition of its full test (a wheel
    prime divides M, and the engine's mask tables need M invertible mod p^D).  Candidates are taken
    by least conditional pass rate, one depth per prime, while their LDS rows (each table held
    twice over, two words per ring-0 word per row) fit the budget beside the ring-0 carry table.
    With `stop` > 0, masks are added only while the expected candidate slots per outer index
    (r0 x the pass rates taken) is at least `stop`: below that, a further mask costs more LDS
    reads than the queued exact re-test it saves.  rows: the row-lane layout's tables, rows
    [0, 3m + 64) of one 8-byte row per ring-0 word (a table per word), instead of [0, 2m) of
    two words per ring-0 word.  Returns [(filter_index, D, modulus, conditional pass rate)]."""
    k = plan.k
    rest = plan.rest if limit is None else plan.rest[:limit]
    if r0 is None:
        r0 = plan.radix[0] if plan.radix else 1
    if q0 is None:
        q0 = plan.rings[0][0] if plan.rings else 1
    words = (r0 + 31) // 32
    cands = []
    for fi, p in enumerate(rest):
        T = plan.wheelT.get(p, 0)
        if T:
            continue
        base = 1.0
        for d in range(1, plan.tp[p] + 1):
            if p ** d > MASK_MOD_CAP:
                break
            rate = (partial_count(k, p, d) / p ** d) / base
            cands.append((rate, fi, d, p ** d))
    cands.sort(key=lambda c: (c[0], c[1], -c[2]))
    # carry table: q0 rows of `words` words, or q0 u16 suffix starts when words > 2; row-lane
    # rows of several words: q0 bytes (the one-word row-lane table is q0 bytes too: the q0
    # words counted for it leave a margin)
    if rows and words > 1:
        carry = (q0 + 3) // 4
    else:
        carry = (q0 + 1) // 2 if words > 2 else q0 * words
    room = lds_budget // 4 - carry - 8

    def cost_of(m):
        return (2 * words * (3 * m + 64) + 4) if rows else (4 * words * m + 4)

    def greedy(left):
        taken, out = set(), []
        dens = float(r0)
        for rate, fi, d, m in cands:
            if len(out) >= n_masks or dens < stop:
                break
            cost = cost_of(m)
            if fi in taken or cost > left:
                continue
            taken.add(fi)
            left -= cost
            dens *= rate
            out.append((fi, d, m, rate))
        return out

    out = greedy(room)
    if stop > 0 and out != greedy(1 << 62):
        # The budget cut the greedy short (a large table taken first can crowd out several
        # small ones): choose within the budget by the kernel's cost per outer index instead.
        alt = _masks_within_budget(cands, n_masks, room, cost_of, r0, words)
        if _mask_cost(alt, r0, words) < _mask_cost(out, r0, words):
            out = alt
    return out


# Cost of a mask set per outer index, in mask-row reads: one read per mask and ring-0 word and
# the carry read, plus ~160 reads' worth of work per candidate slot queued for the exact
# re-test (row-lane counters: k=301 with 0.041 slots per outer index runs its wave-steps 1.67x
# slower than k=304 with 0.009, at the same six reads per step).
_EXACT_PER_SLOT = 160.0


def _mask_cost(masks, r0: int, words: int = 1) -> float:
    dens = float(r0)
    for _, _, _, rate in masks:
        dens *= rate
    return words * len(masks) + 1 + _EXACT_PER_SLOT * dens


def _masks_within_budget(cands, n_masks: int, room: int, cost_of, r0: int, words: int = 1):
    """Least-cost mask set (one depth per prime, at most n_masks, LDS words within room): a
    knapsack over the primes' depth options in units of 64 words, minimising the sum of log
    pass rates for each (count, units), then the cost of each reachable state."""
    unit = 64
    U = max(0, room) // unit
    by_fi = {}
    for rate, fi, d, m in cands:
        by_fi.setdefault(fi, []).append((rate, d, m, -(-cost_of(m) // unit)))
    inf = float("inf")
    best = np.full((n_masks + 1, U + 1), inf)
    best[0, 0] = 0.0
    choice = []                                         # per prime: argmin option per state
    for fi, opts in sorted(by_fi.items()):
        nb = best.copy()
        pick = np.full(best.shape, -1, dtype=np.int32)
        for oi, (rate, d, m, cu) in enumerate(opts):
            if cu > U or rate <= 0.0:
                continue
            lr = math.log(rate)
            cand = np.full(best.shape, inf)
            cand[1:, cu:] = best[:-1, :U + 1 - cu] + lr
            better = cand < nb
            nb[better] = cand[better]
            pick[better] = oi
        choice.append((fi, opts, pick))
        best = nb
    # cost of every reachable (count, units) state; walk the choices back from the best one
    cnt = np.arange(n_masks + 1)[:, None]
    cost = words * cnt + 1 + _EXACT_PER_SLOT * r0 * np.exp(np.minimum(best, 0.0))
    cost[~np.isfinite(best)] = inf
    c, u = np.unravel_index(int(np.argmin(cost)), cost.shape)
    out = []
    for fi, opts, pick in reversed(choice):
        oi = int(pick[c, u])
        if oi < 0:
            continue
        rate, d, m, cu = opts[oi]
        out.append((fi, d, m, rate))
        c, u = c - 1, u - cu
    out.sort(key=lambda x: (x[3], x[0], -x[1]))
    return out


def slow_order(plan: Plan, masks, limit: int | None = None) -> list[int]:
    """Filters re-tested exactly per candidate: all of them (first `limit` if given) except
    those a mask already covers in full, least conditional pass rate (given the masks) first."""
    k = plan.k
    rest = plan.rest if limit is None else plan.rest[:limit]
    by_fi = {fi: D for fi, D, _, _ in masks}
    keyed = []
    for fi, p in enumerate(rest):
        T = plan.wheelT.get(p, 0)
        D = max(T, by_fi.get(fi, 0))
        if D >= plan.tp[p]:
            continue                      # fully decided by the wheel ring or by its mask
        done = partial_count(k, p, D) / p ** D if D else 1.0
        q, res = plan.info[p]
        keyed.append(((len(res) / q) / done, fi))
    keyed.sort()
    return [fi for _, fi in keyed]


class _Desc:
    """es_plan_desc plus the numpy arrays it points into (kept alive together)."""

    def __init__(self, plan: Plan, tun: dict):
        k = plan.k
        M = plan.M
        mods, radix, rterms, self.inner_mod, self.f = kernel_rings(
            plan, int(tun.get("mi_inner_radix", 128)))
        self.lead = 0
        # "auto" reaches here only from callers that build a descriptor without timing (a
        # GpuSieve resolves it to 0 or 1 first): it then follows the fill rule
        mode = swap_mode(tun)
        j = (swap_lead(radix, self.inner_mod, force=mode == 1)
             if mode != 0 and not int(tun.get("mi_global_masks", 0)) else 0)
        if j:
            self.lead = math.prod(radix[:j])
            order = lead_order(len(radix), j)
            mods = [mods[i] for i in order]
            radix = [radix[i] for i in order]
            rterms = [rterms[i] for i in order]
        terms = [t for ring in rterms for t in ring]
        self.r0 = radix[0]
        self.q0 = self.inner_mod or mods[0]
        self.ring_mod = np.asarray(mods, dtype=np.uint32)
        self.ring_radix = np.asarray(radix, dtype=np.uint32)
        self.term_lo = np.asarray([t & ((1 << 64) - 1) for t in terms], dtype=np.uint64)
        self.term_hi = np.asarray([t >> 64 for t in terms], dtype=np.uint64)
        limit = tun.get("mi_filter_limit")
        limit = None if limit is None else int(limit)
        # test-only knob: apply just the first `limit` filters (dense survivor sets)
        self.filters = list(plan.rest if limit is None else plan.rest[:limit])
        qs, offs, words = [], [], []
        nw = 0
        for p in self.filters:
            q, _ = plan.info[p]
            qs.append(q)
            offs.append(nw)
            w = _table_words(k, p, plan.tp[p])
            words.append(w)
            nw += w.size
        self.filt_q = np.asarray(qs or [2], dtype=np.uint32)
        self.filt_off = np.asarray(offs or [0], dtype=np.uint32)
        self.filt_bits = np.concatenate(words) if words else np.zeros(1, dtype=np.uint32)
        nm = tun.get("mi_masks", "auto")
        auto = nm == "auto"
        lead_rings = 0
        if self.lead:
            prod_ = 1
            while prod_ < self.lead:
                lead_rings += 1
                prod_ *= radix[lead_rings]
        self.rows = rows_layout(radix, lead_rings, tun)
        self.masks = choose_masks(plan, (ROWS_MAX_MASKS if self.rows else MAX_MASKS) if auto
                                  else min(MAX_MASKS, int(nm)), limit,
                                  int(tun.get("mi_lds_budget", 65536)), self.r0, self.q0,
                                  AUTO_MASK_STOP if auto else 0.0, rows=self.rows)
        self.slow = slow_order(plan, self.masks, limit)
        mm, ms, mo, mw = [], [], [], []
        nw = 0
        for fi, D, m, _ in self.masks:
            p = plan.rest[fi]
            w = _table_words(k, p, D)
            mm.append(m)
            ms.append(fi)
            mo.append(nw)
            mw.append(w)
            nw += w.size
        self.mask_mod = np.asarray(mm or [2], dtype=np.uint32)
        self.mask_src = np.asarray(ms or [0], dtype=np.uint32)
        self.mask_off = np.asarray(mo or [0], dtype=np.uint32)
        self.mask_bits = np.concatenate(mw) if mw else np.zeros(1, dtype=np.uint32)
        self.slow_arr = np.asarray(self.slow or [0], dtype=np.uint32)
        d = nat.PlanDesc()
        d.abi_version = nat.ES_ABI_VERSION
        d.k = k
        d.n_rings = len(mods)
        d.ring_mod = nat.ptr32(self.ring_mod)
        d.ring_radix = nat.ptr32(self.ring_radix)
        d.term_lo = nat.ptr64(self.term_lo)
        d.term_hi = nat.ptr64(self.term_hi)
        d.M_lo = M & ((1 << 64) - 1)
        d.M_hi = M >> 64
        d.n_filters = len(qs)
        d.filt_q = nat.ptr32(self.filt_q)
        d.filt_off = nat.ptr32(self.filt_off)
        d.filt_bits = nat.ptr32(self.filt_bits)
        d.n_filt_words = int(self.filt_bits.size)
        d.n_slow = len(self.slow)
        d.slow_order = nat.ptr32(self.slow_arr)
        d.n_masks = len(mm)
        d.mask_mod = nat.ptr32(self.mask_mod)
        d.mask_src = nat.ptr32(self.mask_src)
        d.mask_off = nat.ptr32(self.mask_off)
        d.mask_bits = nat.ptr32(self.mask_bits)
        d.n_mask_words = int(self.mask_bits.size)
        d.strip_target = int(tun.get("mi_strip_target", 1024))
        th = tun.get("threads", "auto")
        d.threads = 0 if th == "auto" else int(th)
        d.out_cap = int(tun.get("max_survivors", 1 << 20))
        d.lds_budget = int(tun.get("mi_lds_budget", 65536))
        d.global_masks = int(tun.get("mi_global_masks", 0))
        d.launch_log2 = int(tun.get("mi_launch_log2", 37))
        d.inner_mod = self.inner_mod
        d.lead_radix = self.lead
        d.layout = {"auto": 0, "strips": 1, "rows": 2}[str(tun.get("mi_layout", "auto"))]
        self.desc = d


# --------------------------------------------------------------------------- device proxies

class _Counter:
    """out_n: the device survivor counter (fill(0) resets it, get() reads it)."""

    def __init__(self, eng: "GpuSieve"):
        self._e = eng

    def fill(self, v: int) -> None:
        assert v == 0, "the survivor counter can only be reset to 0"
        nat.check(nat.lib().es_engine_reset(self._e._h))

    def get(self) -> np.ndarray:
        n = C.c_uint64(0)
        nat.check(nat.lib().es_engine_count(self._e._h, C.byref(n)))
        return np.asarray([n.value], dtype=np.uint64)


class _OutSlice:
    def __init__(self, eng: "GpuSieve", lo: int, hi: int):
        self._e, self._lo, self._hi = eng, lo, hi

    def get(self) -> np.ndarray:
        n = max(0, self._hi - self._lo)
        buf = np.zeros(n, dtype=np.uint64)
        if n:
            nat.check(nat.lib().es_engine_read(self._e._h, self._lo, n, nat.ptr64(buf)))
        return buf


class _Out:
    """out: the device survivor buffer; out[a:b].get() copies IDs back."""

    def __init__(self, eng: "GpuSieve"):
        self._e = eng

    def __getitem__(self, sl) -> _OutSlice:
        assert isinstance(sl, slice) and sl.step in (None, 1)
        lo = 0 if sl.start is None else int(sl.start)
        hi = self._e.out_cap if sl.stop is None else min(int(sl.stop), self._e.out_cap)
        return _OutSlice(self._e, lo, hi)


class _TConst:
    """tconst: per-block filter constants; set(arr) uploads arr (u32, one per filter)."""

    def __init__(self, eng: "GpuSieve"):
        self._e = eng

    def set(self, arr) -> None:
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.uint32)[:len(self._e.rest_q)])
        nat.check(nat.lib().es_engine_set_tconst(self._e._h, nat.ptr32(a), len(self._e.rest_q)))
        self._e._tconst_block = None          # constants from the caller: no block known


# --------------------------------------------------------------------------- engine

def timed_rate(scan, R: int, strip_positions: int, max_s: float = 0.05,
               strips_log2: int = 19, lo_log2: int = 28, hi_log2: int = 34,
               floor: float = 0.0, with_window: bool = False):
    """Positions per second of a layout, from scan(lo, hi) -> rate (device time) on windows
    centred in [0, R).  The target window holds 2^strips_log2 strips (about 8 workgroups per CU,
    so long strips are not timed on an under-filled GPU), clamped to [2^lo_log2, 2^hi_log2]
    positions.  Windows grow from 2^24 positions, each to what the latest rate says fits in
    max_s, until the target is reached or the next window would not grow by a quarter; the
    last rate is returned (with its window (lo, hi) when with_window).  A layout measuring
    below `floor` (half the best rate so far, say) after its second window stops there: it
    cannot win."""
    n = max(1, min(R // 2, 1 << hi_log2, max(1 << lo_log2, strip_positions << strips_log2)))
    w = max(1, min(n, 1 << 24))
    rate = scan((R - w) // 2, (R - w) // 2 + w)
    for i in range(8):
        if w >= n or (i >= 1 and rate < floor):
            break
        want = int(rate * max_s)
        if want < w + w // 4:
            break
        w = min(n, want)
        rate = scan((R - w) // 2, (R - w) // 2 + w)
    return (rate, ((R - w) // 2, (R - w) // 2 + w)) if with_window else rate


class GpuSieve:
    """One wheel plan on the current GPU (sieve/gpu_sieve_v9.py:425-568)."""

    def __init__(self, plan: Plan):
        self.plan = plan
        P = plan
        self._h = None
        L = nat.lib()
        self.inner_bound, self.ring_swap, self.lane_layout = self._choose_layout(P)
        self._d = _Desc(P, dict(P.tun, mi_inner_radix=self.inner_bound,
                                mi_ring_swap=self.ring_swap, mi_layout=self.lane_layout))
        h = C.c_void_p()
        nat.check(L.es_engine_create(C.byref(self._d.desc), C.byref(h)))
        self._h = h
        info = nat.EngineInfo()
        nat.check(L.es_engine_info_get(h, C.byref(info)))
        self.info = info
        self.w = len(P.wheel)
        self.rest_q = [P.info[p][0] for p in self._d.filters]
        self.rest_Mmodq = [P.M % q for q in self.rest_q]
        self.nrest = len(self.rest_q)
        self.tconst_h = np.zeros(max(1, self.nrest), dtype=np.uint32)
        self.tconst = _TConst(self)
        self._tconst_block = None              # block whose constants are on the device
        self.out_cap = int(P.tun["max_survivors"])
        self.out = _Out(self)
        self.out_n = _Counter(self)
        self.r0 = P.radix[0] if P.radix else 1
        self.n_fast = info.n_masks_used
        self.threads = int(info.threads)       # the workgroup size the engine uses
        self.chunk = 1 << int(P.tun.get("mi_launch_log2", 37))
        self.R_outer = P.R // self.r0
        # the kernel may sweep ring 0 as digit rings: one plan outer index = f kernel ones
        self._f = self._d.f
        assert self.R_outer * self._f == info.R_outer and self.R_outer * self.r0 == P.R
        self.ochunk = max(1, self.chunk // self.r0)
        self.strip_len = info.strip_len
        self.layout = "rows" if info.layout == 2 else "strips"
        self.mask_mode = f"mi355x-{self.layout}-mask{info.n_masks_used}-l1{info.global_masks}"
        self.masks = self._d.masks
        self.launches = 0
        self.residues_checked = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                nat.lib().es_engine_destroy(h)
            except Exception:
                pass
            self._h = None

    @staticmethod
    def _layout_options(P: Plan) -> list[tuple[int, object, str]]:
        """Distinct kernel layouts of a plan as (mi_inner_radix, mi_ring_swap, mi_layout):
        the digit-split bounds that give distinct kernel inner radices (a ring 0 wider than 64
        slots mod p^T, T >= 2: whole ring, groups <= 128, 64, 32, 16); for each whose ring 0
        stays whole and could trade places with a later ring of <= 32 residues, both orders;
        and where the kernel's ring 0 then has two words (33..64 slots), both lane layouts
        (strip-lane, the "auto" default there, and row-lane).  Tunables other than "auto" fix
        their dimension."""
        inner = P.tun.get("mi_inner_radix", "auto")
        swap = swap_mode(P.tun)
        layout = str(P.tun.get("mi_layout", "auto"))
        glob = int(P.tun.get("mi_global_masks", 0))
        if inner != "auto":
            bounds = [int(inner)]
        elif not P.rings or P.radix[0] <= 64 or P.wheel[0][1] < 2:
            bounds = [128]
        else:
            bounds = [128, 64, 32, 16, 0]
        opts, seen = [], set()
        for bound in bounds:
            mods, radix, _, inner_mod, _ = kernel_rings(P, bound)
            if radix[0] in seen:
                continue
            seen.add(radix[0])
            can = not glob and swap_lead(radix, inner_mod, force=True) > 0
            if swap != "auto" or not can:
                swaps = [swap if swap != "auto" else 0]
            else:
                swaps = [0, 1]
            for sw in swaps:
                j = 0 if glob or sw == 0 else swap_lead(radix, inner_mod, force=sw == 1)
                r0k = radix[j]
                if layout == "auto" and not glob and 32 < r0k <= 64:
                    opts += [(bound, sw, "auto"), (bound, sw, "rows")]
                else:
                    opts.append((bound, sw, layout))
        return opts

    @staticmethod
    def _choose_layout(P: Plan, window_log2: int = 36, rounds: int = 3,
                       max_s: float = 0.05) -> tuple[int, object, str]:
        """(mi_inner_radix, mi_ring_swap, mi_layout) for the engine: with one layout option,
        that one; otherwise each option is timed on windows of block 1 and the fastest kept.
        Every option scans the same positions with the same exact test; only the LDS layout
        (words per row, masks that fit, strip-lane or row-lane lanes, the registers per wave)
        and so the speed differ.  An option the library refuses (a row-lane layout without the
        strip or steps it needs) is skipped.  Each option first runs its window sequence once
        untimed (its first launches load the code and, for an instance that spills, have the
        runtime set up scratch memory for as many waves as the largest window keeps resident,
        which can take longer than the scans); its timed window then grows from
        2^24 positions towards min(R / 2, 2^window_log2), each step to what its latest rate
        scans in max_s (timed_rate), and the options near the best (85 %) are timed again on
        their windows in `rounds` round-robin passes, each keeping its best rate.  Windows of
        2^36 positions keep row-lane grids (one wave per unit of a lane group and a superstrip)
        many waves deep, as in a block scan: smaller ones leave a partial last round of waves
        that penalises the layouts with the largest units (the k=399 confirmation wheel's lead
        block ties strip-lane at 2^33 positions and is 3 % faster at 2^36)."""
        opts = GpuSieve._layout_options(P)
        if len(opts) == 1:
            return opts[0]
        L = nat.lib()

        def scanner(h):
            def scan(lo, hi):
                ms = C.c_double(0)
                la, pos = C.c_uint64(0), C.c_uint64(0)
                nat.check(L.es_engine_set_block(h, 1))
                nat.check(L.es_engine_reset(h))
                nat.check(L.es_engine_scan_window(h, lo, hi))
                nat.check(L.es_engine_sync(h))
                nat.check(L.es_engine_kernel_time(h, C.byref(ms), C.byref(la), C.byref(pos)))
                return pos.value / max(ms.value * 1e-3, 1e-12)
            return scan

        timed = []          # [option, handle, window, best rate]
        try:
            best = 0.0
            for bound, swap, lay in opts:
                try:
                    d = _Desc(P, dict(P.tun, mi_inner_radix=bound, mi_ring_swap=swap,
                                      mi_layout=lay))
                except AssertionError:
                    continue
                h = C.c_void_p()
                if L.es_engine_create(C.byref(d.desc), C.byref(h)) != nat.ES_OK:
                    continue
                timed.append([(bound, swap, lay), h, None, 0.0])
                scan = scanner(h)
                grow = dict(max_s=max_s, strips_log2=window_log2, lo_log2=window_log2,
                            hi_log2=window_log2)
                timed_rate(scan, P.R, 1, **grow)            # warm-up pass, untimed
                rate, win = timed_rate(scan, P.R, 1, floor=0.5 * best, with_window=True, **grow)
                timed[-1][2:] = [win, rate]
                best = max(best, rate)
            if not timed:
                return opts[0]
            close = [t for t in timed if t[3] >= 0.85 * best]
            for _ in range(rounds if len(close) > 1 else 0):
                for t in close:
                    t[3] = max(t[3], scanner(t[1])(*t[2]))
            return max(close, key=lambda t: t[3])[0]
        finally:
            for t in timed:
                L.es_engine_destroy(t[1])

    # -- reference contract ------------------------------------------------------------
    def _launch(self, ostart: int, ocount: int) -> None:
        """ostart/ocount are OUTER indices; each covers r0 positions.  Asynchronous."""
        nat.check(nat.lib().es_engine_scan(self._h, int(ostart) * self._f, int(ocount) * self._f))
        self.launches += 1

    def set_block(self, t: int) -> None:
        """Block constants of block t, computed by the library (which then also knows the
        block: in block 0 the minimum leaves out n <= k+1)."""
        if self._tconst_block == t:            # already uploaded for this block
            return
        for j, q in enumerate(self.rest_q):
            self.tconst_h[j] = ((t % q) * self.rest_Mmodq[j]) % q
        nat.check(nat.lib().es_engine_set_block(self._h, int(t)))
        self._tconst_block = t

    def scan_block(self, t: int) -> list[int]:
        """All positions of block t; one sync per block.  Survivor IDs are block-absolute."""
        self.set_block(t)
        self.out_n.fill(0)
        ostart = 0
        while ostart < self.R_outer:
            ocount = min(self.ochunk, self.R_outer - ostart)
            self._launch(ostart, ocount)
            self.residues_checked += ocount * self.r0
            ostart += ocount
        n = int(self.out_n.get()[0])
        if n > self.out_cap:
            return self._scan_block_chunked(t)
        return [int(x) for x in self.out[:n].get()] if n else []

    def _scan_block_chunked(self, t: int) -> list[int]:
        """Per-launch drain, used only when a block overflows the buffer."""
        survivors = []
        ostart = 0
        while ostart < self.R_outer:
            ocount = min(self.ochunk, self.R_outer - ostart)
            self.out_n.fill(0)
            self._launch(ostart, ocount)
            n = int(self.out_n.get()[0])
            if n > self.out_cap:
                raise RuntimeError(f"survivor buffer overflow ({n} > {self.out_cap}); "
                                   "raise max_survivors")
            if n:
                survivors.extend(int(x) for x in self.out[:n].get())
            ostart += ocount
        return survivors

    def search(self, max_blocks: int | None = None, log=None) -> dict:
        P = self.plan
        t0 = time.time()
        t = 0
        while max_blocks is None or t < max_blocks:
            surv = self.scan_block(t)
            cands = sorted(P.idx_to_n(i, t) for i in surv)
            good = [n for n in cands if is_good(n, P.k)]
            bad = [n for n in cands if n > P.k + 1 and n not in good]
            if bad:
                return {"k": P.k, "status": "DISAGREEMENT", "kernel_survivors": cands[:20],
                        "ref_rejected": bad[:20], "block": t, "wall": time.time() - t0}
            if log and (good or time.time() - getattr(self, "_last_log", 0) > 10):
                self._last_log = time.time()
                log(f"k={P.k} block t={t} survivors={len(surv)} good={len(good)} "
                    f"checked={self.residues_checked:.3e} wall={time.time()-t0:.1f}s")
            if good:
                return {"k": P.k, "status": "FOUND", "g": good[0], "block": t,
                        "survivors_in_block": len(surv), "residues_checked": self.residues_checked,
                        "launches": self.launches, "wall": time.time() - t0}
            t += 1
        return {"k": P.k, "status": "EXHAUSTED_BLOCKS", "blocks": t,
                "residues_checked": self.residues_checked, "wall": time.time() - t0}

    # -- windows (positions) -------------------------------------------------------------
    def scan_window(self, t: int, idx_lo: int, idx_hi: int, collect: bool = True):
        """Survivor IDs with idx in [idx_lo, idx_hi) of block t, and the least n > k+1 among
        them (or None).  Returns (ids, min_n)."""
        self.set_block(t)
        self.out_n.fill(0)
        nat.check(nat.lib().es_engine_scan_window(self._h, int(idx_lo), int(idx_hi)))
        self.residues_checked += int(idx_hi) - int(idx_lo)
        if not collect:
            return None
        return self.collect(t)

    def scan_windows(self, t: int, windows, cap: int | None = None):
        """(ids, min_n) of each window (idx_lo, idx_hi) of block t, in order, as scan_window
        returns them one by one; the library queues each window before reading the previous
        one back, so the device runs them back to back.  cap bounds the IDs of all windows
        together (default: out_cap)."""
        ws = [(int(a), int(b)) for a, b in windows]
        if not ws:
            return []
        self.set_block(t)
        cap = self.out_cap if cap is None else int(cap)
        bounds = np.asarray(ws, dtype=np.uint64).reshape(-1)
        ids = np.zeros(max(cap, 1), dtype=np.uint64)
        counts = np.zeros(len(ws), dtype=np.uint64)
        ms = np.zeros(2 * len(ws), dtype=np.uint64)
        has = np.zeros(len(ws), dtype=np.uint8)
        rc = nat.lib().es_engine_scan_windows(
            self._h, nat.ptr64(bounds), len(ws), nat.ptr64(ids), cap, nat.ptr64(counts),
            nat.ptr64(ms), has.ctypes.data_as(C.POINTER(C.c_uint8)))
        if rc == nat.ES_ERR_OVERFLOW:
            raise RuntimeError(f"survivor buffer overflow (window counts {counts.tolist()}, "
                               f"cap {cap}, out_cap {self.out_cap}); raise max_survivors")
        nat.check(rc)
        self.residues_checked += sum(b - a for a, b in ws)
        out, at = [], 0
        for i, c in enumerate(counts.tolist()):
            mn = t * self.plan.M + (int(ms[2 * i + 1]) << 64 | int(ms[2 * i])) if has[i] else None
            out.append(([int(x) for x in ids[at:at + c]], mn))
            at += c
        return out

    def collect(self, t: int):
        n = C.c_uint64(0)
        nat.check(nat.lib().es_engine_count(self._h, C.byref(n)))
        cnt = n.value
        if cnt > self.out_cap:
            raise RuntimeError(f"survivor buffer overflow ({cnt} > {self.out_cap}); "
                               "raise max_survivors")
        ids = np.zeros(max(cnt, 1), dtype=np.uint64)
        ms = np.zeros(2, dtype=np.uint64)
        has = C.c_int(0)
        nat.check(nat.lib().es_engine_collect(self._h, nat.ptr64(ids), cnt, C.byref(n),
                                              nat.ptr64(ms), C.byref(has)))
        min_n = None
        if has.value:
            min_n = t * self.plan.M + (int(ms[1]) << 64 | int(ms[0]))
        return [int(x) for x in ids[:cnt]], min_n

    def lds_bytes_per_position(self) -> float:
        """Algorithmic LDS bytes the sieve kernel reads per residue position: per outer index
        (r0 positions of the kernel's ring 0) each lane reads one mask row of 8 bytes per
        ring-0 word from each of the masks staged in LDS, and its carry entry -- one byte of the
        j(h) table (every row-lane layout; strip-lane one-word rows with q0 <= 256 and every
        mask in LDS), a row of one 4-byte word per ring-0 word (other rows of one or two
        words), or a u16 j(h) (wider rows).  The candidate queue (a few entries per thousand outer indices) and the
        per-workgroup table staging are left out."""
        info = self.info
        W, F, G = int(info.words), int(info.n_masks_used), int(info.global_masks)
        if info.layout == 2 or (W == 1 and info.q0 <= 256 and G == 0):
            carry = 1
        elif W <= 2:
            carry = 4 * W
        else:
            carry = 2
        return ((F - G) * 8 * W + carry) / int(info.r0)

    def kernel_time(self):
        ms = C.c_double(0)
        la, pos = C.c_uint64(0), C.c_uint64(0)
        nat.check(nat.lib().es_engine_kernel_time(self._h, C.byref(ms), C.byref(la), C.byref(pos)))
        return ms.value, la.value, pos.value

    def sync(self):
        nat.check(nat.lib().es_engine_sync(self._h))


# --------------------------------------------------------------------------- module helpers

def free_vram_mib() -> int | None:
    """Free memory of the current GPU in MiB (sieve/gpu_sieve_v9.py:416-422), None if unknown."""
    try:
        return int(nat.device_query()["free_mib"])
    except Exception:
        return None


def load_known(*paths: str) -> dict[int, tuple[int, str]]:
    """k -> (g, source) from whitespace 'k g' files; later files do not override earlier."""
    known: dict[int, tuple[int, str]] = {}
    for path in paths:
        if not path or not os.path.exists(path):
            continue
        for line in open(path):
            parts = line.split()
            if len(parts) >= 2 and parts[0].isdigit() and parts[1].isdigit():
                known.setdefault(int(parts[0]), (int(parts[1]), os.path.basename(path)))
    return known


def main(argv=None):
    """CLI with the argv and one-JSON-line-per-k output of sieve/gpu_sieve_v9.py:584-618."""
    ap = argparse.ArgumentParser(description="MI355X CRT-wheel search for g(k), partial-power wheel")
    ap.add_argument("k", type=int, nargs="+")
    ap.add_argument("--plan-only", action="store_true")
    ap.add_argument("--tune", type=json.loads, default=None, help="JSON dict overriding TUNABLES")
    ap.add_argument("--wheel", type=json.loads, default=None,
                    help="explicit wheel: JSON list of [p,T] (or bare p = full power)")
    ap.add_argument("--max-blocks", type=int, default=None)
    ap.add_argument("--timeout", type=float, default=None, help="seconds; process exits 124 on expiry")
    ap.add_argument("--min-free-mib", type=int, default=2048)
    ap.add_argument("--bfile", default=es_data.BFILE)
    ap.add_argument("--known", default=None, help="extra 'k g' file of known terms")
    a = ap.parse_args(argv)
    if a.timeout:
        signal.signal(signal.SIGALRM, lambda *_: (print(json.dumps({"status": "TIMEOUT"}), flush=True),
                                                  os._exit(124)))
        signal.setitimer(signal.ITIMER_REAL, a.timeout)
    known = load_known(a.bfile, a.known)
    for k in a.k:
        plan = Plan(k, a.tune, a.wheel)
        if a.plan_only:
            print(json.dumps(plan.summary()))
            continue
        free = free_vram_mib()
        if free is not None and free < a.min_free_mib:
            print(json.dumps({"k": k, "status": "SKIPPED_LOW_VRAM", "free_mib": free}))
            continue
        sieve = GpuSieve(plan)
        res = sieve.search(a.max_blocks, log=lambda s: print(s, file=sys.stderr, flush=True))
        if res.get("status") == "FOUND" and k in known:
            g_known, src = known[k]
            res["matches_bfile"] = (res["g"] == g_known)
            res["bfile"] = g_known
            res["known_source"] = src
        res["plan"] = plan.summary()
        print(json.dumps(res, default=str), flush=True)


if __name__ == "__main__":
    main()
