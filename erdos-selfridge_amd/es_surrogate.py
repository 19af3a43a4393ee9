"""Learned wheel planner: an MLP surrogate of the wheel cost shortlists plans for a GPU probe.

The plain planner (es_plan.plan_wheel, the reference's truecost knapsack of
sieve/gpu_sieve_v9.py:146-227) minimises the expected number of residue positions scanned
before the first nonempty block, E_pos = R / (1 - exp(-M / E)), E = 1 / (density of good n),
and treats every position as equally expensive.  On the GPU the cost of a position depends on
the wheel's shape (ring-0 width, strip rings, the word masks its filters allow, the exact-tail
rate), so the wall time to the first nonempty block is E_pos * ns_per_position(wheel).

Here:
  * PlanSpace(k) enumerates candidate wheels as exponent vectors over the primes <= k (one T
    per prime, M <= 2^126, R < 2^62): the knapsack optimum of every DP cell near the plain
    plan, plus random exponent moves around them, deduplicated;
  * features(E) maps wheels to 64 numbers (ring shapes, small-prime exponents, the leading
    filters' pass rates and moduli, global sizes and the log admissible fractions);
  * the surrogate (es_mlp.Mlp, trained on the GPU by the project's own kernels) predicts
    log2 survivor density per position and log2 ns per position, trained on wheels whose
    ns per position the sieve itself measured (probe());
  * surrogate_plan(k) ranks ~10^5 candidates by log2 E_pos + predicted log2 ns, probes the top
    shortlist and the plain wheel on the GPU, and keeps the least E_pos * measured ns.  The
    plain wheel is the fallback; any wheel scans every position of a block, so the choice
    changes speed, never the completeness of the scan.
"""
from __future__ import annotations

import math
import random
import time

import numpy as np

import es_plan as EP
from gpu_sieve_mi355x import AUTO_MASK_STOP, LEAD_CAP, LEAD_GAIN

N_FEAT = 64
N_RING_FEAT = 8      # rings described by (log2 q, log2 radix)
N_SMALL = 8          # exponent fraction of the 8 smallest primes
SPLIT_BOUNDS = (0, 64, 32, 16)   # digit-split alternatives the engine times (0 = whole ring)
N_LEAD = 8           # leading filters: pass rate and log2 q


class PlanSpace:
    """Per-k tables for vectorised wheel arithmetic over exponent vectors E[n, P]."""

    def __init__(self, k: int, tun: dict | None = None):
        self.k = k
        self.tun = dict(EP.TUNABLES, **(tun or {}))
        self.ps = EP.primes_upto(k)
        P = len(self.ps)
        self.tp = np.array([len(EP.base_digits(k, p)) for p in self.ps], dtype=np.int64)
        tmax = int(self.tp.max())
        self.log2p = np.log2(np.array(self.ps, dtype=np.float64))
        self.cnt = np.ones((P, tmax + 1), dtype=np.float64)       # partial_count(k, p, T)
        self.rate = np.ones((P, tmax + 1), dtype=np.float64)      # filter pass rate given T
        for i, p in enumerate(self.ps):
            full = EP.partial_count(k, p, int(self.tp[i])) / p ** int(self.tp[i])
            for T in range(0, int(self.tp[i]) + 1):
                c = EP.partial_count(k, p, T) if T else 1
                self.cnt[i, T] = c
                self.rate[i, T] = 1.0 if T == self.tp[i] else full / (c / p ** T)
        self.lcnt = np.log2(self.cnt)
        # kernel inner radix of a ring (p, T) placed first under split bound b: its lowest
        # digit group when the radix exceeds b and T >= 2 (the engine's digit-ring split),
        # else the radix (b = 0: never split)
        self.inner = self._inner_table(128)
        self.inner_alt = {b: self._inner_table(b) for b in SPLIT_BOUNDS}
        # word-mask pass rate of p as a filter outside the wheel: the deepest digit sub-filter
        # mod p^D <= 4096 (1 when there is none; a wheel prime divides M and gets no mask)
        self.mrate = np.ones_like(self.cnt)
        self.mmod = np.zeros_like(self.cnt)                       # that sub-filter's modulus
        for i, p in enumerate(self.ps):
            for T in range(0, min(1, int(self.tp[i]))):
                D = T
                while D + 1 <= int(self.tp[i]) and p ** (D + 1) <= 4096:
                    D += 1
                if D > T:
                    self.mrate[i, T] = self.cnt_at(i, D) / p ** D / (self.cnt[i, T] / p ** T)
                    self.mmod[i, T] = p ** D
        dens = float(np.prod(self.rate[:, 0]))
        self.log2E = -math.log2(dens)
        self.small = min(N_SMALL, P)

    def _inner_table(self, bound: int) -> np.ndarray:
        out = self.cnt.copy()
        if bound <= 0:
            return out
        for i, p in enumerate(self.ps):
            digits = EP.base_digits(self.k, p)
            for T in range(2, int(self.tp[i]) + 1):
                if self.cnt[i, T] > bound:
                    g = p - digits[0]
                    for j in range(1, T):
                        if g * (p - digits[j]) > bound:
                            break
                        g *= p - digits[j]
                    out[i, T] = g
        return out

    def cnt_at(self, i: int, D: int) -> float:
        return float(EP.partial_count(self.k, self.ps[i], D))

    # -- wheels <-> exponent vectors ------------------------------------------------------
    def vector(self, wheel) -> np.ndarray:
        e = np.zeros(len(self.ps), dtype=np.int8)
        idx = {p: i for i, p in enumerate(self.ps)}
        for p, T in wheel:
            e[idx[p]] = T
        return e

    def wheel(self, e) -> list[tuple[int, int]]:
        return [(p, int(T)) for p, T in zip(self.ps, e) if T]

    def plain(self) -> np.ndarray:
        return self.vector(EP.plan_wheel(self.k, self.ps, self.log2E, self.tun))

    # -- vectorised wheel quantities --------------------------------------------------------
    def _take(self, table, E):
        """table[p, T_p] for every prime of every wheel row of E."""
        return table[np.arange(len(self.ps))[None, :], E.astype(np.int64)]

    def log2M(self, E):
        return E.astype(np.float64) @ self.log2p

    def log2R(self, E):
        return self._take(self.lcnt, E).sum(1)

    def log2_density(self, E):
        """log2 survivors per position: sum of log2 pass rates of the filters."""
        return np.log2(self._take(self.rate, E)).sum(1)

    def log2_positions(self, E):
        """log2 E_pos = log2 R - log2(1 - exp(-M / E)), the plain planner's objective."""
        lm = self.log2M(E)
        x = np.exp2(np.minimum(lm - self.log2E, 60.0))
        return self.log2R(E) - np.log2(-np.expm1(-x))

    def valid(self, E):
        """M <= 2^126 and R < 2^62 (float test with margin; the engine re-checks exactly), at
        least one ring, and the ring-0 carry rows of the kernel fit in LDS."""
        lm, lr = self.log2M(E), self.log2R(E)
        ok = (lm <= EP.MAX_M_BITS - 1e-6) & (lr < 62 - 1e-6) & ((E > 0).sum(1) > 0)
        rad = np.where(E > 0, self._take(self.cnt, E), 0.0)
        i0 = rad.argmax(1)
        r0 = rad[np.arange(len(E)), i0]
        q0 = np.exp2(E[np.arange(len(E)), i0] * self.log2p[i0])
        W = np.ceil(np.minimum(r0, 128) / 32)
        carry = np.where(W <= 2, 4 * q0 * W, 2 * q0)
        return ok & (carry <= 32768)

    # -- enumeration ----------------------------------------------------------------------
    def enumerate(self, n: int, seed: int = 0, span_bits: float = 16.0) -> np.ndarray:
        """Up to n distinct valid wheels: the DP optimum of each knapsack cell within
        span_bits of log2 M around the plain plan, the plain plan, and random moves (1-3
        exponent changes) around those."""
        scale = int(self.tun.get("cost_scale", 256))
        max_bits = min(int(self.tun.get("m_cap_bits", 120)), EP.MAX_M_BITS)
        best, pick, cells = EP._knapsack_tables(self.k, self.ps, scale, max_bits)
        plain = self.plain()
        c0 = int(round(float(self.log2M(plain[None])[0]) * scale))
        lo = max(1, c0 - int(span_bits * scale))
        hi = min(len(best) - 1, c0 + int(span_bits * scale))
        pool = [plain]
        step = max(1, (hi - lo) // 4096)
        for c in range(lo, hi + 1, step):
            if np.isfinite(best[c]) and best[c] < 62:
                try:
                    pool.append(self.vector(EP._walk_back(self.ps, pick, cells, c)))
                except AssertionError:
                    continue
        pool = np.unique(np.stack(pool), axis=0)
        pool = pool[self.valid(pool)]
        rng = np.random.default_rng(seed)
        out = [plain[None], pool]
        have = len(pool) + 1
        tp = self.tp
        tries = 0
        while have < 2 * n and tries < 64:
            tries += 1
            m = n
            base = pool[rng.integers(len(pool), size=m)].copy()
            moves = rng.integers(1, 4, size=m)
            for j in range(3):
                sel = moves > j
                idx = rng.integers(len(self.ps), size=m)
                newT = (rng.random(m) * (tp[idx] + 1)).astype(np.int8)
                rows = np.nonzero(sel)[0]
                base[rows, idx[rows]] = newT[rows]
            base = base[self.valid(base)]
            out.append(base)
            have += len(base)
        allv = np.unique(np.concatenate(out), axis=0)
        # keep the plain plan, then a seeded sample
        key = np.all(allv == plain[None], axis=1)
        rest = allv[~key]
        if len(rest) > n - 1:
            rest = rest[rng.choice(len(rest), size=n - 1, replace=False)]
        return np.concatenate([plain[None], rest])

    # -- features ---------------------------------------------------------------------------
    def features(self, E) -> np.ndarray:
        """64 features per wheel row of E:
          0-15   rings 0..7 in engine order (radix descending, then p): log2 q, log2 radix;
          16-23  kernel geometry: inner radix (ring 0, or its lowest digit group when wider
                 than 128), words per row, split flag, strip length S (leading radices nearest
                 1024), steps per unit V (row-lane layout) or rings decoded per strip
                 (strip-lane), log2 of the exact-tail fraction left by the 8 best word masks,
                 the row-lane flag, log2 of a per-position cost estimate of the kernel (its
                 LDS row reads per outer index with the mask count the engine's "auto" rule
                 takes within the LDS budget -- fewer bank cycles and padding lanes in the
                 row-lane layout -- plus the queued candidates the masks leave, over the inner
                 radix);
          24-31  exponent fraction T_p / t_p of the 8 smallest primes;
          32-39  the engine's split alternatives for ring 0 (whole, digit groups <= 64, 32,
                 16): log2 inner radix and words per row of each (the engine times them and
                 keeps the fastest);
          40-47  the 8 least filter pass rates (the filter order's head);
          48-55  the 8 least word-mask pass rates (digit sub-filters with modulus <= 4096);
          56-63  k, log2 M, log2 R, log2 density, expected filter evaluations, filter count,
                 log2 E_pos, count of filters with a mask."""
        E = np.asarray(E, dtype=np.int8)
        n, P = E.shape
        F = np.zeros((n, N_FEAT), dtype=np.float32)
        Ei = E.astype(np.int64)
        rows = np.arange(n)
        rad = np.where(E > 0, self._take(self.cnt, E), 0.0)
        lq = Ei * self.log2p[None, :]
        order = np.lexsort((np.broadcast_to(np.arange(P), (n, P)), -rad), axis=1)
        rr_all = np.take_along_axis(rad, order, 1)
        rr = rr_all[:, :N_RING_FEAT]
        qq = np.take_along_axis(lq, order, 1)[:, :N_RING_FEAT]
        has = rr > 0
        F[:, 0:N_RING_FEAT] = np.where(has, qq / 16.0, 0.0)
        F[:, N_RING_FEAT:2 * N_RING_FEAT] = np.where(has, np.log2(np.maximum(rr, 1)) / 8.0, 0.0)
        o = 2 * N_RING_FEAT
        # kernel geometry
        p0 = order[:, 0]
        r0 = rr_all[:, 0]
        r0k = self.inner[p0, Ei[rows, p0]]
        split = r0k < r0
        nr = (rr_all > 0).sum(1)
        # the engine's lead block (gpu_sieve_mi355x.swap_lead): ring 0 of 33..128 residues, not
        # split, ring j the first of <= 32, rings 0..j-1 holding <= LEAD_CAP positions, ring j
        # filling its word LEAD_GAIN times better (or ring 0 > 96) -> the kernel sweeps ring j
        # as its ring 0 and rings 0..j-1 as its first strip rings
        narrow = (rr_all[:, 1:] > 0) & (rr_all[:, 1:] <= 32)
        jl = np.argmax(narrow, axis=1) + 1
        cum = np.cumprod(np.where(rr_all > 0, rr_all, 1.0), axis=1)
        lead = cum[rows, jl - 1]
        rj = rr_all[rows, jl]
        fill0 = r0 / (32.0 * np.ceil(r0 / 32.0))
        swap = (~split & (r0 > 32) & (r0 <= 128) & narrow.any(1) & (lead <= LEAD_CAP)
                & ((r0 > 96) | (rj / 32.0 >= LEAD_GAIN * fill0)))
        r0k = np.where(swap, rj, r0k)
        W = np.ceil(r0k / 32.0)
        # kernel radices: [inner, rest of ring 0 (if split), rings 1..]; 0 = absent
        rest = rr_all[:, 1:]
        kr_split = np.concatenate([r0k[:, None], (r0 / r0k)[:, None], rest], axis=1)
        kr_whole = np.concatenate([r0[:, None], rest, np.zeros((n, 1))], axis=1)
        pos = np.arange(P)[None, :]
        perm = np.where(pos == 0, jl[:, None], np.where(pos <= jl[:, None], pos - 1, pos))
        kr_swap = np.concatenate([np.take_along_axis(rr_all, perm, 1), np.zeros((n, 1))], axis=1)
        kr = np.where(split[:, None], kr_split, np.where(swap[:, None], kr_swap, kr_whole))
        S, m = self._strips(kr, mmin=np.where(swap, jl, 0))
        nk = (kr > 0).sum(1)
        # steps per unit of the row-lane layout: the radices after the strip rings up to 1024
        V = np.ones(n)
        live = np.ones(n, dtype=bool)
        for j in range(1, kr.shape[1]):
            r = kr[:, j]
            sub = live & (j > m) & (r > 0) & (V * np.maximum(r, 1) <= 1024)
            live &= ~((j > m) & ~sub)
            V = np.where(sub, V * r, V)
        rowl = (W == 1) & (S >= 32) & (V >= 16)        # the engine's row-lane layout
        mall = self._take(self.mrate, E)
        mo = np.argsort(mall, axis=1, kind="stable")[:, :N_LEAD]
        mr = np.take_along_axis(mall, mo, 1)
        mq = np.take_along_axis(self._take(self.mmod, E), mo, 1)
        tail = np.log2(mr).sum(1) + np.log2(r0k)
        F[:, o + 0] = np.log2(r0k) / 8.0
        F[:, o + 1] = W / 4.0
        F[:, o + 2] = split
        F[:, o + 3] = np.log2(S) / 16.0
        F[:, o + 4] = np.where(rowl, np.log2(V) / 16.0, np.maximum(nk - m, 0) / 16.0)
        F[:, o + 5] = tail / -32.0
        F[:, o + 6] = rowl
        # masks the "auto" rule stages: least pass rates while r0 x the rates taken >= the stop
        # and their tables fit the LDS budget (row-lane tables: rows [0, 3m + 64) of 8 bytes;
        # strip-lane: [0, 2m) of 8 W bytes)
        pre = r0k[:, None] * np.concatenate([np.ones((n, 1)), np.cumprod(mr, 1)[:, :-1]], 1)
        tbytes = np.where(rowl[:, None], 24.0 * mq + 528.0, 16.0 * W[:, None] * mq + 16.0)
        fits = np.cumsum(np.where(mr < 1.0, tbytes, 0.0), 1) <= 65536.0 - 4.0 * W[:, None] * 512
        take = (pre >= AUTO_MASK_STOP) & (mr < 1.0) & fits
        nm = take.sum(1)
        left = r0k * np.where(take, mr, 1.0).prod(1)   # candidate slots per outer index
        # per-position cost estimate: LDS row reads (row-lane: about two thirds of the
        # strip-lane bank cycles, padding lanes included, plus the unit set-up over V steps)
        # and the queued candidates the masks leave
        pad = np.ceil(S / 64.0) * 64.0 / S
        cost_s = W * (nm + 1.0) + 20.0 * left
        cost_r = (0.67 * (nm + 1.0) + 20.0 * left) * pad + 8.0 / V
        F[:, o + 7] = np.log2(np.where(rowl, cost_r, cost_s) / r0k) / 8.0
        o += 8
        F[:, o:o + self.small] = Ei[:, :self.small] / self.tp[None, :self.small]
        o += N_SMALL
        T0 = Ei[rows, p0]
        for b in SPLIT_BOUNDS:
            rb = self.inner_alt[b][p0, T0]
            F[:, o] = np.log2(rb) / 8.0
            F[:, o + 1] = np.ceil(rb / 32.0) / 4.0
            o += 2
        rate = self._take(self.rate, E)
        isf = Ei < self.tp[None, :]
        rkey = np.where(isf, rate, 2.0)
        forder = np.argsort(rkey, axis=1, kind="stable")
        fr = np.take_along_axis(rkey, forder, 1)
        live = fr <= 1.0
        lead = min(N_LEAD, P)
        F[:, o:o + lead] = np.where(live[:, :lead], fr[:, :lead], 1.0)
        o += N_LEAD
        F[:, o:o + mr.shape[1]] = mr
        o += N_LEAD
        cum = np.cumprod(np.where(live, fr, 1.0), axis=1)
        excl = np.concatenate([np.ones((n, 1)), cum[:, :-1]], axis=1)
        ef = np.where(live, excl, 0.0).sum(1)
        lm, lr = self.log2M(E), self.log2R(E)
        F[:, o + 0] = self.k / 400.0
        F[:, o + 1] = lm / 128.0
        F[:, o + 2] = lr / 64.0
        F[:, o + 3] = self.log2_density(E) / -64.0
        F[:, o + 4] = ef
        F[:, o + 5] = live.sum(1) / 80.0
        F[:, o + 6] = self.log2_positions(E) / 64.0
        F[:, o + 7] = (self._take(self.mrate, E) < 1.0).sum(1) / 80.0
        assert o + 8 == N_FEAT
        return F

    @staticmethod
    def _strips(kr, target: float = 1024.0, cap: float = float(1 << 22), mmin=None):
        """Strip length S and strip ring count m for kernel radices kr[:, 0] (ring 0), kr[:, 1:]
        (0 = absent): the leading product nearest the target in log scale, with at least
        mmin strip rings (a lead block is always whole), as the engine chooses it."""
        n = kr.shape[0]
        mmin = np.zeros(n, dtype=np.int64) if mmin is None else np.asarray(mmin)
        S = np.ones(n)
        m = np.zeros(n, dtype=np.int64)
        done = np.zeros(n, dtype=bool)
        for j in range(1, kr.shape[1]):
            r = kr[:, j]
            live = ~done & (r > 0)
            nxt = S * np.where(live, r, 1.0)
            over = live & (nxt > cap)
            done |= over | ~live
            live &= ~over
            cross = live & (m > 0) & (m >= mmin) & (nxt > target)
            take_cross = cross & (nxt * S < target * target)
            step = live & (~cross | take_cross)
            S = np.where(step, nxt, S)
            m = m + step
            done |= cross
        return S, m

# ----------------------------------------------------------------------------------------- GPU

def probe(k: int, wheel, log2_positions: int = 33, t: int = 1, tun: dict | None = None,
          repeats: int = 2, max_s: float = 0.05) -> float | None:
    """Measured ns per residue position of a wheel on windows of block t (device time, best of
    `repeats`); None if the engine rejects the wheel.  Windows as gpu_sieve_mi355x.timed_rate
    chooses them (GPU-filling, at most 2^log2_positions positions, shrunk for very slow
    wheels)."""
    import gpu_sieve_mi355x as G
    try:
        plan = EP.Plan(k, tun, wheel)
        eng = G.GpuSieve(plan)
    except AssertionError:
        return None

    def scan(lo, hi):
        eng.scan_window(t, lo, hi)
        ms, _, pos = eng.kernel_time()
        return pos / max(ms * 1e-3, 1e-12)
    sp = eng.info.strip_len * eng.info.r0
    best = max(G.timed_rate(scan, plan.R, sp, max_s=max_s, hi_log2=log2_positions)
               for _ in range(repeats))
    return 1e9 / best


class Surrogate:
    """Feature/label scalers plus an ensemble of GPU networks (predictions averaged): the
    networks differ in initialisation and minibatch order, and averaging them steadies the
    ranking the planner takes from them."""

    def __init__(self, seed: int = 0, ensemble: int = 3):
        import es_mlp
        self.nets = [es_mlp.Mlp(batch_cap=8192, seed=seed * 97 + i) for i in range(ensemble)]
        self.seed = seed
        self.fmu = np.zeros(N_FEAT, dtype=np.float32)
        self.fsd = np.ones(N_FEAT, dtype=np.float32)
        self.ymu = np.zeros(2, dtype=np.float32)
        self.ysd = np.ones(2, dtype=np.float32)

    def fit(self, X, Y, epochs: int = 300, batch: int = 128, lr: float = 2e-3, seed: int = 1):
        """Train every network; returns (mean final minibatch loss, total device ms)."""
        X = np.asarray(X, dtype=np.float32)
        Y = np.asarray(Y, dtype=np.float32)
        self.fmu, self.fsd = X.mean(0), X.std(0) + 1e-6
        self.ymu, self.ysd = Y.mean(0), Y.std(0) + 1e-6
        Xn = (X - self.fmu) / self.fsd
        Yn = (Y - self.ymu) / self.ysd
        out = [net.fit(Xn, Yn, epochs=epochs, batch=batch, lr=lr, seed=seed + 7919 * i)
               for i, net in enumerate(self.nets)]
        return float(np.mean([o[0] for o in out])), float(sum(o[1] for o in out))

    def predict(self, X) -> np.ndarray:
        """[n, 2]: log2 survivor density per position, log2 ns per position."""
        Xn = (np.asarray(X, dtype=np.float32) - self.fmu) / self.fsd
        y = np.mean([net.forward(Xn) for net in self.nets], axis=0)
        return y * self.ysd + self.ymu


def training_set(ks, per_k: int = 24, n_candidates: int = 20000, seed: int = 0,
                 log=None, probe_log2: int = 35):
    """Wheels the sieve measures itself: per k, a quarter from the 64 lowest-E_pos candidates,
    half from the 2000 lowest (the band a shortlist is drawn from) and a quarter from the 20000
    lowest; labels (log2 density, log2 measured ns per position on windows of up to
    2^probe_log2 positions -- wheels with long strips fill the GPU only on large windows)."""
    X, Y, meta = [], [], []
    rng = random.Random(seed)
    for k in ks:
        sp = PlanSpace(k)
        E = sp.enumerate(n_candidates, seed=seed + k)
        lp = sp.log2_positions(E)
        order = list(np.argsort(lp, kind="stable"))
        picks = set(rng.sample(order[:64], min(per_k // 4, len(order[:64]))))
        band = order[:2000]
        while len(picks) < min(3 * per_k // 4, len(band)):
            picks.add(rng.choice(band))
        wide = order[:20000]
        while len(picks) < min(per_k, len(E)):
            picks.add(rng.choice(wide))
        picks = sorted(picks)
        feats = sp.features(E[picks])
        dens = sp.log2_density(E[picks])
        for j, i in enumerate(picks):
            ns = probe(k, sp.wheel(E[i]), probe_log2)
            if ns is None:
                continue
            X.append(feats[j])
            Y.append([dens[j], math.log2(ns)])
            meta.append((k, sp.wheel(E[i]), ns))
        if log:
            log(f"training set k={k}: {len(X)} probed wheels so far")
    return np.asarray(X, dtype=np.float32), np.asarray(Y, dtype=np.float32), meta


def surrogate_plan(k: int, model: Surrogate | None, n_candidates: int = 100_000,
                   shortlist: int = 32, probe_log2: int = 36, seed: int = 0, log=None,
                   compare: tuple = ()) -> dict:
    """Rank n_candidates wheels by log2 E_pos + predicted log2 ns, probe the first `shortlist`
    and the plain wheel, and return the least E_pos * measured ns (the plain wheel when no
    shortlisted wheel probes).  model=None ranks by log2 E_pos alone (the analytic-only
    shortlist).  `compare` lists probe budgets s at which both rankings are evaluated from
    one shared pool of probes ("sizes" in the result: the least cost among the plain wheel and
    each ranking's first s wheels other than the plain wheel -- s probes besides the plain
    wheel for either ranking, whose analytic first wheel is the plain one -- and the first
    wheel of each ranking with no fallback)."""
    t0 = time.perf_counter()
    sp = PlanSpace(k)
    E = sp.enumerate(n_candidates, seed=seed)
    lp = sp.log2_positions(E)
    pred = model.predict(sp.features(E)) if model is not None else np.full((len(E), 2), np.nan)
    rank_m = np.argsort(lp + pred[:, 1], kind="stable") if model is not None else None
    rank_a = np.argsort(lp, kind="stable")
    rank = rank_m if model is not None else rank_a
    t_rank = time.perf_counter() - t0
    plain_i = 0                                     # enumerate() puts the plain plan first
    cache = {}

    def probed(i):
        i = int(i)
        if i not in cache:
            ns = probe(k, sp.wheel(E[i]), probe_log2)
            cache[i] = None if ns is None else {
                "wheel": sp.wheel(E[i]), "log2_positions": float(lp[i]), "ns": ns,
                "cost_s": 2.0 ** float(lp[i]) * ns * 1e-9, "plain": i == plain_i,
                "pred_log2_ns": float(pred[i, 1]) if model is not None else None}
        return cache[i]

    def best_of(idx):
        got = [r for r in (probed(i) for i in list(idx) + [plain_i]) if r is not None]
        return min(got, key=lambda r: r["cost_s"])

    def others(r, s_):
        return [i for i in r[:s_ + 1] if i != plain_i][:s_]

    best = best_of(rank[:shortlist])
    plain_r = probed(plain_i)
    sizes = {}
    for s_ in compare:
        row = {"analytic": best_of(others(rank_a, s_))["cost_s"] / plain_r["cost_s"]}
        if rank_m is not None:
            row["surrogate"] = best_of(others(rank_m, s_))["cost_s"] / plain_r["cost_s"]
        sizes[int(s_)] = row
    if compare and rank_m is not None:
        top = probed(rank_m[0])
        sizes["top1_no_probe"] = {"surrogate": (top["cost_s"] if top else math.inf) /
                                  plain_r["cost_s"], "analytic": 1.0}
    if log:
        log(f"k={k}: {len(E)} candidates ranked in {t_rank:.1f} s; best {best['cost_s']:.1f} s "
            f"(plain {plain_r['cost_s']:.1f} s)")
    return {"k": k, "best": best, "plain": plain_r, "first_pick": probed(rank[0]),
            "probed": [v for v in cache.values() if v],
            "n_candidates": len(E), "rank_seconds": t_rank, "sizes": sizes}
