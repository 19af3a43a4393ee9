"""Host-side plan logic of the engine against the oracle and the reference's recorded plans."""
from __future__ import annotations

import json
import math
import os
import random

import pytest

import es_oracle as O
import es_plan as EP
from conftest import GOLDEN


def test_number_theory_matches_oracle():
    for k in (1, 2, 30, 97, 200, 399):
        assert EP.primes_upto(k) == O.primes_upto(k)
        for p in EP.primes_upto(k):
            assert EP.allowed_residues(k, p) == O.allowed_residues(k, p)
            for T in range(1, len(O.digits(k, p)) + 1):
                assert EP.partial_residues(k, p, T) == O.low_residues(k, p, T)
    rng = random.Random(3)
    for _ in range(300):
        k = rng.randrange(2, 401)
        n = rng.randrange(1 << 110)
        p = rng.choice(O.primes_upto(k))
        assert EP.kummer_ok(n, k, p) == O.dominates(n, k, p)


def test_default_plans_k270_275_280():
    """Default planner = the wheels of evidence/logs/2026-08-22/logs/gate_v9_270_280.log."""
    rec = json.load(open(os.path.join(GOLDEN, "gate_v9_270_280.json")))
    for k in (270, 275, 280):
        plan = EP.Plan(k)
        s = plan.summary()
        want = rec[str(k)]["plan"]
        for key in ("wheel", "M", "residues_per_block", "n_filters", "split_primes",
                    "density_wheel", "density_rest", "expected_g", "expected_checks_true"):
            assert s[key] == want[key], (k, key)


def test_default_plans_every_k():
    """The default planner for every k = 2..400 equals the reference planner's: wheel (ring
    order), M, R, radices, filter order and the fp64 summary fields bit for bit.  The vectors
    were written by tests/golden/make_plan_vectors.py, which imports the reference's planner
    module in the build container (recorded plans pin only 39 of these k)."""
    rec = json.load(open(os.path.join(GOLDEN, "plan_vectors.json")))["plans"]
    assert [r["k"] for r in rec] == list(range(2, 401))
    for r in rec:
        plan = EP.Plan(r["k"])
        s = json.loads(json.dumps(plan.summary()))
        for key in ("wheel", "M", "n_filters", "residues_per_block", "density_wheel",
                    "density_rest", "E_g", "expected_checks_true"):
            assert s[key] == r[key], (r["k"], key)
        assert list(plan.radix) == r["radix"] and list(plan.rest) == r["rest"], r["k"]


def test_recorded_frontier_plans():
    """Every recorded primary plan (default planner) and every banned-prime confirmation plan
    (harness/run_k_masks.py:250-258) is reproduced exactly, float summary fields included."""
    plans = json.load(open(os.path.join(GOLDEN, "frontier_plans.json")))
    checked = set()
    for x in plans:
        if x["tag"] not in ("primary", "confirm") or x["log"].endswith("k392_confirm.log"):
            continue                      # target_g plans (next test) or an earlier planner
        k, rec = x["k"], x["plan"]
        key = (k, json.dumps(rec["wheel"]))
        if key in checked:
            continue
        checked.add(key)
        if x["banned"]:
            ban = [int(b) for b in x["banned"].split(",")]
            full = EP.Plan(k)
            ps = [p for p in EP.primes_upto(k) if p not in ban]
            wh = EP.plan_wheel(k, ps, -math.log2(full.density_all), dict(EP.TUNABLES))
            plan = EP.Plan(k, wheel=[[p, T] for p, T in wh])
        else:
            plan = EP.Plan(k)
        s = plan.summary()
        for f in ("wheel", "M", "residues_per_block", "n_filters", "density_wheel",
                  "density_rest", "expected_g"):
            assert s[f] == rec[f], (k, x["tag"], f)
    assert len(checked) >= 35


def test_runs_json_wheels_geometry():
    terms = json.load(open(os.path.join(GOLDEN, "runs_frontier.json")))["terms"]
    for term in terms:
        for role in ("primary", "confirmation"):
            r = term[role]
            plan = EP.Plan(term["k"], None, r["wheel"])
            assert str(plan.M) == r["N"] and plan.R == r["residues_per_block"]
            assert [[p, T] for p, T in plan.wheel] == r["wheel"]   # recorded in ring order
            n = int(term["n"])
            assert n // plan.M == r["block"]


def test_idx_maps_against_oracle():
    for k, wheel in ((30, [[2, 5], [3, 4], [5, 3], [7, 2], [11, 2], [13, 2]]), (200, None), (399, None)):
        plan = EP.Plan(k, None, wheel)
        wm = O.WheelMap(k, plan.wheel)
        assert wm.M == plan.M and wm.R == plan.R
        rng = random.Random(k)
        for _ in range(300):
            idx, t = rng.randrange(plan.R), rng.randrange(1 << 20)
            n = plan.idx_to_n(idx, t)
            assert n == wm.idx_to_n(idx, t)
            assert plan.n_to_idx(n) == (idx, t)


def test_expected_filter_evals_k399():
    """E_f at the k=399 primary wheel: 1.2176 (reference's measured ~1.22)."""
    plan = EP.Plan(399, None, [[107, 1], [3, 6], [211, 1], [139, 1], [83, 1], [59, 1], [103, 1],
                               [137, 1], [41, 1], [2, 9], [37, 1], [29, 1], [101, 1], [13, 1],
                               [31, 1], [67, 1], [5, 2]])
    assert abs(plan.expected_filter_evals() - 1.2176) < 1e-4
    assert plan.radix[0] == 29 and plan.R == 11908694063185920


@pytest.mark.parametrize("k", [30, 100, 200, 281, 399])
def test_masks_are_necessary_conditions(k):
    """Every word-mask prefilter accepts every residue the full filter accepts, and the slow
    list plus the masks cover every filter condition."""
    import gpu_sieve_mi355x as G
    plan = EP.Plan(k)
    masks = G.choose_masks(plan, 12)
    for fi, D, m, rate in masks:
        p = plan.rest[fi]
        q, full = plan.info[p]
        _, low = EP.partial_residues(k, p, D)
        assert q % m == 0 and set(r % m for r in full) <= set(low)
        assert 0 < rate <= 1
    slow = G.slow_order(plan, masks)
    covered = set(slow) | {fi for fi, D, _, _ in masks if D == plan.tp[plan.rest[fi]]}
    assert covered == set(range(len(plan.rest)))


def test_mask_knapsack_is_optimal_on_a_small_instance():
    """The budgeted mask choice equals exhaustive search: 7 primes with up to 3 depths each,
    at most 4 masks, a budget that binds; cost = masks + 1 + 160 x r0 x prod(rates)."""
    import itertools
    import gpu_sieve_mi355x as G
    rng = random.Random(3)
    cands = []
    for fi in range(7):
        rate = 1.0
        for d in range(1, 4):
            rate *= rng.uniform(0.3, 0.9)
            cands.append((rate, fi, d, rng.choice([101, 257, 600, 1100, 2300])))
    cost_of = lambda m: 2 * (3 * m + 64) + 4
    room, r0, n = 14000, 29, 4
    got = G._masks_within_budget(cands, n, room, cost_of, r0)
    unit = 64
    best = None
    by_fi = {}
    for c in cands:
        by_fi.setdefault(c[1], []).append(c)
    for pick in itertools.product(*[[None] + v for v in by_fi.values()]):
        chosen = [c for c in pick if c is not None]
        units = sum(-(-cost_of(c[3]) // unit) for c in chosen)
        if len(chosen) > n or units > room // unit:
            continue
        cost = G._mask_cost([(c[1], c[2], c[3], c[0]) for c in chosen], r0)
        if best is None or cost < best - 1e-12:
            best = cost
    assert abs(G._mask_cost(got, r0) - best) < 1e-9
    assert len({fi for fi, _, _, _ in got}) == len(got) <= n
    assert sum(-(-cost_of(m) // unit) for _, _, m, _ in got) <= room // unit


def test_budgeted_masks_only_where_the_budget_binds(monkeypatch):
    """k=397's default wheel (row-lane tables of 3m + 64 rows) hits the 64 KiB budget with the
    greedy rule, which takes 41^2 = 1681 and is left with 0.02 candidate slots per outer index;
    the knapsack fits six masks within the budget, cheaper by the cost model.  The k=399
    primary wheel does not hit the budget and keeps the greedy's six masks."""
    import gpu_sieve_mi355x as G
    plan = EP.Plan(397)
    d = G._Desc(plan, dict(plan.tun, mi_inner_radix=128))
    assert d.rows
    with monkeypatch.context() as mp:
        mp.setattr(G, "_masks_within_budget", lambda *a, **kw: [(-1, 0, 0, 1.0)])  # greedy only
        g = G._Desc(plan, dict(plan.tun, mi_inner_radix=128))
    assert 1681 in [m for _, _, m, _ in g.masks]
    assert sum(2 * (3 * m + 64) + 4 for _, _, m, _ in d.masks) <= 65536 // 4
    assert len(d.masks) == 6 and G._mask_cost(d.masks, d.r0) < 9.0
    assert G._mask_cost(d.masks, d.r0) < G._mask_cost(g.masks, g.r0) - 3.0
    p399 = EP.Plan(399, None, [[107, 1], [3, 6], [211, 1], [139, 1], [83, 1], [59, 1],
                               [103, 1], [137, 1], [41, 1], [2, 9], [37, 1], [29, 1],
                               [101, 1], [13, 1], [31, 1], [67, 1], [5, 2]])
    d399 = G._Desc(p399, dict(p399.tun, mi_inner_radix=128))
    assert [m for _, _, m, _ in d399.masks] == [529, 223, 227, 229, 233, 149]


def test_fullpower_wheel_k200():
    """Config-2 wheel: full powers for p <= 23 plus the best single-digit extension; no item
    left out fits both caps, and no other subset of the fitting items does better (brute force
    over every subset of the items whose radix fits the R room)."""
    import itertools
    import math
    w = EP.fullpower_wheel(200)
    full = [(p, len(EP.base_digits(200, p))) for p in EP.primes_upto(23)]
    assert all(it in w for it in full)
    P = EP.Plan(200, None, w)
    assert math.log2(P.M) <= 100 and P.R < (1 << 62)
    assert len(P.rest) == 37 and sorted(P.rest)[0] == 29
    R0 = math.prod(EP.partial_count(200, p, T) for p, T in full)
    M0 = math.prod(p ** T for p, T in full)
    items = [(p, T) for p in EP.primes_upto(200) if p > 23
             for T in range(1, len(EP.base_digits(200, p)) + 1)
             if EP.partial_count(200, p, T) * R0 < (1 << 62)]
    best = None
    for r in range(len(items) + 1):
        for sub in itertools.combinations(items, r):
            if len({p for p, _ in sub}) < r:
                continue
            M = M0 * math.prod(p ** T for p, T in sub)
            R = R0 * math.prod(EP.partial_count(200, p, T) for p, T in sub)
            if M <= (1 << 100) and R < (1 << 62):
                key = (round(math.log2(M / R), 9), -R)
                if best is None or key > best[0]:
                    best = (key, sorted(full + list(sub)))
    assert best[1] == w
    assert [it for it in w if it[0] > 23] == [(29, 1), (67, 1), (101, 1)]


def test_fullpower_wheel_reproduces_config1():
    """k = 30 with full powers for p <= 13 and M capped at 2^39 is the config-1 wheel."""
    assert EP.fullpower_wheel(30, p_full=13, m_cap_bits=39) == \
        [(2, 5), (3, 4), (5, 3), (7, 2), (11, 2), (13, 2)]


def test_kernel_digit_rings_keep_idx_order():
    """Ring 0 of the k=200 full-power wheel (13^3, radix 1056) swept as digit rings: for every
    slot, the mixed-radix combination of the digit-ring terms (lowest group fastest) is the
    plan's ring-0 term, and the moduli still multiply to M."""
    import math
    import gpu_sieve_mi355x as G
    P = EP.Plan(200, None, EP.fullpower_wheel(200))
    mods, radix, terms, inner, f = G.kernel_rings(P, 64)
    assert P.wheel[0] == (13, 3) and inner == 13 ** 3
    assert radix[:3] == [8, 11, 12] and f == 132 and math.prod(mods) == P.M
    assert radix[3:] == P.radix[1:] and terms[3:] == P.ring_terms()[1:]
    want = P.ring_terms()[0]
    for slot in range(P.radix[0]):
        a, b, c = slot % 8, (slot // 8) % 11, slot // 88
        assert (terms[0][a] + terms[1][b] + terms[2][c]) % P.M == want[slot]
    M0 = P.M // inner
    assert all(t % M0 == 0 for t in terms[0])
    # a wider inner bound groups digits 0 and 1; no split when ring 0 already fits
    assert G.kernel_rings(P, 128)[1][:2] == [88, 12]
    P399 = EP.Plan(30)
    assert G.kernel_rings(P399, 64)[3:] == (0, 1)


@pytest.mark.parametrize("k,j", [(301, 1), (309, 1), (312, 1), (378, 1), (363, 2), (386, 2)])
def test_swapped_lead_rings_keep_idx(k, j):
    """Wheels whose ring 0 has 33..128 residues are passed with their first ring j of at most
    32 residues in front and rings 0..j-1 (the lead block, `lead` positions) behind it: for
    kernel outer index o = (strip base) + sg and kernel ring-0 digit s, the engine's idx
    (r0 * o + lead * s + X_sg, X_sg = sg % lead + lead * r0 * (sg // lead) - r0 * sg, with a
    strip of whole lead blocks) names the position whose value is the kernel-order CRT sum
    (the plan's idx_to_n of block 0)."""
    import gpu_sieve_mi355x as G
    P = EP.Plan(k)
    d = G._Desc(P, dict(P.tun, mi_inner_radix=128))
    rad = [int(x) for x in d.ring_radix]
    lead, r0 = d.lead, rad[0]
    assert G.swap_lead(P.radix) == j and lead == math.prod(P.radix[:j])
    assert rad == [P.radix[i] for i in G.lead_order(len(rad), j)] and r0 <= 32 < P.radix[0]
    assert not G.swap_lead(rad)
    offs = [0]
    for r in rad:
        offs.append(offs[-1] + r)
    terms = [int(lo) | (int(hi) << 64) for lo, hi in zip(d.term_lo, d.term_hi)]
    S = lead * rad[j + 1]                              # a strip of whole lead blocks
    rng = random.Random(k)
    for _ in range(500):
        o, s = rng.randrange(P.R // r0), rng.randrange(r0)
        digs, x = [s], o
        for r in rad[1:]:
            digs.append(x % r)
            x //= r
        n = sum(terms[offs[j] + digs[j]] for j in range(len(rad))) % P.M
        sg = o % S
        X = sg % lead + lead * r0 * (sg // lead) - r0 * sg
        idx = r0 * o + lead * s + X
        assert idx == o % lead + lead * (s + r0 * (o // lead))
        assert 0 <= idx < P.R and P.idx_to_n(idx, 0) == n


def test_lead_block_rule():
    """The lead block is taken only where ring j fills its mask word clearly better than ring 0
    fills its words (or ring 0 needs four words), within the size cap."""
    import gpu_sieve_mi355x as G
    assert G.swap_lead([42, 30, 25]) == 1            # 42/64 of two words -> 30/32 of one
    assert G.swap_lead([36, 35, 26, 25]) == 2        # lead block 36 x 35
    assert G.swap_lead([64, 45, 32, 26]) == 0        # 64/64 already full
    assert G.swap_lead([77, 36, 24, 18]) == 0        # 77/96 vs 24/32
    assert G.swap_lead([108, 26, 20]) == 1           # four words
    assert G.swap_lead([40, 30], inner_mod=3 ** 6) == 0 and G.swap_lead([30, 20]) == 0
    assert G.swap_lead([100, 90, 90, 20]) == 0       # lead block above the cap
    # forced: wherever a ring of <= 32 can lead within the cap, whatever the fill
    assert G.swap_lead([64, 45, 32, 26], force=True) == 2
    assert G.swap_lead([77, 36, 24, 18], force=True) == 2
    assert G.swap_lead([47, 24, 23], force=True) == 1 and G.swap_lead([47, 24, 23]) == 0
    assert G.swap_lead([100, 90, 90, 20], force=True) == 0
    assert G.swap_lead([40, 30], inner_mod=3 ** 6, force=True) == 0
    assert G.swap_lead([30, 20], force=True) == 0


K399_CONFIRM = [[223, 1], [3, 6], [211, 1], [139, 1], [83, 1], [59, 1], [103, 1], [137, 1],
                [5, 4], [2, 9], [37, 1], [29, 1], [101, 1], [13, 1], [31, 1], [67, 1]]
K399_PRIMARY = [[107, 1], [3, 6], [211, 1], [139, 1], [83, 1], [59, 1], [103, 1], [137, 1],
                [41, 1], [2, 9], [37, 1], [29, 1], [101, 1], [13, 1], [31, 1], [67, 1], [5, 2]]


def test_layout_options_the_engine_times():
    """GpuSieve times the distinct kernel layouts of a plan: both ring orders where ring 0
    (33..128 residues, whole) could trade places with a later ring of <= 32 (the k=399
    confirmation wheel: 47 then 24), and both lane layouts where the kernel's ring 0 has two
    words (strip-lane "auto" and row-lane "rows"); one layout where neither applies (the primary
    wheel, r0 = 29); digit splits of a wide ring 0 mod p^T only unswapped (k=378: 3^6, 108
    residues); a fixed tunable fixes its dimension.  The descriptor follows the choice."""
    import gpu_sieve_mi355x as G
    opts = lambda k, w=None, **t: G.GpuSieve._layout_options(EP.Plan(k, t, w))
    assert opts(399, K399_PRIMARY) == [(128, 0, "auto")]
    assert opts(399, K399_CONFIRM) == [(128, 0, "auto"), (128, 0, "rows"), (128, 1, "auto")]
    assert opts(399, K399_CONFIRM, mi_ring_swap=0) == [(128, 0, "auto"), (128, 0, "rows")]
    assert opts(399, K399_CONFIRM, mi_ring_swap="1") == [(128, 1, "auto")]
    assert opts(399, K399_CONFIRM, mi_ring_swap="rule") == [(128, "rule", "auto"),
                                                             (128, "rule", "rows")]
    assert opts(399, K399_CONFIRM, mi_layout="strips") == [(128, 0, "strips"),
                                                           (128, 1, "strips")]
    assert opts(399, K399_CONFIRM, mi_global_masks=2) == [(128, 0, "auto")]
    o378 = opts(378)
    assert o378[:2] == [(128, 0, "auto"), (128, 1, "auto")] and len(o378) > 2
    assert all(s == 0 for b, s, _ in o378[2:])
    assert all(b == 64 and s == 0 for b, s, _ in opts(378, mi_inner_radix=64))
    with pytest.raises(ValueError):
        opts(399, K399_CONFIRM, mi_ring_swap=2)
    P = EP.Plan(399, None, K399_CONFIRM)
    d1 = G._Desc(P, dict(P.tun, mi_inner_radix=128, mi_ring_swap=1))
    d0 = G._Desc(P, dict(P.tun, mi_inner_radix=128, mi_ring_swap=0))
    dr = G._Desc(P, dict(P.tun, mi_inner_radix=128, mi_ring_swap="rule"))
    assert (d1.lead, d1.r0) == (47, 24) and (d0.lead, d0.r0) == (0, 47) and dr.lead == 0
    assert d1.rows and not d0.rows
    assert G._Desc(P, dict(P.tun, mi_inner_radix=128, mi_ring_swap=0, mi_layout="rows")).rows
    # row-lane for two-word rows only on request, never for three words
    assert not G.rows_layout([84, 24, 23, 18, 16], 0, {"mi_layout": "rows"})
    assert G.rows_layout([64, 24, 23, 18, 16], 0, {"mi_layout": "rows"})
    assert not G.rows_layout([64, 24, 23, 18, 16], 0, {})


def test_target_g_plans_reproduced():
    """The known-g planner (sieve/gpu_sieve_v11.py:199-213): the recorded k=384 v11 gate,
    k=388 v10 check and k=392 confirmation plans bit for bit (float summary fields included),
    and the runs.json confirmation wheels it produced."""
    import es_data
    b = es_data.known_terms()
    plans = json.load(open(os.path.join(GOLDEN, "frontier_plans.json")))
    seen = 0
    for x in plans:
        if x["tag"] in ("v10check", "v11gate") or x["log"].endswith("k392_confirm.log"):
            k, rec = x["k"], x["plan"]
            s = EP.Plan(k, {"target_g": b[k]}).summary()
            for f in ("wheel", "M", "residues_per_block", "n_filters", "density_wheel",
                      "density_rest", "expected_g"):
                assert s[f] == rec[f], (k, x["tag"], f)
            seen += 1
    assert seen == 5
    terms = {t["k"]: t for t in json.load(open(os.path.join(GOLDEN, "runs_frontier.json")))["terms"]}
    for k in (385, 387, 389, 390, 392, 395, 396, 400):
        plan = EP.Plan(k, {"target_g": b[k]})
        assert [[p, T] for p, T in plan.wheel] == terms[k]["confirmation"]["wheel"], k
        assert b[k] // plan.M == terms[k]["confirmation"]["block"]
    # the branch is live: a known g changes the default plan
    assert EP.Plan(385).wheel != EP.Plan(385, {"target_g": b[385]}).wheel


def test_target_g_cost_is_blocks_times_R():
    """Among the planner's candidate cells the chosen wheel covers [0, g) with the fewest
    positions: no other wheel of the same DP (same cells) with a smaller ceil(g/M) * R."""
    import es_data
    k = 300
    g = es_data.known_terms()[k]
    plan = EP.Plan(k, {"target_g": g})
    cost = -(-g // plan.M) * plan.R
    default = EP.Plan(k)
    assert cost <= -(-g // default.M) * default.R
