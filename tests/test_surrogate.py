"""Learned planner: vectorised wheel arithmetic against Plan and the engine's kernel geometry
(CPU); on the GPU, what a surrogate trained on wheels the sieve probes itself gains over the
analytic-only ranking at a fixed probe budget, with identical exact minima."""
from __future__ import annotations

import math

import numpy as np
import pytest

import es_plan as EP
import es_surrogate as S
import gpu_sieve_mi355x as G
from conftest import bfile


@pytest.mark.parametrize("k", [200, 360, 399, 400])
def test_plan_space_matches_plan(k):
    sp = S.PlanSpace(k)
    E = sp.enumerate(3000, seed=k)
    assert np.array_equal(E[0], sp.plain())
    assert len(np.unique(E, axis=0)) == len(E) and sp.valid(E).all()
    lp = sp.log2_positions(E)
    assert lp[0] <= lp.min() + 1e-9                 # the plain plan minimises E_pos
    F = sp.features(E)
    assert F.shape == (len(E), S.N_FEAT) and np.isfinite(F).all()
    rng = np.random.default_rng(k)
    for i in [0] + list(rng.choice(len(E), 6, replace=False)):
        P = EP.Plan(k, None, sp.wheel(E[i]))
        assert abs(sp.log2M(E[i:i + 1])[0] - math.log2(P.M)) < 1e-9
        assert abs(sp.log2R(E[i:i + 1])[0] - math.log2(P.R)) < 1e-9
        assert abs(2 ** sp.log2_density(E[i:i + 1])[0] / P.density_rest - 1) < 1e-9
        assert abs(F[i, 60] - P.expected_filter_evals()) < 1e-5
        assert abs(F[i, S.N_RING_FEAT] * 8 - math.log2(P.radix[0])) < 1e-5
        assert F[i, 0] * 16 == pytest.approx(math.log2(P.rings[0][0]), abs=1e-5)
        # kernel geometry: inner radix and strip length as the engine derives them
        _, kr, _, inner, _ = G.kernel_rings(P, 128)
        j = G.swap_lead(kr, inner)
        kr = [kr[i] for i in G.lead_order(len(kr), j)]
        assert F[i, 16] * 8 == pytest.approx(math.log2(kr[0]), abs=1e-5)
        assert F[i, 19] * 16 == pytest.approx(math.log2(_engine_strip(kr, mmin=j)), abs=1e-5)


def _engine_strip(radix, target=1024, mmin=0):
    """The engine's strip rule (es_engine_create): leading radices after ring 0, nearest the
    target in log scale, with at least mmin rings (a lead block is always whole)."""
    S, m = 1, 0
    while 1 + m < len(radix):
        nxt = S * radix[1 + m]
        if nxt > (1 << 22):
            break
        if m > 0 and m >= mmin and nxt > target:
            if nxt * S < target * target:
                S = nxt
            break
        S, m = nxt, m + 1
    return S


def test_plain_vector_roundtrip():
    sp = S.PlanSpace(380)
    w = EP.Plan(380).wheel
    assert sorted(sp.wheel(sp.vector(w))) == sorted(w)


@pytest.mark.gpu
def test_network_gain_over_analytic_shortlist_at_fixed_probe_budget():
    """What the trained network gains over the method without it.  Trained on 352 wheels the
    sieve probes at even k = 300..320, the planner ranks 10^4 candidate wheels of the ten unseen
    odd k = 301..319 by analytic E_pos + predicted ns; the analytic-only ranking (E_pos alone,
    whose first wheel is the plain one) is the method without the network.  From one shared
    pool of probes (measured ns per position), per k: the network's first wheel with no probe,
    and at s = 1, 2, 4 probes besides the plain wheel the least cost E_pos x ns of each
    ranking, over the plain wheel's cost.

    Where the network gains: the analytic ranking's wheels run at the same few rates on this
    kernel except where ring 0 is narrow -- k=317's plain wheel (ring 0 of 20 residues fills
    under two thirds of its mask word).  There the network's four-wheel shortlist holds a wheel
    costing 0.66 of the plain wheel where the analytic one's best costs 0.80, in every run
    measured (three training seeds, six runs: geomean over the ten k at s = 4, 0.954-0.956
    against 0.973-0.975).  Its unprobed first pick finds that gain in some runs only (k=317:
    0.77-0.91, or the plain wheel; geomean 0.975-1.006).  Elsewhere it keeps the plain wheel
    or one within noise of it: the engine times its own layouts, which removed most of the speed
    differences between near-optimal wheels that the network used to exploit.  Asserted with
    margin: at s = 4 the network's k=317 cost is at most 0.9 x the analytic one and its geomean
    at least 0.005 below; at every budget it is no more than 0.02 behind, and its first pick
    no more than 3 % above the plain wheel.  The chosen and plain wheels return the same exact
    minimum g(k)."""
    X, Y, meta = S.training_set(list(range(300, 321, 2)), per_k=32, n_candidates=10000, seed=2,
                                probe_log2=33)
    assert len(X) >= 300
    model = S.Surrogate(seed=2)
    loss, _ = model.fit(X, Y, epochs=300, batch=32)
    assert np.isfinite(loss)
    rows = []
    for k in range(301, 320, 2):
        res = S.surrogate_plan(k, model, n_candidates=10000, shortlist=1, probe_log2=33, seed=k,
                               compare=(1, 2, 4))
        print(f"k={k} sizes {res['sizes']}", flush=True)
        rows.append(res)

    def gm(key, m):
        return float(np.exp(np.mean([math.log(r["sizes"][key][m]) for r in rows])))
    top1 = gm("top1_no_probe", "surrogate")
    print(f"geomean cost / plain: first pick {top1:.3f}; " + "; ".join(
        f"s={s_} {gm(s_, 'surrogate'):.3f} vs {gm(s_, 'analytic'):.3f}" for s_ in (1, 2, 4)),
        flush=True)
    k317 = next(r for r in rows if r["k"] == 317)["sizes"][4]
    assert k317["surrogate"] <= 0.9 * k317["analytic"]
    assert gm(4, "surrogate") <= gm(4, "analytic") - 0.005
    for s_ in (1, 2, 4):
        assert gm(s_, "surrogate") <= gm(s_, "analytic") + 0.02
    assert top1 <= 1.03
    gs = bfile()
    for k in (317, 311):
        res = next(r for r in rows if r["k"] == k)
        first = res["first_pick"] or res["plain"]
        for wheel in (first["wheel"], res["plain"]["wheel"]):
            plan = EP.Plan(k, None, wheel)
            eng = G.GpuSieve(plan)
            gi, gt = plan.n_to_idx(gs[k])
            lo = max(0, gi - (1 << 30))
            _, mn = eng.scan_window(gt, lo, min(plan.R, lo + (1 << 31)))
            assert mn == gs[k], (k, wheel)


def test_features_follow_digit_ring_split():
    """k = 200 full-power wheel: ring 0 = 13^3 (radix 1056) is swept as 88 x 12, strip 4320."""
    sp = S.PlanSpace(200)
    e = sp.vector(EP.fullpower_wheel(200))
    f = sp.features(e[None])[0]
    assert f[16] * 8 == pytest.approx(math.log2(88), abs=1e-5)
    assert f[17] * 4 == 3 and f[18] == 1
    P = EP.Plan(200, None, EP.fullpower_wheel(200))
    _, kr, _, _, _ = G.kernel_rings(P, 128)
    assert f[19] * 16 == pytest.approx(math.log2(_engine_strip(kr)), abs=1e-5)
    assert _engine_strip(kr) == 4320
