"""Throughput of the CRT-wheel sieve engine on one MI355X.

Workload: the k=399 primary wheel of the published record (17 rings, r0 = 29, M ~ 2^107.6,
R = 1.19e16 positions per block), block t = 6 -- the block that holds g(399).  One step sieves
one window of 2^36 consecutive residue positions of that block with the full filter set (75
filter primes); consecutive steps take consecutive windows, and the last timed window holds
g(399), which the step must return as its minimum.

Prints one JSON line: residue positions sieved per second (whole job), the roofline of the
sieve kernel (device time from HIP events on the engine's stream) and the CPU oracle timed on
the host cores on a bounded sub-window of the same block.
"""
from __future__ import annotations

import argparse
import math
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "erdos-selfridge_amd"))

K = 399
WHEEL = [[107, 1], [3, 6], [211, 1], [139, 1], [83, 1], [59, 1], [103, 1], [137, 1], [41, 1],
         [2, 9], [37, 1], [29, 1], [101, 1], [13, 1], [31, 1], [67, 1], [5, 2]]
BLOCK = 6
# the recorded word-mask confirmation wheel of k=399 (runs.json): r0 = 47, R = 1.755e16, g(399)
# in block 4
WHEEL_CONF = [[223, 1], [3, 6], [211, 1], [139, 1], [83, 1], [59, 1], [103, 1], [137, 1],
              [5, 4], [2, 9], [37, 1], [29, 1], [101, 1], [13, 1], [31, 1], [67, 1]]
BLOCK_CONF = 4
G399 = 1463367633345557449307696345739199
# INT32 VALU peak: 256 CUs x 128 lane-ops per CU-clock (two wave64 VALU instructions per clock,
# the vector rate behind the 157.3 TFLOP/s FP32 figure) x 2.4 GHz.  `--workload valu` measures
# 7.36e13 v_add_u32/s on this device (1.88 wave-instructions per CU-cycle at ~2.35 GHz).
INT32_PEAK = 256 * 128 * 2.4e9


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    # defaults: 50 timed steps after 20 warm-up steps for the 2^36-position windows of the
    # headline and k200 (the device's clocks ramp over its first ~10 such launches, 2.33 -> 1.93
    # ms), 20 after 2 elsewhere
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--window-log2", type=int, default=36)
    ap.add_argument("--masks", type=int, default=None)
    ap.add_argument("--strip", type=int, default=None)
    ap.add_argument("--threads", type=int, default=None)
    ap.add_argument("--layout", choices=("auto", "strips", "rows"), default=None,
                    help="sieve: lane layout of the kernel (speed only)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="target CPU time of the oracle's sample (its positions near g(399) "
                         "run faster than the probe's, so the sample takes ~0.7 of this)")
    ap.add_argument("--cpu-threads", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--part2-candidates", type=int, default=10000,
                    help="sieve: candidate wheels ranked per k for the time to first block")
    ap.add_argument("--part2-shortlist", type=int, default=8,
                    help="sieve: wheels probed per k (besides the plain one) for it")
    ap.add_argument("--no-configs", action="store_true",
                    help="sieve: leave out the short batch (config 3) and filter-pass (config 4) lines")
    ap.add_argument("--no-confirmation", action="store_true",
                    help="sieve: leave out the k=399 confirmation wheel's rate")
    ap.add_argument("--no-part2", action="store_true",
                    help="sieve: leave out the surrogate-vs-plain time to first nonempty block")
    ap.add_argument("--autotune", action="store_true",
                    help="batch, k200: autotune the launch shape of each wheel first")
    ap.add_argument("--workload", choices=("sieve", "fpass", "copy", "stream", "valu", "lds",
                                           "batch", "k200", "surrogate", "mitm"),
                    default="sieve",
                    help="sieve: k=399 wheel windows (headline); fpass: k=397 candidate filter "
                         "pass; copy: 3 device copies of 1 GiB (calibrates HBM byte counters); "
                         "valu: INT32 VALU peak microbenchmark; lds: LDS ds_read_b64 "
                         "peak (conflict-free and random rows); stream: what library "
                         "streaming kernels reach on this device's HBM; "
                         "batch: k=300..320 batched launches; k200: full-power wheel window; "
                         "surrogate: learned planner vs plain planner, k=350..400; "
                         "mitm: bounded n-window search for g(k)")
    ap.add_argument("--fpass-log2", type=int, default=30)
    ap.add_argument("--k-step", type=int, default=1, help="surrogate: every k-step-th k")
    ap.add_argument("--candidates", type=int, default=100_000, help="surrogate: plans per k")
    ap.add_argument("--mitm-k", type=int, default=200)
    ap.add_argument("--mitm-mem", type=float, default=64.0, help="mitm: table budget, GB")
    a = ap.parse_args()
    headline = a.workload in ("sieve", "k200")   # windows of 2^36 positions per step
    if a.steps is None:
        a.steps = 50 if headline else 20
    if a.warmup is None:
        a.warmup = 20 if headline else 2
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the engine has no multi-GPU path "
              "(blocks and outer-index ranges shard by construction; see the sharding notes)",
              file=sys.stderr)
        return 2

    if a.workload == "fpass":
        return bench_fpass(a)
    if a.workload == "batch":
        return bench_batch(a)
    if a.workload == "k200":
        return bench_k200(a)
    if a.workload == "surrogate":
        return bench_surrogate(a)
    if a.workload == "mitm":
        return bench_mitm(a)
    if a.workload == "valu":
        return bench_valu(a)
    if a.workload == "lds":
        lds = lds_read_peak()
        print(json.dumps({
            "metric": "LDS ds_read_b64 bytes/s (conflict-free rows)",
            "value": lds["conflict_free_bytes_per_s"], "unit": "bytes/s", "n_gpus": 1,
            "steps": 3, "warmup": 1, "ms_per_step": None, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
            "config": {"workload": f"{lds['cus']} CUs x 2 workgroups x 1024 threads, 8192 "
                                   "rounds of 8 ds_read_b64 per lane, 32 KiB table"},
            "nominal_peak": LDS_PEAK_NOMINAL, "frac_of_nominal":
                lds["conflict_free_bytes_per_s"] / LDS_PEAK_NOMINAL, **lds}), flush=True)
        return 0
    if a.workload == "stream":
        return bench_stream(a)
    if a.workload == "copy":
        import torch
        src = torch.ones(1 << 28, dtype=torch.int32, device="cuda")       # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"workload": "copy", "bytes_read_per_copy": 1 << 30,
                          "bytes_written_per_copy": 1 << 30, "copies": 3}), flush=True)
        return 0

    import es_plan
    import gpu_sieve_mi355x as G

    tune = {}
    if a.masks is not None:
        tune["mi_masks"] = a.masks
    if a.strip is not None:
        tune["mi_strip_target"] = a.strip
    if a.threads is not None:
        tune["threads"] = a.threads
    if a.layout is not None:
        tune["mi_layout"] = a.layout
    plan = es_plan.Plan(K, tune, WHEEL)
    eng = G.GpuSieve(plan)
    N = 1 << a.window_log2
    g_idx, g_t = plan.n_to_idx(G399)
    assert g_t == BLOCK
    nsteps = a.warmup + a.steps
    base = g_idx - (nsteps - 1) * N - N // 2           # g(399) sits in the last window
    assert base >= 0
    windows = [(base + i * N, base + (i + 1) * N) for i in range(nsteps)]

    # each step: one window, its survivors and minimum read back; the engine queues window
    # i + 1 before reading window i (scan_windows), so the device runs the steps back to back
    eng.scan_windows(BLOCK, windows[:max(a.warmup, 1)])   # also sets up the result slots
    eng.sync()
    eng.kernel_time()                                   # reset the device timers
    t0 = time.perf_counter()
    res = eng.scan_windows(BLOCK, windows[a.warmup:])
    eng.sync()
    t1 = time.perf_counter()
    found = [(len(ids), mn) for ids, mn in res]
    dev_ms, launches, dev_pos = eng.kernel_time()
    wall = t1 - t0
    positions = a.steps * N
    value = positions / wall
    assert found[-1][1] == G399, f"last window must return g(399), got {found[-1]}"
    assert all(c == 0 for c, _ in found[:-1]) and found[-1][0] == 1

    kern_rate = dev_pos / (dev_ms / 1e3)
    F = eng.info.n_masks_used
    lds_peak = lds_read_peak()
    roofline = sieve_roofline(eng, plan, kern_rate, dev_ms, launches, lds_peak)
    cpu = None if a.no_cpu else cpu_baseline(plan, eng, g_idx, a)
    conf = None if a.no_confirmation else confirmation_wheel(a, lds_peak)
    configs = None if a.no_configs else other_configs(a)
    part2 = None if a.no_part2 else first_block_ab(candidates=a.part2_candidates,
                                                   shortlist=a.part2_shortlist)

    print(json.dumps({
        "metric": "residue positions sieved/s (k=399 wheel); time to first nonempty block vs plain",
        "value": value, "unit": "positions/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u128/u32 integer", "data": "synthetic windows of the recorded k=399 wheel",
        "config": {"workload": f"k=399 primary wheel (runs.json), block {BLOCK}, "
                               f"2^{a.window_log2} positions per step, all 75 filters",
                   "k": K, "r0": eng.r0, "R": plan.R, "window": N, "masks": F,
                   "strip": eng.strip_len, "threads": eng.threads, "layout": eng.layout,
                   "steps_per_unit": eng.info.steps if eng.layout == "rows" else None},
        "roofline": roofline, "cpu_baseline": cpu, "confirmation_wheel": conf,
        "time_to_first_block": part2, "other_configs": configs,
    }), flush=True)
    return 0


def other_configs(a) -> dict:
    """Configs 3 and 4 at their full sizes, in short form, so one default run records them:
    the k=300..320 batch (84 jobs of 2^34 positions, every job's survivors checked against its
    engine run alone) and the k=397 filter pass over 2^30 candidates (survivors re-checked by
    the oracle's Kummer test).  `--workload batch` and `--workload fpass` print the full lines."""
    sub = argparse.Namespace(**vars(a))
    sub.steps, sub.warmup, sub.no_cpu, sub.autotune, sub.layout = 3, 1, True, False, None
    b = batch_line(sub)
    sub.steps = 5
    f = fpass_line(sub)
    return {
        "batch_k300_320": {"value": b["value"], "unit": b["unit"], "ms_per_step": b["ms_per_step"],
                           "workload": b["config"]["workload"],
                           "row_lane_plans": b["config"]["row_lane_plans"],
                           "unbatched_positions_per_s": b["unbatched"]["rate"],
                           "roofline": {k: b["roofline"][k] for k in
                                        ("bound", "achieved", "peak", "unit", "frac")}},
        "filter_pass_k397": {"value": f["value"], "unit": f["unit"],
                             "workload": f["config"]["workload"],
                             "survivors": f["config"]["survivors"],
                             "roofline": {k: f["roofline"][k] for k in
                                          ("bound", "achieved", "peak", "unit", "frac",
                                           "kernel_avg_launch_ms")}},
    }


def confirmation_wheel(a, lds_peak: dict, steps: int = 10) -> dict:
    """The second k=399 wheel of the record (the word-mask confirmation run: ring 0 of 47
    residues, which the engine may sweep behind its 24-residue ring as a lead block, whichever
    order it times faster) on consecutive windows of its block 4, the last holding g(399):
    positions per second of wall time and of kernel time, with its LDS roofline."""
    import es_plan
    import gpu_sieve_mi355x as G
    plan = es_plan.Plan(K, None, WHEEL_CONF)
    eng = G.GpuSieve(plan)
    N = 1 << a.window_log2
    g_idx, g_t = plan.n_to_idx(G399)
    assert g_t == BLOCK_CONF
    nsteps = 2 + steps
    base = g_idx - (nsteps - 1) * N - N // 2
    windows = [(base + i * N, base + (i + 1) * N) for i in range(nsteps)]
    eng.scan_windows(BLOCK_CONF, windows[:2])
    eng.sync()
    eng.kernel_time()
    t0 = time.perf_counter()
    found = eng.scan_windows(BLOCK_CONF, windows[2:])
    eng.sync()
    wall = time.perf_counter() - t0
    dev_ms, launches, dev_pos = eng.kernel_time()
    assert found[-1][1] == G399, f"last window must return g(399), got {found[-1][1]}"
    kern_rate = dev_pos / (dev_ms / 1e3)
    roof = sieve_roofline(eng, plan, kern_rate, dev_ms, launches, lds_peak)
    return {"value": steps * N / wall, "unit": "positions/s",
            "config": {"workload": f"k=399 confirmation wheel (runs.json), block {BLOCK_CONF}, "
                                   f"{steps} windows of 2^{a.window_log2} positions",
                       "R": plan.R, "r0_plan": plan.radix[0], "r0_kernel": eng.info.r0,
                       "lead_block": eng._d.lead, "ring_swap": eng.ring_swap,
                       "layout": eng.layout, "masks": eng.info.n_masks_used},
            "kernel_positions_per_s": kern_rate, "kernel_avg_launch_ms": dev_ms / max(1, launches),
            "lds_bytes_per_position": roof["lds_bytes_per_position"],
            "lds_frac": roof["frac"]}


def first_block_ab(train_ks=tuple(range(300, 321, 2)), eval_ks=(301, 305, 309, 313, 317),
                   pairs: int = 2, candidates: int = 10000, shortlist: int = 8) -> dict:
    """Metric part 2 measured end to end: time to exhaust the first nonempty block with the
    surrogate-planned wheel against the plain planner's wheel, on k whose first nonempty block
    scans in about a second (the frontier k = 350..400 take hours per block: bench.py
    --workload surrogate estimates those from measured rates).  The network is trained on
    wheels the sieve probes at even k = 300..320; per odd k (not trained on) it ranks
    `candidates` wheels, the first `shortlist` and the plain wheel are probed, and the least
    E_pos x ns wins.  Both wheels then run the full search (GpuSieve.search: blocks t = 0, 1, ... until
    the referee accepts a survivor) in interleaved pairs (plain, surrogate, surrogate, plain,
    ...; harness/paired_bench.py:22-68), each must return the b-file g(k), and each wheel's
    time is its median over the pairs.  Engine construction and one full-size warm-up window
    per engine are outside the timed searches."""
    import es_data
    import es_plan
    import es_surrogate as S
    import gpu_sieve_mi355x as G
    import numpy as np
    bf = es_data.known_terms()
    t0 = time.perf_counter()
    X, Y, _ = S.training_set(list(train_ks), per_k=24, n_candidates=10000, seed=5, probe_log2=33)
    model = S.Surrogate(seed=3)
    model.fit(X, Y, epochs=300, batch=32)
    t_train = time.perf_counter() - t0
    rows = []
    for k in eval_ks:
        res = S.surrogate_plan(k, model, n_candidates=candidates, shortlist=shortlist,
                               probe_log2=33, seed=k)
        wheels = {"plain": res["plain"]["wheel"], "surrogate": res["best"]["wheel"]}
        engines = {tag: G.GpuSieve(es_plan.Plan(k, None, w)) for tag, w in wheels.items()}
        for eng in engines.values():     # one full-size window each, outside the timed searches
            eng.scan_window(0, 0, min(1 << 34, eng.plan.R), collect=False)
            eng.sync()
        times = {"plain": [], "surrogate": []}
        found = {}
        for i in range(pairs):
            order = ("plain", "surrogate") if i % 2 == 0 else ("surrogate", "plain")
            for tag in order:
                w0 = time.perf_counter()
                r = engines[tag].search(max_blocks=64)
                times[tag].append(time.perf_counter() - w0)
                assert r["status"] == "FOUND" and r["g"] == bf[k], (k, tag, r.get("status"))
                found[tag] = (r["g"], r["block"])
        med = {tag: float(np.median(v)) for tag, v in times.items()}
        rows.append({"k": k, "plain_s": med["plain"], "surrogate_s": med["surrogate"],
                     "ratio": med["surrogate"] / med["plain"],
                     "same_wheel": wheels["plain"] == wheels["surrogate"],
                     "blocks": {tag: found[tag][1] + 1 for tag in found},
                     "R": {tag: engines[tag].plan.R for tag in engines},
                     "expected_ratio": res["best"]["cost_s"] / res["plain"]["cost_s"]})
    ps = sum(r["plain_s"] for r in rows)
    ss = sum(r["surrogate_s"] for r in rows)
    return {"metric": "time to first nonempty block, surrogate-planned / plain-planned wheel",
            "value": ss / ps, "unit": "ratio (sum of median search times over k)",
            "higher_is_better": False,
            "geomean_ratio": float(np.exp(np.mean([math.log(r["ratio"]) for r in rows]))),
            "config": {"workload": f"searches of k={','.join(map(str, eval_ks))} from block 0 "
                                   f"to g(k); network trained on k={train_ks[0]}..{train_ks[-1]}"
                                   f" even, {len(X)} probed wheels", "pairs": pairs,
                       "candidates": candidates, "shortlist": shortlist},
            "train_and_plan_s": t_train, "per_k": rows}


# LDS read peak of gfx950 at the nominal clock: 256 CUs x 256 bytes per CU-clock (64 banks of
# 4 bytes) x 2.4 GHz.  `--workload lds` measures the conflict-free ds_read_b64 rate live.
LDS_PEAK_NOMINAL = 256 * 256 * 2.4e9
# Profiled figures on the k=399 primary wheel (rocprofv3, 2^36-position launches; the counters
# are per wave-step = 64 lanes x one ring-0 word = 1856 positions; HBM bytes = FETCH_SIZE x 2.0
# + WRITE_SIZE, the copy calibration): es_rows_kernel<6, 1, 1> (row-lane layout, the default) and
# es_sieve_kernel<6, 0, 1, 1> (strip-lane).  Other instances are not profiled.
K399_PROFILE = {
    "rows": {"sq_lds_idx_active_per_wave_step": 28.38, "sq_lds_bank_conflict_per_wave_step": 12.60,
             "sq_insts_valu_per_wave_step": 32.78, "hbm_bytes_per_launch": 335e6},
    "strips": {"sq_lds_idx_active_per_wave_step": 39.7, "sq_lds_bank_conflict_per_wave_step": 24.0,
               "sq_insts_valu_per_wave_step": 35.5, "hbm_bytes_per_launch": 1.47e6}}


def lds_read_peak(blocks_per_cu: int = 2, iters: int = 8192) -> dict:
    """Conflict-free and random-row ds_read_b64 rates of this device (es_calib_lds), best of 3."""
    import ctypes as C

    import es_native as nat
    cus = nat.device_query()["cus"] or 256     # no torch: the engine's HIP runtime is loaded
    out = {"cus": cus}
    for mode, name in ((0, "conflict_free"), (1, "random_rows")):
        best = 0.0
        for _ in range(3):
            ms, nb = C.c_double(0), C.c_uint64(0)
            nat.check(nat.lib().es_calib_lds(cus * blocks_per_cu, iters, mode, C.byref(ms),
                                             C.byref(nb)))
            best = max(best, nb.value / (ms.value * 1e-3))
        out[name + "_bytes_per_s"] = best
    out["bytes_per_cu_clock_at_2p4GHz"] = out["conflict_free_bytes_per_s"] / (cus * 2.4e9)
    return out


def sieve_roofline(eng, plan, kern_rate: float, dev_ms: float, launches: int, lds: dict) -> dict:
    """Roofline of the sieve kernel against the LDS read peak: the word-mask kernel reads no HBM
    per position; its per-position work is LDS mask-row reads (lds_bytes_per_position) plus a few
    32-bit VALU ops.  `achieved` = algorithmic LDS bytes per position x positions per second of
    kernel time; `peak` = the conflict-free ds_read_b64 rate measured in this run.  The
    reference algorithm's INT32 op count (12 + 10 E_f per position) is reported apart: this
    kernel does not execute that algorithm, so its rate is no fraction of a hardware peak."""
    bpp = eng.lds_bytes_per_position()
    peak = lds["conflict_free_bytes_per_s"]
    achieved = bpp * kern_rate
    prof = K399_PROFILE.get(eng.layout) \
        if eng.info.n_masks_used == 6 and eng.info.words == 1 and eng.r0 == 29 else None
    profiled = prof is not None
    ef = plan.expected_filter_evals()
    ops_ref = 12 + 10 * ef
    return {
        "bound": "lds", "achieved": achieved / 1e9, "peak": peak / 1e9, "unit": "GB/s",
        "frac": achieved / peak,
        "traffic": prof["hbm_bytes_per_launch"] if profiled else None,
        "lds_bytes_per_position": bpp,
        "lds_peak_nominal_GBps": LDS_PEAK_NOMINAL / 1e9,
        "lds_random_rows_GBps": lds["random_rows_bytes_per_s"] / 1e9,
        "lds_bytes_per_cu_clock": lds["bytes_per_cu_clock_at_2p4GHz"],
        "lds_cycles_per_wave_step_profiled":
            prof["sq_lds_idx_active_per_wave_step"] if profiled else None,
        "lds_conflict_cycles_per_wave_step_profiled":
            prof["sq_lds_bank_conflict_per_wave_step"] if profiled else None,
        "kernel_avg_launch_ms": dev_ms / max(1, launches), "launches": launches,
        "kernel_positions_per_s": kern_rate,
        "reference_algorithm": {
            "int32_ops_per_position": round(ops_ref, 3), "E_f": round(ef, 4),
            "equivalent_int32_ops_per_s": ops_ref * kern_rate,
            "int32_valu_peak": INT32_PEAK,
            "note": "ops the reference's per-position loop would execute; this kernel runs "
                    "the word-mask algorithm instead, so this is work avoided, not a roofline"},
    }


def host_cpu() -> dict:
    """Host cores this process may use, nproc, and the CPU model."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    use = min(aff, cap) if cap > 0 else aff
    return {"threads": max(1, use), "nproc": os.cpu_count(), "affinity": aff,
            "omp_num_threads": cap or None, "cpu_model": model}


def cpu_baseline(plan, eng, g_idx: int, a) -> dict:
    """The oracle's C restatement of the reference CPU path (idx_to_n + Kummer digit test per
    filter prime, tests/test_enumeration.py:37-43) on every host core this process may use (the
    GPU's CPU share where OMP_NUM_THREADS caps it), on a bounded sub-window of block 6 holding
    g(399); the GPU must return the same survivors there."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import es_oracle as O
    import numpy as np
    hc = host_cpu()
    threads = a.cpu_threads or hc["threads"]
    lo = g_idx - 50_000
    # seconds per position from two probes, so that the scan's fixed set-up (filter tables,
    # threads) does not count: the sample then takes about cpu_seconds
    probe = 1 << 23
    dt = []
    for n in (probe, 2 * probe):
        c0 = time.perf_counter()
        O.c_scan(K, plan.wheel, plan.rest, BLOCK, lo, lo + n, threads=threads)
        dt.append(time.perf_counter() - c0)
    per = max(dt[1] - dt[0], dt[1] / 4) / probe
    span = max(probe, min(int(a.cpu_seconds / per), 1 << 36))
    c0 = time.perf_counter()
    n_cpu, ids_cpu = O.c_scan(K, plan.wheel, plan.rest, BLOCK, lo, lo + span, threads=threads)
    cdt = time.perf_counter() - c0
    gpu_ids, _ = eng.scan_window(BLOCK, lo, lo + span)
    assert np.array_equal(np.sort(np.asarray(gpu_ids, dtype=np.uint64)), ids_cpu)
    return {"value": span / cdt, "unit": "positions/s", "cores": threads, "kind": "port",
            "nproc": hc["nproc"], "affinity": hc["affinity"], "cpu_model": hc["cpu_model"],
            "sample": f"positions [{lo}, {lo + span}) of block {BLOCK} (holds g(399)), "
                      f"{cdt:.1f} s on {threads} threads; C restatement of idx_to_n + Kummer "
                      f"digit test per filter prime; survivors identical to the GPU's ({n_cpu})"}


HBM_PEAK = 8.0e12      # bytes/s


def bench_fpass(a) -> int:
    print(json.dumps(fpass_line(a)), flush=True)
    return 0


def fpass_line(a) -> dict:
    """Candidate filter pass: 2^30 u32 IDs into a 2^30-entry u128 stream (k = 397, filter primes
    101..397).  Algorithmic bytes per candidate: 4 (ID) + 16 (value) + 4 x survival (output)."""
    import numpy as np

    import es_filter_pass as FP
    n = 1 << a.fpass_log2
    primes = [p for p in range(101, 398) if all(p % d for d in range(2, int(p ** 0.5) + 1))]
    spec = FP.generator_spec(397, FP.filter_order(397, primes))
    fp = FP.FilterPass(397, FP.conditional_order(397, primes, spec), order="given")
    fp.generate(n, seed=20260901, spec=spec)
    for _ in range(a.warmup):
        fp.run()
    times, surv = [], None
    t0 = time.perf_counter()
    for _ in range(a.steps):
        surv = fp.run()
        times.append(fp.last_ms())
    wall = time.perf_counter() - t0
    dev_ms = sum(times) / len(times)
    s = len(surv) / n
    bytes_per = 20 + 4 * s
    achieved = bytes_per * n / (dev_ms / 1e3)
    # every survivor re-checked by the oracle Kummer test on its value
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import es_oracle as O
    vals = [fp.read_stream(int(i), 1)[0][0] for i in surv[:2048]]
    assert len(O.c_filter_values(397, fp.primes, vals, np.arange(len(vals)))) == len(vals)
    cpu = None
    if not a.no_cpu:
        m = 1 << 22
        vals, ids = fp.read_stream(0, m)
        c0 = time.perf_counter()
        want = O.c_filter_values(397, fp.primes, vals, ids)
        cdt = time.perf_counter() - c0
        sub = surv[surv < m]
        assert np.array_equal(np.sort(want), sub)
        cpu = {"value": m / cdt, "unit": "candidates/s", "cores": 1, "kind": "port",
               "sample": f"first 2^22 stream entries; C Kummer digit test per filter prime; "
                         f"survivors identical to the GPU's ({len(want)})"}
    return {
        "metric": "candidate filter pass throughput (k=397, filter primes 101..397)",
        "value": n / (dev_ms / 1e3), "unit": "candidates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": wall / a.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u128 values, fp64 residues",
        "data": "synthetic: seeded conditioned u128 stream",
        "config": {"workload": f"k=397 filter pass, 2^{a.fpass_log2} candidates, 53 filters",
                   "survivors": int(len(surv)), "survival": s, "spec_survival": spec[3]},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None,
                     "bytes_per_candidate": bytes_per, "kernel_avg_launch_ms": dev_ms},
        "cpu_baseline": cpu}


def bench_valu(a) -> int:
    """INT32 VALU peak of this device: 8 independent v_add_u32 chains per thread, 16 workgroups of
    256 threads per CU (es_calib_valu).  Confirms the peak the sieve roofline is priced at."""
    import ctypes as C

    import es_native as nat
    import torch
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    ms, ops = C.c_double(0), C.c_uint64(0)
    best = 0.0
    for _ in range(max(1, a.steps)):
        nat.check(nat.lib().es_calib_valu(cus * 16, 4096, C.byref(ms), C.byref(ops)))
        best = max(best, ops.value / (ms.value * 1e-3))
    print(json.dumps({
        "metric": "int32 VALU adds/s (v_add_u32 chains)", "value": best, "unit": "ops/s",
        "n_gpus": 1, "steps": a.steps, "warmup": 1, "ms_per_step": ms.value,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic", "config": {"workload": f"{cus} CUs x 16 workgroups x 256 threads, "
                                                    "4096 rounds of 256 adds"},
        "peak": INT32_PEAK, "frac": best / INT32_PEAK,
        "peak_one_instruction_per_cu_clock": INT32_PEAK / 2,
        "cus": cus}), flush=True)
    return 0


def bench_stream(a) -> int:
    """Streaming rates on 16 GiB: library kernels (torch's sum: read only; device copy: read +
    write) and the hand-written dwordx4 read kernel over launch shapes (blocks, threads, loads
    in flight, nontemporal or not); HIP events, best of `steps` -- the HBM read ceiling the
    filter pass is compared with."""
    import torch
    n = 16 * (1 << 30) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    out = {}
    for name, fn, nbytes in (("sum_read", lambda: torch.sum(x), 4 * n),
                             ("copy_read_write", lambda: y.copy_(x), 8 * n)):
        fn()
        best = float("inf")
        for _ in range(max(1, a.steps)):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1))
        out[name] = {"ms": best, "GB_per_s": nbytes / (best * 1e-3) / 1e9}
    # the hand-written dwordx4 streaming read (es_calib_hbm_read), over launch shapes
    import ctypes as C
    import es_native as nat
    sweep = []
    for nt in (0, 1):
        for unroll in (1, 4, 8):
            for threads in (256, 512):
                for blocks in (1024, 2048, 4096, 8192, 16384):
                    ms = C.c_double()
                    nat.check(nat.lib().es_calib_hbm_read(
                        C.c_void_p(x.data_ptr()), 4 * n, blocks, threads, unroll, nt,
                        max(1, a.steps), C.byref(ms)))
                    sweep.append({"nontemporal": nt, "unroll": unroll, "threads": threads,
                                  "blocks": blocks, "ms": ms.value,
                                  "GB_per_s": 4 * n / (ms.value * 1e-3) / 1e9})
    best = max(sweep, key=lambda r: r["GB_per_s"])
    out["dwordx4_read"] = {"best": best, "frac_of_peak": best["GB_per_s"] * 1e9 / HBM_PEAK,
                           "sweep": sweep}
    print(json.dumps({"workload": "stream", "bytes": 4 * n, "peak_GB_per_s": HBM_PEAK / 1e9,
                      **out}), flush=True)
    return 0


def bench_batch(a) -> int:
    print(json.dumps(batch_line(a)), flush=True)
    return 0


def batch_line(a) -> dict:
    """k = 300..320 (default planner wheels), 4 seeded blocks per k, windows of 2^34 positions
    at seeded offsets: one batched run (one grid-mapped launch per kernel shape) against the same
    jobs run one engine launch sequence at a time.  Survivors must agree job by job."""
    import random

    import es_batch
    import es_plan
    import gpu_sieve_mi355x as G
    rng = random.Random(300320)
    N = 1 << 34
    tuned = {}
    if a.autotune:
        import es_autotune
        # batched launches read every mask from LDS: the L1-mask layout is not searched
        space = dict(es_autotune.SPACE, mi_global_masks=[0])
        for k in range(300, 321):
            tuned[k], _ = es_autotune.autotune(k, space=space, probe_log2=34, method="coord")
    lay = {"mi_layout": a.layout} if a.layout is not None else {}
    engines = {k: G.GpuSieve(es_plan.Plan(k, {**tuned.get(k, {}), **lay}))
               for k in range(300, 321)}
    jobs = []
    for k, eng in engines.items():
        for _ in range(4):
            t = rng.randrange(1 << 20)
            lo = rng.randrange(eng.plan.R - N)
            jobs.append((k, t, lo, lo + N))
    b = es_batch.Batch()
    ids = [b.add(engines[k], t, lo, hi) for k, t, lo, hi in jobs]
    b.run()                                              # warm-up
    times = []
    for _ in range(max(1, a.steps)):
        ms, launches = b.run()
        times.append(ms)
    bms = min(times)
    # unbatched: the same jobs, one engine scan (launch sequence) each
    for k, t, lo, hi in jobs[:4]:
        engines[k].scan_window(t, lo, hi)
    for eng in engines.values():
        eng.sync()
        eng.kernel_time()
    import time as _t
    t0 = _t.perf_counter()
    single = []
    for (k, t, lo, hi), jid in zip(jobs, ids):
        got, _ = engines[k].scan_window(t, lo, hi)
        single.append(sorted(got))
    swall = _t.perf_counter() - t0
    sms = sum(engines[k].kernel_time()[0] for k in engines)
    for jid, want in zip(ids, single):
        assert b.collect(jid) == want
    pos = N * len(jobs)
    # block sweep: 21 jobs (one 2^28-position window per plan) moved through 16 blocks, one
    # batched run per block (eager) versus one graph replay per block
    sw = es_batch.Batch()
    sw_jobs = []
    for k, eng in engines.items():
        lo = rng.randrange(eng.plan.R - (1 << 28))
        sw_jobs.append(sw.add(eng, 0, lo, lo + (1 << 28)))
    blocks = [rng.randrange(1 << 20) for _ in range(16)]
    sw.run()
    sweep = {}
    for mode in ("eager", "graph"):
        if mode == "graph":
            sw.capture()
            sw.replay()
        dev = 0.0
        w0 = _t.perf_counter()
        res = []
        for t in blocks:
            sw.set_blocks([t] * len(sw_jobs))
            dev += sw.run()[0] if mode == "eager" else sw.replay()
            res.append([len(sw.collect(j)) for j in sw_jobs])
        wall = _t.perf_counter() - w0
        sweep[mode] = {"device_ms": dev, "wall_s": wall,
                       "positions_per_s_wall": len(blocks) * len(sw_jobs) * (1 << 28) / wall,
                       "survivors": sum(map(sum, res))}
    assert sweep["eager"]["survivors"] == sweep["graph"]["survivors"]
    return {
        "metric": "batched residue positions sieved/s (k=300..320, 4 blocks each)",
        "value": pos / (bms / 1e3), "unit": "positions/s", "n_gpus": 1, "steps": a.steps,
        "warmup": 1, "ms_per_step": bms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u128/u32 integer", "data": "synthetic seeded windows",
        "config": {"workload": "21 plans x 4 blocks x 2^34 positions", "jobs": len(jobs),
                   "launches": launches, "autotuned": bool(a.autotune),
                   "row_lane_plans": sum(e.layout == "rows" for e in engines.values()),
                   "tuned": {str(k): v for k, v in tuned.items()}},
        "roofline": batch_roofline(engines, jobs, N, pos / (bms / 1e3), bms, launches),
        "unbatched": {"device_ms": sms, "wall_s": swall, "rate": pos / (sms / 1e3)},
        "block_sweep": {"jobs": len(sw_jobs), "blocks": len(blocks), "window": 1 << 28,
                        **sweep},
        "survivors_total": sum(len(x) for x in single)}


def batch_roofline(engines, jobs, N: int, rate: float, ms: float, launches: int) -> dict:
    """LDS roofline of a batch: algorithmic LDS bytes per position averaged over the jobs (each
    job's wheel has its own mask count, words per row and ring-0 radix), times the batch's
    positions per second of device time, against the measured conflict-free LDS read rate."""
    lds = lds_read_peak()
    bpp = sum(engines[k].lds_bytes_per_position() for k, _, _, _ in jobs) / len(jobs)
    peak = lds["conflict_free_bytes_per_s"]
    return {"bound": "lds", "achieved": bpp * rate / 1e9, "peak": peak / 1e9, "unit": "GB/s",
            "frac": bpp * rate / peak, "traffic": None, "lds_bytes_per_position": bpp,
            "kernel_ms": ms, "launches": launches}


def bench_k200(a) -> int:
    """k = 200 on its full-power wheel (p <= 23 full powers, extended to M <= 2^100; 37 filter
    primes): one seeded window of 2^36 positions at a seeded block t in [0, 2^20), scanned
    `steps` times; the CPU oracle is timed on the window's first 2^28 positions and must agree
    there; the full window's survivor hash is reported."""
    import random

    import numpy as np

    import es_plan
    import gpu_sieve_mi355x as G
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import es_oracle as O
    k, N = 200, 1 << 36
    wheel = es_plan.fullpower_wheel(k)
    tune = {"max_survivors": 1 << 20}
    if a.layout is not None:
        tune["mi_layout"] = a.layout
    if a.autotune:
        import es_autotune
        best, _ = es_autotune.autotune(k, wheel, probe_log2=36, method="coord")
        tune.update(best)
    plan = es_plan.Plan(k, tune, wheel)
    eng = G.GpuSieve(plan)
    rng = random.Random(200)
    t = rng.randrange(1 << 20)
    lo = rng.randrange(plan.R - N)
    eng.scan_windows(t, [(lo, lo + N)] * max(1, a.warmup))
    eng.sync()
    eng.kernel_time()
    t0 = time.perf_counter()
    res = eng.scan_windows(t, [(lo, lo + N)] * a.steps)
    eng.sync()
    wall = time.perf_counter() - t0
    # survivors are appended in no fixed order: compare each pass's sorted IDs and minimum
    assert all((sorted(i), m) == (sorted(res[0][0]), res[0][1]) for i, m in res), \
        "every pass over the window must agree"
    ids, mn = res[-1]
    dev_ms, launches, dev_pos = eng.kernel_time()
    ns = sorted(plan.idx_to_n(int(i), t) for i in ids)
    cpu = None
    if not a.no_cpu:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        sub = 1 << 28
        c0 = time.perf_counter()
        cnt, want = O.c_scan(k, plan.wheel, plan.rest, t, lo, lo + sub, threads=threads)
        cdt = time.perf_counter() - c0
        got = np.sort(np.asarray([i for i in ids if lo <= int(i) < lo + sub], dtype=np.uint64))
        assert cnt == len(want) and np.array_equal(got, want)
        cpu = {"value": sub / cdt, "unit": "positions/s", "cores": threads, "kind": "port",
               "sample": f"first 2^28 positions of the window ({cdt:.1f} s), survivors agree"}
    kern_rate = dev_pos / (dev_ms / 1e3)
    print(json.dumps({
        "metric": "residue positions sieved/s (k=200 full-power wheel)", "value": a.steps * N / wall,
        "unit": "positions/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u128/u32 integer", "data": "synthetic seeded window",
        "config": {"workload": "k=200 full-power wheel, 2^36 positions at seeded block",
                   "t": t, "idx_lo": lo, "wheel": [list(x) for x in plan.wheel],
                   "R": plan.R, "filters": len(plan.rest), "autotuned": bool(a.autotune),
                   "tune": {x: plan.tun[x] for x in ("mi_masks", "threads", "mi_strip_target")},
                   "inner_bound": eng.inner_bound,
                   "kernel_ring0_radix": eng._d.r0, "strip_len": eng.strip_len,
                   "threads": eng.threads},
        "survivors": len(ns), "survivor_hash": O.survivor_hash(ns),
        "min": str(mn) if mn is not None else None,
        "roofline": sieve_roofline(eng, plan, kern_rate, dev_ms, launches, lds_read_peak()),
        "cpu_baseline": cpu}), flush=True)
    return 0


def bench_surrogate(a) -> int:
    """Learned planner, k = 350..400: train the surrogate on wheels the sieve probes itself,
    then per k rank the candidate wheels, probe the shortlist and the plain wheel, and compare
    the chosen wheel with the plain one by the time to exhaust the first nonempty block:
    (floor(g/M) + 1) * R positions at the wheel's measured rate, where the rate is measured on
    two 2^36-position windows (block 0, and the window of block floor(g/M) holding g), and
    both wheels must return g as that window's minimum."""
    import math

    import numpy as np

    import es_plan
    import es_surrogate as S
    import gpu_sieve_mi355x as G
    import es_data
    bf = es_data.known_terms()
    ks = list(range(350, 401, a.k_step))
    SHORTLISTS = (1, 2, 4, 8, 32)
    t0 = time.perf_counter()
    X, Y, meta = S.training_set(ks, per_k=24, n_candidates=20000, seed=5, probe_log2=35,
                                log=lambda m: print(m, file=sys.stderr, flush=True))
    t_data = time.perf_counter() - t0
    n = len(X)
    rng = np.random.default_rng(7)
    perm = rng.permutation(n)
    hold = perm[: max(1, n // 5)]
    train = perm[max(1, n // 5):]
    model = S.Surrogate(seed=3)
    loss, train_ms = model.fit(X[train], Y[train], epochs=400, batch=128, lr=2e-3)
    pred = model.predict(X[hold])

    def r2(y, p):
        return float(1 - np.sum((y - p) ** 2) / max(1e-30, np.sum((y - y.mean()) ** 2)))
    fit = {"samples": n, "holdout": len(hold), "train_ms": train_ms,
           "r2_log2_density": r2(Y[hold, 0], pred[:, 0]), "r2_log2_ns": r2(Y[hold, 1], pred[:, 1])}
    model.fit(X, Y, epochs=400, batch=128, lr=2e-3)          # final model on every sample
    N = 1 << 36
    rows = []
    for k in ks:
        res = S.surrogate_plan(k, model, n_candidates=a.candidates, shortlist=32, probe_log2=36,
                               seed=k, compare=SHORTLISTS)
        g = bf[k]
        out = {}
        for tag, wheel in (("plain", res["plain"]["wheel"]), ("surrogate", res["best"]["wheel"])):
            if tag == "surrogate" and res["best"]["plain"]:
                out[tag] = dict(out["plain"])
                continue
            plan = es_plan.Plan(k, None, wheel)
            eng = G.GpuSieve(plan)
            gi, gt = plan.n_to_idx(g)
            lo = max(0, min(plan.R - N, gi - N // 2))
            # warm-up at full size: code load, and scratch memory for every resident wave where
            # the kernel instance spills, stay outside the timed scans
            eng.scan_window(0, 0, min(N, plan.R))
            eng.sync()
            eng.kernel_time()
            w0 = time.perf_counter()
            _, mn0 = eng.scan_window(0, 0, min(N, plan.R))
            _, mn = eng.scan_window(gt, lo, lo + N)
            eng.sync()
            wall = time.perf_counter() - w0
            assert mn == g, (k, tag, mn)
            assert mn0 is None or mn0 <= k + 1 or mn0 == g
            rate = (min(N, plan.R) + N) / wall
            blocks = g // plan.M + 1
            out[tag] = {"wheel": [list(x) for x in wheel], "R": plan.R,
                        "log2_M": math.log2(plan.M), "first_block": int(g // plan.M),
                        "rate": rate, "time_to_first_block_s": blocks * plan.R / rate,
                        "expected_s": res["plain" if tag == "plain" else "best"]["cost_s"]}
        out["k"] = k
        out["ratio"] = out["surrogate"]["time_to_first_block_s"] / out["plain"]["time_to_first_block_s"]
        out["expected_ratio"] = out["surrogate"]["expected_s"] / out["plain"]["expected_s"]
        out["same_wheel"] = bool(res["best"]["plain"])
        out["sizes"] = res["sizes"]
        out["analytic_expected_ratio"] = res["sizes"][32]["analytic"]
        out["rank_s"] = res["rank_seconds"]
        rows.append(out)
        print(json.dumps({"k": k, "ratio": round(out["ratio"], 4),
                          "expected_ratio": round(out["expected_ratio"], 4),
                          "analytic_expected_ratio": round(out["analytic_expected_ratio"], 4),
                          "same_wheel": out["same_wheel"]}), file=sys.stderr, flush=True)
    ratios = np.array([r["ratio"] for r in rows])
    eratios = np.array([r["expected_ratio"] for r in rows])
    total_s = sum(r["surrogate"]["time_to_first_block_s"] for r in rows)
    total_p = sum(r["plain"]["time_to_first_block_s"] for r in rows)
    print(json.dumps({
        "metric": "time to first nonempty block, surrogate-planned / plain-planned wheel",
        "value": total_s / total_p, "unit": "ratio (sum over k)", "n_gpus": 1,
        "steps": len(ks), "warmup": 0, "ms_per_step": None, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "u128/u32 integer; fp32 network",
        "data": "b-file g(k); wheels enumerated per k; rates measured on 2^36-position windows",
        "config": {"workload": f"k={ks[0]}..{ks[-1]} step {a.k_step}, {a.candidates} candidates "
                               "per k, shortlist 32 + plain, probes of up to 2^36 positions"},
        "geomean_ratio": float(np.exp(np.mean(np.log(ratios)))),
        "geomean_expected_ratio": float(np.exp(np.mean(np.log(eratios)))),
        "shortlist_sizes": {
            str(sz): {m: float(np.exp(np.mean(np.log([r["sizes"][sz][m] for r in rows]))))
                      for m in ("surrogate", "analytic")}
            for sz in list(SHORTLISTS) + ["top1_no_probe"]},
        "k_with_other_wheel": int(sum(not r["same_wheel"] for r in rows)),
        "training": dict(fit, data_seconds=t_data),
        "per_k": [{"k": r["k"], "ratio": r["ratio"], "expected_ratio": r["expected_ratio"],
                   "analytic_expected_ratio": r["analytic_expected_ratio"],
                   "plain_s": r["plain"]["time_to_first_block_s"],
                   "surrogate_s": r["surrogate"]["time_to_first_block_s"]} for r in rows]}),
        flush=True)
    return 0


def bench_mitm(a) -> int:
    """Meet-in-the-middle search for g(k): windows [0, B) with B = 3E, quadrupled until a good
    n > k+1 appears (the b-file value must be returned).  Reports candidates tested per second
    and the n-range covered per second (device time), the HBM rate of the table reads, and the
    CPU oracle's brute-force range scan (Kummer test of every n) on a bounded n-window."""
    import math

    import es_mitm as MM
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import es_oracle as O
    import es_data
    bf = es_data.known_terms()
    k = a.mitm_k
    t0 = time.perf_counter()
    s = MM.MitmSieve(k, mem_gb=a.mitm_mem)
    build_s = time.perf_counter() - t0
    for _ in range(max(0, a.warmup)):
        MM.mitm_search(k, sieve=s)
    runs = []
    t0 = time.perf_counter()
    for _ in range(max(1, a.steps)):
        runs.append(MM.mitm_search(k, sieve=s))
    wall = (time.perf_counter() - t0) / max(1, a.steps)
    r = runs[-1]
    assert r["status"] == "FOUND" and r["g"] == bf[k], r
    dev_ms = min(x["device_ms"] for x in runs)
    cand = r["candidates"]
    n_range = r["bound"]
    n1, n2 = s.sizes
    swept = s.swept
    # table bytes: 16 per candidate, plus per s1 either one sequential 16-byte read (sorted
    # sweep) or two binary searches of ceil(log2 S2) 16-byte probes
    per_s1 = 16 if swept else 2 * max(1, math.ceil(math.log2(n2))) * 16
    bytes_alg = cand * 16 + n1 * per_s1
    achieved = bytes_alg / (dev_ms / 1e3)
    cpu = None
    if not a.no_cpu:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        span = 1 << 27
        lo = r["g"] - span // 2
        c0 = time.perf_counter()
        _, got = O.c_good_range(k, lo, lo + span, threads=threads)
        cdt = time.perf_counter() - c0
        assert r["g"] in got
        cpu = {"value": span / cdt, "unit": "n/s", "cores": threads, "kind": "port",
               "sample": f"Kummer test of every n in a 2^27-wide range around g({k}) "
                         f"({cdt:.1f} s); contains g"}
    print(json.dumps({
        "metric": f"meet-in-the-middle search for g({k}): n-range covered per second",
        "value": n_range / (dev_ms / 1e3), "unit": "n/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": wall * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u128/u32 integer",
        "data": "exact search; g(k) checked against the b-file",
        "config": {"workload": f"k={k}, table budget {a.mitm_mem} GB", "wheel": s.wheel,
                   "M_bits": s.M.bit_length(), "S1": n1, "S2": n2, "bound": str(n_range),
                   "g": str(r["g"]), "build_s": build_s, "sorted_sweep": swept},
        "candidates": cand, "candidates_per_s": cand / (dev_ms / 1e3), "device_ms": dev_ms,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None,
                     "bytes": "16 B per candidate + per s1: 16 B (sorted sweep) or 2 binary "
                              "searches x ceil(log2 S2) x 16 B"},
        "cpu_baseline": cpu}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
