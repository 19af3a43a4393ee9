"""Crash-resumable block driver for the MI355X engine.

Counterpart of harness/run_k_masks.py:74-224 (and harness/run_k.py:54-186): it runs blocks
t = 0, 1, ... of one plan through any engine with the GpuSieve contract, checkpoints inside a
block (atomic write + fsync), resumes from the checkpoint (refusing a plan or engine revision
mismatch), pauses on SIGTERM/SIGINT or after --max-launches at a launch boundary, and records a
block_audit row per completed block with the survivor hash recipe of harness/run_k.py:38-44.

Block results are exact whatever the interruption pattern: survivor IDs are block-absolute,
the checkpoint stores those found before unit_done, and resuming re-scans only units >=
unit_done, so the block's survivor set is reproduced exactly.
"""
from __future__ import annotations

import argparse
import hashlib
import importlib
import json
import math
import os
import signal
import sys
import time
from pathlib import Path
from typing import Any

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import es_plan  # noqa: E402

SURV_HASH_RECIPE = ("sha256 of the UTF-8 bytes of '\\n'.join(decimal(n) for n in sorted(survivors)), "
                    "no trailing newline; the empty list hashes to sha256(b'')")

STOP_REQUESTED = False


def request_stop(signum: int, frame: Any) -> None:
    global STOP_REQUESTED
    STOP_REQUESTED = True


def survivor_hash(ns) -> str:
    """sha256 of '\\n'.join(sorted decimal n), reproducible by a third party."""
    return hashlib.sha256("\n".join(str(int(n)) for n in sorted(ns)).encode()).hexdigest()


def engine_revision(engine_module) -> dict[str, str]:
    """Content hashes of the engine's sources and its native library."""
    paths = [os.path.join(HERE, "es_plan.py"), os.path.join(HERE, "es_driver.py"),
             os.path.join(HERE, "csrc", "es_sieve.hip")]
    mod_file = getattr(engine_module, "__file__", None)
    if mod_file:
        paths.append(mod_file)
    out = {}
    for p in paths:
        if os.path.exists(p):
            out[os.path.relpath(p, ROOT)] = hashlib.sha256(Path(p).read_bytes()).hexdigest()
    return out


def atomic_write(path: str, obj: dict) -> None:
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(obj, f)
        f.flush()
        os.fsync(f.fileno())
    os.replace(tmp, path)
    parent = os.open(os.path.dirname(os.path.abspath(path)), os.O_DIRECTORY)
    try:
        os.fsync(parent)
    finally:
        os.close(parent)


def run(engine_module, k: int, ckpt_path: str, tun: dict | None, wheel, max_blocks: int | None,
        timeout: float | None, log_every: float, ckpt_every: float, min_free_mib: int,
        resume: bool, start_block: int | None = None, max_launches: int | None = None,
        log=None) -> dict[str, Any]:
    """One k through engine_module (which provides Plan, GpuSieve, free_vram_mib)."""
    log = log or (lambda s: print(s, file=sys.stderr, flush=True))
    t_start = time.time()
    plan = engine_module.Plan(k, tun, wheel)
    wheel_json = [[p, T] for p, T in plan.wheel]
    engine = engine_module.__name__
    rev = engine_revision(engine_module)
    state: dict[str, Any] = {"engine_revision": rev, "k": k, "engine": engine, "wheel": wheel_json,
                             "M": str(plan.M), "R": plan.R, "t": 0, "unit_done": 0, "idx_done": 0,
                             "survivor_idx": [], "residues_checked": 0, "wall_accum": 0.0,
                             "restarts": 0, "block_audit": []}
    if resume and os.path.exists(ckpt_path):
        old = json.load(open(ckpt_path))
        if (old.get("k") == k and old.get("wheel") == wheel_json and old.get("R") == plan.R
                and old.get("engine") == engine and old.get("M") == str(plan.M)
                and old.get("engine_revision") == rev):
            state = old
            state.setdefault("block_audit", [])
            state["restarts"] = state.get("restarts", 0) + 1
            log(f"# RESUME k={k} engine={engine} block t={state['t']} "
                f"residues_done={state['idx_done']}/{plan.R} survivors={len(state['survivor_idx'])} "
                f"restarts={state['restarts']}")
        else:
            raise RuntimeError("Checkpoint plan/engine revision mismatch; explicit qualified migration required")

    if start_block is not None and state["t"] == 0 and state["idx_done"] == 0:
        state["t"] = start_block

    free = engine_module.free_vram_mib()
    if free is not None and free < min_free_mib:
        return {"k": k, "status": "SKIPPED_LOW_VRAM", "free_mib": free}

    sieve = engine_module.GpuSieve(plan)
    sieve.residues_checked = state["residues_checked"]
    checked_at_start = sieve.residues_checked
    r0 = getattr(sieve, "r0", 1)
    units_total = getattr(sieve, "R_outer", plan.R)
    unit_chunk = getattr(sieve, "ochunk", getattr(sieve, "chunk", 1))
    assert units_total * r0 == plan.R, "unit accounting does not cover the block exactly"
    assert 0 <= state["unit_done"] <= units_total
    assert state["idx_done"] == state["unit_done"] * r0
    assert state["residues_checked"] >= state["idx_done"]
    assert len(state["survivor_idx"]) == len(set(state["survivor_idx"]))
    assert all(0 <= i < state["idx_done"] for i in state["survivor_idx"])
    state["mask_mode"] = getattr(sieve, "mask_mode", "scalar")
    last_log = last_ckpt = time.time()
    t0 = time.time()

    def save() -> None:
        atomic_write(ckpt_path, state)

    def wall() -> float:
        return state.get("wall_accum", 0.0) + (time.time() - t0)

    while max_blocks is None or state["t"] < max_blocks:
        t = state["t"]
        for j, q in enumerate(sieve.rest_q):
            sieve.tconst_h[j] = ((t % q) * sieve.rest_Mmodq[j]) % q
        sieve.tconst.set(sieve.tconst_h)

        while state["unit_done"] < units_total:
            if (STOP_REQUESTED or (max_launches is not None and sieve.launches >= max_launches)
                    or (timeout and time.time() - t_start > timeout)):
                state["residues_checked"] = sieve.residues_checked
                state["wall_accum"] = wall()
                save()
                return {"k": k, "status": "PAUSED" if STOP_REQUESTED or max_launches is not None else "TIMEOUT",
                        "block": t, "idx_done": state["idx_done"], "block_audit": state["block_audit"],
                        "surv_hash_recipe": SURV_HASH_RECIPE, "unit_done": state["unit_done"],
                        "units_total": units_total, "R": plan.R,
                        "residues_checked": sieve.residues_checked, "wall": state["wall_accum"],
                        "wall_leg": time.time() - t0, "restarts": state.get("restarts", 0),
                        "checkpoint": ckpt_path}
            ustart = state["unit_done"]
            ucount = min(unit_chunk, units_total - ustart)
            sieve.out_n.fill(0)
            sieve._launch(ustart, ucount)
            nsv = int(sieve.out_n.get()[0])
            if nsv > sieve.out_cap:
                raise RuntimeError(f"survivor buffer overflow ({nsv} > {sieve.out_cap})")
            if nsv:
                state["survivor_idx"].extend(int(x) for x in sieve.out[:nsv].get())
            sieve.residues_checked += ucount * r0
            state["unit_done"] = ustart + ucount
            state["idx_done"] = state["unit_done"] * r0
            state["residues_checked"] = sieve.residues_checked
            now = time.time()
            if now - last_log >= log_every:
                el = now - t0
                rate = (sieve.residues_checked - checked_at_start) / el if el else 0
                eta = (plan.R - state["idx_done"]) / rate if rate else float("inf")
                log(f"k={k} t={t} {100 * state['unit_done'] / units_total:6.2f}%  "
                    f"idx={state['idx_done']:.6e}/{plan.R:.6e} surv={len(state['survivor_idx'])} "
                    f"checked={sieve.residues_checked:.4e} rate={rate:.3e}/s "
                    f"elapsed={el / 3600:.2f}h block_eta={eta / 3600:.2f}h")
                last_log = now
            if now - last_ckpt >= ckpt_every:
                save()
                last_ckpt = now

        cands = sorted(plan.idx_to_n(i, t) for i in state["survivor_idx"])
        good = [n for n in cands if es_plan.is_good(n, k)]
        bad = [n for n in cands if n > k + 1 and n not in good]
        state["block_audit"].append({
            "t": t, "residues_enumerated": plan.R, "kernel_survivors": len(cands),
            "after_n_gt_k_plus_1": sum(1 for n in cands if n > k + 1),
            "passed_ref_is_good": len(good),
            "survivors_n": [str(n) for n in cands] if len(cands) <= 64 else None,
            "sha256_survivors": survivor_hash(cands)})
        if bad:
            state["residues_checked"] = sieve.residues_checked
            save()
            return {"k": k, "status": "DISAGREEMENT", "block": t, "block_audit": state["block_audit"],
                    "surv_hash_recipe": SURV_HASH_RECIPE, "kernel_survivors": [str(x) for x in cands[:20]],
                    "ref_rejected": [str(x) for x in bad[:20]], "residues_checked": sieve.residues_checked,
                    "wall": wall(), "wall_leg": time.time() - t0}
        if good:
            state["residues_checked"] = sieve.residues_checked
            save()
            return {"k": k, "status": "FOUND", "g": str(good[0]), "block": t, "engine": engine,
                    "block_audit": state["block_audit"], "surv_hash_recipe": SURV_HASH_RECIPE,
                    "blocks_scanned": t + 1, "survivors_in_block": len(state["survivor_idx"]),
                    "all_good_in_block": [str(x) for x in good],
                    "residues_checked": sieve.residues_checked, "wall": wall(),
                    "wall_leg": time.time() - t0, "restarts": state.get("restarts", 0)}
        log(f"k={k} block t={t} complete: no survivor passed the referee; advancing")
        state["t"] = t + 1
        state["unit_done"] = 0
        state["idx_done"] = 0
        state["survivor_idx"] = []
        save()

    return {"k": k, "status": "EXHAUSTED_BLOCKS", "blocks": state["t"], "block_audit": state["block_audit"],
            "surv_hash_recipe": SURV_HASH_RECIPE, "residues_checked": sieve.residues_checked,
            "wall": wall(), "wall_leg": time.time() - t0, "restarts": state.get("restarts", 0)}


def main(argv=None) -> None:
    """argv of harness/run_k_masks.py:230-248; ES_ENGINE selects the engine module."""
    signal.signal(signal.SIGTERM, request_stop)
    signal.signal(signal.SIGINT, request_stop)
    ap = argparse.ArgumentParser(description="resumable block driver for the MI355X engine")
    ap.add_argument("k", type=int)
    ap.add_argument("--tag", default="primary")
    ap.add_argument("--ckpt-dir", default=None)
    ap.add_argument("--tune", type=json.loads, default=None)
    ap.add_argument("--wheel", type=json.loads, default=None)
    ap.add_argument("--ban", default=None, help="comma-separated primes banned from the wheel")
    ap.add_argument("--max-blocks", type=int, default=None)
    ap.add_argument("--start-block", type=int, default=None)
    ap.add_argument("--max-launches", type=int, default=None)
    ap.add_argument("--timeout", type=float, default=None)
    ap.add_argument("--log-every", type=float, default=60.0)
    ap.add_argument("--ckpt-every", type=float, default=120.0)
    ap.add_argument("--min-free-mib", type=int, default=2048)
    ap.add_argument("--no-resume", action="store_true")
    ap.add_argument("--bfile", default=os.path.join(ROOT, "tests", "golden", "b003458_1to400.txt"))
    ap.add_argument("--known", default=None)
    a = ap.parse_args(argv)
    mod = importlib.import_module(os.environ.get("ES_ENGINE", "gpu_sieve_mi355x"))
    wheel = a.wheel
    if a.ban:
        ban = [int(x) for x in a.ban.split(",")]
        ps = [p for p in es_plan.primes_upto(a.k) if p not in ban]
        full = mod.Plan(a.k, a.tune)
        w = mod.plan_wheel(a.k, ps, -math.log2(full.density_all), dict(mod.TUNABLES, **(a.tune or {})))
        wheel = [[p, T] for p, T in w]
    ckpt_dir = a.ckpt_dir or os.path.join(ROOT, "checkpoints")
    os.makedirs(ckpt_dir, exist_ok=True)
    ckpt = os.path.join(ckpt_dir, f"k{a.k}_{a.tag}.json")
    res = run(mod, a.k, ckpt, a.tune, wheel, a.max_blocks, a.timeout, a.log_every, a.ckpt_every,
              a.min_free_mib, not a.no_resume, a.start_block, a.max_launches)
    res["plan"] = mod.Plan(a.k, a.tune, wheel).summary()
    res["engine_module"] = mod.__name__
    res["tag"] = a.tag
    if a.ban:
        res["banned"] = a.ban
    known = mod.load_known(a.bfile, a.known)
    if res.get("status") == "FOUND" and a.k in known:
        g_known, src = known[a.k]
        res["matches_bfile"] = int(res["g"]) == g_known
        res["bfile"] = str(g_known)
        res["known_source"] = src
    res["engine_revision"] = engine_revision(mod)
    print(json.dumps(res, default=str), flush=True)
    if res.get("status") in ("DISAGREEMENT", "SKIPPED_LOW_VRAM"):
        raise SystemExit(2)


if __name__ == "__main__":
    main()
