"""Crash-resumable block driver: runs blocks t = 0, 1, ... of one wheel through an engine module
until a block holds a valid survivor, checkpointing inside blocks.

Role of harness/run_k_masks.py:74-224 (and harness/run_k.py:54-186) for any module with the
GpuSieve contract (gpu_sieve_mi355x here).  Shared with the reference only where callers and
records depend on it: the checkpoint and result keys, the block_audit row schema, the survivor
hash recipe (harness/run_k.py:38-44), the status strings and the CLI flags.

Design.  A `Checkpoint` owns the persisted state (one JSON object, written atomically with
fsync) and knows when it may be resumed: same k, wheel, M, R, engine module and engine source
revision, and internally consistent (units done within the block, survivor IDs unique and all
below the scanned prefix).  `Budget` decides, before each launch, whether the run must stop
(signal, launch budget, deadline).  `scan_block` advances the current block launch by launch
and records each launch's survivor IDs before the unit counter moves past it, so a checkpoint
never claims a unit whose survivors it lacks.  `conclude_block` turns a finished block's IDs
into n, applies the referee, and writes the audit row.  Survivor IDs are block-absolute, so
resuming re-scans exactly the units not yet recorded and the block's survivor set is the same
whatever the interruption pattern.
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
from typing import Any, Callable

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import es_plan  # noqa: E402
import es_data  # noqa: E402

SURV_HASH_RECIPE = ("sha256 of the UTF-8 bytes of '\\n'.join(decimal(n) for n in sorted(survivors)), "
                    "no trailing newline; the empty list hashes to sha256(b'')")
MISMATCH = "Checkpoint plan/engine revision mismatch; explicit qualified migration required"

STOP_REQUESTED = False


def request_stop(signum: int, frame: Any) -> None:
    """Signal handler: the run pauses at the next launch boundary."""
    global STOP_REQUESTED
    STOP_REQUESTED = True


def survivor_hash(ns) -> str:
    text = "\n".join(str(int(n)) for n in sorted(ns))
    return hashlib.sha256(text.encode()).hexdigest()


def engine_revision(engine_module) -> dict[str, str]:
    """Content hashes of what decides the survivor sets: the engine module, the planner, this
    driver, and (for the HIP engine) the kernel sources."""
    files = [Path(HERE) / "es_plan.py", Path(HERE) / "es_driver.py"]
    src = getattr(engine_module, "__file__", None)
    if src:
        files.append(Path(src))
    if getattr(engine_module, "__name__", "") == "gpu_sieve_mi355x":
        files += sorted((Path(HERE) / "csrc").glob("es_sieve*"))
    rev = {}
    for f in files:
        if f.exists():
            key = os.path.relpath(f, ROOT) if str(f).startswith(ROOT) else f.name
            rev[key] = hashlib.sha256(f.read_bytes()).hexdigest()
    return rev


class Checkpoint:
    """The persisted state of one (k, wheel, engine) search."""

    def __init__(self, path: str, state: dict):
        self.path = path
        self.s = state

    @classmethod
    def fresh(cls, path: str, k: int, plan, engine: str, revision: dict) -> "Checkpoint":
        return cls(path, {
            "engine_revision": revision, "k": k, "engine": engine,
            "wheel": [[p, T] for p, T in plan.wheel], "M": str(plan.M), "R": plan.R,
            "t": 0, "unit_done": 0, "idx_done": 0, "survivor_idx": [], "residues_checked": 0,
            "wall_accum": 0.0, "restarts": 0, "block_audit": []})

    def identity(self) -> tuple:
        s = self.s
        return (s.get("k"), s.get("wheel"), s.get("R"), s.get("engine"), s.get("M"),
                s.get("engine_revision"))

    @classmethod
    def resume_or_fresh(cls, path: str, k: int, plan, engine: str, revision: dict,
                        resume: bool) -> "Checkpoint":
        new = cls.fresh(path, k, plan, engine, revision)
        if not (resume and os.path.exists(path)):
            return new
        with open(path) as fh:
            old = cls(path, json.load(fh))
        if old.identity() != new.identity():
            raise RuntimeError(MISMATCH)
        old.s.setdefault("block_audit", [])
        old.s["restarts"] = int(old.s.get("restarts", 0)) + 1
        return old

    def check(self, units_total: int, unit: int) -> None:
        """Invariants a resumable state must satisfy (any violation is a corrupt checkpoint)."""
        s = self.s
        ids = s["survivor_idx"]
        problems = [
            (0 <= s["unit_done"] <= units_total, "units done outside the block"),
            (s["idx_done"] == s["unit_done"] * unit, "idx_done != unit_done * unit"),
            (s["residues_checked"] >= s["idx_done"], "fewer residues checked than scanned"),
            (len(set(ids)) == len(ids), "repeated survivor ID"),
            (all(0 <= i < s["idx_done"] for i in ids), "survivor ID beyond the scanned prefix"),
        ]
        for ok, what in problems:
            if not ok:
                raise AssertionError(f"corrupt checkpoint {self.path}: {what}")

    def save(self) -> None:
        tmp = self.path + ".tmp"
        with open(tmp, "w") as fh:
            json.dump(self.s, fh)
            fh.flush()
            os.fsync(fh.fileno())
        os.replace(tmp, self.path)
        d = os.open(os.path.dirname(os.path.abspath(self.path)), os.O_DIRECTORY)
        try:
            os.fsync(d)
        finally:
            os.close(d)

    def next_block(self) -> None:
        self.s.update(t=self.s["t"] + 1, unit_done=0, idx_done=0, survivor_idx=[])


class Budget:
    """Stop conditions checked before every launch: a stop signal or an exhausted launch budget
    pause the run (PAUSED), an expired deadline ends it (TIMEOUT)."""

    def __init__(self, max_launches: int | None, timeout: float | None, started: float):
        self.max_launches = max_launches
        self.deadline = started + timeout if timeout else None

    def verdict(self, launches: int) -> str | None:
        if STOP_REQUESTED or (self.max_launches is not None and launches >= self.max_launches):
            return "PAUSED"
        if self.deadline is not None and time.time() > self.deadline:
            return "TIMEOUT"
        return None


def scan_block(sieve, ck: Checkpoint, units_total: int, unit: int, chunk: int, budget: Budget,
               on_launch: Callable[[], None]) -> str | None:
    """Launch the rest of the current block; returns a stop verdict, or None once every unit
    of the block has been scanned and its survivors recorded."""
    s = ck.s
    while s["unit_done"] < units_total:
        stop = budget.verdict(sieve.launches)
        if stop:
            return stop
        first = s["unit_done"]
        count = min(chunk, units_total - first)
        sieve.out_n.fill(0)
        sieve._launch(first, count)
        found = int(sieve.out_n.get()[0])            # waits for the launch
        if found > sieve.out_cap:
            raise RuntimeError(f"survivor buffer overflow ({found} > {sieve.out_cap})")
        if found:
            s["survivor_idx"] += [int(x) for x in sieve.out[:found].get()]
        sieve.residues_checked += count * unit
        s.update(unit_done=first + count, idx_done=(first + count) * unit,
                 residues_checked=sieve.residues_checked)
        on_launch()
    return None


def conclude_block(plan, k: int, t: int, ids) -> tuple[dict, list[int], list[int]]:
    """Audit row of a fully scanned block, its referee-approved n, and the n > k+1 the referee
    rejects (a disagreement between kernel and referee)."""
    ns = sorted(plan.idx_to_n(i, t) for i in ids)
    good = [n for n in ns if es_plan.is_good(n, k)]
    valid = [n for n in ns if n > k + 1]
    rejected = sorted(set(valid) - set(good))
    row = {"t": t, "residues_enumerated": plan.R, "kernel_survivors": len(ns),
           "after_n_gt_k_plus_1": len(valid), "passed_ref_is_good": len(good),
           "survivors_n": [str(n) for n in ns] if len(ns) <= 64 else None,
           "sha256_survivors": survivor_hash(ns)}
    return row, good, rejected


def run(engine_module, k: int, ckpt_path: str, tun: dict | None, wheel, max_blocks: int | None,
        timeout: float | None, log_every: float, ckpt_every: float, min_free_mib: int,
        resume: bool, start_block: int | None = None, max_launches: int | None = None,
        log: Callable[[str], None] | None = None) -> dict[str, Any]:
    """Search blocks of one wheel until a valid survivor is found, the block limit is reached,
    or the run is paused / times out (state saved).  Result keys follow the reference driver."""
    log = log or (lambda m: print(m, file=sys.stderr, flush=True))
    started = time.time()
    plan = engine_module.Plan(k, tun, wheel)
    engine = engine_module.__name__
    ck = Checkpoint.resume_or_fresh(ckpt_path, k, plan, engine, engine_revision(engine_module),
                                    resume)
    s = ck.s
    if s["restarts"]:
        log(f"# resumed k={k} at block {s['t']}, {s['idx_done']}/{plan.R} positions, "
            f"{len(s['survivor_idx'])} survivors, restart {s['restarts']}")
    if start_block is not None and s["t"] == 0 and s["idx_done"] == 0:
        s["t"] = int(start_block)
        log(f"# k={k}: fresh run starting at block {start_block}")
    free = engine_module.free_vram_mib()
    if free is not None and free < min_free_mib:
        return {"k": k, "status": "SKIPPED_LOW_VRAM", "free_mib": free}

    sieve = engine_module.GpuSieve(plan)
    unit = getattr(sieve, "r0", 1)
    units_total = getattr(sieve, "R_outer", plan.R)
    chunk = getattr(sieve, "ochunk", getattr(sieve, "chunk", 1))
    assert units_total * unit == plan.R, "launch units do not tile the block"
    ck.check(units_total, unit)
    sieve.residues_checked = s["residues_checked"]
    s["mask_mode"] = getattr(sieve, "mask_mode", None)
    leg0 = time.time()
    checked0 = sieve.residues_checked
    budget = Budget(max_launches, timeout, started)
    clock = {"log": leg0, "save": leg0}

    def wall() -> float:
        return s.get("wall_accum", 0.0) + (time.time() - leg0)

    def common(status: str) -> dict:
        return {"k": k, "status": status, "block_audit": s["block_audit"],
                "surv_hash_recipe": SURV_HASH_RECIPE, "residues_checked": sieve.residues_checked,
                "wall": wall(), "wall_leg": time.time() - leg0, "restarts": s.get("restarts", 0)}

    def on_launch() -> None:
        now = time.time()
        if now - clock["log"] >= log_every:
            rate = (sieve.residues_checked - checked0) / max(now - leg0, 1e-9)
            left = (plan.R - s["idx_done"]) / rate if rate else math.inf
            log(f"k={k} block {s['t']}: {s['idx_done']:.4e}/{plan.R:.4e} positions "
                f"({100 * s['unit_done'] / units_total:.2f} %), {len(s['survivor_idx'])} survivors, "
                f"{rate:.3e} positions/s, block done in {left / 3600:.2f} h")
            clock["log"] = now
        if now - clock["save"] >= ckpt_every:
            ck.save()
            clock["save"] = now

    while max_blocks is None or s["t"] < max_blocks:
        t = s["t"]
        sieve.tconst_h[:len(sieve.rest_q)] = [((t % q) * mq) % q
                                              for q, mq in zip(sieve.rest_q, sieve.rest_Mmodq)]
        sieve.tconst.set(sieve.tconst_h)
        stop = scan_block(sieve, ck, units_total, unit, chunk, budget, on_launch)
        if stop:
            s["wall_accum"] = wall()
            ck.save()
            return dict(common(stop), block=t, idx_done=s["idx_done"], unit_done=s["unit_done"],
                        units_total=units_total, R=plan.R, wall=s["wall_accum"],
                        checkpoint=ckpt_path)
        row, good, rejected = conclude_block(plan, k, t, s["survivor_idx"])
        s["block_audit"].append(row)
        if rejected or good:
            ck.save()
            if rejected:
                return dict(common("DISAGREEMENT"), block=t,
                            kernel_survivors=[str(n) for n in sorted(
                                plan.idx_to_n(i, t) for i in s["survivor_idx"])][:20],
                            ref_rejected=[str(n) for n in rejected[:20]])
            return dict(common("FOUND"), g=str(good[0]), block=t, engine=engine,
                        blocks_scanned=t + 1, survivors_in_block=len(s["survivor_idx"]),
                        all_good_in_block=[str(n) for n in good])
        log(f"k={k} block {t}: every position scanned, no survivor passes the referee")
        ck.next_block()
        ck.save()
    return dict(common("EXHAUSTED_BLOCKS"), blocks=s["t"])


def main(argv=None) -> None:
    """CLI with the flags of harness/run_k_masks.py:253-292 (one JSON result line)."""
    signal.signal(signal.SIGTERM, request_stop)
    signal.signal(signal.SIGINT, request_stop)
    ap = argparse.ArgumentParser(description="resumable block driver for an engine module")
    ap.add_argument("k", type=int)
    ap.add_argument("--engine", default=os.environ.get("ES_ENGINE", "gpu_sieve_mi355x"))
    ap.add_argument("--tag", default="primary")
    ap.add_argument("--ckpt-dir", default=None)
    ap.add_argument("--tune", type=json.loads, default=None)
    ap.add_argument("--wheel", type=json.loads, default=None)
    ap.add_argument("--ban", default=None, help="comma-separated primes kept out of the wheel")
    ap.add_argument("--max-blocks", type=int, default=None)
    ap.add_argument("--start-block", type=int, default=None)
    ap.add_argument("--max-launches", type=int, default=None)
    ap.add_argument("--timeout", type=float, default=None)
    ap.add_argument("--log-every", type=float, default=60.0)
    ap.add_argument("--ckpt-every", type=float, default=120.0)
    ap.add_argument("--min-free-mib", type=int, default=2048)
    ap.add_argument("--no-resume", action="store_true")
    ap.add_argument("--bfile", default=es_data.BFILE)
    ap.add_argument("--known", default=None)
    a = ap.parse_args(argv)
    eng = importlib.import_module(a.engine)

    wheel = a.wheel
    if a.ban:
        banned = {int(x) for x in a.ban.split(",")}
        base = eng.Plan(a.k, a.tune)
        keep = [p for p in es_plan.primes_upto(a.k) if p not in banned]
        wheel = [[p, T] for p, T in eng.plan_wheel(a.k, keep, -math.log2(base.density_all),
                                                   dict(eng.TUNABLES, **(a.tune or {})))]
        print(f"# wheel without {sorted(banned)}: {wheel}", file=sys.stderr, flush=True)
    ckpt_dir = a.ckpt_dir or os.path.join(os.getcwd(), "checkpoints")
    os.makedirs(ckpt_dir, exist_ok=True)
    ckpt = os.path.join(ckpt_dir, f"k{a.k}_{a.tag}.json")
    res = run(eng, a.k, ckpt, a.tune, wheel, a.max_blocks, a.timeout, a.log_every, a.ckpt_every,
              a.min_free_mib, not a.no_resume, a.start_block, a.max_launches)
    res.update(plan=eng.Plan(a.k, a.tune, wheel).summary(), engine_module=eng.__name__,
               engine_file=os.path.relpath(eng.__file__, ROOT), tag=a.tag,
               engine_revision=engine_revision(eng))
    if a.ban:
        res["banned"] = a.ban
    known = eng.load_known(a.bfile, a.known)
    if res.get("status") == "FOUND" and a.k in known:
        g, src = known[a.k]
        res.update(matches_bfile=int(res["g"]) == g, bfile=str(g), known_source=src)
    print(json.dumps(res, default=str), flush=True)
    if res.get("status") in ("DISAGREEMENT", "SKIPPED_LOW_VRAM"):
        raise SystemExit(2)


if __name__ == "__main__":
    main()
