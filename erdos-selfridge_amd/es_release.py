"""Consistency checks of recorded block scans (runs.json schema) and of b-file tables.

Counterpart of harness/verify_release.py:33-88: ``read_table`` validates a 'k g' table and
``check_run`` validates one recorded run of a term -- wheel validity, modulus and residue
count recomputed from the wheel, consecutive complete blocks, per-block survivor interval,
uniqueness, hash, count and admissibility, acceptance counts, and that the terminal block's
least valid survivor is the term.  Same error messages, so callers can match on them.
Block records produced by es_driver.run (block_audit rows) use the same schema.
"""
from __future__ import annotations

import hashlib
import json
import re
from math import prod
from pathlib import Path
from typing import Any

from es_plan import base_digits, kummer_ok, primes_upto


def require(condition: bool, message: str) -> None:
    if not condition:
        raise ValueError(message)


def read_table(path, count: int) -> dict[int, int]:
    """Rows 'k g' for k = 1..count in order, LF-terminated ASCII."""
    path = Path(path)
    raw = path.read_bytes()
    require(raw.endswith(b"\n") and b"\r" not in raw, f"{path}: expected LF-terminated text")
    rows: dict[int, int] = {}
    for line in raw.decode("ascii").splitlines():
        if not line.strip() or line.startswith("#"):
            continue
        require(re.fullmatch(r"[1-9][0-9]* [1-9][0-9]*", line) is not None, f"{path}: invalid row")
        k, n = map(int, line.split())
        require(k == len(rows) + 1, f"{path}: missing, repeated, or unordered index {k}")
        rows[k] = n
    require(len(rows) == count, f"{path}: expected {count} rows, got {len(rows)}")
    return rows


def check_run(k: int, n: int, run: dict[str, Any]) -> int:
    """Validate one recorded run of term g(k) = n; returns the number of block records."""
    ps = primes_upto(k)
    wheel = run["wheel"]
    require(len({p for p, _ in wheel}) == len(wheel), f"k={k}: repeated wheel prime")
    require(all(p in ps and 1 <= e <= len(base_digits(k, p)) for p, e in wheel), f"k={k}: invalid wheel")
    modulus = prod(p ** e for p, e in wheel)
    residues = prod(prod(p - d for d in base_digits(k, p)[:e]) for p, e in wheel)
    require(modulus == int(run["N"]), f"k={k}: wheel modulus mismatch")
    require(residues == int(run["residues_per_block"]), f"k={k}: wheel residue count mismatch")
    require(n // modulus == run["block"], f"k={k}: answer outside terminal block")
    if "reported_g" in run:
        require(n == int(run["reported_g"]), f"k={k}: terminal answer mismatch")
    blocks = run["block_audit"]
    require([b["t"] for b in blocks] == list(range(run["block"] + 1)), f"k={k}: missing or repeated block")
    for block in blocks:
        t = block["t"]
        require(block["residues_enumerated"] == residues, f"k={k}, t={t}: incomplete residue count")
        survivors = sorted(map(int, block["survivors_n"]))
        require(len(set(survivors)) == len(survivors), f"k={k}: repeated survivor")
        require(all(t * modulus <= v < (t + 1) * modulus for v in survivors), f"k={k}: survivor outside block")
        digest = hashlib.sha256("\n".join(map(str, survivors)).encode()).hexdigest()
        require(digest == block["sha256_survivors"], f"k={k}: survivor hash mismatch")
        require(len(survivors) == block["kernel_survivors"], f"k={k}: survivor count mismatch")
        require(all(all(kummer_ok(v, k, p) for p in ps) for v in survivors),
                f"k={k}: inadmissible recorded survivor")
        valid = [v for v in survivors if v > k + 1]
        require(block["after_n_gt_k_plus_1"] == block["passed_ref_is_good"] == len(valid),
                f"k={k}: acceptance count mismatch")
        if t < run["block"]:
            require(not valid, f"k={k}: earlier valid survivor")
        else:
            require(bool(valid) and min(valid) == n, f"k={k}: terminal minimum mismatch")
    prov = run.get("provenance", {})
    if prov.get("record_kind") == "completed block checkpoint":
        require(prov["idx_done"] == prov["R"] == residues, f"k={k}: checkpoint incomplete")
    return len(blocks)


def check_runs_file(path, table: dict[int, int]) -> dict:
    """Every run of a runs.json-style file against the term table."""
    terms = json.loads(Path(path).read_text())["terms"]
    blocks = 0
    for term in terms:
        k, n = term["k"], int(term["n"])
        require(n == table[k], f"k={k}: evidence/table mismatch")
        for role in ("primary", "confirmation"):
            if role in term and term[role]:
                blocks += check_run(k, n, term[role])
    return {"status": "PASS", "terms": len(terms), "complete_unique_block_records": blocks}


def run_record_from_driver(result: dict, plan) -> dict:
    """runs.json run record from an es_driver FOUND result."""
    return {"wheel": [[p, T] for p, T in plan.wheel], "N": str(plan.M), "block": result["block"],
            "residues_per_block": plan.R, "block_audit": result["block_audit"],
            "reported_g": str(result["g"])}
