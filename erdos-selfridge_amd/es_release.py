"""Consistency checks of recorded block scans (runs.json schema) and of 'k g' term tables.

Role of harness/verify_release.py:33-88 (`read_table`, `check_run`).  A recorded run claims a
wheel, its modulus N and positions per block, and one audit row per scanned block.  The checker
re-derives every fact it can from (k, wheel, the recorded survivor lists) alone -- the wheel's
modulus and residue count, the block sequence, each survivor list's interval, uniqueness,
hash, count and Kummer admissibility, the count of n > k+1, and where the first valid
survivor lies -- and compares the record with them.  The first disagreement raises ValueError
with the reference's message text (callers match on it).  Shared with the reference: the
schema keys and those messages.

The reference also re-hashes the engine sources a run names (engine_revision) against its own
tree; here that check runs only when `source_root` is given, since the recorded sources are the
reference's, not this package's.
"""
from __future__ import annotations

import hashlib
import json
import re
from dataclasses import dataclass
from math import prod
from pathlib import Path
from typing import Any, Iterator

from es_plan import base_digits, kummer_ok, primes_upto

_ROW = re.compile(r"[1-9][0-9]* [1-9][0-9]*")


def require(condition: bool, message: str) -> None:
    if not condition:
        raise ValueError(message)


def _rows(path: Path, text: str) -> Iterator[tuple[int, int]]:
    for line in text.splitlines():
        if line.strip() and not line.startswith("#"):
            require(_ROW.fullmatch(line) is not None, f"{path}: invalid row")
            k, n = line.split()
            yield int(k), int(n)


def read_table(path, count: int) -> dict[int, int]:
    """Terms g(1..count) of an LF-terminated ASCII 'k g' table, indices consecutive from 1."""
    path = Path(path)
    raw = path.read_bytes()
    require(raw.endswith(b"\n") and b"\r" not in raw, f"{path}: expected LF-terminated text")
    table: dict[int, int] = {}
    for expect, (k, n) in enumerate(_rows(path, raw.decode("ascii")), start=1):
        require(k == expect, f"{path}: missing, repeated, or unordered index {k}")
        table[k] = n
    require(len(table) == count, f"{path}: expected {count} rows, got {len(table)}")
    return table


@dataclass(frozen=True)
class WheelFacts:
    """What a wheel fixes: the primes <= k, the block modulus, positions per block."""
    primes: tuple
    modulus: int
    residues: int

    @classmethod
    def derive(cls, k: int, wheel) -> "WheelFacts":
        ps = primes_upto(k)
        items = [(int(p), int(e)) for p, e in wheel]
        require(len({p for p, _ in items}) == len(items), f"k={k}: repeated wheel prime")
        require(all(p in ps and 1 <= e <= len(base_digits(k, p)) for p, e in items),
                f"k={k}: invalid wheel")
        residues = 1
        for p, e in items:
            for digit in base_digits(k, p)[:e]:
                residues *= p - digit
        return cls(tuple(ps), prod(p ** e for p, e in items), residues)


@dataclass(frozen=True)
class BlockFacts:
    """What a block's recorded survivor list fixes on its own."""
    survivors: tuple
    unique: bool
    inside: bool
    digest: str
    admissible: bool
    valid: tuple            # survivors n > k+1

    @classmethod
    def derive(cls, k: int, t: int, wheel: WheelFacts, listed) -> "BlockFacts":
        ns = tuple(sorted(int(x) for x in listed))
        lo, hi = t * wheel.modulus, (t + 1) * wheel.modulus
        return cls(
            survivors=ns, unique=len(set(ns)) == len(ns),
            inside=all(lo <= n < hi for n in ns),
            digest=hashlib.sha256("\n".join(str(n) for n in ns).encode()).hexdigest(),
            admissible=all(kummer_ok(n, k, p) for n in ns for p in wheel.primes),
            valid=tuple(n for n in ns if n > k + 1))


def _check_block(k: int, n: int, terminal_t: int, wheel: WheelFacts, row: dict) -> None:
    t = row["t"]
    f = BlockFacts.derive(k, t, wheel, row["survivors_n"])
    require(row["residues_enumerated"] == wheel.residues, f"k={k}, t={t}: incomplete residue count")
    require(f.unique, f"k={k}: repeated survivor")
    require(f.inside, f"k={k}: survivor outside block")
    require(f.digest == row["sha256_survivors"], f"k={k}: survivor hash mismatch")
    require(len(f.survivors) == row["kernel_survivors"], f"k={k}: survivor count mismatch")
    require(f.admissible, f"k={k}: inadmissible recorded survivor")
    require(row["after_n_gt_k_plus_1"] == row["passed_ref_is_good"] == len(f.valid),
            f"k={k}: acceptance count mismatch")
    if t < terminal_t:
        require(not f.valid, f"k={k}: earlier valid survivor")
    else:
        require(bool(f.valid) and min(f.valid) == n, f"k={k}: terminal minimum mismatch")


def check_run(k: int, n: int, run: dict[str, Any], source_root=None) -> int:
    """Validate one recorded run of the term g(k) = n; returns its number of block records."""
    wheel = WheelFacts.derive(k, run["wheel"])
    require(wheel.modulus == int(run["N"]), f"k={k}: wheel modulus mismatch")
    require(wheel.residues == int(run["residues_per_block"]), f"k={k}: wheel residue count mismatch")
    terminal_t = run["block"]
    require(n // wheel.modulus == terminal_t, f"k={k}: answer outside terminal block")
    if "reported_g" in run:
        require(int(run["reported_g"]) == n, f"k={k}: terminal answer mismatch")
    rows = run["block_audit"]
    require([row["t"] for row in rows] == list(range(terminal_t + 1)),
            f"k={k}: missing or repeated block")
    for row in rows:
        _check_block(k, n, terminal_t, wheel, row)
    if source_root is not None:
        for rel, digest in run.get("engine_revision", {}).items():
            src = Path(source_root) / rel
            require(src.exists() and hashlib.sha256(src.read_bytes()).hexdigest() == digest,
                    f"k={k}: source revision mismatch: {rel}")
    prov = run.get("provenance", {})
    if prov.get("record_kind") == "completed block checkpoint":
        require(prov["idx_done"] == prov["R"] == wheel.residues, f"k={k}: checkpoint incomplete")
    return len(rows)


def check_runs_file(path, table: dict[int, int], source_root=None) -> dict:
    """Every run (primary and confirmation) of a runs.json-style file against the term table."""
    terms = json.loads(Path(path).read_text())["terms"]
    blocks = 0
    for term in terms:
        k, n = term["k"], int(term["n"])
        require(table.get(k) == n, f"k={k}: evidence/table mismatch")
        for role in ("primary", "confirmation"):
            if term.get(role):
                blocks += check_run(k, n, term[role], source_root)
    return {"status": "PASS", "terms": len(terms), "complete_unique_block_records": blocks}


def run_record_from_driver(result: dict, plan) -> dict:
    """A runs.json run record from an es_driver FOUND result."""
    return {"wheel": [[p, T] for p, T in plan.wheel], "N": str(plan.M), "block": result["block"],
            "residues_per_block": plan.R, "block_audit": result["block_audit"],
            "reported_g": str(result["g"])}
