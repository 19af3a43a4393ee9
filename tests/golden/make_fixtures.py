"""Extract golden fixtures from the reference's DATA files (run in the build container only).

Reads, as text/JSON, never importing reference code:
  data/b003458_1to400.txt                       g(k) for k = 1..400
  evidence/through400/runs.json                 frontier wheels, per-block survivor lists, hashes
  evidence/logs/2026-08-22/logs/gate_v9_270_280.log   default-planner plans + block-0 results
  evidence/logs/*/logs/k3*_{primary,confirm}.log      recorded plans (JSON result lines)
Writes small JSON fixtures next to this script.
"""
from __future__ import annotations

import glob
import json
import os
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def main() -> None:
    shutil.copyfile(os.path.join(REF, "data", "b003458_1to400.txt"),
                    os.path.join(HERE, "b003458_1to400.txt"))

    runs = json.load(open(os.path.join(REF, "evidence", "through400", "runs.json")))
    keep = []
    for term in runs["terms"]:
        row = {"k": term["k"], "n": term["n"]}
        for role in ("primary", "confirmation"):
            r = term[role]
            row[role] = {"wheel": r["wheel"], "N": r["N"], "block": r["block"],
                         "residues_per_block": r["residues_per_block"],
                         "block_audit": r["block_audit"]}
        keep.append(row)
    json.dump({"schema_version": runs.get("schema_version"), "terms": keep},
              open(os.path.join(HERE, "runs_frontier.json"), "w"), indent=1)

    plans = {}
    gate = os.path.join(REF, "evidence/logs/2026-08-22/logs/gate_v9_270_280.log")
    for line in open(gate):
        if line.startswith("{"):
            r = json.loads(line)
            plans[str(r["k"])] = {"source": "gate_v9_270_280.log", "g": str(r["g"]),
                                  "survivors_in_block": r["survivors_in_block"],
                                  "residues_checked": r["residues_checked"], "plan": r["plan"]}
    json.dump(plans, open(os.path.join(HERE, "gate_v9_270_280.json"), "w"), indent=1)

    frontier = []
    for path in sorted(glob.glob(os.path.join(REF, "evidence/logs/*/logs/k3*_*.log"))):
        for line in open(path, errors="replace"):
            if not line.startswith("{"):
                continue
            try:
                r = json.loads(line)
            except json.JSONDecodeError:
                continue
            if "plan" not in r:
                continue
            frontier.append({"log": os.path.relpath(path, REF), "k": r["plan"]["k"],
                             "tag": r.get("tag"), "banned": r.get("banned"),
                             "status": r.get("status"), "plan": r["plan"]})
    json.dump(frontier, open(os.path.join(HERE, "frontier_plans.json"), "w"), indent=1)
    print(f"fixtures: {len(keep)} frontier terms, {len(plans)} gate plans, "
          f"{len(frontier)} recorded frontier plans")


if __name__ == "__main__":
    main()
