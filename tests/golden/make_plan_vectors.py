"""Write golden vectors of the default wheel planner (run in the build container only).

The planner's output is pinned by recorded plans at k = 270, 275, 280 and at the frontier runs
(frontier_plans.json, gate_v9_270_280.json).  For every other k nothing recorded fixes it, so
this script imports the reference's planner module (sieve/gpu_sieve_v9.py under
/root/reference: pure Python and NumPy, CuPy optional at import) and records, for each k, the
summary of its default Plan(k) and its filter order into plan_vectors.json.  The file holds
data only: each k (the input) and the plan it yields (the output).

    python3 tests/golden/make_plan_vectors.py [k_lo k_hi]
"""
from __future__ import annotations

import json
import os
import sys

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "plan_vectors.json")
# the summary's exact fields (ints, strings) and its fp64 fields, compared bit for bit
KEYS = ("wheel", "M", "n_filters", "residues_per_block", "density_wheel", "density_rest",
        "E_g", "expected_checks_true")


def main() -> int:
    lo, hi = (int(sys.argv[1]), int(sys.argv[2])) if len(sys.argv) == 3 else (2, 400)
    sys.path.insert(0, os.path.join(REF, "sieve"))
    import gpu_sieve_v9 as v9
    out = []
    for k in range(lo, hi + 1):
        P = v9.Plan(k)
        s = P.summary()
        rec = {key: s[key] for key in KEYS}
        rec["k"] = k
        rec["radix"] = list(P.radix)
        rec["rest"] = list(P.rest)
        out.append(rec)
        print(k, rec["wheel"], flush=True)
    with open(OUT, "w") as f:
        json.dump({"source": "sieve/gpu_sieve_v9.py Plan(k).summary(), default TUNABLES",
                   "plans": out}, f, separators=(",", ":"))
        f.write("\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
