"""Package data: the known terms g(k), k = 1..400 (OEIS A003458 b-file, 'k g' per line), which
the engine CLI and the block driver compare their results with."""
from __future__ import annotations

import os

HERE = os.path.dirname(os.path.abspath(__file__))
BFILE = os.path.join(HERE, "data", "b003458_1to400.txt")


def known_terms(path: str = BFILE) -> dict[int, int]:
    """k -> g(k) from a 'k g' file (blank and '#' lines skipped)."""
    out: dict[int, int] = {}
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) == 2 and not line.startswith("#"):
                out[int(parts[0])] = int(parts[1])
    return out
