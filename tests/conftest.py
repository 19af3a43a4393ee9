"""Shared pytest setup: import paths, the `gpu` marker, and helpers."""
from __future__ import annotations

import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "erdos-selfridge_amd")
for p in (PKG, os.path.join(ROOT, "oracle"), ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libes_sieve.so)")
    config.addinivalue_line("markers", "slow: long CPU test")


def bfile() -> dict[int, int]:
    out = {}
    for line in open(os.path.join(GOLDEN, "b003458_1to400.txt")):
        if line.strip():
            k, n = line.split()
            out[int(k)] = int(n)
    return out
