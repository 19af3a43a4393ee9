"""Engine CLI wire contract on the CPU: --plan-only output and the timeout exit code
(sieve/gpu_sieve_v9.py:584-621: one JSON object per k; exit 124 when --timeout expires)."""
from __future__ import annotations

import json
import os
import subprocess
import sys

from conftest import PKG


def _cli(*args, timeout=120):
    env = dict(os.environ, PYTHONPATH=PKG)
    return subprocess.run([sys.executable, os.path.join(PKG, "gpu_sieve_mi355x.py"), *args],
                          capture_output=True, text=True, timeout=timeout, env=env)


def test_plan_only_prints_one_summary_per_k():
    p = _cli("270", "275", "--plan-only")
    assert p.returncode == 0, p.stderr
    rows = [json.loads(x) for x in p.stdout.splitlines()]
    assert [r["k"] for r in rows] == [270, 275]
    assert rows[0]["wheel"] and rows[0]["residues_per_block"] > 0


def test_timeout_exits_124_with_a_timeout_line():
    """The alarm fires while the k=399 planner runs, before any device work."""
    p = _cli("399", "--timeout", "0.01")
    assert p.returncode == 124
    assert json.loads(p.stdout.splitlines()[-1]) == {"status": "TIMEOUT"}
