"""The block driver and the engine CLI on the HIP engine (harness/gate_driver.py:49-72 applied to
es_driver.run with gpu_sieve_mi355x; the CLI argv of harness/replay_release.py:21-39).

Every case prints its k and mode before touching the GPU."""
from __future__ import annotations

import json
import os

import pytest

import es_driver
import es_release
import gpu_sieve_mi355x as G
from conftest import bfile

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def drive(tmp_path, k, tag="gate", **kw):
    os.makedirs(tmp_path, exist_ok=True)
    args = dict(k=k, ckpt_path=str(tmp_path / f"k{k}_{tag}.json"), tun=None, wheel=None,
                max_blocks=None, timeout=None, log_every=1e9, ckpt_every=1e9, min_free_mib=0,
                resume=True)
    args.update(kw)
    return es_driver.run(G, log=lambda s: None, **args)


def test_driver_matches_bfile_k2_to_120(tmp_path):
    b = bfile()
    for k in range(2, 121):
        say(f"case driver k={k}")
        res = drive(tmp_path / str(k), k, resume=False)
        assert res["status"] == "FOUND" and int(res["g"]) == b[k], (k, res.get("status"))
        plan = G.Plan(k)
        rec = es_release.run_record_from_driver(res, plan)
        assert es_release.check_run(k, b[k], rec) == res["block"] + 1


def test_driver_max_blocks_then_resume(tmp_path):
    """k=156: g(156) lies in block 7 of the default wheel.  Two blocks, stop, resume."""
    k = 156
    say(f"case driver k={k} max-blocks 2 then resume")
    full = drive(tmp_path / "full", k, resume=False)
    assert full["status"] == "FOUND" and full["block"] == 7
    first = drive(tmp_path / "leg", k, max_blocks=2, resume=False)
    assert first["status"] == "EXHAUSTED_BLOCKS" and first["blocks"] == 2
    ck = json.load(open(tmp_path / "leg" / f"k{k}_gate.json"))
    assert ck["t"] == 2 and ck["unit_done"] == 0
    again = drive(tmp_path / "leg", k)
    assert again["status"] == "FOUND" and again["restarts"] > 0
    assert again["g"] == full["g"] and again["block"] == full["block"]
    assert again["block_audit"] == full["block_audit"]


def test_driver_max_launches_mid_block_then_resume(tmp_path):
    """k=309 (its wheel's ring 0 of 42 gives way to a lead block, forced): pause after a few
    launches inside block 0, resume twice, and require the uninterrupted run's audit rows and
    hashes."""
    k = 309
    say(f"case driver k={k} max-launches mid-block")
    tun = {"mi_launch_log2": 40, "mi_ring_swap": 1}
    full = drive(tmp_path / "full", k, tun=tun, resume=False)
    assert full["status"] == "FOUND" and int(full["g"]) == bfile()[k]
    leg = drive(tmp_path / "leg", k, tun=tun, max_launches=3, resume=False)
    assert leg["status"] == "PAUSED" and 0 < leg["unit_done"] < leg["units_total"]
    leg2 = drive(tmp_path / "leg", k, tun=tun, max_launches=3)
    assert leg2["status"] == "PAUSED" and leg2["unit_done"] > leg["unit_done"]
    done = drive(tmp_path / "leg", k, tun=tun)
    assert done["status"] == "FOUND" and done["restarts"] == 2
    assert done["block_audit"] == full["block_audit"]
    assert [r["sha256_survivors"] for r in done["block_audit"]] == \
           [r["sha256_survivors"] for r in full["block_audit"]]


def test_driver_refuses_a_mismatched_checkpoint(tmp_path):
    k = 70
    say(f"case driver k={k} refused checkpoint")
    res = drive(tmp_path, k, max_blocks=1, resume=False)
    assert res["status"] == "EXHAUSTED_BLOCKS"
    path = tmp_path / f"k{k}_gate.json"
    ck = json.load(open(path))
    ck["engine_revision"] = {"gpu_sieve_mi355x.py": "0" * 64}
    json.dump(ck, open(path, "w"))
    with pytest.raises(RuntimeError, match="mismatch"):
        drive(tmp_path, k)


def test_engine_cli_replay_argv(capsys):
    """`gpu_sieve_mi355x k --wheel JSON --max-blocks N`, as harness/replay_release.py:21-39 runs a
    recorded wheel: one JSON object per k with the v9 keys and the b-file verdict; with one
    block too few the search reports EXHAUSTED_BLOCKS."""
    b = bfile()
    k = 117                                     # g(117) lies in block 10 of its default wheel
    plan = G.Plan(k)
    wheel = [[p, T] for p, T in plan.wheel]
    blocks = b[k] // plan.M + 1
    say(f"case engine CLI k={k} wheel={wheel} max-blocks {blocks} and {blocks - 1}")
    G.main([str(k), "--wheel", json.dumps(wheel), "--max-blocks", str(blocks)])
    G.main([str(k), "--wheel", json.dumps(wheel), "--max-blocks", str(blocks - 1)])
    lines = [json.loads(x) for x in capsys.readouterr().out.splitlines() if x.startswith("{")]
    assert len(lines) == 2
    r = lines[0]
    for key in ("k", "status", "g", "block", "survivors_in_block", "residues_checked",
                "launches", "wall", "matches_bfile", "bfile", "known_source", "plan"):
        assert key in r, key
    assert r["k"] == k and r["status"] == "FOUND" and r["g"] == b[k] and r["block"] == blocks - 1
    assert r["matches_bfile"] is True and r["bfile"] == b[k]
    assert r["residues_checked"] == blocks * plan.R
    assert r["plan"]["wheel"] == wheel
    assert lines[1]["status"] == "EXHAUSTED_BLOCKS" and lines[1]["blocks"] == blocks - 1
