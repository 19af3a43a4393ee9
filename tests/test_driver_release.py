"""Block driver (checkpoint, resume, pause, audit rows) and the release-record checker, on the
CPU: the driver runs against an oracle-backed engine double."""
from __future__ import annotations

import json
import os
import signal
from copy import deepcopy

import pytest

import es_driver
import es_release
from conftest import GOLDEN, bfile
from support import oracle_engine as OE


WHEEL = [[2, 2], [3, 2], [5, 2]]     # k=30: M = 900, R = 240, g(30) = 58782 in block 65


def _run(tmp_path, **kw):
    os.makedirs(tmp_path, exist_ok=True)
    args = dict(k=30, ckpt_path=str(tmp_path / "k30.json"), tun={"test_ochunk": 2},
                wheel=WHEEL, max_blocks=None, timeout=None, log_every=1e9, ckpt_every=0.0,
                min_free_mib=0, resume=True)
    args.update(kw)
    return es_driver.run(OE, log=lambda s: None, **args)


def test_driver_finds_g_and_audits(tmp_path):
    res = _run(tmp_path)
    assert res["status"] == "FOUND" and int(res["g"]) == bfile()[30]
    plan = OE.Plan(30, {"test_ochunk": 2}, WHEEL)
    rec = es_release.run_record_from_driver(res, plan)
    assert es_release.check_run(30, int(res["g"]), rec) == res["block"] + 1


def test_resume_after_max_launches_reproduces_the_block(tmp_path):
    full = _run(tmp_path / "a")
    paused = _run(tmp_path / "b", max_launches=3)
    assert paused["status"] == "PAUSED" and paused["unit_done"] > 0
    resumed = _run(tmp_path / "b")
    assert resumed["status"] == "FOUND" and resumed["restarts"] == 1
    assert resumed["g"] == full["g"]
    assert resumed["block_audit"] == full["block_audit"]


def test_resume_after_block_boundary(tmp_path):
    """--max-blocks then resume (harness/gate_driver.py:56-66 semantics)."""
    full = _run(tmp_path / "a")
    assert full["block"] == 65
    first = _run(tmp_path / "b", max_blocks=10)
    assert first["status"] == "EXHAUSTED_BLOCKS"
    again = _run(tmp_path / "b")
    assert again["g"] == full["g"] and again["restarts"] == 1


def test_mismatched_checkpoint_is_refused(tmp_path):
    _run(tmp_path, max_launches=1)
    ck = json.load(open(tmp_path / "k30.json"))
    ck["M"] = "12345"
    json.dump(ck, open(tmp_path / "k30.json", "w"))
    with pytest.raises(RuntimeError, match="mismatch"):
        _run(tmp_path)


def test_sigterm_pauses_at_launch_boundary(tmp_path, monkeypatch):
    monkeypatch.setattr(es_driver, "STOP_REQUESTED", False)
    orig = OE.GpuSieve._launch

    def launch_then_signal(self, a, b):
        orig(self, a, b)
        es_driver.request_stop(signal.SIGTERM, None)
    monkeypatch.setattr(OE.GpuSieve, "_launch", launch_then_signal)
    res = _run(tmp_path)
    assert res["status"] == "PAUSED" and res["unit_done"] > 0
    monkeypatch.setattr(es_driver, "STOP_REQUESTED", False)


# ----------------------------------------------------------------- release checker

@pytest.fixture
def term397():
    terms = json.load(open(os.path.join(GOLDEN, "runs_frontier.json")))["terms"]
    return next(t for t in terms if t["k"] == 397)


def test_recorded_runs_pass():
    table = es_release.read_table(os.path.join(GOLDEN, "b003458_1to400.txt"), 400)
    res = es_release.check_runs_file(os.path.join(GOLDEN, "runs_frontier.json"), table)
    assert res["status"] == "PASS" and res["terms"] == 23 and res["complete_unique_block_records"] >= 100


def test_missing_completed_block_is_rejected(term397):
    run = deepcopy(term397["confirmation"])
    del run["block_audit"][1]
    with pytest.raises(ValueError, match="missing or repeated block"):
        es_release.check_run(397, int(term397["n"]), run)


def test_partial_residue_count_is_rejected(term397):
    run = deepcopy(term397["confirmation"])
    run["block_audit"][0]["residues_enumerated"] -= 1
    with pytest.raises(ValueError, match="incomplete residue count"):
        es_release.check_run(397, int(term397["n"]), run)


def test_changed_survivor_is_rejected(term397):
    run = deepcopy(term397["confirmation"])
    run["block_audit"][-1]["survivors_n"][0] = str(int(term397["n"]) + 1)
    with pytest.raises(ValueError, match="survivor hash mismatch"):
        es_release.check_run(397, int(term397["n"]), run)


def test_duplicate_bfile_index_is_rejected(tmp_path):
    path = tmp_path / "b.txt"
    path.write_text("1 3\n1 5\n")
    with pytest.raises(ValueError, match="missing, repeated, or unordered"):
        es_release.read_table(path, 2)


def test_timeout_saves_and_reports(tmp_path):
    res = _run(tmp_path, timeout=1e-9)
    assert res["status"] == "TIMEOUT" and os.path.exists(tmp_path / "k30.json")
    again = _run(tmp_path)
    assert again["status"] == "FOUND" and again["restarts"] == 1


def test_corrupt_checkpoint_is_refused(tmp_path):
    _run(tmp_path, max_launches=2)
    ck = json.load(open(tmp_path / "k30.json"))
    ck["survivor_idx"] = ck["survivor_idx"] + [10 ** 9]          # beyond the scanned prefix
    json.dump(ck, open(tmp_path / "k30.json", "w"))
    with pytest.raises(AssertionError, match="scanned prefix"):
        _run(tmp_path)


def test_source_revision_checked_against_a_tree(tmp_path, term397):
    """engine_revision hashes are verified only against a given source tree."""
    run = deepcopy(term397["primary"])
    src = tmp_path / "src"
    (src / "sieve").mkdir(parents=True)
    (src / "sieve" / "engine.py").write_text("kernel\n")
    import hashlib
    good = hashlib.sha256(b"kernel\n").hexdigest()
    run["engine_revision"] = {"sieve/engine.py": good}
    n = int(term397["n"])
    assert es_release.check_run(397, n, run, source_root=src) == len(run["block_audit"])
    run["engine_revision"] = {"sieve/engine.py": "0" * 64}
    assert es_release.check_run(397, n, run) == len(run["block_audit"])   # no tree: not checked
    with pytest.raises(ValueError, match="source revision mismatch"):
        es_release.check_run(397, n, run, source_root=src)


def test_wheel_facts_closed_form():
    """Residues per block = prod over wheel items of prod_{i<T} (p - a_i)
    (harness/verify_release.py:56-59), and an exponent beyond k's digit count is invalid."""
    f = es_release.WheelFacts.derive(30, [[2, 5], [3, 4], [5, 3], [7, 2], [11, 2], [13, 2]])
    assert f.modulus == 2**5 * 3**4 * 5**3 * 7**2 * 11**2 * 13**2
    assert f.residues == 230947200
    with pytest.raises(ValueError, match="invalid wheel"):
        es_release.WheelFacts.derive(30, [[2, 6]])
    with pytest.raises(ValueError, match="repeated wheel prime"):
        es_release.WheelFacts.derive(30, [[2, 1], [2, 2]])
