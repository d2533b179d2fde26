"""CPU tests of the run controller and the cuckoo_bench command line.

The controller's C++ tests run through tests/cpp/test_run_controller.cpp (a
fake device, no GPU).  The CLI is exercised in --status mode, which reads journals
without touching a device: a journal written by the reference's own cpu_bench
(oracle/_ref, when built) must resume here and export the golden failures."""
import json
import subprocess
from pathlib import Path

import pytest

from paths import PKG, ROOT

HOST_TESTS = PKG / "lib" / "ck_run_tests"
CLI = PKG / "lib" / "cuckoo_bench"
REF_CPU_BENCH = ROOT / "oracle" / "_ref" / "cpu_bench"

SMOKE = ["--et", "13", "--bt", "8", "--pnum", "16", "--ell-i", "8", "--seed", "123"]


def test_controller_unit_tests(tmp_path):
    out = subprocess.run([str(HOST_TESTS), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr + out.stdout
    assert "0 failed" in out.stdout and "FAIL" not in out.stdout


def read_csv(d: Path):
    lines = (d / "failures.csv").read_text().splitlines()
    assert lines[0] == "batch_id,offset"
    return [tuple(map(int, x.split(","))) for x in lines[1:]]


def test_cli_status_on_empty_directory(tmp_path):
    out = subprocess.run([str(CLI), "--mode", "bgh25", *SMOKE, "--out", str(tmp_path), "--status"],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    s = json.loads((tmp_path / "summary.json").read_text())
    assert s["status"] == "ready" and s["completed"] == 0 and s["planned"] == 8192
    assert s["BUCKETS_NUM"] == 32 and s["MAX_KICK_TIMES"] == 48 and s["seed_decimal"] == "123"
    assert read_csv(tmp_path) == []


def test_cli_rejects_bad_options(tmp_path):
    for bad in (["--et", "41"], ["--pnum", "2"], ["--mode", "x"], ["--k", "5"], ["--scratch-mib", "0"]):
        out = subprocess.run([str(CLI), *bad, "--out", str(tmp_path), "--status"], capture_output=True,
                             text=True, timeout=60)
        assert out.returncode == 1 and "FATAL" in out.stderr, bad


@pytest.mark.skipif(not REF_CPU_BENCH.exists(), reason="oracle/_ref not built")
@pytest.mark.parametrize("mode", ["bgh25", "drrt18"])
def test_reads_reference_journal(tmp_path, fixtures, mode):
    d = tmp_path / mode
    ref = subprocess.run([str(REF_CPU_BENCH), "--mode", mode, *SMOKE, "--workers", "2", "--out", str(d)],
                         capture_output=True, text=True, timeout=300)
    assert ref.returncode == 0, ref.stderr
    ours = subprocess.run([str(CLI), "--mode", mode, *SMOKE, "--out", str(d), "--status"],
                          capture_output=True, text=True, timeout=60)
    assert ours.returncode == 0, ours.stderr
    s = json.loads((d / "summary.json").read_text())
    assert s["status"] == "complete" and s["completed"] == 8192 and s["completed_batches"] == 32
    want = fixtures["smoke_expected"]["failures"][mode]
    assert sorted(b * 256 + o for b, o in read_csv(d)) == want
    # A different seed is a different experiment: resume must refuse the journal.
    bad = subprocess.run([str(CLI), "--mode", mode, *SMOKE[:-1], "124", "--out", str(d), "--status"],
                         capture_output=True, text=True, timeout=60)
    assert bad.returncode == 1 and "different experiment" in bad.stderr and "seed" in bad.stderr


@pytest.mark.parametrize("extra,enabled,planned", [
    ([], True, "cached-global-residual"),
    (["--prp-cache", "off"], False, "direct-global-residual"),
    (["--prp-cache", "on", "--table-layout", "global"], True, "cached-global-residual"),
    (["--table-layout", "shared", "--shared-variant", "map"], True, "shared-map"),
    (["--table-layout", "shared", "--shared-variant", "round"], True, "shared-trial"),
    (["--tier", "lds"], True, "shared-map"),
])
def test_cli_summary_reports_the_placement(tmp_path, extra, enabled, planned):
    """prp_cache_enabled and table_layout_planned follow the placement flags
    (the reference writes round_cache_enabled(o,p), cuckoo_bench.cu:545-553,762,
    and compare_cpu_gpu.py:55 checks prp_cache_enabled == (--prp-cache == 'on'))."""
    out = subprocess.run([str(CLI), "--mode", "bgh25", *SMOKE, *extra, "--out", str(tmp_path), "--status"],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    s = json.loads((tmp_path / "summary.json").read_text())
    assert s["prp_cache_enabled"] is enabled and s["table_layout_planned"] == planned
    assert s["backend"] == "hip" and s["scratch_mib"] == 0


def test_cli_rejects_impossible_placements(tmp_path):
    # ell_i 32 puts half_bits at 17: no round cache fits, so --prp-cache on is refused.
    out = subprocess.run([str(CLI), "--mode", "bgh25", "--pnum", "16384", "--ell-i", "32", "--et", "4", "--bt", "4",
                          "--prp-cache", "on", "--out", str(tmp_path), "--status"],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "round cache" in out.stderr
