"""End-to-end runs of the cuckoo_bench command line on the GPU (the
counterpart of the reference's smoke_test.py): golden failure sets through the
journals, several workers, a mid-run PAUSE and resume with other workers and
micro size, a replay of a complete directory, and journals the reference's own
cpu_bench accepts (oracle/_ref, when present)."""
import json
import subprocess
import time
from pathlib import Path

import pytest

from paths import PKG, ROOT

pytestmark = pytest.mark.gpu

CLI = PKG / "lib" / "cuckoo_bench"
REF_CPU_BENCH = ROOT / "oracle" / "_ref" / "cpu_bench"
SMOKE = ["--et", "13", "--bt", "8", "--pnum", "16", "--ell-i", "8", "--seed", "123"]


def run(args, timeout=120):
    out = subprocess.run([str(CLI), *args], capture_output=True, text=True, timeout=timeout)
    assert out.returncode == 0, out.stderr[-2000:]
    return out


def failures(d: Path):
    lines = (d / "failures.csv").read_text().splitlines()
    assert lines[0] == "batch_id,offset"
    rows = [tuple(map(int, x.split(","))) for x in lines[1:]]
    assert len(rows) == len(set(rows))
    return sorted(b * 256 + o for b, o in rows)


def summary(d: Path):
    return json.loads((d / "summary.json").read_text())


@pytest.mark.parametrize("mode", ["bgh25", "drrt18"])
def test_cli_golden_and_workers(tmp_path, fixtures, mode):
    want = fixtures["smoke_expected"]["failures"][mode]
    for name, extra in [("baseline", ["--workers", "1", "--micro", "257"]),
                        ("multi", ["--workers", "2", "--micro", "37", "--tier", "hbm"]),
                        ("lds", ["--workers", "1", "--micro", "100", "--tier", "lds"]),
                        ("lane", ["--workers", "1", "--micro", "100", "--tier", "lane"]),
                        ("tuned", ["--workers", "1", "--micro", "256", "--tune"]),
                        ("graph", ["--workers", "2", "--micro", "100", "--graph"])]:
        d = tmp_path / name
        out = run(["--mode", mode, *SMOKE, *extra, "--stop-after", "8192", "--out", str(d)])
        assert "RESULT status=complete" in out.stdout
        assert ("tuned=" in out.stderr) == (name == "tuned")
        assert ("graph=on" in out.stderr) == (name == "graph")
        s = summary(d)
        assert s["status"] == "complete" and s["completed"] == 8192 and s["completed_batches"] == 32
        assert s["failed"] == len(want)
        assert failures(d) == want, name
        if name == "lane":
            assert s["table_layout_planned"] == "cached-lane-shared"


@pytest.mark.parametrize("mode", ["bgh25", "drrt18"])
def test_cli_pause_and_resume(tmp_path, fixtures, mode):
    d = tmp_path / "resumed"
    d.mkdir()
    cmd = [str(CLI), "--mode", mode, *SMOKE, "--workers", "1", "--micro", "1", "--stop-after", "8192",
           "--out", str(d)]
    proc = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    try:
        deadline = time.monotonic() + 120
        while proc.poll() is None:
            # a 56-byte v3 header plus one complete empty record (24 bytes)
            if any(p.stat().st_size >= 80 for p in d.glob("worker_*.bin")):
                (d / "PAUSE").touch()
                break
            assert time.monotonic() < deadline
            time.sleep(0.005)
        proc.wait(timeout=120)
    finally:
        if proc.poll() is None:
            proc.kill()
            proc.wait()
    assert proc.returncode == 0
    s = summary(d)
    if s["status"] == "paused":
        assert 0 < s["completed"] < 8192 and s["completed"] % 256 == 0
    (d / "PAUSE").unlink(missing_ok=True)
    run(["--mode", mode, *SMOKE, "--workers", "2", "--micro", "113", "--stop-after", "8192", "--out", str(d)])
    s = summary(d)
    assert s["status"] == "complete" and s["completed"] == 8192
    want = fixtures["smoke_expected"]["failures"][mode]
    assert failures(d) == want
    # Replaying a complete directory performs no work and changes nothing.
    before = {p.name: p.stat().st_size for p in d.glob("worker_*.bin")}
    run(["--mode", mode, *SMOKE, "--workers", "1", "--micro", "19", "--stop-after", "8192", "--out", str(d)])
    assert {p.name: p.stat().st_size for p in d.glob("worker_*.bin")} == before
    assert failures(d) == want
    if REF_CPU_BENCH.exists():
        ref = subprocess.run([str(REF_CPU_BENCH), "--mode", mode, *SMOKE, "--status", "--out", str(d)],
                             capture_output=True, text=True, timeout=60)
        assert ref.returncode == 0, ref.stderr
        assert json.loads((d / "summary.json").read_text())["completed"] == 8192


def test_cli_extension_journal(tmp_path):
    # Non-power-of-two plan, k = 4, stash 1: v4 journal, resume to completion.
    d = tmp_path / "ext"
    args = ["--mode", "bgh25", "--pnum", "200", "--ell-i", "12", "--k", "4", "--buckets", "230", "--kick", "5",
            "--stash", "1", "--trials", "1000", "--bt", "8", "--seed", "9", "--out", str(d)]
    run(args)
    s = summary(d)
    assert s["status"] == "complete" and s["completed"] == 1000 and s["journal_version"] == 4
    assert s["K"] == 4 and s["STASH"] == 1 and s["planned"] == 1000
    import sys
    sys.path.insert(0, str(ROOT / "oracle"))
    import ck_oracle as ock
    op = ock.parameters("bgh25", 200, 12, k=4, buckets=230, kick=5, stash=1)
    want = []
    for b in range(4):
        n = min(256, 1000 - b * 256)
        f, _, _ = ock.run_range(ock.batch_key(9, b), 0, n, op)
        want += [b * 256 + int(x) for x in f]
    assert failures(d) == want


def test_cli_checkpoint_resume_n2_20(tmp_path):
    """The rare-event HBM shape (n = 2^20, m = 1.3n, k = 3, stash 2, kick 1000,
    seed 4) through the journal: a run paused after its first committed batch
    and resumed with two workers ends with the same summary and failures as an
    uninterrupted run."""
    shape = ["--mode", "bgh25", "--pnum", str(1 << 20), "--ell-i", "20", "--k", "3", "--buckets", "1363149",
             "--kick", "1000", "--stash", "2", "--seed", "4", "--trials", "1536", "--bt", "9", "--micro", "512"]
    whole = tmp_path / "whole"
    run([*shape, "--workers", "1", "--out", str(whole)], timeout=300)
    d = tmp_path / "resumed"
    d.mkdir()
    proc = subprocess.Popen([str(CLI), *shape, "--workers", "1", "--out", str(d)], stdout=subprocess.PIPE,
                            stderr=subprocess.PIPE, text=True)
    try:
        deadline = time.monotonic() + 300
        while proc.poll() is None:
            if any(p.stat().st_size > 64 for p in d.glob("worker_*.bin")):
                (d / "PAUSE").touch()
                break
            assert time.monotonic() < deadline
            time.sleep(0.01)
        proc.wait(timeout=300)
    finally:
        if proc.poll() is None:
            proc.kill()
            proc.wait()
    assert proc.returncode == 0
    first = summary(d)
    (d / "PAUSE").unlink(missing_ok=True)
    run([*shape, "--workers", "2", "--out", str(d)], timeout=300)
    a, b = summary(whole), summary(d)
    assert a["status"] == b["status"] == "complete"
    assert a["completed"] == b["completed"] == 1536 and a["journal_version"] == 4
    assert a["failed"] == b["failed"]
    assert (whole / "failures.csv").read_text().splitlines()[1:] == (d / "failures.csv").read_text().splitlines()[1:]
    if first["status"] == "paused":
        assert 0 < first["completed"] < 1536 and first["completed"] % 512 == 0


# The reference harness's preflight cases (compare_cpu_gpu.py:27-58): name,
# --prp-cache, --micro, --block-threads, --table-layout, --warps-per-block,
# --shared-variant, and the table= value the Worker line must show.
PREFLIGHT = [
    ("gpu_direct", "off", 257, 128, "global", 0, "auto", "direct-global-residual"),
    ("gpu_cached128", "on", 257, 128, "global", 0, "auto", "cached-global-residual"),
    ("gpu_shared_round", "on", 257, 256, "shared", 0, "round", "shared-trial"),
    ("gpu_shared_map", "on", 257, 256, "shared", 0, "map", "shared-map"),
    ("gpu_cached_tail", "on", 37, 256, "global", 0, "auto", "cached-global-residual"),
    ("gpu_shared_tail", "on", 5, 256, "shared", 2, "map", "shared-map"),
]


@pytest.mark.parametrize("mode", ["bgh25", "drrt18"])
def test_cli_reference_harness_keys(tmp_path, fixtures, mode):
    """Each preflight case of the reference's harness: the summary keys it
    checks (with backend 'hip', the one key a harness run against this build
    expects differently), the golden failures, and the Worker-line fields it
    parses (micro=, table=, warps=, shared_bytes=, scratch_bytes=)."""
    import re
    want = fixtures["smoke_expected"]["failures"][mode]
    for name, cache, micro, block, layout, warps, variant, table in PREFLIGHT:
        d = tmp_path / name
        args = ["--mode", mode, *SMOKE, "--input-mode", "sequential", "--workers", "1", "--micro", str(micro),
                "--scratch-mib", "64", "--prp-cache", cache, "--block-threads", str(block), "--table-layout", layout,
                "--shared-variant", variant, "--out", str(d)]
        if warps:
            args += ["--warps-per-block", str(warps)]
        out = run(args)
        s = summary(d)
        expected = dict(status="complete", completed=8192, planned=8192, ET=13, BT=8, mode=mode, PNUM_T=16, ELL_I=8,
                        seed_decimal="123", input_mode="sequential", seed_scheme="aes-batch-v1", backend="hip",
                        prp_cache_enabled=cache == "on", table_layout_requested=layout,
                        shared_variant_requested=variant, completed_batches=32, pending_batches=0, workers=1,
                        scratch_mib=64)
        for key, value in expected.items():
            assert s.get(key) == value, (name, key, s.get(key))
        assert s["table_layout_planned"] == table, name
        assert s["work_elapsed_seconds"] > 0
        assert failures(d) == want and s["failed"] == len(want), name
        text = out.stderr
        # trials per launch: the requested micro, at most one batch (2^8 here)
        assert [int(x) for x in re.findall(r"Worker .*? micro=(\d+)", text)] == [min(micro, 256)], name
        assert re.findall(r"table=(\S+)", text)[:1] == [table], name
        assert len(re.findall(r"\bwarps=(\d+)", text)) == 1
        assert int(re.findall(r"shared_bytes=(\d+)", text)[0]) > 0 if layout == "shared" else True
        scratch = int(re.findall(r"scratch_bytes=(\d+)", text)[0])
        assert 0 < scratch <= 64 << 20, (name, scratch)
