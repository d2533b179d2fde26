#!/usr/bin/env python3
"""Writes tests/golden/headline_batch0.json: the outcome of every trial of
batch 0 (2^20 trials) at the headline shape -- bgh25, n = 16384, m = 24576,
k = 3, stash 0, ell_i 16, sequential input, seed 1, BT 20.

- `reference`: the failing offsets that the reference's own cpu_bench
  (oracle/_ref/cpu_bench, built from its sources by `make -C oracle ref`)
  writes to failures.csv for `--mode bgh25 --pnum 16384 --ell-i 16 --seed 1
  --et 20 --bt 20`.  Its kick bound is derived from the mode: 3n = 49152.
- `kick500`: the failing offsets and the total probe count of the oracle
  (oracle/ck_oracle.c) over the same trials with the kick bound 500.

The two failure lists are equal when no insertion of the batch reaches 500
evictions.  Run in a container that holds the reference build; takes a few
minutes on 8 cores.
"""
from __future__ import annotations

import json
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
import ck_oracle as ock  # noqa: E402

SEED, PNUM, ELL, COUNT = 1, 16384, 16, 1 << 20


def reference_failures():
    exe = ROOT / "oracle" / "_ref" / "cpu_bench"
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([str(exe), "--mode", "bgh25", "--pnum", str(PNUM), "--ell-i", str(ELL), "--seed", str(SEED),
                        "--et", "20", "--bt", "20", "--workers", "8", "--out", d],
                       check=True, capture_output=True, text=True)
        summary = json.loads((Path(d) / "summary.json").read_text())
        assert summary["status"] == "complete" and summary["completed"] == COUNT, summary
        rows = (Path(d) / "failures.csv").read_text().split()  # "batch_id,offset" header, then rows
    return sorted(int(r.split(",")[1]) for r in rows[1:]), summary["MAX_KICK_TIMES"]


def main():
    ref, ref_kick = reference_failures()
    op = ock.parameters("bgh25", PNUM, ELL, kick=500)
    f, _, st = ock.run_range(ock.batch_key(SEED, 0), 0, COUNT, op, threads=8)
    out = {
        "shape": {"mode": "bgh25", "pnum": PNUM, "ell_i": ELL, "buckets": op.buckets, "k": 3, "stash": 0,
                  "input_mode": 0, "seed": SEED, "batch": 0, "count": COUNT},
        "reference": {"kick": ref_kick, "failures": ref, "source": "oracle/_ref/cpu_bench failures.csv"},
        "kick500": {"kick": 500, "failures": [int(x) for x in f], "probes": int(st.probes),
                    "source": "oracle/ck_oracle.c run_range"},
    }
    path = HERE / "headline_batch0.json"
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {path}: reference {len(ref)} failures, kick 500 {len(f)} failures, {st.probes} probes")


if __name__ == "__main__":
    main()
