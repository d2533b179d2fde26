#!/usr/bin/env python3
"""Writes tests/golden/reference_fixtures.json.

Run in a container that holds the reference sources and the reference build
(`make -C oracle ref`).  The fixture is data only: failure-offset lists and
known answers taken from the reference's own data files, plus outcomes and
table digests computed by the reference's own host functions through
oracle/_ref/libref_shim.so.  Nothing of the reference's source text is stored.

Sources:
  cpu_golden.hpp:2-4         32768-trial failure lists (seed 123, pnum 16, ell_i 8, BT 8)
  smoke_expected.json        the ET 13 prefix of the same
  host_tests.cpp:37-46       batch-key vectors host_aes128(batch_key(123,id),0,0,0,0)
  cpu_tests.cpp:30-39        affine vectors (seed 123, pnum 3, ell_i 8, BT 20, batch 0)
  host_tests.cpp:158-159     bucket counts at pnum 16 / ell_i 8
  cuckoo_bench.cu:332-344    FIPS-197 AES-128 known answer
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import re
import sys
from pathlib import Path

REF = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference")
HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
import ck_oracle as ock  # noqa: E402


def parse_golden(text: str, name: str):
    m = re.search(name + r"\s*=\s*\{([^}]*)\}", text)
    return [int(x) for x in m.group(1).split(",") if x.strip()]


def table_digest(bins) -> str:
    return hashlib.sha256(b"".join(int(v).to_bytes(4, "little") for v in bins)).hexdigest()


def main():
    shim = ock.ref_shim()
    if shim is None:
        raise SystemExit("build the reference first: make -C oracle ref")
    golden_text = (REF / "cpu_golden.hpp").read_text()
    smoke = json.loads((REF / "smoke_expected.json").read_text())
    fx = {
        "golden_32768": {
            "seed": 123, "pnum": 16, "ell_i": 8, "et": 15, "bt": 8, "input_mode": "sequential",
            "failures": {"bgh25": parse_golden(golden_text, "golden_bgh25"),
                         "drrt18": parse_golden(golden_text, "golden_drrt18")},
            "source": "cpu_golden.hpp:2-4",
        },
        "smoke_expected": dict(smoke, source="smoke_expected.json"),
        "fips197": {"key_lo": 0x0706050403020100, "key_hi": 0x0F0E0D0C0B0A0908,
                    "plain": [0x00112233, 0x44556677, 0x8899AABB, 0xCCDDEEFF],
                    "cipher": [0x69C4E0D8, 0x6A7B0430, 0xD8CDB780, 0x70B4C55A],
                    "source": "cuckoo_bench.cu:341, cpu_trials.hpp:153"},
        "batch_key_vectors": {"seed": 123,
                              "cases": [{"batch": 0, "aes_zero": [0x992DC0B2, 0xECC207F0, 0x5F9A7017, 0x1FB84A2F]},
                                        {"batch": 7, "aes_zero": [0x58F85B99, 0xFC3CBD02, 0x022BF19C, 0x10DAD53D]}],
                              "source": "host_tests.cpp:37-46"},
        "affine_vectors": {"seed": 123, "pnum": 3, "ell_i": 8, "bt": 20, "batch": 0, "count": 256,
                           "failures": {"bgh25": [4, 174, 197], "drrt18": [166, 227]},
                           "source": "cpu_tests.cpp:30-39"},
        "bucket_counts": {"pnum": 16, "ell_i": 8, "bgh25": 32, "drrt18": 23, "source": "host_tests.cpp:158-159"},
    }

    # Reference parameters() over a pnum sweep (exercises the float log2 quirk).
    params = []
    for mode in (1, 2):
        for pnum in (3, 4, 5, 16, 29, 30, 31, 100, 1000, 1024, 1493, 4096, 10000, 16000, 16383, 16384,
                     32768, 65534):
            for ell_i in (8, 14, 16, 20, 32, 61):
                if ell_i < 16 and (1 << ell_i) < pnum:
                    continue
                b, k = C.c_uint32(), C.c_uint32()
                bs, d = C.c_uint64(), C.c_uint64()
                h, e = C.c_uint32(), C.c_uint32()
                rc = shim.ref_parameters(mode, pnum, ell_i, C.byref(b), C.byref(k), C.byref(bs),
                                         C.byref(d), C.byref(h), C.byref(e))
                if rc:
                    continue
                params.append({"mode": mode, "pnum": pnum, "ell_i": ell_i, "buckets": b.value, "kick": k.value,
                               "bucket_size": bs.value, "domain": d.value, "half_bits": h.value,
                               "ell_d": e.value})
    fx["reference_parameters"] = params

    # Per-trial outcomes and final tables from cpu_trial_cached (cpu_trials.hpp:128-149),
    # entry = ordinal | choice<<16, sha256 over little-endian u32 entries.
    trials = []
    points = [(1, 0, 16, 8, 123, 7), (2, 0, 16, 8, 123, 7), (1, 1, 16, 8, 123, 7), (2, 1, 16, 8, 123, 7),
              (1, 0, 1024, 16, 0, 0), (2, 0, 1024, 16, 0, 0), (1, 1, 1024, 16, 5, 3), (2, 1, 1024, 16, 5, 3),
              (1, 0, 4096, 16, 123, 0), (2, 1, 4096, 16, 123, 1),
              (1, 0, 16384, 16, 1, 0), (2, 0, 16384, 16, 1, 0), (1, 1, 16000, 14, 9, 2)]
    for mode, im, pnum, ell_i, seed, batch in points:
        b = C.c_uint32()
        dummy32, dummy64 = C.c_uint32(), C.c_uint64()
        shim.ref_parameters(mode, pnum, ell_i, C.byref(b), C.byref(dummy32), C.byref(dummy64), C.byref(dummy64),
                            C.byref(dummy32), C.byref(dummy32))
        offsets = list(range(6)) + [255, 1000]
        for off in offsets:
            bins = (C.c_uint32 * b.value)()
            failed = shim.ref_trial(mode, im, pnum, ell_i, seed, batch, off, 1, bins)
            trials.append({"mode": mode, "input_mode": im, "pnum": pnum, "ell_i": ell_i, "seed": seed,
                           "batch": batch, "offset": off, "failed": failed, "table_sha256": table_digest(bins)})
    fx["reference_trials"] = trials

    # Failure lists over ranges with frequent failures (cpu_trial_cached decisions).
    ranges = []
    for mode, im, pnum, ell_i, seed, batch, count in [(2, 0, 1024, 16, 0, 0, 2048), (1, 0, 30, 8, 77, 1, 4096),
                                                      (2, 1, 64, 12, 5, 2, 4096), (1, 1, 5, 8, 7, 0, 8192),
                                                      (1, 0, 4, 4, 7, 3, 8192), (2, 1, 8, 6, 7, 0, 8192),
                                                      (2, 0, 32, 8, 9, 5, 8192)]:
        fails = [off for off in range(count) if shim.ref_trial(mode, im, pnum, ell_i, seed, batch, off, 1, None)]
        ranges.append({"mode": mode, "input_mode": im, "pnum": pnum, "ell_i": ell_i, "seed": seed,
                       "batch": batch, "count": count, "failures": fails})
    fx["reference_ranges"] = ranges

    out = HERE / "reference_fixtures.json"
    out.write_text(json.dumps(fx, indent=1) + "\n")
    print(f"wrote {out} ({out.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
