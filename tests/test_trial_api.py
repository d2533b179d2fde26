"""The device-level trial API (include/cuckoo_trial.hpp) against the
reference's own single-trial outcomes and final-table sha256 digests
(cpu_trial_cached, cpu_trials.hpp:128-149, mode-derived buckets and kick):
insert_core over caller-side Bins (StridedBins u32, SharedBins u16) and
Buckets (CachedBuckets over a round cache from ck_round_cache, and a
PrecomputedBuckets map), with per-trial inputs from ck_trial_inputs_of.

The CPU test runs the templates on the host (lib/ck_trial_api_tests); the GPU
test runs the same templates in a HIP kernel (tests/hip/trial_api_kernels.hip)."""
import hashlib
import subprocess

import numpy as np
import pytest

from paths import PKG

TOOL = PKG / "lib" / "ck_trial_api_tests"


def fixture_rows(fixtures):
    return fixtures["reference_trials"]


def rows_text(rows):
    return "".join(f"{r['mode']} {r['input_mode']} {r['pnum']} {r['ell_i']} {r['seed']} {r['batch']} {r['offset']}\n"
                   for r in rows)


def parse_blocks(words: np.ndarray):
    """[(combo, failed, probes, table)] from the tool's u32 stream."""
    out, at = [], 0
    while at < len(words):
        combo, failed, probes, n = (int(x) for x in words[at:at + 4])
        out.append((combo, failed, probes, words[at + 4:at + 4 + n]))
        at += 4 + n
    return out


def check_against_reference(rows, blocks_per_row):
    for row, blocks in zip(rows, blocks_per_row):
        combos = {c for c, _, _, _ in blocks}
        assert combos == ({0, 1, 2, 3} if row["pnum"] < (1 << 14) else {0, 1}), row
        probes = {p for _, _, p, _ in blocks}
        assert len(probes) == 1, (row, probes)  # every placement makes the same probes
        for combo, failed, _, table in blocks:
            assert failed == row["failed"], (row, combo)
            digest = hashlib.sha256(table.astype("<u4").tobytes()).hexdigest()
            assert digest == row["table_sha256"], (row, combo)


def test_trial_api_on_host(tmp_path, fixtures):
    rows = fixture_rows(fixtures)
    (tmp_path / "rows.txt").write_text(rows_text(rows))
    out = subprocess.run([str(TOOL), str(tmp_path / "rows.txt"), str(tmp_path / "out.bin")],
                         capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr
    blocks = parse_blocks(np.fromfile(tmp_path / "out.bin", dtype="<u4"))
    per_row, at = [], 0
    for row in rows:
        n = 4 if row["pnum"] < (1 << 14) else 2
        per_row.append(blocks[at:at + n])
        at += n
    assert at == len(blocks)
    check_against_reference(rows, per_row)


@pytest.mark.gpu
def test_trial_api_in_a_kernel(fixtures):
    """The same templates in a HIP kernel (tests/hip/trial_api_kernels.hip),
    fed with host-derived inputs (ck_trial_inputs_of, ck_round_cache) and the
    library's own device bucket map (ck_trial_map); the host derivation must
    equal the device's, and the probe counts the library's own chains'."""
    import ctypes as C
    from collections import defaultdict

    from paths import ROOT, load_product

    ck = load_product()
    lib = C.CDLL(str(ROOT / "tests" / "hip" / "lib" / "libtrial_api_kernels.so"))
    fn = lib.trial_api_device_run
    P = C.POINTER
    fn.restype = C.c_int
    fn.argtypes = [C.c_uint32] * 7 + [C.c_uint64] * 4 + [C.c_uint32, P(C.c_uint64), P(C.c_uint16), P(C.c_uint32),
                                                         P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
    groups = defaultdict(list)
    for r in fixture_rows(fixtures):
        groups[(r["mode"], r["input_mode"], r["pnum"], r["ell_i"])].append(r)
    with ck.Context(0) as ctx:
        for (mode, input_mode, pnum, ell), rows in groups.items():
            p = ck.params(mode, pnum, ell, input_mode=input_mode)
            n = len(rows)
            inputs = np.zeros((n, 4), dtype=np.uint64)
            rounds = np.zeros((n, 8 << p.half_bits), dtype=np.uint16)
            maps = np.zeros((n, pnum * 3), dtype=np.uint32)
            lib_probes = np.zeros(n, dtype=np.uint32)
            for j, r in enumerate(rows):
                rk = ck.batch_key(r["seed"], r["batch"])
                ti = ck.trial_inputs(p, rk, r["offset"])
                inputs[j] = (ti.walk, ti.randomness, ti.shift, ti.step)
                rounds[j] = ck.round_cache(p, list(ti.prp_key))
                key, walk, rnd, m = ctx.trial_map(p, rk, r["offset"])
                assert list(key) == list(ti.prp_key) and (walk, rnd) == (ti.walk, ti.randomness), r
                maps[j] = m
                _, pr, _, _ = ctx.launch_debug(p, rk, r["offset"], 1)
                lib_probes[j] = pr[0]
            per_row = [[] for _ in rows]
            for combo in (0, 1, 2, 3) if pnum < (1 << 14) else (0, 1):
                tables = np.zeros((n, p.buckets), dtype=np.uint32)
                failed = np.zeros(n, dtype=np.uint32)
                probes = np.zeros(n, dtype=np.uint32)
                rc = fn(combo, 1 if mode == 1 else 0, pnum, p.kick, p.buckets, p.half_bits, 1 if input_mode == 0 else 0,
                        p.domain, p.bucket_size, p.mask, p.range_mask, n,
                        inputs.ctypes.data_as(P(C.c_uint64)), rounds.ctypes.data_as(P(C.c_uint16)),
                        maps.ctypes.data_as(P(C.c_uint32)), tables.ctypes.data_as(P(C.c_uint32)),
                        failed.ctypes.data_as(P(C.c_uint32)), probes.ctypes.data_as(P(C.c_uint32)))
                assert rc == 0, rc
                assert (probes == lib_probes).all(), (mode, pnum, combo)
                for j in range(n):
                    per_row[j].append((combo, int(failed[j]), int(probes[j]), tables[j]))
            check_against_reference(rows, per_row)
