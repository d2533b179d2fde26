#!/usr/bin/env python3
"""Headline benchmark: cuckoo trials/s at n = 2^14, k = 3, bit-exact.

Workload: bgh25 insertion, n = 16384 items, m = 24576 single-slot buckets,
k = 3, stash 0, max 500 kicks, ell_i 16, master seed 1, batches of 2^20
trials.  A step runs the next --trials trials of that sequence (per-trial keys
are AES-derived from (seed, batch, offset), so inputs are generated on the
device).  Timing brackets exactly --steps steps with device synchronisation on
both sides.  The JSON line carries the roofline of the dominant kernel (the
insertion chains, timed with HIP events on the library's stream) and the CPU
oracle's rate on a bounded sample of the same workload, whose failure set is
also checked against the GPU's.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "tests"))

T_START = time.perf_counter()


def log(msg: str) -> None:
    """Progress line on stderr, stamped with the seconds since the process started."""
    print(f"[{time.perf_counter() - T_START:7.1f} s] {msg}", file=sys.stderr, flush=True)


METRIC = "cuckoo trials/sec at n=2^14,k=3 (bit-exact); sweep jobs/sec + p50/p99"
BT = 20
SEED = 1
SHAPE = dict(mode="bgh25", pnum=16384, ell=16, kick=500)
HBM_PEAK_GBS = 8000.0
# Integer VALU peak of MI355X: 256 CUs x 64 lanes x 2.4 GHz.
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
# Algorithmic integer work per unit, the constants used for the int-op model:
AES_OPS = 336          # one AES-128 block in T-table form
KEYSCHED_OPS = 160     # one AES-128 key schedule
SPLITMIX_OPS = 22      # one SplitMix64 draw (int32-equivalent)
CACHED_ROUND_OPS = 4   # one cached Feistel round (lookup index, xor, mask, swap)
EVAL_FIXED_OPS = 6     # element, y = k*e+w, cycle-walk test, division
PROBE_OPS = 20         # choice draw (div/mod by k), table read, compare, write, advance
# Algorithmic bytes of the insertion chain: a probe reads one u32 slot of the
# reference's table (ordinal | choice << 16) and writes at most one back, 8 B;
# a trial's table starts as buckets u32 empties, 4 B per bucket (written by the
# build, or skipped through pass tags, so not part of k_chain's launch).
ALGO_PROBE_BYTES = 8
ALGO_BUCKET_INIT_BYTES = 4
# Per table layout at the headline shape: the chain kernel, the bytes this
# build's own layout moves per probe (its slot read and written back) and per
# inserted item (the item's candidate word, read once from the insertion-order
# stream), and the profile-derived HBM bytes and DRAM requests
# (TCC_EA0_RDREQ + TCC_EA0_WRREQ) per probe from rocprofv3 --pmc.  The chain's
# reads are 64-byte requests, which FETCH_SIZE counts exactly.
# The figures are those of tables placed over the three memory regions (as the
# headline context places them); in one region a probe makes 1.876 / 1.864
# requests (residual / descriptor), in a two-region straddle 1.836 / 1.833.
CHAIN_PROFILES = {
    "desc-plain": dict(kernel="k_chain<3, bgh, u64 descriptor stream, u64 entries>", slot_bytes=16, item_bytes=8,
                       read=64.60, write=27.46, requests=1.0092 + 0.8577,
                       profile="profiles/k_chain_hbm_n16384_region_placement.md"),
    "residual": dict(kernel="k_chain_res<3, bgh, u64 candidate stream, u32 entries>", slot_bytes=8, item_bytes=8,
                     read=65.08, write=27.42, requests=1.0170 + 0.8560,
                     profile="profiles/k_chain_hbm_n16384_region_placement.md"),
}
# The n = 2^20 chain (44-bit residual entries, windowed): HBM bytes and DRAM
# requests per probe and the L2 hit rate, from its counter profile.
N20_PROFILE = dict(kernel="k_chain_res<3, bgh, u64 candidate stream, 44-bit entries>", read=57.4, write=30.4,
                   requests=0.897 + 0.885, l2_hit_rate=0.686, profile="profiles/k_chain_res_hbm_n2_20_counters.md")
# DRAM requests per step of ck_randbench (profiles/randbench_192k_traffic.md).
RB_READ_WRITE_REQ_PER_STEP = 0.998 + 1.0005
RB_READ_ONLY_REQ_PER_STEP = 0.998
RB_READ_ONLY_W4_REQ_PER_STEP = 0.9636  # four walks share a lane's table: a few more L2 hits
RB_READ_WRITE_W4_REQ_PER_STEP = 0.9989 + 0.9998
RB_READ_WRITE_7OF8_REQ_PER_STEP = 0.9977 + 0.8062  # seven steps in eight write back
# Engines of the sweep and serving legs take their arenas spread over the
# device's memory regions (--no-spread: one allocation).
SPREAD_ARENA = True


def int_ops_per_trial(p, probes_per_trial: float) -> float:
    """Modelled integer ops per trial: two AES blocks, a key schedule and two
    SplitMix64 draws of the derivation, the 8 * 2^half AES blocks of the round
    cache, and per bucket evaluation (E = P for bgh25) 8 cached Feistel rounds
    per cycle-walk pass plus the fixed part, plus the probe itself.  A pass
    lands in the domain with probability domain / 2^(2 * half_bits)."""
    accept = p.domain / float(1 << (2 * p.half_bits))
    cache = (8 << p.half_bits) * AES_OPS
    chain = probes_per_trial * (8 / accept * CACHED_ROUND_OPS + EVAL_FIXED_OPS) + probes_per_trial * PROBE_OPS
    return 2 * AES_OPS + KEYSCHED_OPS + 2 * SPLITMIX_OPS + cache + chain


def host_cores() -> int:
    """Every host core this process may run on (its affinity mask)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, n)


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_limits() -> dict:
    """What bounds the CPU legs on this host: logical CPUs, the affinity mask
    and the cgroup CPU quota (cgroup v2 cpu.max or v1 cfs quota / period)."""
    out = {"nproc": os.cpu_count()}
    try:
        out["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    for path in ("/sys/fs/cgroup/cpu.max", "/sys/fs/cgroup/cpu/cpu.cfs_quota_us"):
        try:
            text = Path(path).read_text().strip()
        except OSError:
            continue
        if path.endswith("cpu.max"):
            quota, _, period = text.partition(" ")
        else:
            quota = text
            try:
                period = Path("/sys/fs/cgroup/cpu/cpu.cfs_period_us").read_text().strip()
            except OSError:
                period = ""
        out["cgroup_cpu_quota"] = None if quota in ("max", "-1") or not period else int(quota) / int(period)
        break
    return out


def cpu_baseline(ctx, ck, p, seconds: float):
    """Oracle rate on a bounded prefix of the workload; also the parity check."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import ck_oracle as ock

    op = ock.parameters(SHAPE["mode"], SHAPE["pnum"], SHAPE["ell"], kick=SHAPE["kick"])
    cores = host_cores()
    bk = ock.batch_key(SEED, 0)
    t0 = time.perf_counter()
    ock.run_range(bk, 0, cores * 2, op, threads=cores)
    est = (cores * 2) / max(time.perf_counter() - t0, 1e-6)
    sample = int(max(cores * 4, min(est * seconds, 1 << 16)))
    t0 = time.perf_counter()
    of, _, ost = ock.run_range(bk, 0, sample, op, threads=cores)
    dt = time.perf_counter() - t0
    gf, _ = ctx.launch(p, ck.batch_key(SEED, 0), 0, sample)
    parity = list(gf) == list(of)
    one = max(4, sample // (cores * 2))
    t0 = time.perf_counter()
    ock.run_range(bk, 0, one, op, threads=1)
    one_rate = one / (time.perf_counter() - t0)
    g15, _ = ctx.launch(p, ck.batch_key(SEED, 0), 0, 1 << 15)
    return {
        "value": sample / dt, "unit": "trials/s", "cores": cores, "kind": "port",
        "sample": f"first {sample} trials of batch 0 (seed 1), oracle/ck_oracle.c with AES-NI round cache, "
                  f"{cores} threads on {cpu_model()}",
        "seconds": dt, "failures": len(of), "gpu_failures_equal": parity,
        "single_core_trials_per_s": one_rate, "host": host_limits(),
        "reference_cpu_bench": cpu_reference(cores, len(g15)),
    }, parity


def cpu_reference(cores: int, gpu_failures_prefix):
    """The reference's own cpu_bench (oracle/_ref, built from its sources) on the
    headline shape: first 2^15 trials of batch 0.  Its CLI derives the kick
    bound from the mode (3n = 49152 rather than 500); no insertion at this load
    comes near 500 kicks, so the work and the failures are the same."""
    import re
    import subprocess
    import tempfile
    exe = ROOT / "oracle" / "_ref" / "cpu_bench"
    if not exe.exists():
        return None
    with tempfile.TemporaryDirectory() as d:
        out = subprocess.run([str(exe), "--mode", "bgh25", "--pnum", str(SHAPE["pnum"]), "--ell-i", str(SHAPE["ell"]),
                              "--seed", str(SEED), "--et", "15", "--bt", "15", "--workers", str(cores), "--out", d],
                             capture_output=True, text=True, timeout=600)
    m = re.search(r"RESULT .*completed=(\d+) failed=(\d+) elapsed_s=([0-9.e+-]+) work_s=([0-9.e+-]+)", out.stdout)
    if out.returncode != 0 or not m:
        return {"error": out.stderr[-300:]}
    done, failed, work = int(m.group(1)), int(m.group(2)), float(m.group(4))
    return {"value": done / work, "unit": "trials/s", "cores": cores, "kind": "reference",
            "sample": f"reference cpu_bench (oracle/_ref) bgh25 pnum 16384 ell_i 16 seed 1, trials 0..{done - 1} "
                      f"of batch 0, kick 3n", "seconds": work, "failures": failed,
            "gpu_failures_equal": failed == gpu_failures_prefix}


def random_access_ceiling(probes_per_s: float, requests_per_probe, stream_requests_per_probe: float = 0.0):
    """Runs lib/ck_randbench: the chain's access pattern (each lane a dependent
    random read + write of one 8-byte slot of its own 192 KB table, with one
    XOR and one 32-bit multiply-high between a load and the next address)
    without the chain's logic, at 32K..256K lanes, in 16 candidate buffers and
    in one range spread over the device's memory regions (three pieces, as the
    library places the chains' tables).  The ceiling is the best random-read
    request rate measured, with one walk per lane or four independent walks per
    lane (latency out of the way).  The walks' requests are all random lines;
    frac sets against the ceiling the chain's random requests, its total less
    `stream_requests_per_probe` (the candidate stream, read in order, one
    64-byte line per eight inserted items)."""
    import re
    import subprocess
    from paths import PKG
    exe = PKG / "lib" / "ck_randbench"
    if not exe.exists():
        return None
    out = subprocess.run([str(exe), "--search", "16", "--regions", "196608", "16384", "32768", "65536", "131072",
                          "262144"], capture_output=True, text=True, timeout=300)
    forms = r"(read_write_7of8|read_write_w4|read_write|read_only_w4|read_only)"
    rows = re.findall(r"^" + forms + r" lanes=(\d+) .*steps_per_s=([0-9.e+]+)", out.stdout, re.M)
    found = re.findall(r"^search_" + forms + r" buffer=(\d+) .*steps_per_s=([0-9.e+]+)", out.stdout, re.M)
    spread = {form: float(rate) for form, rate in
              re.findall(r"^regions_" + forms + r" lanes=65536 .*steps_per_s=([0-9.e+]+)", out.stdout, re.M)}
    nreg = re.findall(r"^regions_found=(\d+)", out.stdout, re.M)
    if out.returncode != 0 or not rows:
        return {"error": out.stderr[-300:]}
    sweep = {}
    for form, lanes, rate in rows:
        sweep.setdefault(int(lanes), {})[form] = float(rate)
    search = {}
    for form, buf, rate in found:
        search.setdefault(int(buf), {})[form] = float(rate)
    best_lanes = max(sweep, key=lambda n: sweep[n].get("read_write", 0))
    # The best over the lane sweep and the 65,536-lane buffers of the search:
    # the chain's pass buffers are the fastest of its own search.
    rw = max([sweep[best_lanes]["read_write"]] + [v.get("read_write", 0) for v in search.values()] +
             [spread.get("read_write", 0)])
    rows_all = list(sweep.values()) + list(search.values()) + [spread]
    ro1 = max(v.get("read_only", 0) for v in rows_all)
    ro4 = max(v.get("read_only_w4", 0) for v in rows_all)
    ro = max(ro1, ro4)
    rw4 = max(v.get("read_write_w4", 0) for v in rows_all)
    rw78 = max(v.get("read_write_7of8", 0) for v in rows_all)
    chain_req = probes_per_s * requests_per_probe if requests_per_probe else None
    chain_rand = probes_per_s * (requests_per_probe - stream_requests_per_probe) if requests_per_probe else None
    # The most DRAM requests per second any of the walks got through.
    ceiling_req = max(ro1 * RB_READ_ONLY_REQ_PER_STEP, ro4 * RB_READ_ONLY_W4_REQ_PER_STEP,
                      rw * RB_READ_WRITE_REQ_PER_STEP, rw4 * RB_READ_WRITE_W4_REQ_PER_STEP,
                      rw78 * RB_READ_WRITE_7OF8_REQ_PER_STEP)
    return {"pattern": "lanes x 192 KB tables, dependent random 8-byte read (+ write) per step "
                       "(load -> xor -> mul_hi -> address)",
            "read_write_steps_per_s": rw, "best_lanes": best_lanes, "read_only_steps_per_s": ro,
            "read_only_one_walk_steps_per_s": ro1, "read_only_four_walks_steps_per_s": ro4,
            "read_write_four_walks_steps_per_s": rw4, "read_write_7of8_steps_per_s": rw78,
            "sweep": [{"lanes": n, **sweep[n]} for n in sorted(sweep)],
            "search_65536_lanes": [{"buffer": i, **search[i]} for i in sorted(search)],
            # One range spread over the memory regions, as the library places
            # the chains' tables.
            "regions_65536_lanes": {"regions": int(nreg[0]) if nreg else None, **spread},
            "chain_probes_per_s": probes_per_s,
            "chain_over_read_write_steps": probes_per_s / rw,
            # The bound is the DRAM random-request rate: the chain's requests per
            # second against the most any of the walks got through (read-only or
            # read + write, one or four walks per lane, any buffer or range).
            "chain_dram_requests_per_s": chain_req,
            "chain_stream_requests_per_probe": stream_requests_per_probe,
            "chain_random_requests_per_s": chain_rand,
            "ceiling_requests_per_s": ceiling_req,
            "read_only_requests_per_s": max(ro1 * RB_READ_ONLY_REQ_PER_STEP, ro4 * RB_READ_ONLY_W4_REQ_PER_STEP),
            "read_write_requests_per_s": rw * RB_READ_WRITE_REQ_PER_STEP,
            "frac": chain_rand / ceiling_req if chain_rand else None,
            "frac_all_requests": chain_req / ceiling_req if chain_req else None,
            "note": f"requests per probe / per step are profile-derived (TCC_EA0_RDREQ + WRREQ): chain "
                    f"{requests_per_probe:.4f}, read + write step 1.999 (one or four walks per lane), seven in "
                    f"eight written back 1.804, read-only 0.998 (one walk per lane) or 0.9636 (four). The walks' "
                    f"requests are all random lines; frac counts the chain's random requests only (its total less "
                    f"{stream_requests_per_probe:.4f} per probe for the candidate stream, read in order); "
                    f"frac_all_requests counts them all"}


def load_sweep():
    import importlib.util
    from paths import PKG
    spec = importlib.util.spec_from_file_location("ck_sweep", PKG / "sweep.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def pct(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    i = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[i]


def prewarm(ck, ctx, pts, ps):
    """Grows the context's workspace to the largest shape once, so that the
    timed per-point launches do not include device allocation."""
    big = max(range(len(pts)), key=lambda i: (ps[i].pnum + ps[i].buckets))
    ctx.launch(ps[big], ck.batch_key(pts[big]["seed"], 0), 0, min(pts[big]["trials"], 1 << 16))


def wait_all(eng, ids, label, t0):
    """Results of every job (they finish in the scheduler's order), with a
    progress line every 20 s."""
    res = [None] * len(ids)
    left, last = len(ids), time.perf_counter()
    while left:
        for n, i in enumerate(ids):
            if res[n] is None:
                r = eng.poll(i)
                if r is not None:
                    res[n] = r
                    left -= 1
        if left:
            time.sleep(0.005)
        if time.perf_counter() - last > 20:
            last = time.perf_counter()
            log(f"{label}: {len(ids) - left} of {len(ids)} jobs done at {last - t0:.1f} s")
    return res


def plain_loop(ck, pts, ps, trials_of):
    """One single-shape launch per point, in order (the per-point loop of
    tune_bench.py:104-151): wall time, per-point results, and the device time
    the launches spanned (first kernel start to last kernel end, per launch)."""
    ctx = ck.Context(0)
    prewarm(ck, ctx, pts, ps)
    t0 = time.perf_counter()
    out, span_ms, probes = [], 0.0, 0
    for pt, p in zip(pts, ps):
        f, st = ctx.launch(p, ck.batch_key(pt["seed"], 0), 0, trials_of(pt))
        out.append((len(f), st.probes))
        span_ms += st.span_ms
        probes += st.probes
    wall = time.perf_counter() - t0
    ctx.close()
    return out, wall, span_ms / 1e3, probes


def sweep_run(ck, pts, trials_of, plain_too, label):
    ps = [sw_params(ck, pt) for pt in pts]
    plain = plain_s = busy_s = plain_probes = None
    if plain_too:
        plain, plain_s, busy_s, plain_probes = plain_loop(ck, pts, ps, trials_of)
        log(f"{label}: per-point launches {plain_s:.2f} s (device busy {busy_s:.2f} s)")
    eng = ck.Engine(0, policy=ck.SCHED_SRPT, spread=SPREAD_ARENA)
    regions = eng.arena_regions()
    t0 = time.perf_counter()
    ids = [eng.submit(p, pt["seed"], trials_of(pt), bt=pt["bt"]) for pt, p in zip(pts, ps)]
    res = wait_all(eng, ids, label, t0)
    eng_s = time.perf_counter() - t0
    st = eng.stats()
    eng.close()
    log(f"{label}: engine {eng_s:.2f} s")
    equal = plain is None or all((r.failed, r.probes) == pl for r, pl in zip(res, plain))
    total = sum(trials_of(pt) for pt in pts)
    probes = sum(r.probes for r in res)
    return {
        "jobs_per_s": len(pts) / eng_s, "trials_per_s": total / eng_s,
        "plain_jobs_per_s": len(pts) / plain_s if plain_s else None,
        "plain_trials_per_s": total / plain_s if plain_s else None,
        "speedup_vs_plain": plain_s / eng_s if plain_s else None,
        "engine_s": eng_s, "plain_s": plain_s, "passes": st.passes,
        "device_build_ms": st.build_ms, "device_chain_ms": st.chain_ms,
        # The per-point loop's device occupancy: the fraction of its wall time
        # inside a launch's kernels, and its probe rate while there.
        "plain_device_busy_s": busy_s, "plain_device_busy_frac": busy_s / plain_s if plain_s else None,
        "plain_probes_per_busy_s": plain_probes / busy_s if busy_s else None,
        "failed": sum(r.failed for r in res), "probes": probes,
        "probes_per_s": probes / eng_s, "plain_probes_per_s": plain_probes / plain_s if plain_s else None,
        "results_equal_plain": equal,
        # The engine's arena: 1 GiB pieces per memory region (0, 1, 2, unclear).
        "arena_regions": list(regions),
    }, equal


def sw_params(ck, pt):
    return load_sweep().params_of(ck, pt)


SWEEP_DESC = ("points from seed 2 (n log-uniform [2^8,2^16], k 2..4, m/n U[1.2,1.6], stash 0..4, "
              "kick 100/500/1000, bgh25, ell_i 16)")


def sweep_leg(ck, points: int, trials: int, plain_too: bool = True):
    """Mixed sweep (seed 2): one engine run over all points against one
    single-shape launch per point, same trials; failure counts must agree."""
    pts = load_sweep().sweep_points(points, seed=2, trials=trials)
    rec, equal = sweep_run(ck, pts, lambda pt: trials, plain_too, "sweep")
    return {"workload": f"{points} {SWEEP_DESC}, {trials} trials each", **rec}, equal


def sweep_full_leg(ck, points: int, trials: int = 10**6):
    """The sweep at its stated size, 10^6 trials per point, on the first
    `points` points of the seed-2 sequence (an i.i.d. sample of the 512)."""
    pts = load_sweep().sweep_points(points, seed=2, trials=trials)
    rec, equal = sweep_run(ck, pts, lambda pt: trials, True, "sweep-full")
    return {"workload": f"first {points} of the 512 {SWEEP_DESC}, {trials} trials each", "points": points,
            **rec}, equal


def serve_leg(ck, rates, max_jobs: int, budget_lo: float, budget_hi: float, plain_jobs: int, window_s: float):
    """Poisson arrivals (seed 3) of sweep-shaped jobs: the engine (continuous
    batching, least-remaining-work first) over the arrivals of `window_s`
    seconds per rate (at least 200 and at most `max_jobs` jobs), and a plain
    loop that launches one job at a time in arrival order over the first
    `plain_jobs` of them.  Latency = completion - arrival."""
    sw = load_sweep()
    shapes_all = sw.sweep_points(max_jobs, seed=3)
    ps_all = [sw.params_of(ck, pt) for pt in shapes_all]
    out = []
    ok = True
    for rate in rates:
        jobs = max(1, min(max_jobs, max(200, int(rate * window_s))))
        shapes, ps = shapes_all[:jobs], ps_all[:jobs]
        arr = sw.arrivals(jobs, rate, seed=3, budget_lo=budget_lo, budget_hi=budget_hi)
        eng = ck.Engine(0, policy=ck.SCHED_SRPT, spread=SPREAD_ARENA)
        t0 = time.perf_counter()
        ids = []
        for (at, budget), pt, p in zip(arr, shapes, ps):
            d = at - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            ids.append(eng.submit(p, pt["seed"], budget))
        res = [eng.wait(i) for i in ids]
        eng.close()
        lat = [r.end_ms - r.submit_ms for r in res]
        span = (max(r.end_ms for r in res) - min(r.submit_ms for r in res)) / 1e3
        # Plain loop: one launch at a time, in arrival order.
        n_plain = min(plain_jobs, jobs)
        ctx = ck.Context(0)
        prewarm(ck, ctx, [dict(pt, trials=b) for pt, (_, b) in zip(shapes[:n_plain], arr[:n_plain])], ps[:n_plain])
        t0 = time.perf_counter()
        plat, pres = [], []
        for (at, budget), pt, p in zip(arr[:n_plain], shapes[:n_plain], ps[:n_plain]):
            d = at - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            f, st = ctx.launch(p, ck.batch_key(pt["seed"], 0), 0, budget)
            pres.append((len(f), st.probes))
            plat.append((time.perf_counter() - t0 - at) * 1e3)
        pspan = time.perf_counter() - t0
        ctx.close()
        equal = all((r.failed, r.probes) == q for r, q in zip(res[:n_plain], pres))
        ok = ok and equal
        out.append({"rate": rate, "jobs": jobs, "jobs_per_s": jobs / span, "p50_ms": pct(lat, 50),
                    "p99_ms": pct(lat, 99), "plain_jobs": n_plain, "plain_jobs_per_s": n_plain / pspan,
                    "plain_p50_ms": pct(plat, 50), "plain_p99_ms": pct(plat, 99), "results_equal_plain": equal})
        log(f"serve: rate {rate}: {jobs} jobs, p50 {pct(lat, 50):.1f} ms p99 {pct(lat, 99):.1f} ms "
              f"(plain, first {n_plain}: {pct(plat, 50):.1f} / {pct(plat, 99):.1f})")
    return {"workload": f"Poisson arrivals (seed 3), shapes as the sweep (seed 3), trial budgets log-uniform "
                        f"[{budget_lo:g}, {budget_hi:g}]", "rates": out}, ok


def serve_stated_leg(ck, rate: float, window_s: float, budget_lo: float = 1e5, budget_hi: float = 1e9):
    """Serving at the stated budgets (log-uniform over [1e5, 1e9], mean
    ~1.1e8 trials per job): jobs arrive at `rate` per second (seed 3) for
    `window_s` seconds; the engine serves them least-remaining-work first and
    the row reports what it finished inside the window.  The offered load is
    the submitted trials per second of the window against the trials per
    second the engine executed: far above 1, so the queue only grows and the
    latency of the finished jobs is that of the small ones SRPT lets through."""
    sw = load_sweep()
    n = int(rate * window_s * 1.5) + 16
    shapes = sw.sweep_points(n, seed=3)
    arr = [a for a in sw.arrivals(n, rate, seed=3, budget_lo=budget_lo, budget_hi=budget_hi) if a[0] < window_s]
    eng = ck.Engine(0, policy=ck.SCHED_SRPT, spread=SPREAD_ARENA)
    t0 = time.perf_counter()
    ids, done = [], {}
    for (at, budget), pt in zip(arr, shapes):
        d = at - (time.perf_counter() - t0)
        if d > 0:
            time.sleep(d)
        ids.append(eng.submit(sw.params_of(ck, pt), pt["seed"], budget))
    while time.perf_counter() - t0 < window_s:
        for i in ids:
            if i not in done:
                r = eng.poll(i)
                if r is not None:
                    done[i] = r
        time.sleep(0.01)
    for i in ids:
        if i not in done:
            r = eng.poll(i)
            if r is not None:
                done[i] = r
    st = eng.stats()
    elapsed = time.perf_counter() - t0
    eng.close()
    lat = [r.end_ms - r.submit_ms for r in done.values()]
    submitted_trials = sum(b for _, b in arr)
    executed = st.slots  # trial slots run in passes, finished jobs or not
    fin_budgets = sorted(b for (_, b), i in zip(arr, ids) if i in done)
    log(f"serve-stated: {len(done)} of {len(ids)} jobs finished in {elapsed:.1f} s")
    return {
        "workload": f"Poisson arrivals at {rate:g} jobs/s (seed 3) for {window_s:g} s, shapes as the sweep "
                    f"(seed 3), trial budgets log-uniform [{budget_lo:g}, {budget_hi:g}]",
        "rate": rate, "window_s": window_s, "jobs_submitted": len(ids), "jobs_finished": len(done),
        "p50_ms_finished": pct(lat, 50), "p99_ms_finished": pct(lat, 99),
        "largest_finished_budget": fin_budgets[-1] if fin_budgets else None,
        "jobs_per_s": len(done) / elapsed, "executed_trials_per_s": executed / elapsed,
        "offered_trials_per_s": submitted_trials / window_s,
        "offered_over_capacity": (submitted_trials / window_s) / (executed / elapsed) if executed else None,
    }


def config1_leg(ck):
    """The reference's CPU-runnable point: bgh25 n=1024 m=1536 k=3 stash 0 kick
    500, ell_i 16, seed 0, trials [0, 100000) of batch 0 -- GPU and CPU oracle
    (all host cores) on the whole range, failure lists compared."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import ck_oracle as ock
    n = 100000
    p = ck.params("bgh25", 1024, 16, kick=500)
    op = ock.parameters("bgh25", 1024, 16, kick=500)
    with ck.Context(0) as ctx:
        ctx.set_place_search(32)  # tables spread over the memory regions, placed by the untimed launch
        ctx.launch(p, ck.batch_key(0, 1), 0, n)  # untimed: allocates the workspace of this trial count
        t0 = time.perf_counter()
        f, st = ctx.launch(p, ck.batch_key(0, 0), 0, n)
        gdt = time.perf_counter() - t0
    cores = host_cores()
    t0 = time.perf_counter()
    of, _, _ = ock.run_range(ock.batch_key(0, 0), 0, n, op, threads=cores)
    cdt = time.perf_counter() - t0
    return {"workload": "bgh25 n=1024 m=1536 k=3 stash=0 kick=500 ell_i=16 seed=0, trials [0, 100000) of batch 0",
            "gpu_trials_per_s": n / gdt, "cpu_trials_per_s": n / cdt, "cpu_cores": cores, "failed": len(f),
            "failures_equal": list(f) == list(of)}, list(f) == list(of)


def tiers_leg(ck):
    """The on-chip lane tier (a trial per lane, its table in LDS) against the
    HBM tier on small tables, where it could win: per shape one pass of
    65,536 trials (no build/chain overlap, so each kernel runs alone; chain
    rate = probes / chain time) and 2^20 trials in overlapped passes (end to
    end).  Failure lists and probe totals of the two tiers must agree."""
    shapes = [("n=256 m=384 k=3", dict(pnum=256, buckets=384, k=3)),
              ("n=1024 m=1536 k=3", dict(pnum=1024, buckets=1536, k=3)),
              ("n=1024 m=1331 k=2", dict(pnum=1024, buckets=1331, k=2))]
    out, ok = [], True
    for name, kw in shapes:
        p = ck.params("bgh25", kw["pnum"], 16, k=kw["k"], buckets=kw["buckets"], kick=500)
        row = {"workload": f"bgh25 {name} kick 500 ell_i 16 seed 0"}
        res = {}
        for tier, tname in ((ck.TIER_HBM, "hbm"), (ck.TIER_LANE, "lane")):
            with ck.Context(0) as ctx:
                ctx.set_tier(tier)
                ctx.set_place_search(32)  # HBM tier: tables spread over the memory regions where placeable
                if tier == ck.TIER_LANE:
                    row["lane_lds_per_block"] = ctx.plan(p, 1 << 16).chain_lds_bytes
                ctx.set_chunk(1 << 16)
                ctx.launch(p, ck.batch_key(0, 1), 0, 1 << 16)  # untimed: workspace
                f1, st1 = ctx.launch(p, ck.batch_key(0, 0), 0, 1 << 16)
                ctx.set_chunk(0)
                ctx.launch(p, ck.batch_key(0, 1), 0, 1 << 20)
                t0 = time.perf_counter()
                f, st = ctx.launch(p, ck.batch_key(0, 0), 0, 1 << 20)
                dt = time.perf_counter() - t0
            res[tname] = (list(map(int, f)), st.probes)
            row[f"{tname}_one_pass_chain_gprobes_per_s"] = st1.probes / st1.chain_ms / 1e6 if st1.chain_ms else None
            row[f"{tname}_one_pass_build_ms"] = st1.build_ms
            row[f"{tname}_trials_per_s"] = (1 << 20) / dt
        row["lane_vs_hbm"] = row["lane_trials_per_s"] / row["hbm_trials_per_s"]
        row["results_equal"] = res["hbm"] == res["lane"]
        ok = ok and row["results_equal"]
        out.append(row)
        log(f"tiers: {name}: lane {row['lane_trials_per_s']:.0f} vs hbm {row['hbm_trials_per_s']:.0f} trials/s")
    return out, ok


def graph_leg(ck, p, steps: int, trials: int, place: int):
    """Graph mode against kernel-by-kernel launches on three launch sizes:
    the headline steps (2^20 trials), the n = 1024 point (1e5 trials per
    launch) and small n = 256 launches of 1,024 trials, where host launch work
    weighs most.  Each form runs in its own context, with the headline's table
    placement, after one untimed launch (which places the tables and builds
    the graph); the failure lists of every launch must agree."""
    cases = [("headline n=16384, 2^20 trials per launch", p, trials, min(steps, 6)),
             ("n=1024 m=1536 kick 500, 1e5 trials per launch", ck.params("bgh25", 1024, 16, kick=500), 100000, 20),
             ("n=256 m=384 kick 500, 1024 trials per launch", ck.params("bgh25", 256, 16, kick=500), 1024, 200)]
    out = []
    for name, shp, count, reps in cases:
        rate, fails = {}, {}
        for mode in ("kernels", "graph"):
            with ck.Context(0) as ctx:
                ctx.set_graph(mode == "graph")
                ctx.set_place_search(place)
                ctx.launch(shp, ck.batch_key(SEED, 0), 0, count)
                got = []
                t0 = time.perf_counter()
                for r in range(reps):
                    f, st = ctx.launch(shp, ck.batch_key(SEED, 1 + r), 0, count)
                    got.append(f)
                rate[mode] = reps * count / (time.perf_counter() - t0)
                fails[mode] = [list(map(int, f)) for f in got]
        out.append({"workload": name, "launches": reps, "kernel_by_kernel_trials_per_s": rate["kernels"],
                    "graph_trials_per_s": rate["graph"], "graph_vs_kernels": rate["graph"] / rate["kernels"],
                    "results_equal": fails["kernels"] == fails["graph"]})
        log(f"graph: {name}: {rate['graph']:.0f} vs {rate['kernels']:.0f} trials/s")
    return out


def hbm_leg(ck, trials: int, layout: int = 0, warm: int = 0):
    """Rare-event HBM shape: n = 2^20, m = 1.3n, k = 3, stash 2, kick 1000,
    ell_i 20, seed 4 -- the first `trials` trials of batch 0, after an untimed
    launch of `warm` trials of batch 1 (default: 2^16, more than one pass
    holds, so that the workspace, 95 % of the free memory, is allocated outside
    the timed region).  trials = 0 times two full passes of the size that
    launch used."""
    p = ck.params("bgh25", 1 << 20, 20, k=3, buckets=1363149, kick=1000, stash=2)
    ctx = ck.Context(0)
    ctx.set_layout(layout)
    _, ws = ctx.launch(p, ck.batch_key(4, 1), 0, warm or max(trials, 1 << 16))
    trials = trials or 2 * ws.chunk
    t0 = time.perf_counter()
    f, st = ctx.launch(p, ck.batch_key(4, 0), 0, trials)
    dt = time.perf_counter() - t0
    ctx.close()
    chain_s = st.chain_ms / 1e3
    passes = -(-trials // st.chunk)
    return {
        "workload": f"bgh25 n=2^20 m=1363149 k=3 stash=2 kick=1000 ell_i=20 seed=4, trials [0, {trials}) of batch 0",
        "trials_per_s": trials / dt, "probes_per_s": st.probes / dt, "probes_per_trial": st.probes / trials,
        "failed": len(f), "trials_per_pass": st.chunk, "passes": passes,
        # Each pass inserts its trials' items in windows (build + chain launches)
        # when memory caps the pass; launches / passes / 2 is the window count.
        "windows_per_pass": st.launches // (2 * passes),
        # Chains are latency-bound, so a pass takes about as long whatever its trial
        # count: a long run of full passes goes at trials_per_pass / pass time.
        "full_pass_trials_per_s": st.chunk * passes / dt,
        # Algorithmic bytes of the shape (8 per probe, 4 per bucket of each
        # trial's initial table) over the wall time, and the chain's share
        # (8 per probe) over the chains' device time.
        "algorithmic_gbs": (ALGO_PROBE_BYTES * st.probes + ALGO_BUCKET_INIT_BYTES * p.buckets * trials) / dt / 1e9,
        "chain_algorithmic_gbs": ALGO_PROBE_BYTES * st.probes / chain_s / 1e9 if chain_s > 0 else None,
        "device_ms": {"build": st.build_ms, "chain": st.chain_ms},
        # HBM traffic and L2 hit rate of the chains: per-probe figures of the
        # counter profile times this run's probes, over the chains' device time.
        "chain_hbm_gbs": (N20_PROFILE["read"] + N20_PROFILE["write"]) * st.probes / chain_s / 1e9 if chain_s > 0 else None,
        "chain_dram_requests_per_s": N20_PROFILE["requests"] * st.probes / chain_s if chain_s > 0 else None,
        "chain_l2_hit_rate": N20_PROFILE["l2_hit_rate"],
        "traffic_source": "profile-derived, not measured in this run: per-probe FETCH_SIZE + WRITE_SIZE, "
                          "TCC_EA0_RDREQ + WRREQ and TCC_HIT / (HIT + MISS) of " + N20_PROFILE["profile"],
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--trials", type=int, default=1 << BT, help="trials per step (power of two <= 2^20); "
                    "the default is one whole batch")
    ap.add_argument("--tier", choices=["auto", "hbm", "lds"], default="auto")
    ap.add_argument("--chunk", type=int, default=0, help="trials per device pass (0 = library default)")
    ap.add_argument("--layout", choices=["auto", "map", "desc", "map-xchg", "desc-plain", "residual"], default="auto")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sweep-points", type=int, default=512)
    ap.add_argument("--sweep-trials", type=int, default=1 << 15,
                    help="trials per sweep point (the full sweep is 10^6)")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--sweep-no-plain", action="store_true", help="skip the per-point baseline of the sweep")
    ap.add_argument("--sweep-full-points", type=int, default=16,
                    help="points of the sweep run at 10^6 trials each, engine and per-point (0: skip)")
    ap.add_argument("--serve-rates", default="10,100,1000")
    ap.add_argument("--serve-jobs", type=int, default=1000, help="at most this many arrivals per rate at the small "
                    "budgets")
    ap.add_argument("--serve-window", type=float, default=30.0,
                    help="seconds of arrivals per rate at the small budgets (at least 200 jobs)")
    ap.add_argument("--serve-plain-jobs", type=int, default=100,
                    help="arrivals per rate for the one-at-a-time baseline loop")
    ap.add_argument("--serve-budget", default="10,1e5", help="small trial budget range lo,hi (log-uniform)")
    ap.add_argument("--serve-stated-window", type=float, default=30.0,
                    help="seconds of arrivals at the stated budgets [1e5, 1e9] (0: skip)")
    ap.add_argument("--serve-stated-rate", type=float, default=10.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--hbm-trials", type=int, default=0,
                    help="trials of the n=2^20 leg (default 0: two full memory-capped passes)")
    ap.add_argument("--no-hbm", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="run the headline with the rule-based placement")
    ap.add_argument("--place-search", type=int, default=32,
                    help="candidate table regions the headline context times before keeping two (0: no search)")
    ap.add_argument("--no-graph", action="store_true", help="skip the graph-mode comparison")
    ap.add_argument("--no-tiers", action="store_true", help="skip the lane-tier comparison")
    ap.add_argument("--no-spread", action="store_true",
                    help="engines take their arena as one allocation, not spread over the memory regions")
    ap.add_argument("--tune-sample", type=int, default=1 << 17, help="trials per autotuner candidate")
    ap.add_argument("--hbm-only", action="store_true", help="run only the n=2^20 leg (for counter profiles)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="before warmup, copy a buffer of this many MiB once (calibrates HBM byte counters)")
    args = ap.parse_args()
    global SPREAD_ARENA
    SPREAD_ARENA = not args.no_spread
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path "
              "(trial batches would shard across GPUs by batch id)", file=sys.stderr)
        return 2
    t = args.trials
    if t <= 0 or t & (t - 1) or t > (1 << BT):
        print("--trials must be a power of two <= 2^20", file=sys.stderr)
        return 2

    import torch
    from paths import load_product

    ck = load_product()
    layout = {"auto": ck.LAYOUT_AUTO, "map": ck.LAYOUT_MAP, "desc": ck.LAYOUT_DESC, "map-xchg": ck.LAYOUT_MAP_XCHG,
              "desc-plain": ck.LAYOUT_DESC_PLAIN, "residual": ck.LAYOUT_RESIDUAL}[args.layout]
    if args.hbm_only:
        # Counter profiles: a 64-trial warm-up keeps the dispatches to profile few.
        print(json.dumps({"hbm_tier": hbm_leg(ck, args.hbm_trials or (1 << 15), layout, warm=64)}), flush=True)
        return 0
    p = ck.params(SHAPE["mode"], SHAPE["pnum"], SHAPE["ell"], kick=SHAPE["kick"])
    ctx = ck.Context(0)
    ctx.self_test()
    log("device context ready (AES known-answer test passed)")
    ctx.set_tier({"auto": ck.TIER_AUTO, "hbm": ck.TIER_HBM, "lds": ck.TIER_LDS}[args.tier])
    if args.chunk:
        ctx.set_chunk(args.chunk)
    ctx.set_layout(layout)
    # Tables spread over the device's memory regions (the chain rate belongs
    # to the memory its tables land in).
    ctx.set_place_search(args.place_search)
    # With every placement knob left to the library, the context's placement
    # (tier, table layout, trials per pass) is the autotuner's pick on a sample
    # of batch 0, measured here before the warmup steps: which layout is
    # fastest at this shape differs from box to box (the descriptor and the
    # u32 residual entries have each been measured 7-9 % ahead of the other).
    names = {0: "auto", 1: "map", 2: "desc", 3: "map-xchg", 4: "desc-plain", 5: "residual"}
    tuned = args.layout == "auto" and args.tier == "auto" and not args.chunk and not args.no_tune
    tune_rec = None
    if tuned:
        ents = ctx.tune(p, ck.batch_key(SEED, 0), args.tune_sample)
        tune_rec = {"sample_trials": args.tune_sample, "candidates": [
            {"tier": {1: "hbm", 2: "lds", 3: "lane"}.get(e["tier"], "auto"), "layout": names[e["layout"]],
             "chunk": e["chunk"], "trials_per_s": round(e["trials_per_s"])} for e in ents]}
        best = max(ents, key=lambda e: e["trials_per_s"])
        log(f"autotune: best tier {best['tier']} layout {names[best['layout']]} chunk {best['chunk']} "
            f"{best['trials_per_s']:.0f} trials/s")

    if args.copy_probe_mib:
        n = args.copy_probe_mib << 18  # float32 elements
        src = torch.ones(n, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log(f"copy probe: {args.copy_probe_mib} MiB read + written")

    # The table layout the library plans for a step (auto picks by shape).
    layout_name = {ck.LAYOUT_MAP: "map", ck.LAYOUT_DESC: "desc", ck.LAYOUT_MAP_XCHG: "map-xchg",
                   ck.LAYOUT_DESC_PLAIN: "desc-plain", ck.LAYOUT_RESIDUAL: "residual"}.get(ctx.plan(p, t).layout, "?")
    keys = {}

    def step(i):
        g = i * t
        batch, first = g >> BT, g & ((1 << BT) - 1)
        if batch not in keys:
            keys[batch] = ck.batch_key(SEED, batch)
        return ctx.launch(p, keys[batch], first, t)

    for i in range(args.warmup):
        step(i)
        log(f"warmup step {i} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    agg = dict(failed=0, probes=0, build_ms=0.0, chain_ms=0.0, launches=0, chunk=0, tier=0)
    for i in range(args.warmup, args.warmup + args.steps):
        f, st = step(i)
        agg["failed"] += len(f)
        agg["probes"] += st.probes
        agg["build_ms"] += st.build_ms
        agg["chain_ms"] += st.chain_ms
        agg["launches"] += st.launches
        agg["chunk"], agg["tier"] = st.chunk, st.tier
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    log(f"timed {args.steps} steps in {elapsed:.3f} s")

    trials = args.steps * t
    value = trials / elapsed
    chain_launches = max(1, agg["launches"] // 2)  # build + chain per pass
    chain_avg_s = agg["chain_ms"] / 1e3 / chain_launches
    probes_per_launch = agg["probes"] / chain_launches
    bytes_per_launch = ALGO_PROBE_BYTES * probes_per_launch
    achieved = bytes_per_launch / chain_avg_s / 1e9 if chain_avg_s > 0 else 0.0
    prof = CHAIN_PROFILES.get(layout_name)
    own_bytes_per_launch = ((prof["slot_bytes"] * agg["probes"] + prof["item_bytes"] * p.pnum * trials) / chain_launches
                            if prof else None)
    traffic = ((prof["read"] + prof["write"]) * probes_per_launch if prof and prof["read"] is not None else None)
    probes_per_trial = agg["probes"] / trials
    ops = int_ops_per_trial(p, probes_per_trial)
    rec = {
        "metric": METRIC,
        "value": value,
        "unit": "trials/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic: per-trial keys AES-derived on device from (seed 1, batch, offset)",
        "config": {"workload": "bgh25 n=16384 m=24576 k=3 stash=0 kick=500 ell_i=16 seed=1 BT=20 sequential",
                   "trials_per_step": t, "tier": {1: "hbm", 2: "lds", 3: "lane"}.get(agg["tier"], "auto"),
                   "layout": args.layout, "layout_used": layout_name,
                   "placement": "autotuned before the warmup (autotune)" if tuned else "as requested",
                   "trials_per_pass": agg["chunk"]},
        "insertions_per_s": value * p.pnum,
        "probes_per_s": agg["probes"] / elapsed,
        "probes_per_trial": probes_per_trial,
        "failed": agg["failed"],
        "roofline": {
            "kernel": (prof["kernel"] if prof else layout_name) + " (insertion chains, one lane per trial)",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "avg_launch_ms": chain_avg_s * 1e3, "probes_per_launch": probes_per_launch,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "achieved_note": "8 algorithmic bytes per probe: one u32 slot of the reference's table read and at "
                             "most one written",
            "own_slot_bytes_per_launch": own_bytes_per_launch,
            "own_slot_gbs": own_bytes_per_launch / chain_avg_s / 1e9 if own_bytes_per_launch and chain_avg_s > 0
            else None,
            "own_slot_note": (f"this build's {layout_name} layout: {prof['slot_bytes']} B per probe (its slot read "
                              f"and written back) and {prof['item_bytes']} B per inserted item (its candidate word)"
                              if prof else None),
            "traffic_source": ("profile-derived, not measured in this run: (FETCH_SIZE + WRITE_SIZE) per probe from "
                               f"rocprofv3 --pmc ({prof['profile']}) times this run's probes per launch"
                               if traffic else None),
            "traffic_rate_gbs": traffic / chain_avg_s / 1e9 if traffic and chain_avg_s > 0 else None,
        },
        "int_ops_model": {
            "ops_per_trial": ops, "achieved_tops": ops * value / 1e12, "peak_tops": VALU_PEAK_TOPS,
            "frac": ops * value / 1e12 / VALU_PEAK_TOPS,
        },
        "device_ms": {"build": agg["build_ms"], "chain": agg["chain_ms"]},
    }
    # The candidate stream: item_bytes per inserted item, read in order, one
    # 64-byte line per 64 / item_bytes items.
    stream_req = (p.pnum / probes_per_trial) * prof["item_bytes"] / 64 if prof and probes_per_trial else 0.0
    rec["roofline"]["random_access_ceiling"] = random_access_ceiling(probes_per_launch / chain_avg_s,
                                                                     prof["requests"] if prof else None, stream_req)
    # The table placement of this context: by region (pieces per region, the
    # regions of each pass buffer's three pieces, walk ms of each candidate
    # table range, the two kept) or by search (walk ms per candidate, the two
    # kept).
    ms, kept = ctx.place_report()
    regions, b0, b1 = ctx.place_regions()
    if regions:
        rec["config"]["placement_regions"] = {
            "pieces": len(regions), "per_region": [regions.count(r) for r in range(3)],
            "buffer0_regions": [regions[i] for i in b0], "buffer1_regions": [regions[i] for i in b1],
            "candidate_walk_ms": [round(x, 2) for x in ms], "kept": list(kept)}
    else:
        rec["config"]["placement_search"] = {"candidate_walk_ms": [round(x, 2) for x in ms], "kept": list(kept)}
    ok = True
    if not args.no_cpu:
        rec["cpu_baseline"], ok = cpu_baseline(ctx, ck, p, args.cpu_seconds)
        log(f"cpu baseline: {rec['cpu_baseline']['value']:.0f} trials/s on {rec['cpu_baseline']['cores']} cores")
    ctx.close()
    if tune_rec:
        rec["autotune"] = tune_rec
    if not args.no_cpu:
        rec["config1"], good = config1_leg(ck)
        log(f"config 1: {rec['config1']['gpu_trials_per_s']:.0f} trials/s")
        ok = ok and good
    if not args.no_tiers:
        rec["lane_tier"], good = tiers_leg(ck)
        ok = ok and good
    if not args.no_hbm:
        rec["hbm_tier"] = hbm_leg(ck, args.hbm_trials, layout)
        log(f"hbm leg: {rec['hbm_tier']['trials_per_s']:.1f} trials/s")
    if not args.no_sweep:
        rec["sweep"], good = sweep_leg(ck, args.sweep_points, args.sweep_trials, not args.sweep_no_plain)
        ok = ok and good
        if args.sweep_full_points:
            rec["sweep_full"], good = sweep_full_leg(ck, args.sweep_full_points)
            ok = ok and good
            rec["config"]["sweep_full_points"] = args.sweep_full_points
    if not args.no_serve:
        lo, hi = (float(x) for x in args.serve_budget.split(","))
        rec["serving"], good = serve_leg(ck, [float(r) for r in args.serve_rates.split(",")], args.serve_jobs, lo, hi,
                                         args.serve_plain_jobs, args.serve_window)
        ok = ok and good
        if args.serve_stated_window > 0:
            rec["serving"]["stated_budgets"] = serve_stated_leg(ck, args.serve_stated_rate, args.serve_stated_window)
    if not args.no_graph:
        # Last: a graph-mode context leaves its workspace allocated until the
        # process ends, which would shrink the passes of any later leg.
        rec["graph"] = graph_leg(ck, p, args.steps, t, args.place_search)
        ok = ok and all(g["results_equal"] for g in rec["graph"])
    print(json.dumps(rec), flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
