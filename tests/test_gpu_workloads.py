"""The two multi-job workloads against the oracle, job by job.

Mixed sweep (seed 2): at least 32 of its 512 points chosen to span every
octave of n from 2^8 to 2^16, every k, every table-layout group and the
stash / kick values, run as one engine mix and as one single-shape launch per
point (the per-point loop of tune_bench.py:104-151 that the engine replaces).

Serving (seed 3): sweep-shaped jobs with the seed-3 shapes, submitted to a
running engine at their Poisson arrival times (100 jobs/s), with the stated
trial budgets -- log-uniform over [1e5, 1e9] -- scaled by 2e-6 so that the
oracle can check every job (budgets 1 .. 2000, the same draws).  Each job's
failing trials and probe total must equal the oracle's, i.e. what the batch
controller commits for it alone (cuckoo_bench.cu:724-745), and later jobs must
have been admitted while earlier ones were on the device."""
import importlib.util
import math
import time
from concurrent.futures import ThreadPoolExecutor

import pytest

import ck_oracle as ock
from paths import PKG, load_product

pytestmark = pytest.mark.gpu

ck = load_product()
GiB = 1 << 30


def load_sweep():
    spec = importlib.util.spec_from_file_location("ck_sweep", PKG / "sweep.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


sw = load_sweep()


def index_bits(buckets: int) -> int:
    bb = 1
    while bb < 32 and (1 << bb) < buckets:
        bb += 1
    return bb


def layout_group(pt) -> str:
    """The engine's table-layout group of a point: packed descriptors when
    k * bits + 2 <= 63, residual entries over a u64 candidate stream when
    k * bits <= 64, else the u32 bucket map."""
    bb = index_bits(pt["buckets"])
    if pt["k"] * bb + 2 <= 63:
        return "desc"
    if pt["k"] * bb <= 64:
        return "residual"
    return "map"


def covering_points(count=512, seed=2, want=36):
    pts = sw.sweep_points(count, seed=seed)
    chosen, seen = [], set()

    def take(pt):
        if pt["index"] not in seen:
            seen.add(pt["index"])
            chosen.append(pt)

    for octave in range(8, 16):
        for k in (2, 3, 4):
            hit = [pt for pt in pts if k == pt["k"] and min(15, int(math.log2(pt["pnum"]))) == octave]
            if hit:
                take(hit[0])
    for group in ("desc", "residual", "map"):
        hit = [pt for pt in pts if layout_group(pt) == group]
        if hit and group not in {layout_group(p) for p in chosen}:
            take(hit[0])
    combos = {(p["stash"], p["kick"]) for p in chosen}
    for pt in pts:
        if len(chosen) >= want:
            break
        if (pt["stash"], pt["kick"]) not in combos:
            combos.add((pt["stash"], pt["kick"]))
            take(pt)
    largest = max(pts, key=lambda p: p["pnum"])
    take(largest)
    for pt in pts:  # then the sweep's own order
        if len(chosen) >= want:
            break
        take(pt)
    return chosen


def oracle_job(pt, trials, threads=4):
    op = ock.parameters(pt["mode"], pt["pnum"], pt["ell_i"], k=pt["k"], buckets=pt["buckets"], kick=pt["kick"],
                        stash=pt["stash"])
    f, _, st = ock.run_range(ock.batch_key(pt["seed"], 0), 0, trials, op, threads=threads)
    return [int(x) for x in f], st.probes


def test_sweep_points_cover_every_octave_and_group():
    pts = covering_points()
    assert len(pts) >= 32
    octaves = {min(15, int(math.log2(p["pnum"]))) for p in pts}
    assert octaves == set(range(8, 16))
    assert {p["k"] for p in pts} == {2, 3, 4}
    assert {layout_group(p) for p in pts} == {"desc", "residual", "map"}
    assert max(p["pnum"] for p in pts) > 60000

    trials = [128 if p["pnum"] <= 16384 else 48 for p in pts]
    with ThreadPoolExecutor(8) as pool:
        want = list(pool.map(lambda a: oracle_job(*a), zip(pts, trials)))
    with ck.Engine(0, arena_bytes=16 * GiB, max_slots=8192) as eng:
        ids = [eng.submit(sw.params_of(ck, p), p["seed"], n, bt=p["bt"], keep_failures=True)
               for p, n in zip(pts, trials)]
        probes = [eng.wait(i).probes for i in ids]
        got = [(list(map(int, eng.failures(i))), pr) for i, pr in zip(ids, probes)]
    with ck.Context(0) as ctx:
        for p, n, (gf, gp), (of, oprobes) in zip(pts, trials, got, want):
            assert gf == of and gp == oprobes, (p, "engine")
            f, st = ctx.launch(sw.params_of(ck, p), ck.batch_key(p["seed"], 0), 0, n)
            assert list(map(int, f)) == of and st.probes == oprobes, (p, "launch")


SERVE_JOBS, SERVE_RATE, BUDGET_SCALE = 64, 100.0, 2e-6


def test_serving_poisson_arrivals_seed3():
    shapes = sw.sweep_points(SERVE_JOBS, seed=3)
    arrivals = sw.arrivals(SERVE_JOBS, SERVE_RATE, seed=3, budget_lo=1e5, budget_hi=1e9)
    budgets = [max(1, round(b * BUDGET_SCALE)) for _, b in arrivals]
    assert max(budgets) > 1000 and min(budgets) < 10  # the scaled draws still span the range
    with ck.Engine(0, arena_bytes=32 * GiB, max_slots=16384, buffers=4, policy=ck.SCHED_SRPT) as eng:
        t0 = time.perf_counter()
        ids = []
        for (at, _), pt, n in zip(arrivals, shapes, budgets):
            delay = at - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            ids.append(eng.submit(sw.params_of(ck, pt), pt["seed"], n, bt=pt["bt"], keep_failures=True))
        res = [eng.wait(i) for i in ids]
        fails = [list(map(int, eng.failures(i))) for i in ids]
    # Continuous batching: jobs arrived while earlier ones were on the device,
    # and some of them started before an earlier-submitted job finished.
    overlapped = sum(1 for j in range(1, len(res)) if any(res[j].submit_ms < res[i].end_ms for i in range(j)))
    jumped = sum(1 for j in range(1, len(res)) if any(res[j].start_ms < res[i].end_ms for i in range(j)))
    assert overlapped >= SERVE_JOBS // 4 and jumped >= 1, (overlapped, jumped)
    with ThreadPoolExecutor(8) as pool:
        want = list(pool.map(lambda a: oracle_job(a[0], a[1], threads=4 if a[1] >= 64 else 1),
                             zip(shapes, budgets)))
    for pt, n, r, f, (of, oprobes) in zip(shapes, budgets, res, fails, want):
        assert r.trials == n and f == of and r.failed == len(of) and r.probes == oprobes, (pt, n)
