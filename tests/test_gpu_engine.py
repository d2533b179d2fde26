"""The batched / served engine through the C ABI: every job's failing trials
and probe total equal the oracle's (and the single-shape launch's), whatever
the mix, the pass size, the buffer count or the scheduling policy; jobs that
arrive while others run are admitted into later passes and get the results
they get alone."""
import time

import numpy as np
import pytest

import ck_oracle as ock
from paths import load_product

pytestmark = pytest.mark.gpu

ck = load_product()
GiB = 1 << 30

# (mode, pnum, ell_i, extension kwargs, seed, trials, bt)
JOBS = [
    ("bgh25", 100, 12, dict(kick=50), 11, 1000, 20),
    ("bgh25", 257, 16, dict(k=2, buckets=310, kick=100, stash=2), 12, 777, 20),
    ("drrt18", 500, 16, dict(k=4, buckets=600, kick=100, stash=1, input_mode=1), 13, 640, 20),
    ("bgh25", 2000, 16, dict(buckets=2400, kick=500), 14, 300, 20),
    ("bgh25", 1000, 16, dict(k=4, buckets=1200, kick=1000, stash=4), 15, 321, 20),
    ("bgh25", 16, 32, dict(), 16, 200, 20),                          # half 17: direct-AES group
    ("bgh25", 3000, 16, dict(buckets=3900, kick=100), 17, 500, 6),    # 64-trial batches
    ("drrt18", 30000, 16, dict(k=4, buckets=40000, kick=100), 18, 70, 20),  # 4*16 + 2 > 63: bucket-map group
    ("bgh25", 40000, 16, dict(k=4, buckets=50000, kick=200, stash=1), 21, 40, 20),
    ("drrt18", 16, 8, dict(), 19, 999, 8),
    ("bgh25", 64, 10, dict(k=3, buckets=70, kick=30, stash=3, input_mode=1), 20, 1500, 9),
]


def oracle_job(mode, pnum, ell, kw, seed, trials, bt):
    op = ock.parameters(mode, pnum, ell, **kw)
    cached = op.half_bits <= 16
    fails, probes = [], 0
    per = 1 << bt
    for batch in range((trials + per - 1) // per):
        n = min(per, trials - batch * per)
        f, _, st = ock.run_range(ock.batch_key(seed, batch), 0, n, op, cached=cached)
        fails += [batch * per + int(x) for x in f]
        probes += st.probes
    return fails, probes


@pytest.fixture(scope="module")
def expected():
    return [oracle_job(*j) for j in JOBS]


def run_jobs(eng, jobs):
    ids = [eng.submit(ck.params(m, n, ell, **kw), seed, trials, bt=bt, keep_failures=True)
           for m, n, ell, kw, seed, trials, bt in jobs]
    out = []
    for jid, j in zip(ids, jobs):
        r = eng.wait(jid)
        out.append((list(map(int, eng.failures(jid))), r))
    return out


CONFIGS = [
    dict(policy=ck.SCHED_BALANCE, buffers=3, max_slots=2048),
    dict(policy=ck.SCHED_FIFO, buffers=2, max_slots=128),
    dict(policy=ck.SCHED_SRPT, buffers=4, max_slots=640, window=3),
]


@pytest.mark.parametrize("cfg", CONFIGS, ids=["balance", "fifo-small-passes", "srpt-window"])
def test_engine_matches_oracle(expected, cfg):
    with ck.Engine(0, arena_bytes=2 * GiB, **cfg) as eng:
        got = run_jobs(eng, JOBS)
        st = eng.stats()
    for (fails, r), (ofails, oprobes), j in zip(got, expected, JOBS):
        assert fails == ofails, j
        assert r.failed == len(ofails) and r.trials == j[5], j
        assert r.probes == oprobes, j
        assert r.start_ms >= r.submit_ms and r.end_ms >= r.start_ms
    assert st.jobs_done == len(JOBS)
    assert st.trials == sum(j[5] for j in JOBS)
    assert st.passes >= 2


def test_engine_golden_32768(fixtures):
    """The reference's golden failure lists (seed 123, pnum 16, ell_i 8,
    2^15 trials in batches of 2^8) as two jobs of one engine."""
    g = fixtures["golden_32768"]
    with ck.Engine(0, arena_bytes=GiB, max_slots=4096) as eng:
        ids = {m: eng.submit(ck.params(m, g["pnum"], g["ell_i"]), g["seed"], 1 << g["et"], bt=g["bt"],
                             keep_failures=True) for m in ("bgh25", "drrt18")}
        for m, jid in ids.items():
            eng.wait(jid)
            assert list(map(int, eng.failures(jid))) == g["failures"][m]


def test_engine_matches_launch_at_n16384():
    """Engine and single-shape launch agree on failures and probe totals at the
    headline shape and two near-threshold variants."""
    shapes = [dict(kick=500), dict(buckets=19661, kick=500), dict(k=2, buckets=20000, kick=100, stash=2)]
    trials = 1000
    with ck.Context(0) as ctx, ck.Engine(0, arena_bytes=8 * GiB, max_slots=1024) as eng:
        ids = []
        for i, kw in enumerate(shapes):
            p = ck.params("bgh25", 16384, 16, **kw)
            ids.append((eng.submit(p, 100 + i, trials, keep_failures=True), p, 100 + i))
        for jid, p, seed in ids:
            r = eng.wait(jid)
            f, st = ctx.launch(p, ck.batch_key(seed, 0), 0, trials)
            assert list(map(int, eng.failures(jid))) == list(map(int, f))
            assert r.probes == st.probes and r.failed == len(f)


def test_serving_results_alone_and_in_mix():
    """Continuous batching: jobs submitted while a long job is on the device
    start before it ends, and every result equals the job's solo launch."""
    long = (ck.params("bgh25", 16384, 16, kick=500), 41, 24576)
    short = [(ck.params("bgh25", 300, 16, k=2, buckets=400, kick=100, stash=1), 42, 3000),
             (ck.params("drrt18", 1000, 16, k=4, buckets=1250, kick=100), 43, 2000),
             (ck.params("bgh25", 5000, 16, buckets=6000, kick=500, stash=2), 44, 1500)]
    with ck.Engine(0, arena_bytes=16 * GiB, max_slots=4096, buffers=3, policy=ck.SCHED_SRPT) as eng:
        lid = eng.submit(*long, keep_failures=True)
        sids = []
        for p, seed, n in short:
            time.sleep(0.02)
            sids.append(eng.submit(p, seed, n, keep_failures=True))
        lres = eng.wait(lid)
        sres = [eng.wait(s) for s in sids]
        got = [(list(map(int, eng.failures(i))), r.probes) for i, r in zip([lid] + sids, [lres] + sres)]
    for r in sres:
        assert r.start_ms < lres.end_ms  # admitted while the long job was on the device
    with ck.Context(0) as ctx:
        for (p, seed, n), (fails, probes) in zip([long] + short, got):
            f, st = ctx.launch(p, ck.batch_key(seed, 0), 0, n)
            assert fails == list(map(int, f))
            assert probes == st.probes


def test_engine_rejects_bad_jobs():
    with ck.Engine(0, arena_bytes=64 << 20, max_slots=1024) as eng:
        p = ck.params("bgh25", 100, 12)
        with pytest.raises(ck.CkError):
            eng.submit(p, 1, 100, bt=5)
        with pytest.raises(ck.CkError):
            eng.submit(p, 1, 0)
        with pytest.raises(ck.CkError):  # 64 trials need more than one pass buffer
            eng.submit(ck.params("bgh25", 1 << 20, 20, buckets=1363149, kick=1000, stash=2), 1, 10)
        jid = eng.submit(p, 1, 100)
        eng.wait(jid)
        with pytest.raises(ck.CkError):
            eng.failures(jid)  # submitted without keep_failures
        with pytest.raises(ck.CkError):
            eng.wait(jid + 1000)
        eng.release(jid)
        with pytest.raises(ck.CkError):
            eng.poll(jid)


def test_sweep_points_run_bit_exact():
    """Eight points of the mixed sweep (seed 2), 256 trials each, against the
    oracle: every group the sweep produces (k 2..4, both table layouts)."""
    import importlib.util
    from paths import PKG
    spec = importlib.util.spec_from_file_location("ck_sweep", PKG / "sweep.py")
    sw = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(sw)
    pts = [pt for pt in sw.sweep_points(64, seed=2, trials=256) if pt["pnum"] <= 8192][:8]
    with ck.Engine(0, arena_bytes=4 * GiB, max_slots=1024) as eng:
        ids = [eng.submit(sw.params_of(ck, pt), pt["seed"], pt["trials"], bt=pt["bt"], keep_failures=True)
               for pt in pts]
        res = [(eng.wait(i), list(map(int, eng.failures(i)))) for i in ids]
    for pt, (r, fails) in zip(pts, res):
        kw = dict(k=pt["k"], buckets=pt["buckets"], kick=pt["kick"], stash=pt["stash"])
        of, op = oracle_job(pt["mode"], pt["pnum"], pt["ell_i"], kw, pt["seed"], pt["trials"], pt["bt"])
        assert fails == of, pt
        assert r.probes == op, pt


def test_engine_bucket_map_group():
    """n = 2^20, k = 4 with 2^21 buckets: the bucket-map group at its largest
    entries (21-bit bucket indices, 16 MB of map per trial)."""
    job = ("bgh25", 1 << 20, 21, dict(k=4, buckets=1 << 21, kick=100, stash=1), 23, 12, 20)
    ofails, oprobes = oracle_job(*job)
    m, n, ell, kw, seed, trials, bt = job
    with ck.Engine(0, arena_bytes=8 * GiB, buffers=2, max_slots=1024) as eng:
        jid = eng.submit(ck.params(m, n, ell, **kw), seed, trials, bt=bt, keep_failures=True)
        r = eng.wait(jid)
        assert list(map(int, eng.failures(jid))) == ofails
        assert r.probes == oprobes


def test_engine_spread_arena_matches_oracle(expected):
    """A spread arena (1 GiB pieces taken from the memory regions in turn):
    the pieces cover two regions or more, and every job's results are the
    oracle's."""
    with ck.Engine(0, arena_bytes=24 * GiB, buffers=3, max_slots=2048, spread=True) as eng:
        counts = eng.arena_regions()
        got = run_jobs(eng, JOBS)
    assert sum(counts) == 24 and sum(1 for c in counts[:3] if c) >= 2, counts
    assert max(counts[:3]) - min(c for c in counts[:3] if c) <= 1, counts  # the regions in turn
    for (fails, r), (ofails, oprobes), j in zip(got, expected, JOBS):
        assert fails == ofails and r.probes == oprobes, j


def test_engine_without_spread_reports_no_pieces():
    with ck.Engine(0, arena_bytes=GiB, max_slots=1024) as eng:
        assert eng.arena_regions() == (0, 0, 0, 0)
    with pytest.raises(ck.CkError):
        ck.Engine(0, arena_bytes=GiB, spread=2)
