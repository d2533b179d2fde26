"""Residual-layout item windows.

A pass may insert each trial's items in windows: the build writes the
candidates of one window's items only, and the chain stops at the window's end
(an item boundary) and resumes in the next launch from the saved walk, random
word, stash count and probe count.  These tests force windows of several sizes
and compare with the oracle: failure offsets, per-trial probe counts and table
digests.  Memory-capped shapes choose windows by themselves; the n = 2^20 test
in the parity suite covers that choice."""
import pytest

import ck_oracle as ock
from paths import load_product

pytestmark = pytest.mark.gpu

ck = load_product()

SHAPES = [
    # mode, pnum, ell_i, overrides, trials
    ("bgh25", 1024, 16, {"kick": 500}, 256),                                        # u32 entries, u64 stream
    ("bgh25", 300, 12, {"k": 2, "buckets": 600, "kick": 30, "stash": 2}, 512),       # 162 failures; stash across windows
    ("drrt18", 600, 12, {"k": 4, "buckets": 1000, "kick": 50, "stash": 1}, 256),    # 44-bit entries
    ("drrt18", 200, 10, {"buckets": 230, "kick": 8, "stash": 3}, 512),
    ("bgh25", 30000, 16, {"buckets": 40000, "kick": 300, "stash": 1}, 16),          # 44-bit entries
    ("bgh25", 20000, 16, {"k": 4, "buckets": 30000, "kick": 200, "stash": 1}, 16),  # 48-bit entries
    ("bgh25", 60000, 16, {"k": 4, "buckets": 70000, "kick": 200, "stash": 2}, 12),  # u32 map-row stream, u64 entries
]


def windows_for(pnum):
    if pnum <= 1024:
        return [1, 7, 64, pnum // 8 + 1, pnum - 1]
    return [pnum // 8 + 1, 9999, pnum - 1]


@pytest.fixture(scope="module")
def ctx():
    c = ck.Context(0)
    c.set_layout(ck.LAYOUT_RESIDUAL)
    yield c
    c.close()


@pytest.mark.parametrize("mode,pnum,ell,kw,count", SHAPES, ids=lambda v: str(v) if not isinstance(v, dict) else "")
def test_windows_match_oracle(ctx, mode, pnum, ell, kw, count):
    p = ck.params(mode, pnum, ell, **kw)
    op = ock.parameters(mode, pnum, ell, **kw)
    seed, batch, first = 41, 2, 9
    bk = ock.batch_key(seed, batch)
    of, oprobes, _ = ock.run_range(bk, first, count, op, probes=True)
    digests = {i: ock.table_digest(bins, ock.bucket_map(bk, first + i, op), op.k)
               for i in range(0, count, max(1, count // 6))
               for _, bins, _ in [ock.trial_run(bk, first + i, op)]}
    runs = [(w, 0) for w in windows_for(pnum)] + [(pnum // 8 + 1, max(1, count // 3))]
    try:
        for window, chunk in runs:
            ctx.set_window(window)
            ctx.set_chunk(chunk)
            f, probes, dig, st = ctx.launch_digest(p, ck.batch_key(seed, batch), first, count)
            nwin = -(-pnum // window)
            passes = -(-count // st.chunk)
            assert st.launches == 2 * nwin * passes, (window, chunk)
            assert list(f) == list(of), (window, chunk)
            assert (probes == oprobes).all(), (window, chunk)
            for i, d in digests.items():
                assert int(dig[i]) == d, (window, chunk, i)
    finally:
        ctx.set_window(0)
        ctx.set_chunk(0)


def test_windows_at_n2_20(ctx):
    """The rare-event HBM shape (n = 2^20, m = 1.3n, k = 3, stash 2, kick
    1000, seed 4) in 8 windows, against the oracle."""
    p = ck.params("bgh25", 1 << 20, 20, k=3, buckets=1363149, kick=1000, stash=2)
    op = ock.parameters("bgh25", 1 << 20, 20, k=3, buckets=1363149, kick=1000, stash=2)
    ctx.set_window(1 << 17)
    try:
        f, probes, _, st = ctx.launch_digest(p, ck.batch_key(4, 0), 100, 8)
    finally:
        ctx.set_window(0)
    assert st.launches == 16
    of, oprobes, _ = ock.run_range(ock.batch_key(4, 0), 100, 8, op, probes=True)
    assert list(f) == list(of) and (probes == oprobes).all()


def test_windows_leave_other_layouts_and_outputs_alone(ctx):
    """Windows apply to residual-layout launches only, and not to launches that
    return the bucket map, walks or random words."""
    p = ck.params("bgh25", 1024, 16, kick=500)
    ctx.set_window(100)
    try:
        ctx.set_layout(ck.LAYOUT_DESC_PLAIN)
        f, st = ctx.launch(p, ck.batch_key(3, 0), 0, 64)
        assert st.launches == 2
        ctx.set_layout(ck.LAYOUT_RESIDUAL)
        key, walk, rnd, m = ctx.trial_map(p, ck.batch_key(3, 0), 5)
        t = ock.trial_setup(ock.batch_key(3, 0), 5, ock.parameters("bgh25", 1024, 16, kick=500))
        assert (walk, rnd) == (t.walk, t.randomness)
        g, st = ctx.launch(p, ck.batch_key(3, 0), 0, 64)
        assert st.launches == 2 * 11
        assert list(map(int, f)) == list(map(int, g))
    finally:
        ctx.set_window(0)
        ctx.set_layout(ck.LAYOUT_RESIDUAL)


# n = 2^20, m = ceil(1.3 n), seed 4 with kick bounds low enough that trials
# fail: (kick 75, stash 0) fails 12 of the first 32 trials, (kick 60, stash 3)
# fails 25 of them with the stash in use (2.9 stashed items per trial).
N2_20_FAILING = [dict(buckets=1363149, kick=75, stash=0), dict(buckets=1363149, kick=60, stash=3)]


@pytest.mark.parametrize("kw", N2_20_FAILING, ids=["kick75", "kick60-stash3"])
def test_n2_20_failure_detection(kw):
    """Failure detection at the rare-event size, against the oracle trial by
    trial (failing offsets and per-trial probes): the default launch, item
    windows forced small, and an engine job."""
    trials = 32
    p = ck.params("bgh25", 1 << 20, 20, **kw)
    of, oprobes, _ = ock.run_range(ock.batch_key(4, 0), 0, trials, ock.parameters("bgh25", 1 << 20, 20, **kw),
                                   probes=True, threads=8)
    assert 0 < len(of) < trials
    with ck.Context(0) as c:
        f, pr, _, _ = c.launch_debug(p, ck.batch_key(4, 0), 0, trials)
        assert list(f) == list(of) and (pr == oprobes).all()
        c.set_window(100000)  # 11 windows per trial
        f, st = c.launch(p, ck.batch_key(4, 0), 0, trials)
        assert list(f) == list(of) and st.probes == int(oprobes.sum())
    with ck.Engine(0, arena_bytes=16 << 30, max_slots=1024) as eng:
        jid = eng.submit(p, 4, trials, keep_failures=True)
        r = eng.wait(jid)
        assert list(map(int, eng.failures(jid))) == list(map(int, of)) and r.probes == int(oprobes.sum())
