"""Seeded random shapes against the oracle: every table layout and both tiers
for single-shape launches, and random job mixes through the engine under
random pass sizes, buffer counts and scheduling policies.  Failure offsets and
probe counts must match exactly; the shapes cover both insertion policies,
k = 2..4, stash 0..4, explicit kick and bucket counts, affine input and ell_i
from 10 to 24."""
import random

import numpy as np
import pytest

import ck_oracle as ock
from paths import load_product

pytestmark = pytest.mark.gpu

ck = load_product()


def random_shape(rng):
    mode = rng.choice(["bgh25", "drrt18"])
    k = rng.choice([2, 3, 4])
    pnum = rng.choice([3, 17, 64, 100, 257, 640, 1000, 2049, 3000])
    ell = max(rng.choice([10, 12, 16, 20, 24]), (pnum - 1).bit_length())
    buckets = max(pnum + 1, int(pnum * rng.uniform(1.05, 2.0)))
    kw = dict(k=k, buckets=buckets, kick=rng.choice([1, 5, 30, 100, 500]), stash=rng.choice([0, 1, 2, 4]),
              input_mode=rng.choice([0, 1]))
    return mode, pnum, ell, kw


SHAPES = [random_shape(random.Random(1000 + i)) for i in range(24)]
LAYOUTS = {"auto": ck.LAYOUT_AUTO, "map": ck.LAYOUT_MAP, "map-xchg": ck.LAYOUT_MAP_XCHG,
           "desc-plain": ck.LAYOUT_DESC_PLAIN, "residual": ck.LAYOUT_RESIDUAL}


@pytest.fixture(scope="module")
def ctx():
    c = ck.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("i", range(len(SHAPES)))
def test_random_shape_all_placements(ctx, i):
    mode, pnum, ell, kw = SHAPES[i]
    p = ck.params(mode, pnum, ell, **kw)
    op = ock.parameters(mode, pnum, ell, **kw)
    seed, batch, first, count = 500 + i, i % 5, 31 * i, 200 + 7 * i
    of, oprobes, _ = ock.run_range(ock.batch_key(seed, batch), first, count, op, probes=True,
                                   cached=op.half_bits <= 16)
    rk = ck.batch_key(seed, batch)
    ran = 0
    try:
        for name, layout in LAYOUTS.items():
            ctx.set_tier(ck.TIER_HBM)
            ctx.set_layout(layout)
            try:
                f, probes, _, _ = ctx.launch_debug(p, rk, first, count)
            except ck.CkError:
                assert name in ("desc-plain", "residual")  # only these can refuse a shape (entry too narrow)
                continue
            assert list(f) == list(of), name
            assert (probes == oprobes).all(), name
            ran += 1
        if op.half_bits <= 12:
            ctx.set_tier(ck.TIER_LDS)
            ctx.set_layout(ck.LAYOUT_AUTO)
            f, probes, _, _ = ctx.launch_debug(p, rk, first, count)
            assert list(f) == list(of) and (probes == oprobes).all(), "lds"
    finally:
        ctx.set_tier(ck.TIER_AUTO)
        ctx.set_layout(ck.LAYOUT_AUTO)
    assert ran >= 3


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_engine_mixes(seed):
    rng = random.Random(seed)
    jobs = []
    for j in range(10):
        mode, pnum, ell, kw = random_shape(rng)
        jobs.append((mode, pnum, ell, kw, rng.getrandbits(64), rng.choice([1, 63, 64, 65, 300, 1000]),
                     rng.choice([6, 8, 20])))
    cfg = dict(policy=rng.choice([ck.SCHED_BALANCE, ck.SCHED_FIFO, ck.SCHED_SRPT]), buffers=rng.choice([2, 3, 5]),
               max_slots=rng.choice([64, 256, 1024, 4096]), window=rng.choice([0, 2]))
    with ck.Engine(0, arena_bytes=1 << 30, **cfg) as eng:
        ids = [eng.submit(ck.params(m, n, ell, **kw), s, t, bt=bt, keep_failures=True)
               for m, n, ell, kw, s, t, bt in jobs]
        res = [eng.wait(i) for i in ids]
        got = [(list(map(int, eng.failures(i))), r) for i, r in zip(ids, res)]
    for (fails, r), (m, n, ell, kw, s, t, bt) in zip(got, jobs):
        op = ock.parameters(m, n, ell, **kw)
        want, probes = [], 0
        per = 1 << bt
        for b in range((t + per - 1) // per):
            cnt = min(per, t - b * per)
            f, _, st = ock.run_range(ock.batch_key(s, b), 0, cnt, op, cached=op.half_bits <= 16)
            want += [b * per + int(x) for x in f]
            probes += st.probes
        assert fails == want, (cfg, m, n, kw)
        assert r.probes == probes and r.trials == t
