"""The lane tier (one trial per lane, its residual table in LDS, its new items'
candidates streamed through an LDS ring) against the oracle, through the C ABI.

Bit-exact: failing offsets, per-trial probe counts and probe totals, and the
final-table digests against the HBM residual layout's.  The shapes cover both
policies, k = 2..4, stash slots, affine input, 16- and 32-bit entries, blocks
of 64 lanes and of fewer, partial last blocks, and runs of one-probe
insertions (sparse tables), which drain a lane's ring fastest."""
import pytest

import ck_oracle as ock
from paths import load_product

pytestmark = pytest.mark.gpu

ck = load_product()


@pytest.fixture(scope="module")
def ctx():
    c = ck.Context(0)
    yield c
    c.close()


def both(mode, pnum, ell, **kw):
    return ck.params(mode, pnum, ell, **kw), ock.parameters(mode, pnum, ell, **kw)


def lane_launch(ctx, p, rk, first, count, debug=True):
    ctx.set_tier(ck.TIER_LANE)
    try:
        if debug:
            f, probes, _, st = ctx.launch_debug(p, rk, first, count)
        else:
            f, st = ctx.launch(p, rk, first, count)
            probes = None
    finally:
        ctx.set_tier(ck.TIER_AUTO)
    assert st.tier == ck.TIER_LANE
    return f, probes, st


LANE_SHAPES = [
    # mode, pnum, ell_i, extra params, trials
    ("bgh25", 16, 8, {}, 2048),                                          # golden shape, u16 entries
    ("drrt18", 16, 8, {}, 2048),
    ("bgh25", 3, 8, {"input_mode": 1}, 1024),                            # fewer items than one ring chunk
    ("bgh25", 64, 12, {"input_mode": 1, "kick": 20}, 2048),
    ("bgh25", 1024, 16, {"kick": 500}, 1024),                            # config 1 shape, u32 entries, 25 lanes
    ("bgh25", 1024, 16, {"kick": 6}, 1024),                              # most trials fail
    ("drrt18", 1024, 16, {"kick": 3, "buckets": 1100}, 512),
    ("bgh25", 300, 12, {"k": 2, "buckets": 420, "stash": 2, "kick": 30}, 1024),
    ("bgh25", 300, 12, {"k": 4, "buckets": 320, "stash": 1, "kick": 10}, 1024),
    ("drrt18", 300, 12, {"k": 2, "buckets": 420, "stash": 3, "kick": 10}, 1024),
    ("drrt18", 300, 12, {"k": 4, "buckets": 310, "stash": 1, "kick": 5}, 1024),
    ("bgh25", 500, 16, {"buckets": 4000, "kick": 100}, 700),             # sparse: one-probe insertions
    ("drrt18", 500, 16, {"buckets": 4000, "kick": 100}, 700),
    ("bgh25", 777, 16, {"k": 2, "buckets": 1001, "stash": 4, "kick": 1000}, 333),
    ("bgh25", 2000, 16, {"buckets": 2600, "kick": 500}, 200),            # 14 lanes per block
]


@pytest.mark.parametrize("mode,pnum,ell,kw,count", LANE_SHAPES)
def test_lane_tier_against_oracle(ctx, mode, pnum, ell, kw, count):
    p, op = both(mode, pnum, ell, **kw)
    seed, batch, first = 91, 2, 517
    rk = ck.batch_key(seed, batch)
    f, probes, st = lane_launch(ctx, p, rk, first, count)
    of, oprobes, ost = ock.run_range(ock.batch_key(seed, batch), first, count, op, probes=True)
    assert list(f) == list(of)
    assert (probes == oprobes).all()
    assert st.probes == ost.probes == int(oprobes.sum())
    g, _, st2 = lane_launch(ctx, p, rk, first, count, debug=False)
    assert list(g) == list(of) and st2.probes == ost.probes


@pytest.mark.parametrize("mode,pnum,ell,kw,count", LANE_SHAPES)
def test_lane_tier_table_digests(ctx, mode, pnum, ell, kw, count):
    """Final tables: the digest of every trial's LDS table equals the digest
    of the HBM residual layout's table for the same trial."""
    p = ck.params(mode, pnum, ell, **kw)
    rk = ck.batch_key(5, 1)
    ctx.set_tier(ck.TIER_LANE)
    try:
        f, _, d_lane, st = ctx.launch_digest(p, rk, 0, count)
    finally:
        ctx.set_tier(ck.TIER_AUTO)
    assert st.tier == ck.TIER_LANE
    ctx.set_tier(ck.TIER_HBM)
    ctx.set_layout(ck.LAYOUT_RESIDUAL)
    try:
        g, _, d_hbm, st2 = ctx.launch_digest(p, rk, 0, count)
    finally:
        ctx.set_layout(ck.LAYOUT_AUTO)
        ctx.set_tier(ck.TIER_AUTO)
    assert st2.tier == ck.TIER_HBM
    assert list(f) == list(g)
    assert (d_lane == d_hbm).all()


def test_lane_tier_golden_32768(ctx, fixtures):
    g = fixtures["golden_32768"]
    for mode in ("bgh25", "drrt18"):
        p = ck.params(mode, g["pnum"], g["ell_i"])
        fails = []
        for batch in range(1 << (g["et"] - g["bt"])):
            f, _, _ = lane_launch(ctx, p, ck.batch_key(g["seed"], batch), 0, 1 << g["bt"], debug=False)
            fails += [batch * (1 << g["bt"]) + int(x) for x in f]
        assert fails == g["failures"][mode]


def test_lane_tier_affine_and_reference_ranges(ctx, fixtures):
    v = fixtures["affine_vectors"]
    for mode, want in v["failures"].items():
        p = ck.params(mode, v["pnum"], v["ell_i"], input_mode=1)
        f, _, _ = lane_launch(ctx, p, ck.batch_key(v["seed"], v["batch"]), 0, v["count"], debug=False)
        assert list(f) == want
    for row in fixtures["reference_ranges"]:
        p = ck.params(row["mode"], row["pnum"], row["ell_i"], input_mode=row["input_mode"])
        f, _, _ = lane_launch(ctx, p, ck.batch_key(row["seed"], row["batch"]), 0, row["count"], debug=False)
        assert list(f) == row["failures"], row


def test_auto_tier_and_tuner(ctx):
    """Auto stays on the HBM tier; the autotuner measures the lane tier beside
    the HBM layouts (all candidates must agree on the failing trials) and a
    plan forced onto the lane tier reports its one-wave blocks and LDS."""
    small = ck.params("bgh25", 1024, 16, kick=500)
    assert ctx.plan(small, 1 << 16).tier == ck.TIER_HBM
    ctx.set_tier(ck.TIER_LANE)
    try:
        pl = ctx.plan(small, 1 << 16)
    finally:
        ctx.set_tier(ck.TIER_AUTO)
    assert pl.tier == ck.TIER_LANE and pl.layout == ck.LAYOUT_RESIDUAL
    assert pl.chain_waves == 1 and 0 < pl.chain_lds_bytes <= 160 * 1024
    with ck.Context(0) as t:
        ents = t.tune(small, ck.batch_key(0, 0), 1 << 14)
    assert any(e["tier"] == ck.TIER_LANE for e in ents)


def test_lane_tier_refuses_tables_and_wide_entries(ctx):
    ctx.set_tier(ck.TIER_LANE)
    try:
        with pytest.raises(ck.CkError):
            ctx.launch_debug(ck.params("bgh25", 1024, 16, kick=500), ck.batch_key(0, 0), 0, 8, tables=True)
        with pytest.raises(ck.CkError):  # n = 2^20: 44-bit residual entries
            ctx.plan(ck.params("bgh25", 1 << 20, 20, buckets=1363149, kick=1000, stash=2), 1024)
    finally:
        ctx.set_tier(ck.TIER_AUTO)


def test_config1_prefix_on_lane_tier(ctx):
    """Config 1 (n = 1024, m = 1536, kick 500, ell_i 16, seed 0): the first
    4096 trials of batch 0 and their probe total equal the oracle's."""
    p, op = both("bgh25", 1024, 16, kick=500)
    f, _, st = lane_launch(ctx, p, ck.batch_key(0, 0), 0, 4096, debug=False)
    of, _, ost = ock.run_range(ock.batch_key(0, 0), 0, 4096, op)
    assert list(f) == list(of) and st.probes == ost.probes
