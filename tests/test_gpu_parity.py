"""Parity of the HIP path with the oracle, through the C ABI.

Bit-exact: failure offsets, per-trial probe counts, final tables, the device
key derivation and bucket map.  Sizes are chosen so the oracle finishes in
seconds."""
import hashlib
import json
from pathlib import Path

import numpy as np
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


TIERS = {"hbm": ck.TIER_HBM, "lds": ck.TIER_LDS}


@pytest.fixture(params=["hbm", "lds"])
def tier(request, ctx):
    ctx.set_tier(TIERS[request.param])
    yield request.param
    ctx.set_tier(ck.TIER_AUTO)


def both(mode, pnum, ell, **kw):
    return ck.params(mode, pnum, ell, **kw), ock.parameters(mode, pnum, ell, **kw)


def test_device_aes_self_test(ctx):
    ctx.self_test()


SHAPES = [
    ("bgh25", 16, 8, {}),
    ("drrt18", 16, 8, {"input_mode": 1}),
    ("bgh25", 1024, 16, {"kick": 500}),
    ("drrt18", 4096, 16, {"input_mode": 1}),
    ("bgh25", 16384, 16, {"kick": 500}),
    ("bgh25", 16, 32, {}),                      # half 17: direct-AES build
    ("drrt18", 300, 24, {"input_mode": 1}),     # half 13: direct-AES build
    ("bgh25", 200, 12, {"k": 2, "buckets": 260, "stash": 1}),
    ("drrt18", 200, 12, {"k": 4, "buckets": 230, "stash": 2}),
]


@pytest.mark.parametrize("mode,pnum,ell,kw", SHAPES)
def test_trial_derivation_and_bucket_map(ctx, mode, pnum, ell, kw):
    p, op = both(mode, pnum, ell, **kw)
    for seed, batch, off in [(1, 0, 0), (123, 7, 255), (2**63 + 5, 2**40 + 3, 123456)]:
        rk = ck.batch_key(seed, batch)
        key, walk, rnd, m = ctx.trial_map(p, rk, off)
        bk = ock.batch_key(seed, batch)
        t = ock.trial_setup(bk, off, op)
        assert list(key) == t.key.words()
        assert (walk, rnd) == (t.walk, t.randomness)
        assert (m == ock.bucket_map(bk, off, op)).all()


@pytest.mark.parametrize("mode", ["bgh25", "drrt18"])
def test_golden_32768(ctx, tier, fixtures, mode):
    g = fixtures["golden_32768"]
    p = ck.params(mode, g["pnum"], g["ell_i"])
    fails = []
    for batch in range(1 << (g["et"] - g["bt"])):
        f, _ = ctx.launch(p, ck.batch_key(g["seed"], batch), 0, 1 << g["bt"])
        fails += [batch * (1 << g["bt"]) + int(x) for x in f]
    assert fails == g["failures"][mode]


def test_affine_vectors(ctx, tier, fixtures):
    v = fixtures["affine_vectors"]
    for mode, want in v["failures"].items():
        p = ck.params(mode, v["pnum"], v["ell_i"], input_mode=1)
        f, _ = ctx.launch(p, ck.batch_key(v["seed"], v["batch"]), 0, v["count"])
        assert list(f) == want


def test_reference_ranges(ctx, tier, fixtures):
    for row in fixtures["reference_ranges"]:
        p = ck.params(row["mode"], row["pnum"], row["ell_i"], input_mode=row["input_mode"])
        f, _ = ctx.launch(p, ck.batch_key(row["seed"], row["batch"]), 0, row["count"])
        assert list(f) == row["failures"], row


def test_reference_trial_tables(ctx, tier, fixtures):
    """Outcome and final table of single trials as the reference's
    cpu_trial_cached leaves them (cpu_trials.hpp:128-149), with its own
    mode-derived buckets and kick (3n for bgh25: 49152 at n = 2^14).  The
    fixture holds sha256 over the reference's u32 entries ordinal | choice<<16."""
    for row in fixtures["reference_trials"]:
        p = ck.params(row["mode"], row["pnum"], row["ell_i"], input_mode=row["input_mode"])
        f, _, tables, _ = ctx.launch_debug(p, ck.batch_key(row["seed"], row["batch"]), row["offset"], 1,
                                           tables=True)
        assert len(f) == row["failed"], row
        t = tables[0].copy()
        occ = t != 0xFFFFFFFF
        t[occ] = (t[occ] & 0x3FFFFFFF) | ((t[occ] >> 30) << 16)
        assert hashlib.sha256(t.astype("<u4").tobytes()).hexdigest() == row["table_sha256"], row


def test_headline_shape_reference_kick(ctx):
    """The headline shape with the reference's own kick bound (3n = 49152)
    instead of 500: no insertion there nears 500 kicks, so failures and
    probe counts are those of the 500-kick run, and both equal the oracle."""
    p_ref, op_ref = both("bgh25", 16384, 16)
    p_500 = ck.params("bgh25", 16384, 16, kick=500)
    assert p_ref.kick == 49152 and p_ref.buckets == 24576
    count = 2048
    f, probes, _, _ = ctx.launch_debug(p_ref, ck.batch_key(1, 0), 0, count)
    g, probes500, _, _ = ctx.launch_debug(p_500, ck.batch_key(1, 0), 0, count)
    of, oprobes, _ = ock.run_range(ock.batch_key(1, 0), 0, count, op_ref, probes=True)
    assert list(f) == list(of) == list(g)
    assert (probes == oprobes).all() and (probes == probes500).all()


PROBE_SHAPES = [
    ("bgh25", 16, 8, {}, 2048),
    ("drrt18", 16, 8, {}, 2048),
    ("bgh25", 64, 12, {"input_mode": 1, "kick": 20}, 2048),
    ("drrt18", 1024, 16, {"kick": 3, "buckets": 1100}, 512),
    ("bgh25", 1024, 16, {"kick": 500}, 512),
    ("bgh25", 300, 12, {"k": 2, "buckets": 420, "stash": 2, "kick": 30}, 512),
    ("bgh25", 300, 12, {"k": 4, "buckets": 320, "stash": 1, "kick": 10}, 512),
    ("drrt18", 300, 12, {"k": 2, "buckets": 420, "stash": 3, "kick": 10}, 512),
    ("drrt18", 300, 12, {"k": 4, "buckets": 310, "stash": 1, "kick": 5}, 512),
]


@pytest.mark.parametrize("mode,pnum,ell,kw,count", PROBE_SHAPES)
def test_failures_probes_tables(ctx, tier, mode, pnum, ell, kw, count):
    p, op = both(mode, pnum, ell, **kw)
    seed, batch, first = 77, 3, 1000
    f, probes, tables, st = ctx.launch_debug(p, ck.batch_key(seed, batch), first, count, tables=True)
    assert st.tier == TIERS[tier]
    bk = ock.batch_key(seed, batch)
    of, oprobes, ost = ock.run_range(bk, first, count, op, probes=True)
    assert list(f) == list(of)
    assert (probes == oprobes).all()
    assert int(probes.sum()) == ost.probes
    for i in range(0, count, max(1, count // 16)):
        failed, bins, _ = ock.trial_run(bk, first + i, op)
        if not failed:
            assert (tables[i] == bins).all(), i


def test_headline_shape_prefix(ctx, tier):
    # n = 2^14, m = 24576, k = 3, kick 500, ell_i 16, seed 1, batch 0.
    p, op = both("bgh25", 16384, 16, kick=500)
    count = 1024
    f, probes, _, st = ctx.launch_debug(p, ck.batch_key(1, 0), 0, count)
    of, oprobes, _ = ock.run_range(ock.batch_key(1, 0), 0, count, op, probes=True)
    assert list(f) == list(of)
    assert (probes == oprobes).all()


HEADLINE_LAYOUTS = {"auto": ck.LAYOUT_AUTO, "residual": ck.LAYOUT_RESIDUAL, "desc-plain": ck.LAYOUT_DESC_PLAIN,
                    "map": ck.LAYOUT_MAP, "map-xchg": ck.LAYOUT_MAP_XCHG}


@pytest.mark.parametrize("kick", [48, 64])
def test_headline_shape_failure_detection(ctx, kick):
    """The headline shape (n = 2^14, m = 24576, k = 3; seed 1, batch 0) with a
    kick bound low enough that trials fail -- about half of them at 48, a few
    at 64 -- so that failure detection is pinned at this size, not only the
    absence of failures.  Failing offsets and per-trial probes equal the
    oracle's under every HBM layout (auto is u32 residual entries), on the LDS
    tier and through the engine."""
    p, op = both("bgh25", 16384, 16, kick=kick)
    count = 256
    of, oprobes, _ = ock.run_range(ock.batch_key(1, 0), 0, count, op, probes=True)
    assert 0 < len(of) < count
    for name, lay in HEADLINE_LAYOUTS.items():
        ctx.set_layout(lay)
        try:
            f, probes, _, _ = ctx.launch_debug(p, ck.batch_key(1, 0), 0, count)
            g, st = ctx.launch(p, ck.batch_key(1, 0), 0, count)
        finally:
            ctx.set_layout(ck.LAYOUT_AUTO)
        assert list(f) == list(of) == list(g), name
        assert (probes == oprobes).all() and st.probes == int(oprobes.sum()), name
    ctx.set_tier(ck.TIER_LDS)
    try:
        f, probes, _, st = ctx.launch_debug(p, ck.batch_key(1, 0), 0, count)
    finally:
        ctx.set_tier(ck.TIER_AUTO)
    assert st.tier == ck.TIER_LDS and list(f) == list(of) and (probes == oprobes).all()
    with ck.Engine(0, arena_bytes=8 << 30) as eng:
        jid = eng.submit(p, 1, count, bt=20, keep_failures=True)
        res = eng.wait(jid)
        assert list(map(int, eng.failures(jid))) == list(of) and res.probes == int(oprobes.sum())


def test_headline_shape_whole_batch0(ctx):
    """All 2^20 trials of batch 0 at the headline shape (seed 1).  With the
    reference's own kick bound (3n = 49152) the failing offsets equal those
    the reference's cpu_bench wrote; with kick 500 the failing offsets and the
    probe total equal the oracle's."""
    h = json.loads((Path(__file__).parent / "golden" / "headline_batch0.json").read_text())
    count = h["shape"]["count"]
    p_ref = ck.params("bgh25", 16384, 16)
    assert p_ref.kick == h["reference"]["kick"]
    f, _ = ctx.launch(p_ref, ck.batch_key(1, 0), 0, count)
    assert list(map(int, f)) == h["reference"]["failures"]
    g, st = ctx.launch(ck.params("bgh25", 16384, 16, kick=500), ck.batch_key(1, 0), 0, count)
    assert list(map(int, g)) == h["kick500"]["failures"]
    assert st.probes == h["kick500"]["probes"]


def test_chunking_does_not_change_results(ctx, tier):
    p, _ = both("drrt18", 16, 8)
    rk = ck.batch_key(123, 3)
    ref, _ = ctx.launch(p, rk, 0, 4096)
    ctx.set_chunk(37)
    try:
        got, st = ctx.launch(p, rk, 0, 4096)
        assert st.chunk == 37
    finally:
        ctx.set_chunk(0)
    assert (got == ref).all()


def test_direct_build_shapes_reject_lds_tier(ctx):
    p = ck.params("bgh25", 16, 32)  # half 17: no round cache, HBM tier only
    ctx.set_tier(ck.TIER_LDS)
    try:
        with pytest.raises(ck.CkError):
            ctx.launch(p, ck.batch_key(1, 0), 0, 4)
    finally:
        ctx.set_tier(ck.TIER_AUTO)
    f, st = ctx.launch(p, ck.batch_key(1, 0), 0, 64)
    assert st.tier == ck.TIER_HBM


@pytest.mark.parametrize("mode,pnum,ell,kw", [("bgh25", 16, 32, {}), ("drrt18", 300, 24, {"input_mode": 1})])
def test_direct_build_outcomes(ctx, mode, pnum, ell, kw):
    p, op = both(mode, pnum, ell, **kw)
    f, probes, _, _ = ctx.launch_debug(p, ck.batch_key(5, 1), 0, 256)
    of, oprobes, _ = ock.run_range(ock.batch_key(5, 1), 0, 256, op, cached=False, probes=True)
    assert list(f) == list(of) and (probes == oprobes).all()


def test_tiers_agree_on_mixed_shapes(ctx):
    for mode, pnum, ell, kw in [("bgh25", 500, 16, {"kick": 7, "buckets": 560}),
                                ("drrt18", 2000, 14, {"input_mode": 1, "kick": 4})]:
        p = ck.params(mode, pnum, ell, **kw)
        rk = ck.batch_key(9, 9)
        res = {}
        for name, t in TIERS.items():
            ctx.set_tier(t)
            res[name] = ctx.launch_debug(p, rk, 17, 1500)
        ctx.set_tier(ck.TIER_AUTO)
        assert list(res["hbm"][0]) == list(res["lds"][0])
        assert (res["hbm"][1] == res["lds"][1]).all()


# ------------------------------------------- layouts and table digests

COMBOS = {"hbm-map": (ck.TIER_HBM, ck.LAYOUT_MAP), "hbm-desc": (ck.TIER_HBM, ck.LAYOUT_DESC),
          "hbm-map-xchg": (ck.TIER_HBM, ck.LAYOUT_MAP_XCHG), "hbm-desc-plain": (ck.TIER_HBM, ck.LAYOUT_DESC_PLAIN),
          "hbm-residual": (ck.TIER_HBM, ck.LAYOUT_RESIDUAL), "lds": (ck.TIER_LDS, ck.LAYOUT_AUTO)}

DIGEST_SHAPES = [
    ("bgh25", 16, 8, {}, 1024),
    ("drrt18", 16, 8, {"input_mode": 1}, 1024),
    ("bgh25", 300, 12, {"k": 2, "buckets": 420, "stash": 2, "kick": 30}, 512),
    ("bgh25", 300, 12, {"k": 4, "buckets": 320, "stash": 1, "kick": 10}, 512),
    ("drrt18", 300, 12, {"k": 2, "buckets": 420, "stash": 3, "kick": 10}, 512),
    ("drrt18", 300, 12, {"k": 4, "buckets": 310, "stash": 1, "kick": 5}, 512),
    ("bgh25", 1024, 16, {"kick": 500}, 512),
    ("drrt18", 2000, 14, {"input_mode": 1, "kick": 4}, 256),
]


@pytest.mark.parametrize("combo", list(COMBOS))
@pytest.mark.parametrize("mode,pnum,ell,kw,count", DIGEST_SHAPES)
def test_table_digests(ctx, combo, mode, pnum, ell, kw, count):
    tier, layout = COMBOS[combo]
    ctx.set_tier(tier)
    ctx.set_layout(layout)
    try:
        p, op = both(mode, pnum, ell, **kw)
        seed, batch, first = 31, 4, 77
        f, probes, dig, st = ctx.launch_digest(p, ck.batch_key(seed, batch), first, count)
    finally:
        ctx.set_tier(ck.TIER_AUTO)
        ctx.set_layout(ck.LAYOUT_AUTO)
    bk = ock.batch_key(seed, batch)
    of, oprobes, _ = ock.run_range(bk, first, count, op, probes=True)
    assert list(f) == list(of)
    assert (probes == oprobes).all()
    for i in range(count):
        _, bins, _ = ock.trial_run(bk, first + i, op)
        assert int(dig[i]) == ock.table_digest(bins, ock.bucket_map(bk, first + i, op), op.k), i


def test_headline_shape_descriptor_layout(ctx):
    p, op = both("bgh25", 16384, 16, kick=500)
    ctx.set_layout(ck.LAYOUT_DESC)
    try:
        f, probes, dig, st = ctx.launch_digest(p, ck.batch_key(1, 0), 0, 128)
    finally:
        ctx.set_layout(ck.LAYOUT_AUTO)
    bk = ock.batch_key(1, 0)
    of, oprobes, _ = ock.run_range(bk, 0, 128, op, probes=True)
    assert list(f) == list(of) and (probes == oprobes).all()
    for i in range(0, 128, 9):
        _, bins, _ = ock.trial_run(bk, i, op)
        assert int(dig[i]) == ock.table_digest(bins, ock.bucket_map(bk, i, op), 3)


def test_descriptor_trial_map(ctx):
    p, op = both("drrt18", 4096, 16, input_mode=1)
    ctx.set_layout(ck.LAYOUT_DESC)
    try:
        key, walk, rnd, m = ctx.trial_map(p, ck.batch_key(8, 2), 99)
    finally:
        ctx.set_layout(ck.LAYOUT_AUTO)
    assert (m == ock.bucket_map(ock.batch_key(8, 2), 99, op)).all()


def test_descriptor_layout_rejects_oversized_shapes(ctx):
    p = ck.params("bgh25", 1 << 20, 20, buckets=1363149, kick=1000, stash=2)  # 3 x 21 bits + 2 > 63
    ctx.set_layout(ck.LAYOUT_DESC)
    try:
        with pytest.raises(ck.CkError):
            ctx.launch(p, ck.batch_key(4, 0), 0, 1)
    finally:
        ctx.set_layout(ck.LAYOUT_AUTO)


def test_n2_20_one_and_two_buffer_passes_agree(ctx):
    """The rare-event HBM shape (n = 2^20, m = 1.3n, k = 3, stash 2, kick 1000,
    seed 4).  One 32768-trial call is memory-capped (a slot holding every
    item's candidates is 16.6 MB), so it runs single-buffered passes that
    insert the items in 64 windows; 8192-trial calls run double-buffered
    without windows.  Failures and probe totals agree, and the first 24 trials
    match the oracle probe for probe."""
    p, op = both("bgh25", 1 << 20, 20, k=3, buckets=1363149, kick=1000, stash=2)
    rk = ck.batch_key(4, 0)
    f1, st1 = ctx.launch(p, rk, 0, 32768)
    assert st1.launches == 2 * 64 * -(-32768 // st1.chunk)
    f2, probes2 = [], 0
    for off in range(0, 32768, 8192):
        f, st = ctx.launch(p, rk, off, 8192)
        assert st.launches == 2 * -(-8192 // st.chunk)
        f2 += list(map(int, f))
        probes2 += st.probes
    assert list(map(int, f1)) == f2
    assert st1.probes == probes2
    f, pr, _, _ = ctx.launch_debug(p, rk, 0, 24)
    of, opr, _ = ock.run_range(ock.batch_key(4, 0), 0, 24, op, probes=True)
    assert list(f) == list(of) and (pr == opr).all()


RESIDUAL_SHAPES = [
    ("bgh25", 16384, 16, {"kick": 500}, 64),                                   # u32 entries, 2*15 + 2 = 32 bits
    ("drrt18", 600, 12, {"k": 4, "buckets": 1000, "kick": 50, "stash": 1}, 256),  # 3*10 + 2 = 32 bits, k = 4: 44-bit
    ("bgh25", 600, 12, {"k": 4, "buckets": 1000, "kick": 50}, 256),             # 44-bit entries
    ("bgh25", 30000, 16, {"buckets": 40000, "kick": 300, "stash": 1}, 24),      # 2*16 + 2 = 34 bits: 44-bit
    ("bgh25", 20000, 16, {"k": 4, "buckets": 30000, "kick": 200, "stash": 1}, 24),  # 3*15 + 2 = 47 bits: 48-bit
    ("drrt18", 20000, 16, {"k": 4, "buckets": 30000, "kick": 60}, 24),             # 48-bit entries
    ("bgh25", 60000, 16, {"k": 4, "buckets": 70000, "kick": 200, "stash": 2}, 24),  # 4*17 > 64: u32 map stream, u64
    ("drrt18", 60000, 16, {"k": 4, "buckets": 70000, "kick": 100}, 24),            # 3*17 + 2 = 53 bits: u64 entries
    ("bgh25", 1 << 20, 20, {"buckets": 1363149, "kick": 1000, "stash": 2}, 6),   # n = 2^20: 2*21 + 2 = 44 bits
]


@pytest.mark.parametrize("mode,pnum,ell,kw,count", RESIDUAL_SHAPES)
def test_residual_layout(ctx, mode, pnum, ell, kw, count):
    """Residual entries (the occupant's other candidates and its choice) at the
    entry-width boundaries and with both candidate-stream forms."""
    p, op = both(mode, pnum, ell, **kw)
    ctx.set_layout(ck.LAYOUT_RESIDUAL)
    try:
        f, probes, dig, st = ctx.launch_digest(p, ck.batch_key(21, 1), 5, count)
        with pytest.raises(ck.CkError):  # ordinals are not kept
            ctx.launch_debug(p, ck.batch_key(21, 1), 5, 1, tables=True)
    finally:
        ctx.set_layout(ck.LAYOUT_AUTO)
    bk = ock.batch_key(21, 1)
    of, oprobes, _ = ock.run_range(bk, 5, count, op, probes=True)
    assert list(f) == list(of)
    assert (probes == oprobes).all()
    for i in range(0, count, max(1, count // 8)):
        _, bins, _ = ock.trial_run(bk, 5 + i, op)
        assert int(dig[i]) == ock.table_digest(bins, ock.bucket_map(bk, 5 + i, op), op.k), i


@pytest.mark.parametrize("mode,pnum,ell,kw,sample", [
    ("bgh25", 1024, 16, {"kick": 500}, 4096),                     # both tiers, four layouts
    ("drrt18", 600, 12, {"k": 4, "buckets": 1000, "kick": 50}, 2048),
])
def test_autotuner(mode, pnum, ell, kw, sample):
    """Every candidate reports the same failures (checked inside the tuner);
    the chosen setting keeps giving the oracle's results."""
    p, op = both(mode, pnum, ell, **kw)
    rk = ck.batch_key(77, 3)
    with ck.Context(0) as c:
        ents = c.tune(p, rk, sample)
        assert len(ents) >= 4 and all(e["trials_per_s"] > 0 for e in ents)
        assert {e["layout"] for e in ents if e["tier"] == ck.TIER_HBM} >= {ck.LAYOUT_MAP, ck.LAYOUT_MAP_XCHG,
                                                                          ck.LAYOUT_RESIDUAL}
        f, _ = c.launch(p, rk, 100, 700)
    of, _, _ = ock.run_range(ock.batch_key(77, 3), 100, 700, op)
    assert list(f) == list(of)


def test_config1_cpu_point(ctx):
    """The reference's CPU-runnable point (bgh25 n=1024 m=1536 k=3 kick 500,
    ell_i 16, seed 0, 10^5 trials of batch 0): failures and probe total."""
    p, op = both("bgh25", 1024, 16, kick=500)
    f, st = ctx.launch(p, ck.batch_key(0, 0), 0, 100000)
    of, _, ost = ock.run_range(ock.batch_key(0, 0), 0, 100000, op)
    assert list(f) == list(of)
    assert st.probes == ost.probes
