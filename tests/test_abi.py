"""CPU checks of the C ABI: the library loads, exports every declared symbol,
and its host-side pieces (shapes, batch keys, host AES) agree with the oracle.
No device calls."""
import ctypes as C
import random
import re

import numpy as np
import pytest

import ck_oracle as ock
from paths import HEADER, load_product

ck = load_product()


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(ck_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared_symbols()
    assert names, "no declarations parsed"
    for n in names:
        assert hasattr(ck.LIB, n), n
    assert sorted(ck.EXPORTS) == names


def test_params_match_oracle():
    rng = random.Random(4)
    for _ in range(300):
        mode = rng.choice([1, 2])
        k = rng.choice([2, 3, 4])
        pnum = rng.choice([3, 16, 100, 1024, 16384, 65534, 1 << 20])
        ell = max(rng.choice([8, 16, 20, 32, 61]), (pnum - 1).bit_length())
        b = rng.choice([ck.DERIVED, int(pnum * rng.uniform(1.2, 1.6))])
        kick = rng.choice([ck.DERIVED, 0, 100, 500, 1000])
        stash = rng.choice([0, 1, 4])
        im = rng.choice([0, 1])
        a = ck.params(mode, pnum, ell, k=k, buckets=b, kick=kick, stash=stash, input_mode=im).as_dict()
        o = ock.parameters(mode, pnum, ell, k=k, buckets=b, kick=kick, stash=stash, input_mode=im).as_dict()
        assert a == o


def test_invalid_params_rejected():
    with pytest.raises(ck.CkError):
        ck.params("bgh25", 16, 8, k=5)
    with pytest.raises(ck.CkError):
        ck.params("bgh25", 1000, 8)  # 2^8 < pnum
    with pytest.raises(ck.CkError):
        ck.params(7, 16, 8)


def test_batch_key_and_host_aes(fixtures):
    rng = random.Random(8)
    for _ in range(50):
        seed, batch = rng.getrandbits(64), rng.getrandbits(64)
        assert list(ck.batch_key(seed, batch)) == ock.batch_key(seed, batch).words()
    v = fixtures["batch_key_vectors"]
    for case in v["cases"]:
        assert ck.host_aes(ck.batch_key(v["seed"], case["batch"]), [0, 0, 0, 0]) == case["aes_zero"]
    f = fixtures["fips197"]
    rk = np.array(ock.expand_key(f["key_lo"], f["key_hi"]).words(), dtype=np.uint32)
    assert ck.host_aes(rk, f["plain"]) == f["cipher"]


def test_ctx_create_without_device_fails_cleanly():
    import paths
    if paths.gpu_available():
        pytest.skip("a device is present")
    h = C.c_void_p()
    rc = ck.LIB.ck_ctx_create(0, C.byref(h))
    assert rc != 0 and not h.value


def test_engine_create_without_device_fails_cleanly():
    import paths
    if paths.gpu_available():
        pytest.skip("a device is present")
    h = C.c_void_p()
    rc = ck.LIB.ck_engine_create(0, None, C.byref(h))
    assert rc != 0 and not h.value
    assert ck.LIB.ck_last_error()


def test_engine_config_validation():
    h = C.c_void_p()
    for bad in (dict(buffers=1), dict(buffers=9), dict(max_slots=100), dict(policy=3)):
        cfg = ck.EngineConfig(**bad)
        assert ck.LIB.ck_engine_create(0, C.byref(cfg), C.byref(h)) == -1, bad
        assert not h.value


def test_struct_layouts():
    # ck_job = ck_params (72 B) + seed + trials + bt + keep_failures
    assert C.sizeof(ck.Params) == 72
    assert C.sizeof(ck.Job) == 96
    assert C.sizeof(ck.JobResult) == 48
    assert C.sizeof(ck.EngineConfig) == 32
    assert C.sizeof(ck.EngineStats) == 88


def test_plan_describe_follows_the_placement_request():
    """The planner the CLI reports from (no device): round cache or direct
    AES as asked, the LDS tier's map or round-cache form, the budget."""
    p = ck.params("bgh25", 16384, 16, kick=500)
    pl = ck.plan_describe(p, 1 << 17)
    assert (pl.tier, pl.layout, pl.round_cache) == (ck.TIER_HBM, ck.LAYOUT_RESIDUAL, 1)  # 2 x 15 + 2 bits: u32
    assert pl.trials_per_pass == 65536 and pl.buffers == 2 and pl.chain_waves == 4 and pl.chain_lds_bytes == 0
    assert ck.plan_describe(p, 1 << 17, prp=ck.PRP_DIRECT).round_cache == 0
    small = ck.plan_describe(p, 1 << 17, budget_bytes=256 << 20)
    # A 256 MiB budget caps the pass: one buffer, the items in 64 windows.
    assert small.workspace_bytes <= 256 << 20 and 0 < small.trials_per_pass < 4096
    assert small.buffers == 1 and small.windows == 64
    assert ck.plan_describe(ck.params("bgh25", 30000, 16, buckets=40000), 1024).layout == ck.LAYOUT_DESC_PLAIN
    wide = ck.params("bgh25", 16384, 32)  # half_bits 17: no round cache fits
    assert ck.plan_describe(wide, 1024).round_cache == 0
    with pytest.raises(ck.CkError):
        ck.plan_describe(wide, 1024, prp=ck.PRP_CACHE)
    g = ck.params("drrt18", 16, 8)
    lds = ck.plan_describe(g, 256, tier=ck.TIER_LDS, lds_variant=ck.LDS_ROUND)
    assert (lds.tier, lds.lds_variant, lds.round_cache) == (ck.TIER_LDS, ck.LDS_ROUND, 1) and lds.chain_lds_bytes > 0
    assert ck.plan_describe(g, 256, tier=ck.TIER_LDS).lds_variant == ck.LDS_MAP
    with pytest.raises(ck.CkError):
        ck.plan_describe(g, 256, tier=ck.TIER_LDS, prp=ck.PRP_DIRECT)


def test_plan_describe_lane_tier():
    """The lane tier's plan (no device): one-wave blocks, W lanes each with
    the whole table (16-bit residual entries where (k-1)*bits + 2 <= 16, else
    32-bit) and a 64-word ring in LDS, W chosen for the most trials per CU of
    160 KB; shapes whose entries need more than 32 bits are refused."""
    def best(buckets, eb):
        cands = []
        for w in range(1, 65):
            blk = ((w * buckets * eb + 15) // 16) * 16 + w * 64 * 8
            if blk <= 160 << 10:
                cands.append(((160 << 10) // blk * w, w, blk))
        return max(cands)[2]
    for (n, m, k, eb, blk) in [(256, 384, 3, 4, 81920), (1024, 1536, 3, 4, 159744), (1024, 1331, 2, 2, 161888)]:
        p = ck.params("bgh25", n, 16, k=k, buckets=m, kick=500)
        pl = ck.plan_describe(p, 1 << 16, tier=ck.TIER_LANE)
        assert (pl.tier, pl.layout, pl.round_cache) == (ck.TIER_LANE, ck.LAYOUT_RESIDUAL, 1)
        assert pl.chain_waves == 1 and pl.chain_lds_bytes == blk == best(m, eb)
        assert pl.trials_per_pass == 1 << 16 and pl.windows == 1
        assert ck.plan_describe(p, 1 << 16).tier == ck.TIER_HBM  # auto
    with pytest.raises(ck.CkError):
        ck.plan_describe(ck.params("bgh25", 1 << 20, 20, buckets=1363149, kick=1000, stash=2), 1024, tier=ck.TIER_LANE)


def test_trial_inputs_match_the_oracle():
    """ck_trial_inputs_of / ck_round_cache (the trial API's host inputs) equal
    the oracle's trial setup and round cache."""
    for mode, pnum, ell, im, seed, batch, off in [("bgh25", 16, 8, 0, 123, 7, 3), ("drrt18", 4096, 16, 1, 5, 2, 99),
                                                  ("bgh25", 1 << 20, 20, 1, 4, 0, 12)]:
        p = ck.params(mode, pnum, ell, input_mode=im)
        op = ock.parameters(mode, pnum, ell, input_mode=im)
        ti = ck.trial_inputs(p, ck.batch_key(seed, batch), off)
        ot = ock.trial_setup(ock.batch_key(seed, batch), off, op)
        assert list(ti.prp_key) == ot.key.words()
        assert (ti.walk, ti.randomness, ti.shift, ti.step) == (ot.walk, ot.randomness, ot.shift, ot.step)
        assert (ck.round_cache(p, list(ti.prp_key)) == ock.build_rounds(ot, op)).all()
