"""Pins the CPU oracle: known answers, invariants, the reference's golden data,
and (when oracle/_ref is built) the reference's own host functions."""
import ctypes as C
import hashlib
import random

import numpy as np
import pytest

import ck_oracle as ock


def ref_table(bins: np.ndarray) -> np.ndarray:
    """Oracle entry ordinal|choice<<30 -> reference entry ordinal|choice<<16."""
    out = bins.copy()
    occ = bins != 0xFFFFFFFF
    out[occ] = (bins[occ] & 0x3FFFFFFF) | ((bins[occ] >> 30) << 16)
    return out


def digest(bins: np.ndarray) -> str:
    return hashlib.sha256(bins.astype("<u4").tobytes()).hexdigest()


# ------------------------------------------------------------------ AES

def test_sbox_fips197():
    s = ock.sbox()
    assert s[0x00] == 0x63 and s[0x01] == 0x7C and s[0x53] == 0xED and s[0xFF] == 0x16
    assert sorted(s) == list(range(256))


def test_fips197_c1_both_paths(fixtures):
    f = fixtures["fips197"]
    key = ock.expand_key(f["key_lo"], f["key_hi"])
    assert ock.aes(key, f["plain"], bytewise=True) == f["cipher"]
    assert ock.aes(key, f["plain"]) == f["cipher"]


def test_fips197_appendix_a_and_b():
    # FIPS-197 A.1 key 2b7e1516 28aed2a6 abf71588 09cf4f3c: last round key words.
    key = ock.expand_words([0x2B7E1516, 0x28AED2A6, 0xABF71588, 0x09CF4F3C])
    assert key.words()[40:44] == [0xD014F9A8, 0xC9EE2589, 0xE13F0CC8, 0xB6630CA6]
    assert key.words()[4] == 0xA0FAFE17
    # Appendix B example block.
    out = ock.aes(key, [0x3243F6A8, 0x885A308D, 0x313198A2, 0xE0370734], bytewise=True)
    assert out == [0x3925841D, 0x02DC09FB, 0xDC118597, 0x196A0B32]


def test_aesni_equals_bytewise():
    rng = random.Random(5)
    for _ in range(300):
        key = ock.expand_words([rng.getrandbits(32) for _ in range(4)])
        blk = [rng.getrandbits(32) for _ in range(4)]
        assert ock.aes(key, blk) == ock.aes(key, blk, bytewise=True)


def test_batch_key_vectors(fixtures):
    v = fixtures["batch_key_vectors"]
    for case in v["cases"]:
        k = ock.batch_key(v["seed"], case["batch"])
        assert ock.aes(k, [0, 0, 0, 0], bytewise=True) == case["aes_zero"]


def test_splitmix64_published_vector():
    # SplitMix64 (Steele, Lea, Flood 2014; Vigna's splitmix64.c), seed 1234567.
    s = 1234567
    got = []
    for _ in range(5):
        s, v = ock.splitmix(s)
        got.append(v)
    assert got == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                   4593380528125082431, 16408922859458223821]


# ----------------------------------------------------------- parameters

def test_bucket_counts(fixtures):
    b = fixtures["bucket_counts"]
    assert ock.parameters("bgh25", b["pnum"], b["ell_i"]).buckets == b["bgh25"]
    assert ock.parameters("drrt18", b["pnum"], b["ell_i"]).buckets == b["drrt18"]


def test_bgh25_closed_forms():
    for pnum in (3, 16, 29, 30, 1024, 16384, 65534):
        p = ock.parameters("bgh25", pnum, 16 if pnum <= 65536 else 20)
        assert p.buckets == (2 * pnum if pnum < 30 else pnum * 3 // 2)
        assert p.kick == 3 * pnum
        assert p.bucket_size == -(-(3 << p.ell_i) // p.buckets)
        assert p.domain == p.buckets * p.bucket_size
        bits = max(2, (p.domain - 1).bit_length())
        assert p.half_bits == (bits + (bits & 1)) // 2


def test_config_shapes():
    # n = 2^14, m = 24576, ell_i 16: half 9, bucket_size 8, domain 196608.
    p = ock.parameters("bgh25", 16384, 16, kick=500)
    assert (p.buckets, p.kick, p.bucket_size, p.domain, p.half_bits) == (24576, 500, 8, 196608, 9)
    # n = 2^20 with m = ceil(1.3 n) = 1,363,149 (1.3 n = 1,363,148.8).
    p = ock.parameters("bgh25", 1 << 20, 20, buckets=1363149, kick=1000, stash=2)
    assert (p.bucket_size, p.domain, p.half_bits) == (3, 4089447, 11)


def test_parameters_match_reference(fixtures):
    for row in fixtures["reference_parameters"]:
        p = ock.parameters(row["mode"], row["pnum"], row["ell_i"])
        got = {k: getattr(p, k) for k in ("buckets", "kick", "bucket_size", "domain", "half_bits", "ell_d")}
        want = {k: row[k] for k in got}
        assert got == want, row


# ------------------------------------------------------------ Feistel PRP

@pytest.mark.parametrize("mode", ["bgh25", "drrt18"])
def test_prp_is_bijection_on_domain(mode):
    # pnum 16, ell_i 8: y = 3*element+which runs over all of [0, 3*2^8).  Where
    # the domain equals that source range (bgh25: 32*24 = 768) every bucket
    # receives exactly bucket_size preimages; otherwise at most bucket_size.
    p = ock.parameters(mode, 16, 8)
    t = ock.trial_setup(ock.batch_key(123, 7), 3, p)
    hist = np.zeros(p.buckets, dtype=np.int64)
    for o in range(1 << p.ell_i):
        for w in range(3):
            hist[ock.bucket_direct(t, p, o, w)] += 1
    assert hist.max() <= p.bucket_size
    if p.domain == 3 << p.ell_i:
        assert (hist == p.bucket_size).all()
    assert hist.sum() == 3 << p.ell_i


def test_cached_equals_direct():
    rng = random.Random(2)
    for mode, pnum, ell, im in [("bgh25", 16, 8, 0), ("drrt18", 1024, 16, 1), ("bgh25", 16384, 16, 0),
                                ("bgh25", 3, 2, 1)]:
        p = ock.parameters(mode, pnum, ell, input_mode=im)
        t = ock.trial_setup(ock.batch_key(9, 2), rng.getrandbits(20), p)
        rounds = ock.build_rounds(t, p)
        for _ in range(200):
            o, w = rng.randrange(pnum), rng.randrange(3)
            b = ock.bucket_direct(t, p, o, w)
            assert b == ock.bucket_cached(t, p, rounds, o, w) and b < p.buckets


# ------------------------------------------------------ golden outcomes

@pytest.mark.parametrize("mode", ["bgh25", "drrt18"])
def test_golden_32768(fixtures, mode):
    g = fixtures["golden_32768"]
    p = ock.parameters(mode, g["pnum"], g["ell_i"])
    fails = []
    for batch in range(1 << (g["et"] - g["bt"])):
        f, _, _ = ock.run_range(ock.batch_key(g["seed"], batch), 0, 1 << g["bt"], p)
        fails += [batch * (1 << g["bt"]) + int(x) for x in f]
    assert fails == g["failures"][mode]
    prefix = fixtures["smoke_expected"]
    assert [x for x in fails if x < (1 << prefix["ET"])] == prefix["failures"][mode]


@pytest.mark.parametrize("mode", ["bgh25", "drrt18"])
def test_golden_direct_path(fixtures, mode):
    g = fixtures["golden_32768"]
    p = ock.parameters(mode, g["pnum"], g["ell_i"])
    fails = []
    for batch in range(16):
        f, _, _ = ock.run_range(ock.batch_key(g["seed"], batch), 0, 256, p, cached=False)
        fails += [batch * 256 + int(x) for x in f]
    assert fails == [x for x in g["failures"][mode] if x < 16 * 256]


def test_affine_vectors(fixtures):
    v = fixtures["affine_vectors"]
    for mode, want in v["failures"].items():
        p = ock.parameters(mode, v["pnum"], v["ell_i"], input_mode=1)
        f, _, _ = ock.run_range(ock.batch_key(v["seed"], v["batch"]), 0, v["count"], p)
        assert list(f) == want


def test_reference_trial_tables(fixtures):
    for row in fixtures["reference_trials"]:
        p = ock.parameters(row["mode"], row["pnum"], row["ell_i"], input_mode=row["input_mode"])
        failed, bins, _ = ock.trial_run(ock.batch_key(row["seed"], row["batch"]), row["offset"], p)
        assert int(failed) == row["failed"]
        assert digest(ref_table(bins)) == row["table_sha256"], row


def test_reference_ranges(fixtures):
    for row in fixtures["reference_ranges"]:
        p = ock.parameters(row["mode"], row["pnum"], row["ell_i"], input_mode=row["input_mode"])
        f, _, _ = ock.run_range(ock.batch_key(row["seed"], row["batch"]), 0, row["count"], p)
        assert list(f) == row["failures"], row


# ------------------------------------------------ the reference build itself

REF = ock.ref_shim()
needs_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref not built (reference sources absent)")


@needs_ref
def test_against_reference_random_points():
    rng = random.Random(11)
    for _ in range(60):
        mode = rng.choice([1, 2])
        im = rng.choice([0, 1])
        pnum = rng.choice([3, 5, 16, 40, 300, 1024, 2500])
        ell = max(rng.choice([4, 8, 12, 16, 20, 24]), (pnum - 1).bit_length())
        seed, batch, off = rng.getrandbits(64), rng.getrandbits(40), rng.getrandbits(20)
        p = ock.parameters(mode, pnum, ell, input_mode=im)
        bins_ref = (C.c_uint32 * p.buckets)()
        fr = REF.ref_trial(mode, im, pnum, ell, seed, batch, off, 0 if p.half_bits > 10 else 1, bins_ref)
        failed, bins, _ = ock.trial_run(ock.batch_key(seed, batch), off, p)
        assert int(failed) == fr
        assert (ref_table(bins) == np.frombuffer(bins_ref, dtype=np.uint32)).all()


@needs_ref
def test_against_reference_aes_and_keys():
    rng = random.Random(3)
    for _ in range(100):
        seed, batch = rng.getrandbits(64), rng.getrandbits(64)
        w = (C.c_uint32 * 44)()
        REF.ref_batch_key(seed, batch, w)
        assert list(w) == ock.batch_key(seed, batch).words()
        blk = (C.c_uint32 * 4)(*[rng.getrandbits(32) for _ in range(4)])
        out = (C.c_uint32 * 4)()
        REF.ref_host_aes(w, blk, out)
        assert list(out) == ock.aes(ock.batch_key(seed, batch), list(blk), bytewise=True)
    for n in [1, 2, 3, 100, 1000, 1024, 4095, 16383, 65534, 1 << 40]:
        assert REF.ref_original_log2(n) == ock.LIB.ock_original_log2(n)


# ------------------------------------------------------------- extension

def brute_force_trial(p, bmap, walk, randomness):
    """Independent statement of the extended insertion over a bucket map."""
    M = (1 << 64) - 1

    def sm(s):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return s, z ^ (z >> 31)

    k = p.k
    table = [None] * p.buckets
    stash = 0
    bucket = lambda o, w: int(bmap[o * k + w])  # noqa: E731
    for item in range(p.pnum):
        if p.mode == 1:
            cur, col = item, 0
            while True:
                if randomness < k:
                    walk, randomness = sm(walk)
                w = randomness % k
                randomness //= k
                t = bucket(cur, w)
                if table[t] is None:
                    table[t] = (cur, w)
                    break
                col += 1
                if col > p.kick:
                    if stash < p.stash:
                        stash += 1
                        break
                    return True, table
                old = table[t]
                table[t] = (cur, w)
                cur = old[0]
        else:
            cands = [bucket(item, w) for w in range(k)]
            free = [w for w in range(k) if table[cands[w]] is None]
            if free:
                table[cands[free[0]]] = (item, free[0])
                continue
            cur, choice, target, kicks = item, 0, cands[0], 0
            while True:
                if table[target] is None:
                    table[target] = (cur, choice)
                    break
                if kicks >= p.kick:
                    if stash < p.stash:
                        stash += 1
                        break
                    return True, table
                disp = table[target]
                table[target] = (cur, choice)
                cur, choice = disp
                nxt = None
                for _ in range(k):
                    choice = (choice + 1) % k
                    b = bucket(cur, choice)
                    if b != target:
                        nxt = b
                        break
                if nxt is None:
                    if stash < p.stash:
                        stash += 1
                        break
                    return True, table
                target = nxt
                kicks += 1
    return False, table


@pytest.mark.parametrize("k", [2, 3, 4])
@pytest.mark.parametrize("mode", [1, 2])
def test_extension_against_independent_statement(k, mode):
    rng = random.Random(k * 10 + mode)
    for case in range(25):
        pnum = rng.choice([4, 7, 12, 30])
        buckets = max(k, int(pnum * rng.uniform(1.05, 1.6)))
        p = ock.parameters(mode, pnum, 8, k=k, buckets=buckets, kick=rng.choice([0, 1, 3, 10]),
                           stash=rng.choice([0, 1, 2]), input_mode=rng.choice([0, 1]))
        key = ock.batch_key(case, k)
        off = rng.getrandbits(16)
        t = ock.trial_setup(key, off, p)
        bmap = ock.bucket_map(key, off, p)
        want_failed, want_table = brute_force_trial(p, bmap, t.walk, t.randomness)
        failed, bins, st = ock.trial_run(key, off, p)
        assert failed == want_failed
        if not failed:
            got = [None if v == 0xFFFFFFFF else (int(v) & 0x3FFFFFFF, int(v) >> 30) for v in bins]
            assert got == want_table
            placed = [e[0] for e in got if e is not None]
            assert len(placed) == len(set(placed)) and len(placed) + st.stashed == pnum


def test_reduction_to_reference_region():
    # k=3, stash 0, explicit buckets/kick equal to the derived ones: identical.
    for mode in ("bgh25", "drrt18"):
        d = ock.parameters(mode, 16, 8)
        e = ock.parameters(mode, 16, 8, k=3, buckets=d.buckets, kick=d.kick, stash=0)
        assert d.as_dict() == e.as_dict()
