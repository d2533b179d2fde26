"""Descriptor tables reused across passes without a clear.

In the descriptor layout each pass tags the entries it writes, and the tables
are cleared only on a buffer's first use, when the plan of the tables in the
workspace changes, or when the tag field wraps.  These tests compare such a
context with a bucket-map-layout context (tables cleared every pass) on the
same trials.  Both layouts are bit-exact to the oracle in the parity tests.
The cases:
- a shape with 2 tag bits (k*bits + 2 = 62), so every pass buffer wraps and is
  cleared every third pass;
- shapes with many tag bits, over several calls on one context;
- contexts that alternate shapes, layouts and digest launches, which must
  clear where the plan changes.
"""
import pytest

from paths import load_product

pytestmark = pytest.mark.gpu

ck = load_product()

CHUNK = 1024

SHAPES = [
    ("bgh25", 20000, 16, dict(k=4, buckets=25000, kick=100, stash=1)),  # 4*15 + 2 = 62: 2 tag bits
    ("drrt18", 20000, 16, dict(k=4, buckets=25000, kick=100)),          # 2 tag bits
    ("bgh25", 16384, 16, dict(kick=500)),                               # 17 tag bits
    ("bgh25", 3000, 16, dict(k=2, buckets=3900, kick=100, stash=2)),
    ("drrt18", 16, 8, dict()),
]


def reference(p, rk, first, count):
    with ck.Context(0) as m:
        m.set_chunk(CHUNK)
        m.set_layout(ck.LAYOUT_MAP)
        return m.launch(p, rk, first, count)


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}-n{s[1]}-k{s[3].get('k', 3)}")
def test_tagged_tables_match_cleared_tables(shape):
    m_, n, ell, kw = shape
    p = ck.params(m_, n, ell, **kw)
    count = 7 * CHUNK - 5  # seven passes, the last one short
    with ck.Context(0) as d, ck.Context(0) as m:
        d.set_chunk(CHUNK)
        d.set_layout(ck.LAYOUT_DESC_PLAIN)  # the tagged layout (auto prefers u32 residual entries where they fit)
        m.set_chunk(CHUNK)
        m.set_layout(ck.LAYOUT_MAP)
        for call in range(3):
            rk = ck.batch_key(3, call)
            f, st = d.launch(p, rk, call * 10000, count)
            g, sm = m.launch(p, rk, call * 10000, count)
            assert list(map(int, f)) == list(map(int, g)), call
            assert st.probes == sm.probes, call


def test_tables_cleared_where_the_plan_changes():
    a = ck.params("bgh25", 2000, 16, buckets=2600, kick=500, stash=1)
    b = ck.params("drrt18", 3000, 16, k=4, buckets=3600, kick=100)
    T = ck.LAYOUT_DESC_PLAIN  # tagged tables
    steps = [(a, T, False, 3000), (a, T, False, 3000), (b, T, False, 2500), (a, T, False, 3000),
             (a, ck.LAYOUT_MAP, False, 3000), (a, T, False, 3000), (a, T, True, 3000), (a, T, False, 3000),
             (a, ck.LAYOUT_AUTO, False, 3000), (a, T, False, 4000), (b, ck.LAYOUT_DESC, False, 2500), (b, T, False, 2500)]
    with ck.Context(0) as d:
        d.set_chunk(CHUNK)
        for i, (p, layout, digest, count) in enumerate(steps):
            d.set_layout(layout)
            rk = ck.batch_key(11, i)
            if digest:
                f, probes, _, st = d.launch_digest(p, rk, 0, count)
            else:
                f, st = d.launch(p, rk, 0, count)
            g, sm = reference(p, rk, 0, count)
            assert list(map(int, f)) == list(map(int, g)), i
            assert st.probes == sm.probes, i
    # The LDS tier uses the same workspace as scratch.
    with ck.Context(0) as d:
        d.set_chunk(CHUNK)
        for i, tier in enumerate([ck.TIER_HBM, ck.TIER_LDS, ck.TIER_HBM, ck.TIER_HBM]):
            d.set_tier(tier)
            rk = ck.batch_key(12, i)
            f, st = d.launch(a, rk, 0, 2048)
            g, sm = reference(a, rk, 0, 2048)
            assert list(map(int, f)) == list(map(int, g)), tier
            assert st.probes == sm.probes, tier


def test_graph_replays_across_tag_wraps():
    """Graph mode rewrites each pass's clear flag and tag on every replay: with
    2 tag bits a replay's passes alternate between clearing and reusing."""
    p = ck.params("bgh25", 20000, 16, k=4, buckets=25000, kick=100, stash=1)
    count = 3 * CHUNK + 100  # four passes: buffers 0, 1, 0, 1
    with ck.Context(0) as g:
        g.set_graph(True)
        g.set_chunk(CHUNK)
        for rep in range(5):
            rk = ck.batch_key(21, rep)
            f, st = g.launch(p, rk, rep * 5000, count)
            assert st.graph == (2 if rep == 0 else 1)
            r, sr = reference(p, rk, rep * 5000, count)
            assert list(map(int, f)) == list(map(int, r)), rep
            assert st.probes == sr.probes, rep
