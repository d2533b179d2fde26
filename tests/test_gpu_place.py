"""Table placement: the tables of a placed plan's pass buffers are held apart
from them, spread over the device's memory regions (pieces sorted into regions
by pair walks, every third slot's table in each), or, where the memory budget
holds too few pieces, the two fastest of candidate allocations.  Results must
not depend on it."""
import pytest

from paths import load_product

pytestmark = pytest.mark.gpu

ck = load_product()


def _check_regions(ctx):
    """The region report of a context whose tables were placed by region."""
    ms, kept = ctx.place_report()
    regions, b0, b1 = ctx.place_regions()
    assert len(regions) >= 6 and set(regions) >= {0, 1} and regions[0] == 0
    assert len(set(b0 + b1)) == 6 and min(b0 + b1) >= 0 and max(b0 + b1) < len(regions)
    assert min(regions[i] for i in b0 + b1) >= 0  # no piece left out as unclear
    for pieces in (b0, b1):
        assert len({regions[i] for i in pieces}) >= 2  # every table range spans two regions or more
    # The candidate table ranges were timed; the two fastest became the tables.
    assert 2 <= len(ms) <= 6 and all(x > 0 for x in ms)
    order = sorted(range(len(ms)), key=lambda i: ms[i])
    assert (ms[kept[0]], ms[kept[1]]) == (ms[order[0]], ms[order[1]])
    return regions, b0, b1


def test_region_placement_spreads_tables_and_changes_no_result():
    p = ck.params("bgh25", 16384, 16, kick=500)  # tables of 6.4 GB per pass buffer: pieces of 2.1 GB
    rk = ck.batch_key(1, 0)
    with ck.Context(0) as a:
        a.set_place_search(16)
        f1, s1 = a.launch(p, rk, 0, 1 << 17)
        report = _check_regions(a)
        f1b, s1b = a.launch(p, rk, 0, 1 << 17)  # same tables, no new placement
        assert a.place_regions() == report
    with ck.Context(0) as b:  # the default: no placement
        f2, s2 = b.launch(p, rk, 0, 1 << 17)
        assert b.place_report()[1] == (-1, -1) and b.place_regions()[0] == []
    assert list(f1) == list(f2) == list(f1b)
    assert s1.probes == s2.probes == s1b.probes


@pytest.mark.parametrize("layout", [ck.LAYOUT_DESC_PLAIN, ck.LAYOUT_RESIDUAL])
def test_region_placement_follows_the_table_size(layout):
    """A plan with tables of another size places them again; every placed run
    matches the unplaced one (u64 descriptor entries and u32 residual entries,
    two shapes)."""
    shapes = [ck.params("bgh25", 16384, 16, kick=500), ck.params("bgh25", 8192, 16, kick=500)]
    rk = ck.batch_key(1, 0)
    with ck.Context(0) as a, ck.Context(0) as b:
        a.set_place_search(8)
        a.set_layout(layout)
        b.set_layout(layout)
        reports = []
        for p in shapes:
            fa, sa = a.launch(p, rk, 0, 1 << 17)
            reports.append(_check_regions(a))
            fb, sb = b.launch(p, rk, 0, 1 << 17)
            assert list(fa) == list(fb) and sa.probes == sb.probes
        assert reports[0] != reports[1]


@pytest.mark.parametrize("pnum", [1024, 4096])
def test_region_placement_of_small_tables(pnum):
    """Tables of 0.4 and 1.6 GB per pass buffer: each of a range's three 1 GiB
    pieces holds the tables of every third slot, from its start; results equal
    the unplaced context's."""
    p = ck.params("bgh25", pnum, 16, kick=500)
    rk = ck.batch_key(3, 0)
    with ck.Context(0) as a, ck.Context(0) as b:
        a.set_place_search(8)
        fa, sa = a.launch(p, rk, 0, 1 << 17)
        _check_regions(a)
        fb, sb = b.launch(p, rk, 0, 1 << 17)
        assert list(fa) == list(fb) and sa.probes == sb.probes


def test_search_where_the_budget_holds_few_pieces():
    """A memory budget of 10 GiB holds five 1 GiB pieces beside the pass
    buffers at n = 4096 (2.2 GB of maps and 1.6 GB of tables each), too few to
    sort: the library searches three whole candidate table allocations instead
    and keeps the two fastest."""
    p = ck.params("bgh25", 4096, 16, kick=500)
    rk = ck.batch_key(3, 0)
    with ck.Context(0) as a, ck.Context(0) as b:
        a.set_placement(layout=ck.LAYOUT_RESIDUAL, budget_bytes=10 << 30)
        a.set_place_search(8)
        b.set_layout(ck.LAYOUT_RESIDUAL)
        fa, sa = a.launch(p, rk, 0, 1 << 17)
        ms, kept = a.place_report()
        assert a.place_regions()[0] == []
        assert 2 <= len(ms) <= 3 and all(x > 0 for x in ms)
        order = sorted(range(len(ms)), key=lambda i: ms[i])
        assert (ms[kept[0]], ms[kept[1]]) == (ms[order[0]], ms[order[1]])
        fb, sb = b.launch(p, rk, 0, 1 << 17)
        assert list(fa) == list(fb) and sa.probes == sb.probes


def test_place_search_skips_small_plans():
    small = ck.params("bgh25", 256, 16, kick=500)  # 0.25 GB pass buffers
    with ck.Context(0) as a:
        a.set_place_search(16)
        a.launch(small, ck.batch_key(0, 0), 0, 1 << 17)
        assert a.place_report() == ([], (-1, -1))


def test_graph_replays_run_in_placed_tables():
    """Graph mode places the tables like kernel-by-kernel launches do; the
    graph's replays (new batch keys and offsets) give the unplaced eager
    context's failing trials and probe totals."""
    p = ck.params("bgh25", 4096, 16, kick=500)
    with ck.Context(0) as g, ck.Context(0) as e:
        g.set_place_search(16)
        g.set_graph(True)
        for rep, (batch, first) in enumerate([(0, 0), (1, 777), (2, 5)]):
            rk = ck.batch_key(6, batch)
            f, st = g.launch(p, rk, first, 1 << 17)
            assert st.graph == (2 if rep == 0 else 1)
            assert g.place_report()[1] != (-1, -1)
            ef, est = e.launch(p, rk, first, 1 << 17)
            assert list(f) == list(ef), (batch, first)
            assert st.probes == est.probes and st.failed == est.failed


def test_place_search_rejects_one_candidate():
    with ck.Context(0) as a:
        with pytest.raises(ck.CkError):
            a.set_place_search(1)


@pytest.mark.parametrize("layout", [ck.LAYOUT_RESIDUAL, ck.LAYOUT_DESC_PLAIN, ck.LAYOUT_MAP])
def test_region_placed_tables_hold_the_same_entries(layout):
    """Per-trial table digests and probe counts of a region-placed context
    (every third slot's table in each piece) equal the unplaced context's, for
    residual, descriptor and bucket-map entries (n = 2048)."""
    p = ck.params("bgh25", 2048, 16, kick=500)
    rk = ck.batch_key(5, 0)
    got = []
    for place in (8, 0):
        with ck.Context(0) as c:
            c.set_layout(layout)
            c.set_place_search(place)
            c.launch(p, rk, 0, 1 << 17)  # places the tables (the digests launch runs in them)
            assert (len(c.place_regions()[0]) >= 6) == bool(place)
            got.append(c.launch_digest(p, rk, 65536, 1 << 17))
    (fa, pa, da, _), (fb, pb, db, _) = got
    assert list(fa) == list(fb) and (pa == pb).all() and (da == db).all()
