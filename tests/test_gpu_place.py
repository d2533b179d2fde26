"""Table-region placement search: the library times a random walk over each
candidate table region and keeps the two fastest as the tables of its pass
buffers.  Results must not depend on it."""
import pytest

from paths import load_product

pytestmark = pytest.mark.gpu

ck = load_product()


def test_place_search_keeps_the_two_fastest_and_changes_no_result():
    p = ck.params("bgh25", 16384, 16, kick=500)  # 6-13 GB table regions: searched
    rk = ck.batch_key(1, 0)
    with ck.Context(0) as a:
        a.set_place_search(16)
        f1, s1 = a.launch(p, rk, 0, 1 << 17)
        ms, kept = a.place_report()
        assert len(ms) >= 3 and all(x > 0 for x in ms)
        order = sorted(range(len(ms)), key=lambda i: ms[i])
        assert (ms[kept[0]], ms[kept[1]]) == (ms[order[0]], ms[order[1]])
        f1b, s1b = a.launch(p, rk, 0, 1 << 17)  # same buffers, no new search
        assert a.place_report() == (ms, kept)
    with ck.Context(0) as b:  # the default: no search
        f2, s2 = b.launch(p, rk, 0, 1 << 17)
        assert b.place_report()[1] == (-1, -1)
    assert list(f1) == list(f2) == list(f1b)
    assert s1.probes == s2.probes == s1b.probes


@pytest.mark.parametrize("layout", [ck.LAYOUT_DESC_PLAIN, ck.LAYOUT_RESIDUAL])
def test_place_search_follows_the_table_size(layout):
    """A plan with tables of another size searches again; every searched run
    matches the unsearched one (u64 descriptor entries and u32 residual
    entries, two shapes)."""
    shapes = [ck.params("bgh25", 16384, 16, kick=500), ck.params("bgh25", 8192, 16, kick=500)]
    rk = ck.batch_key(1, 0)
    with ck.Context(0) as a, ck.Context(0) as b:
        a.set_place_search(8)
        a.set_layout(layout)
        b.set_layout(layout)
        reports = []
        for p in shapes:
            fa, sa = a.launch(p, rk, 0, 1 << 17)
            reports.append(a.place_report())
            fb, sb = b.launch(p, rk, 0, 1 << 17)
            assert list(fa) == list(fb) and sa.probes == sb.probes
        for ms, kept in reports:
            assert 3 <= len(ms) <= 8 and -1 not in kept
        assert reports[0] != reports[1]


def test_place_search_skips_small_and_graph_plans():
    small = ck.params("bgh25", 1024, 16, kick=500)  # 0.4 GB pass buffers
    with ck.Context(0) as a:
        a.set_place_search(16)
        a.launch(small, ck.batch_key(0, 0), 0, 1 << 17)
        assert a.place_report() == ([], (-1, -1))
    head = ck.params("bgh25", 16384, 16, kick=500)
    with ck.Context(0) as g:
        g.set_place_search(16)
        g.set_graph(True)
        g.launch(head, ck.batch_key(1, 0), 0, 1 << 17)
        assert g.place_report()[1] == (-1, -1)


def test_place_search_rejects_one_candidate():
    with ck.Context(0) as a:
        with pytest.raises(ck.CkError):
            a.set_place_search(1)
