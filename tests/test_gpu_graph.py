"""Graph mode through the C ABI.

ck_launch calls replayed as one hipGraph per (shape, trial count, placement)
must give the failing trials and probe totals of the kernel-by-kernel form:
- on the first call, which builds the graph;
- on every replay, with new batch keys and offsets;
- with other launches (other shapes, debug launches) in between on the same
  context, and after a change of trial count, which rebuilds the graph.
"""
import pytest

from paths import load_product

pytestmark = pytest.mark.gpu

ck = load_product()

# (mode, pnum, ell_i, extension kwargs); every HBM-tier table layout the automatic choice makes
SHAPES = [
    ("bgh25", 16384, 16, dict(kick=500)),                                       # descriptor entries
    ("drrt18", 1000, 16, dict(k=4, buckets=1250, kick=100, stash=1)),
    ("bgh25", 300, 16, dict(k=2, buckets=400, kick=100, stash=2, input_mode=1)),
    ("bgh25", 30000, 16, dict(k=4, buckets=40000, kick=100)),                   # u32 map, exchange probes
    ("drrt18", 16, 8, dict()),                                                  # golden shape
]
RANGES = [(1, 0, 0), (1, 1, 5000), (7, 3, 123), (2, 0, 0), (9, 9, 99999)]  # (seed, batch, first)
CHUNK = 2048
COUNT = 3 * CHUNK + 77  # four passes, the last one short


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}-n{s[1]}")
def test_graph_replays_match_kernel_by_kernel(shape):
    m, n, ell, kw = shape
    p = ck.params(m, n, ell, **kw)
    other = ck.params("bgh25", 100, 12, kick=50)
    with ck.Context(0) as g, ck.Context(0) as e:
        g.set_graph(True)
        for c in (g, e):
            c.set_chunk(CHUNK)
        # Grow the workspace for the interleaved shape first: a workspace that
        # grows later moves, and every graph of the context is then rebuilt.
        g.launch_debug(other, ck.batch_key(0, 0), 0, 300)
        for rep, (seed, batch, first) in enumerate(RANGES):
            rk = ck.batch_key(seed, batch)
            f, st = g.launch(p, rk, first, COUNT)
            assert st.graph == (2 if rep == 0 else 1)
            # Eager work on the graph's own context between replays.
            g.launch_debug(other, rk, 0, 300)
            g.launch(other, rk, 0, 200)
            ef, est = e.launch(p, rk, first, COUNT)
            assert est.graph == 0
            assert list(map(int, f)) == list(map(int, ef)), (seed, batch, first)
            assert st.probes == est.probes and st.failed == est.failed == len(ef)
            assert st.trials == COUNT


def test_graph_cache_rebuilds_on_new_count_and_shape():
    """A context keeps its four most recently used graphs: a repeat replays,
    a new count or shape builds, and the least recently used one goes."""
    p = ck.params("bgh25", 2000, 16, buckets=2400, kick=500)
    q = ck.params("drrt18", 500, 16, k=4, buckets=600, kick=100)
    calls = [(p, 3000, 2), (p, 3000, 1), (p, 2999, 2), (q, 2999, 2), (q, 2999, 1), (p, 3000, 1),
             (q, 3001, 2), (q, 3002, 2),  # fifth key: replaces (p, 2999), the least recently used
             (p, 2999, 2),                # replaces (q, 2999)
             (p, 3000, 1), (q, 2999, 2)]
    with ck.Context(0) as g, ck.Context(0) as e:
        g.set_graph(True)
        for c in (g, e):
            c.set_chunk(1024)
        rk = ck.batch_key(5, 0)
        for i, (shp, count, want) in enumerate(calls):
            f, st = g.launch(shp, rk, 17 + i, count)
            ef, est = e.launch(shp, rk, 17 + i, count)
            assert list(map(int, f)) == list(map(int, ef)), i
            assert st.probes == est.probes, i
            assert st.graph == want, i
        g.set_graph(False)
        f, st = g.launch(p, rk, 17, 3000)
        assert st.graph == 0


def test_graph_full_batch_matches_at_headline_shape():
    """A 2^16-trial range at the headline shape in four double-buffered passes
    of 16,384 trials, replayed twice.  (The pass size is below the library's
    default to bound the workspace that a graph-mode context leaves allocated
    in this process.)"""
    p = ck.params("bgh25", 16384, 16, kick=500)
    with ck.Context(0) as g, ck.Context(0) as e:
        g.set_graph(True)
        g.set_chunk(1 << 14)
        for batch in (0, 1):
            rk = ck.batch_key(1, batch)
            f, st = g.launch(p, rk, 0, 1 << 16)
            ef, est = e.launch(p, rk, 0, 1 << 16)
            assert st.graph == (2 if batch == 0 else 1)
            assert list(map(int, f)) == list(map(int, ef))
            assert st.probes == est.probes



def device_free_bytes():
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("place", [0, 16], ids=["unplaced", "placed"])
def test_graph_contexts_do_not_accumulate_memory(place):
    """The runtime keeps a workspace that graph nodes pointed into after it is
    freed; the library keeps such workspaces for reuse instead, so a sequence
    of graph-mode contexts (and an eager one after them) holds one workspace,
    not one per context.  Placed tables (virtual-memory pieces) go back to the
    device.  Results stay those of kernel-by-kernel launches."""
    p = ck.params("bgh25", 4096, 16, kick=500)
    count = 1 << 17  # two passes: double-buffered, so placeable
    with ck.Context(0) as ctx:
        want, _ = ctx.launch(p, ck.batch_key(3, 0), 0, count)
    free_after = []
    for graph in (True, True, True, False):
        with ck.Context(0) as ctx:
            ctx.set_place_search(place)
            ctx.set_graph(graph)
            for _ in range(2):
                f, st = ctx.launch(p, ck.batch_key(3, 0), 0, count)
                assert list(f) == list(want)
            assert st.graph == (1 if graph else 0)
            assert (ctx.place_report()[1] != (-1, -1)) == (place > 0)
        free_after.append(device_free_bytes())
    slack = 1 << 30
    assert all(f >= free_after[0] - slack for f in free_after[1:]), [f >> 30 for f in free_after]
