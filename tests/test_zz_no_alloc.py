"""ck_launch_static -- the allocation-free launch -- against ck_launch through
the C ABI: the same failing trials and probe totals, on two calls in a row
(the second smaller, over the array the first left filled), with the device's
free memory unchanged across the calls, and a clear error for a shape whose
single trial does not fit the array."""
import pytest

from paths import load_product

pytestmark = pytest.mark.gpu

ck = load_product()

# (mode, pnum, ell_i, extension kwargs, seed, first call's trials, second call's trials)
SHAPES = [
    ("bgh25", 16384, 16, dict(kick=500), 1, 1500, 300),          # headline shape: descriptor entries
    ("drrt18", 1024, 16, dict(), 7, 3000, 700),
    ("bgh25", 3000, 16, dict(k=4, buckets=3700, kick=100, stash=2), 9, 2000, 65),
    ("bgh25", 1 << 20, 20, dict(buckets=1363149, kick=1000, stash=2), 4, 40, 7),  # residual entries, windows
]


def free_bytes():
    """hipMemGetInfo of the current device (the library's own runtime)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("mode,pnum,ell,kw,seed,n1,n2", SHAPES)
def test_static_launch_matches_launch(mode, pnum, ell, kw, seed, n1, n2):
    p = ck.params(mode, pnum, ell, **kw)
    with ck.Context(0) as ctx:
        want = [ctx.launch(p, ck.batch_key(seed, 0), 0, n1), ctx.launch(p, ck.batch_key(seed, 1), 17, n2)]
        before = free_bytes()
        got = [ctx.launch_static(p, ck.batch_key(seed, 0), 0, n1),
               ctx.launch_static(p, ck.batch_key(seed, 1), 17, n2)]
        after = free_bytes()
    for (wf, wst), (gf, gst) in zip(want, got):
        assert list(gf) == list(wf)
        assert gst.probes == wst.probes and gst.failed == wst.failed and gst.trials == wst.trials
        assert gst.graph == 0
    assert after == before  # nothing was allocated


def test_static_launch_rejects_oversized_trials():
    # n = 2^26: one trial's table alone is hundreds of MiB.
    p = ck.params("bgh25", 1 << 26, 26, kick=100)
    with ck.Context(0) as ctx:
        with pytest.raises(ck.CkError, match="ck_launch_static"):
            ctx.launch_static(p, ck.batch_key(1, 0), 0, 1)
        # The context still launches normally afterwards.
        q = ck.params("bgh25", 256, 16, kick=100)
        f, st = ctx.launch_static(q, ck.batch_key(1, 0), 0, 64)
        g, st2 = ctx.launch(q, ck.batch_key(1, 0), 0, 64)
        assert list(f) == list(g) and st.probes == st2.probes
