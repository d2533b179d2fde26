"""The sweep and arrival generators (host side of the batched / served runs)."""
import importlib.util
import math

from paths import PKG, load_product

spec = importlib.util.spec_from_file_location("ck_sweep", PKG / "sweep.py")
sw = importlib.util.module_from_spec(spec)
spec.loader.exec_module(sw)
ck = load_product()


def test_splitmix_published_vector():
    # SplitMix64 from seed 1234567 (the generator's reference output).
    g = sw.SplitMix64(1234567)
    assert [g.next() for _ in range(5)] == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                            4593380528125082431, 16408922859458223821]


def test_sweep_points_ranges_and_determinism():
    pts = sw.sweep_points(512, seed=2)
    assert pts == sw.sweep_points(512, seed=2)
    assert pts != sw.sweep_points(512, seed=3)
    assert len(pts) == 512
    for pt in pts:
        assert 256 <= pt["pnum"] <= 65536
        assert pt["k"] in (2, 3, 4)
        assert 1.2 * pt["pnum"] <= pt["buckets"] <= math.ceil(1.6 * pt["pnum"])
        assert pt["stash"] in range(5)
        assert pt["kick"] in (100, 500, 1000)
        assert pt["trials"] == 10**6 and pt["bt"] == 20
        p = sw.params_of(ck, pt)  # every point is a valid shape
        assert p.buckets == pt["buckets"] and p.kick == pt["kick"] and p.k == pt["k"]
    # log-uniform n: every octave of [2^8, 2^16] is populated, none dominates
    octaves = [0] * 8
    for pt in pts:
        octaves[min(7, int(math.log2(pt["pnum"])) - 8)] += 1
    assert min(octaves) > 30 and max(octaves) < 100
    assert len({pt["seed"] for pt in pts}) == 512


def test_arrivals():
    a = sw.arrivals(2000, rate=100.0, seed=3)
    gaps = [a[0][0]] + [b[0] - x[0] for x, b in zip(a, a[1:])]
    assert all(g > 0 for g in gaps)
    assert abs(sum(gaps) / len(gaps) - 0.01) < 0.001  # mean gap 1/rate
    budgets = [b for _, b in a]
    assert min(budgets) >= 1e5 and max(budgets) <= 1e9
    logs = [math.log10(b) for b in budgets]
    assert abs(sum(logs) / len(logs) - 7.0) < 0.1  # log-uniform over [1e5, 1e9]
    assert a == sw.arrivals(2000, rate=100.0, seed=3)
