"""Parameter sweeps and arrival processes for the batched / served engine.

sweep_points(): mixed-shape sweep points -- n log-uniform over [2^8, 2^16],
k in {2, 3, 4}, buckets/n uniform over [1.2, 1.6], stash in {0..4}, max kicks
in {100, 500, 1000}, BGH25 random-walk insertion, ell_i 16, sequential items.
Every draw comes from one SplitMix64 stream started at the sweep seed, so a
(seed, count) pair names the same points everywhere; each point also draws its
own 64-bit master seed for its trial keys (batch bits 20: trial t is offset t
of batch 0 for t < 2^20).

arrivals(): Poisson arrival times (exponential gaps at `rate` jobs/s) and
log-uniform trial budgets, from a SplitMix64 stream of their own seed.

Host-side job generation only; the trial arithmetic stays integer on the device.
"""
from __future__ import annotations

import math

MASK64 = (1 << 64) - 1
KICKS = (100, 500, 1000)


class SplitMix64:
    """SplitMix64 (Steele, Lea & Flood, 2014)."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Uniform on [0, 1) with 53 random bits."""
        return (self.next() >> 11) / float(1 << 53)


def sweep_points(count: int = 512, seed: int = 2, trials: int = 10**6, ell_i: int = 16, mode: str = "bgh25",
                 bt: int = 20):
    g = SplitMix64(seed)
    pts = []
    for j in range(count):
        n = int(round(2.0 ** (8.0 + 8.0 * g.uniform())))
        n = min(max(n, 1 << 8), 1 << 16)
        k = 2 + g.next() % 3
        ratio = 1.2 + 0.4 * g.uniform()
        buckets = int(math.ceil(ratio * n))
        stash = g.next() % 5
        kick = KICKS[g.next() % 3]
        pts.append(dict(index=j, mode=mode, pnum=n, ell_i=ell_i, k=k, buckets=buckets, kick=kick, stash=stash,
                        seed=g.next(), trials=trials, bt=bt))
    return pts


def arrivals(count: int, rate: float, seed: int = 3, budget_lo: float = 1e5, budget_hi: float = 1e9):
    """[(arrival_s, budget)]: Poisson arrivals at `rate` jobs/s and trial
    budgets log-uniform over [budget_lo, budget_hi]."""
    g = SplitMix64(seed)
    t = 0.0
    out = []
    lo, hi = math.log10(budget_lo), math.log10(budget_hi)
    for _ in range(count):
        t += -math.log(1.0 - g.uniform()) / rate
        budget = int(round(10.0 ** (lo + (hi - lo) * g.uniform())))
        out.append((t, max(1, budget)))
    return out


def params_of(ck, pt):
    """ck_params of a sweep point through the library's parameter derivation."""
    return ck.params(pt["mode"], pt["pnum"], pt["ell_i"], k=pt["k"], buckets=pt["buckets"], kick=pt["kick"],
                     stash=pt["stash"])
