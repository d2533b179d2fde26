"""CPU checks of the measurement model in bench.py (no GPU)."""
import importlib.util
from types import SimpleNamespace

import pytest

from paths import ROOT


def load_bench():
    spec = importlib.util.spec_from_file_location("ck_bench", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_int_ops_at_headline_shape():
    b = load_bench()
    # n = 2^14, m = 24576, ell_i 16: bucket_size 8, domain 196,608, half_bits 9.
    p = SimpleNamespace(domain=196608, half_bits=9, ell_i=16, pnum=16384, buckets=24576)
    # The cycle walk lands in the domain with probability 196608 / 2^18 = 0.75,
    # so a bucket evaluation costs 8 / 0.75 = 32 / 3 cached rounds on average.
    fixed = 2 * 336 + 160 + 2 * 22 + (8 << 9) * 336
    per_probe = (32 / 3) * 4 + 6 + 20
    assert fixed == 1377132
    assert b.int_ops_per_trial(p, 0.0) == pytest.approx(fixed)
    assert b.int_ops_per_trial(p, 1000.0) - b.int_ops_per_trial(p, 0.0) == pytest.approx(1000 * per_probe)
    # The headline batch averages 33,924.44 probes per trial: 3.71 M ops.
    assert b.int_ops_per_trial(p, 33924.4415) == pytest.approx(3.7066e6, rel=1e-4)


def test_int_ops_direct_domain_acceptance():
    b = load_bench()
    # Config 5 shape (n = 2^20, m = 1,363,149, ell_i 20): bucket_size 3,
    # domain 4,089,447, half_bits 11 -> acceptance 4,089,447 / 2^22.
    p = SimpleNamespace(domain=4089447, half_bits=11, ell_i=20, pnum=1 << 20, buckets=1363149)
    accept = 4089447 / float(1 << 22)
    per_probe = 8 / accept * 4 + 6 + 20
    assert b.int_ops_per_trial(p, 10.0) - b.int_ops_per_trial(p, 0.0) == pytest.approx(10 * per_probe)


def test_algorithmic_bytes_constants():
    b = load_bench()
    assert b.ALGO_PROBE_BYTES == 8 and b.ALGO_BUCKET_INIT_BYTES == 4
    assert b.HBM_PEAK_GBS == 8000.0
